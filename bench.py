"""Benchmark of the MI355X SIMBA/VFB search.

Default line (one JSON line on stdout): the largest single-instance SIMBA
shape -- leaves {a, b, c, d, 1}, uint64 words, 64 sample points, a uniform
random size-15 target -- searched with the semantics of an exhaustive search
to size 20 (every size s is checked with the unary-chain windows of span
20 - s); one step sweeps sizes 1..15 (3.90e13 candidate trees) and stops
there, so a step is bounded.  Inputs are resident on the device before the
timed region.  `value` is candidates per second over the K timed steps.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload ...]

roofline (bound "valu", kernel k_pair_check): the algorithmic work of the
pair kernel is 10 32-bit lane-ops per (A, B) value pair of a split -- the
fused low-half window test of the pair's 6 or 7 operator instances (two
xor-add distances, the product and its two distances, two distances of
fold(b), three 3-way minima) -- summed over every pair of sizes 3..15,
divided by the kernel's HIP-event time on the library stream.  Peak:
256 CU x 64 lanes x 2.4 GHz = 39.3 T lane-ops/s.  The naive-equivalent rate
(every candidate evaluated node by node) is reported beside it, and the
executed-VALU counters and HBM traffic of the same command come from the
rocprofv3 summary under profiles/ (labelled with their source).

cpu_baseline: the CPU oracle (multithreaded C restatement of the reference
semantics, threads pinned to the cores this process may use, at most 16:
the host share of one GPU) on the same instance, every size 1..11; the first
solution must be the GPU's.

The same line carries compact records of the other configurations
(`configs`): C1 time to first solution, C2 VFB levels 1..13 (with the
insert pass against its measured random-access ceiling), C3 as the
reference runs it (one search to size 20 under a 60 s cap), C4 batched
solves/s, C5 learned vs untrained proposal.  `--no-configs` skips them; each
has its own workload for a full line.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

VALU_PEAK = 256 * 64 * 2.4e9  # integer lane-ops/s
HBM_PEAK_GBS = 8000.0
PAIR_OPS = 10  # algorithmic lane-ops per (A, B) pair in window mode
M64 = (1 << 64) - 1


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_info():
    """(threads to use, CPU model): the CPUs this process may run on, at most
    16 (a one-GPU box's host share)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return max(1, min(16, n)), model


def naive_ops_per_size(L, max_size, W):
    """C[s] = sum over all trees of size s of sum_{internal nodes} c(op)."""
    from simba_amd.synth import count_table
    T = count_table(L, max_size)
    c = [2, 2, 2, 2, 2, 2, 2, 4] if W == 64 else [1] * 8
    C = [0] * (max_size + 1)
    for s in range(2, max_size + 1):
        tot = 0
        for op in (0, 4):
            tot += c[op] * T[s - 1][8] + C[s - 1]
        for op in (1, 2, 3, 5, 6, 7):
            last = s - 2 if op == 6 else (s - 1) // 2
            for j in range(1, last + 1):
                m = s - 1 - j
                tot += c[op] * T[j][8] * T[m][8] + C[j] * T[m][8] + T[j][8] * C[m]
        C[s] = tot
    return T, C


def naive_pair_kernel_ops(T, C, L, max_size, W):
    """Naive-equivalent lane-ops of the trees the pair kernel decides: every
    tree of sizes 3..max_size except the unary chains over a single leaf."""
    c_un = 2 if W == 64 else 1
    return sum(C[s] - L * (1 << (s - 1)) * (s - 1) * c_un for s in range(3, max_size + 1))


def pair_count(T, s_lo, s_hi):
    """(A, B) value pairs the pair kernel checks for sizes s_lo..s_hi: one per
    pair of trees of an unordered split {j, s-1-j}, j < s-1-j; an equal-size
    split {j, j} (both sides tabulated) one per unordered pair of its trees,
    T[j] (T[j] + 1) / 2, every operator instance giving both orders."""
    tot = 0
    for s in range(max(s_lo, 3), s_hi + 1):
        for j in range(1, (s - 1) // 2 + 1):
            n, m = T[j][8], T[s - 1 - j][8]
            tot += n * (n + 1) // 2 if j == s - 1 - j else n * m
    return tot


def profile_fields(name):
    """Figures recorded by rocprofv3 for the pair kernel on this workload:
    the `bench-fields:` JSON line of profiles/<name>, or {}."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("bench-fields:"):
                    return json.loads(line.split(":", 1)[1])
    except OSError:
        pass
    return {}


def _verified(r, inst):
    """Re-evaluates a returned SIMBA RPN on every sample."""
    from simba_amd.synth import TOK_OP, evaluate
    L = inst.n_variables + len(inst.consts)
    toks = [ord(c) - 48 if c.isdigit() and ord(c) - 48 < L else TOK_OP + "~&|^#+-*".index(c)
            for c in r["rpn"]]
    for i in range(inst.n_examples):
        lv = [int(inst.inputs[i * inst.n_variables + v]) for v in range(inst.n_variables)]
        lv += [int(c) for c in inst.consts]
        if evaluate(toks, lv, inst.width) != int(inst.outputs[i]):
            return False
    return True


def _key(r):
    return (r["status"], r["size"], r["rank"], r["expr"])


def _safe(fn, args):
    """A configuration record, or the error that stopped it."""
    t0 = time.perf_counter()
    try:
        rec = fn(args)
    except Exception as e:  # a failing record must not hide the headline
        rec = {"error": "%s: %s" % (type(e).__name__, e)}
    rec["record_seconds"] = round(time.perf_counter() - t0, 2)
    return rec


# ---- default line: the largest single-instance SIMBA shape -----------------
def run_simba(args):
    import simba_amd
    from simba_amd.synth import make_instance
    from oracle import oracle as orc

    inst = make_instance(args.seed, 4, 64, 15, 64, (1,))
    S, MX, L = args.size, args.window_max, 5
    T, C = naive_ops_per_size(L, MX, 64)
    cands = sum(T[s][8] for s in range(1, S + 1))
    pairs = pair_count(T, 3, S)
    ctx = simba_amd.Context(0)
    kw = dict(min_size=1, max_size=MX, stop_size=S, exhaustive=True,
              table_bytes=args.table_gb << 30)
    tuned = {k: v for k, v in (("lanes_b", args.lanes_b), ("wg_per_cu", args.wg_per_cu)) if v}
    kw.update(tuned)
    if args.autotune:
        from simba_amd import autotune
        tuned, _ = autotune.tune_simba(ctx, inst, S - 1)
        kw.update(tuned)
    log("simba: warmup")
    for _ in range(args.warmup):
        ctx.simba(inst, **kw)
    ktimes = []
    log("simba: timed steps")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = ctx.simba(inst, **kw)
        ktimes.append(r["kernel_seconds"])
        assert r["size_reached"] == S and r["candidates"] == cands, r
    wall = time.perf_counter() - t0
    ctx.close()
    ms_per_step = 1e3 * wall / args.steps
    kavg = sum(ktimes) / len(ktimes)
    launches = r["kernel_launches"]
    achieved = pairs * PAIR_OPS / kavg
    naive = naive_pair_kernel_ops(T, C, L, S, 64) / kavg
    prof = profile_fields("k_pair_check_simba_c3_windows20_sizes1to%d.md" % S)
    out = {
        "metric": "Gcand/s evaluated",
        "value": round(cands * args.steps / wall / 1e9, 3),
        "unit": "Gcand/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64", "data": "synthetic",
        "config": {"workload": "simba_4vars_const1_u64_64samples_size15target_"
                               "maxsize%d_sizes1to%d" % (MX, S),
                   "search": "exhaustive search to size %d (windows of span %d - s), "
                             "sweep stopped after size %d" % (MX, MX, S),
                   "target": inst.original_formula, "target_size": 15,
                   "candidates_per_step": cands, "pairs_per_step": pairs,
                   "first_solution": r["expr"], "first_solution_size": r["size"],
                   "first_solution_rank": r["rank"], "launch": tuned or "default"},
        "roofline": {
            "bound": "valu", "kernel": "k_pair_check",
            "achieved": round(achieved / 1e12, 3), "peak": round(VALU_PEAK / 1e12, 3),
            "unit": "TOP/s", "frac": round(achieved / VALU_PEAK, 4),
            "traffic": prof.get("traffic_bytes_per_launch"),
            "definition": "algorithmic 32-bit lane-ops of the pair kernel: %d per (A, B) "
                          "value pair (the fused low-half window test of the pair's 6-7 "
                          "operator instances) x %d pairs per step / pair-kernel HIP-event "
                          "time per step" % (PAIR_OPS, pairs),
            "launches_per_step": launches,
            "avg_launch_ms": round(1e3 * kavg / max(launches, 1), 4),
            "kernel_seconds_per_step": round(kavg, 4),
            "naive_equivalent_TOPs": round(naive / 1e12, 1),
            "counters": {k: prof.get(k) for k in ("valu_issue_util", "valu_busy",
                                                  "executed_lane_ops_per_step",
                                                  "kernel_avg_launch_us", "hbm_GBs")},
            "counters_source": prof.get("source"),
        },
    }
    if not args.no_cpu:
        threads, model = cpu_info()
        log("simba: cpu baseline")
        cb = orc.simba_search(inst.inputs, inst.outputs, 4, [1], 64, 1, args.cpu_size, threads,
                              exhaustive=True)
        # the headline's first solution is the oracle's when it lies within
        # the sizes the oracle swept (else the oracle finds none there)
        if r["size"] <= args.cpu_size:
            assert (cb["status"], cb["size"], cb["rank"], cb["expr"]) == \
                (0, r["size"], r["rank"], r["expr"]), (cb, r)
        else:
            assert cb["status"] != 0, cb
        out["config"]["first_solution_matches_oracle"] = True
        out["cpu_baseline"] = {
            "value": round(cb["candidates"] / cb["seconds"] / 1e9, 5), "unit": "Gcand/s",
            "cores": threads, "kind": "port", "cpu_model": model,
            "sample": "oracle SIMBA sweep of the same instance, every size 1..%d "
                      "(%d candidates, %.1f s, threads pinned); per-candidate cost is "
                      "about constant over sizes, so the rate stands for sizes 1..%d"
                      % (args.cpu_size, cb["candidates"], cb["seconds"], S)}
    if not args.no_configs:
        out["configs"] = {}
        for name, fn in (("c1_latency", rec_c1), ("c2_vfb", rec_c2), ("c3_capped", rec_c3_cap),
                         ("c4_batch", rec_c4), ("c5_learned", rec_c5)):
            log("config record: " + name)
            out["configs"][name] = _safe(fn, args)
    return out


# ---- C1: three leaves, time to the first solution --------------------------
def rec_c1(args):
    """Leaves {x, y, 1}, u64, 16 samples, a uniform size-7 target, sizes 1..9:
    median time to the first solution over repeated calls (latency-bound:
    at most 6.3e6 candidates).  The seed draws a target whose first solution
    has size 7."""
    import numpy as np

    import simba_amd
    from simba_amd.synth import make_instance
    from oracle import oracle as orc

    inst = make_instance(args.small_seed, 2, 16, 7, 64, (1,))
    ctx = simba_amd.Context(0)
    kw = dict(min_size=1, max_size=9)
    for _ in range(3):
        r = ctx.simba(inst, **kw)
    times = []
    for _ in range(100):
        t0 = time.perf_counter()
        r = ctx.simba(inst, **kw)
        times.append(time.perf_counter() - t0)
    ctx.close()
    threads, _ = cpu_info()
    cb = orc.simba_search(inst.inputs, inst.outputs, 2, [1], 64, 1, 9, threads)
    assert (cb["size"], cb["rank"], cb["expr"]) == (r["size"], r["rank"], r["expr"])
    return {"workload": "simba_2vars_const1_u64_16samples_size7target_sizes1to9",
            "p50_ms": round(1e3 * float(np.median(times)), 4),
            "p95_ms": round(1e3 * float(np.percentile(times, 95)), 4),
            "first_solution": r["expr"], "size": r["size"], "rank": r["rank"],
            "matches_oracle": True,
            "cpu_baseline_ms": round(1e3 * cb["seconds"], 4), "cpu_threads": threads}


# ---- C2: VFB levels 1..13 ---------------------------------------------------
def _vfb_instance(args):
    from simba_amd.synth import make_instance
    return make_instance(args.seed, 3, 32, 11, 64, (1, M64))


def rec_c2(args):
    """VFB, vars {x, y, z} + constants {1, -1}, u64, 32 samples, a size-11
    target, every level 1..13 swept (the hash-only last level included):
    candidates/s, algorithmic and atomic rates, kept rows per level; the CPU
    baseline is the multithreaded oracle VFB to level 11 (same counts)."""
    import simba_amd
    from oracle import oracle as orc

    inst = _vfb_instance(args)
    ctx = simba_amd.Context(0)
    kw = dict(max_size=args.vfb_levels, exhaustive=True, cache_bytes=args.cache_gb << 30)
    ctx.vfb(inst, **kw)
    t0 = time.perf_counter()
    reps = 3
    for _ in range(reps):
        r = ctx.vfb(inst, **kw)
    wall = (time.perf_counter() - t0) / reps
    ctx.close()
    nb = sum(r["level_bytes"])
    ks = r["kernel_seconds"]
    rec = {"workload": "vfb_3vars_consts1m1_u64_32samples_levels1to%d" % args.vfb_levels,
           "ms_per_sweep": round(1e3 * wall, 2), "kernel_ms": round(1e3 * ks, 2),
           "Gcand_s": round(r["candidates"] / wall / 1e9, 3),
           "algorithmic_GBs": round(nb / ks / 1e9, 1),
           "hbm_frac": round(nb / ks / 1e9 / HBM_PEAK_GBS, 4),
           "level_new": r["level_new"][1:args.vfb_levels + 1],
           "level_cand": r["level_cand"][1:args.vfb_levels + 1]}
    for k in ("level_cas", "level_atomics"):
        if k in r:
            rec[k.replace("level_", "") + "_per_s"] = round(sum(r[k]) / ks, 1)
    rec["insert_roofline"] = vfb_insert_ceiling(r, args.vfb_levels)
    if not args.no_cpu:
        threads, model = cpu_info()
        lv = args.vfb_cpu_levels
        w = orc.vfb_search_mt(inst.inputs, inst.outputs, 3, [1, M64], 64, lv, threads=threads)
        assert w["level_new"][1:lv + 1] == r["level_new"][1:lv + 1]
        rec["cpu_baseline"] = {"Gcand_s": round(w["candidates"] / w["seconds"] / 1e9, 5),
                               "cores": threads, "kind": "port", "cpu_model": model,
                               "sample": "oracle VFB (level-synchronous, concurrent "
                                         "fingerprint set), same instance, levels 1..%d "
                                         "(%.1f s), kept rows equal to the GPU's"
                                         % (lv, w["seconds"])}
    return rec


def vfb_insert_ceiling(r, last):
    """The hash-only last level's insert against the random-access ceiling
    of its table code: the same table size, as many keys as the level has
    candidates, as many distinct keys as make the level's count of new keys,
    after as many other keys as the earlier levels added (the ceiling's own
    insert and probe-only passes, no fingerprint work)."""
    import simba_amd
    count, new = r["level_cand"][last], r["level_new"][last]
    prefill = sum(r["level_new"][1:last])
    lo, hi = float(new), 8.0 * count  # distinct values whose draws give `new` keys
    for _ in range(60):
        mid = (lo + hi) / 2
        if mid * -math.expm1(-count / mid) < new:
            lo = mid
        else:
            hi = mid
    slots = r["table_slots"]
    c = simba_amd.Context(0)
    try:
        ce = c.vfb_probe_ceiling(slots.bit_length() - 1, count, int(lo), prefill)
    finally:
        c.close()
    t_level = r["level_seconds"][last]
    return {"bound": "random access: a probe of a random 16-byte slot per candidate, "
                     "a 64-bit compare-and-swap per new key",
            "level": last, "candidates": count, "new_keys": new,
            "level_ms": round(1e3 * t_level, 2),
            "ceiling_insert_ms": round(1e3 * ce["insert_seconds"], 2),
            "ceiling_probe_ms": round(1e3 * ce["probe_seconds"], 2),
            "ceiling_fresh": ce["fresh"],
            "achieved_Gprobes_s": round(count / t_level / 1e9, 2),
            "peak_Gprobes_s": round(count / ce["insert_seconds"] / 1e9, 2),
            "frac": round(ce["insert_seconds"] / t_level, 3)}


# ---- C4: batched solving ----------------------------------------------------
def rec_c4(args):
    """4096 instances (vars 2..5, target size 5..17, samples 16..64) by one
    batched call vs one plain call per instance, with the autotuned launch:
    (a) the same size budget for both (every size up to --budget-size), and
    (b) the same time budget per instance (--time-budget-ms: the plain
    call's wall time; the batched instance's share of the device time) with
    sizes up to 17, the configuration's largest target.  Solves/s, solve rate
    and exact match on the instances both solved."""
    import simba_amd
    from simba_amd import autotune
    from simba_amd.learn import sample_instances

    insts = sample_instances(args.seed, args.count, (5, 17), (2, 5), (16, 64))
    ctx = simba_amd.Context(0)
    if args.wg_per_cu:
        tuned = {"wg_per_cu": args.wg_per_cu}
    else:
        tuned, _ = autotune.tune_batch(ctx, insts[:1024], args.budget_size, store=False)

    def compare(kw, reps):
        ctx.batch(insts, **kw)
        t0 = time.perf_counter()
        for _ in range(reps):
            rb = ctx.batch(insts, **kw)
        t_batch = (time.perf_counter() - t0) / reps
        sb = sum(r["status"] == 0 for r in rb)
        if args.batch_only:  # the batched calls alone (for profiling them)
            return {"solves_per_s": round(sb / t_batch, 1),
                    "solve_rate": round(sb / len(insts), 4), "ms_per_call": round(1e3 * t_batch, 2)}
        t0 = time.perf_counter()
        rp = [ctx.simba(x, **{k: v for k, v in kw.items() if k != "wg_per_cu"}) for x in insts]
        t_plain = time.perf_counter() - t0
        sp = sum(r["status"] == 0 for r in rp)
        both = [i for i in range(len(insts)) if rb[i]["status"] == 0 and rp[i]["status"] == 0]
        return {"solves_per_s": round(sb / t_batch, 1), "solve_rate": round(sb / len(insts), 4),
                "ms_per_call": round(1e3 * t_batch, 2),
                "plain_per_instance_solves_per_s": round(sp / t_plain, 1),
                "plain_solve_rate": round(sp / len(insts), 4),
                "plain_seconds": round(t_plain, 3),
                "speedup_vs_plain": round((sb / t_batch) / (sp / t_plain), 2) if sp else None,
                "exact_match_solved_by_both": round(
                    sum(_key(rb[i]) == _key(rp[i]) for i in both) / max(len(both), 1), 4),
                "solved_by_both": len(both)}

    size_rec = compare(dict(max_size=args.budget_size, **tuned), 3)
    out = {"workload": "batch_%d_vars2to5_size5to17_samples16to64" % args.count,
           "launch": tuned,
           "solves_per_s": size_rec["solves_per_s"], "solve_rate": size_rec["solve_rate"],
           "size_budget_%d" % args.budget_size: size_rec}
    if not args.batch_only:
        out["time_budget_%gms_sizes_to_17" % args.time_budget_ms] = compare(
            dict(max_size=17, budget_s=args.time_budget_ms / 1e3, **tuned), 1)
    ctx.close()
    return out


# ---- C5: learned proposal ---------------------------------------------------
def rec_c5(args):
    """Learned block proposal: trained on 200k instances (sizes 3..13, 2..4
    vars, 32 samples) labelled by the enumerator (the plain search's solution
    within the labelling budget, else the generating tree at its size),
    evaluated on held-out instances (sizes 9..19) against the same network
    untrained and the plain search: time to solution per instance, overall
    and on the instances whose solution size the network ordered, the
    held-out ranking of the solution block, and batched solves/s."""
    import numpy as np

    import simba_amd
    from simba_amd import learn

    ctx = simba_amd.Context(0)
    t0 = time.perf_counter()
    X, Y, Z = learn.build_dataset(ctx, args.train, args.seed, max_size=args.label_size, log=log)
    t_label = time.perf_counter() - t0
    Xe, Ye, Ze = learn.build_dataset(ctx, 8000, args.seed + 2 * 10**6, max_size=args.label_size)
    mlp = simba_amd.Mlp(ctx, seed=args.seed)
    untrained = simba_amd.Mlp(ctx, seed=args.seed + 1)
    t0 = time.perf_counter()
    losses = learn.train(mlp, X, Y, epochs=args.epochs, log=log)
    t_train = time.perf_counter() - t0
    big = Ze >= 10
    ranking = {"held_out_rows": int(len(Ye)),
               "top1": {"trained": round(learn.top1_accuracy(mlp, Xe, Ye), 3),
                        "untrained": round(learn.top1_accuracy(untrained, Xe, Ye), 3)},
               "mean_position": {"trained": round(learn.mean_label_position(mlp, Xe, Ye), 2),
                                 "untrained": round(learn.mean_label_position(untrained, Xe, Ye),
                                                    2)},
               "mean_position_sizes_10_13": {
                   "trained": round(learn.mean_label_position(mlp, Xe[big], Ye[big]), 2),
                   "untrained": round(learn.mean_label_position(untrained, Xe[big], Ye[big]),
                                      2)}}
    ev = learn.sample_instances(args.seed + 10**6, args.eval_count, (9, 19), (2, 4), (32, 32))
    kw = dict(max_size=args.learned_budget)
    # the batched calls' launch, autotuned on a part of the set at a smaller budget
    if args.wg_per_cu:
        tuned = {"wg_per_cu": args.wg_per_cu}
    else:
        from simba_amd import autotune
        tuned, _ = autotune.tune_batch(ctx, ev[:1024], min(11, args.learned_budget), store=False)
    bkw = dict(kw, **tuned)
    runs = {}
    for name, fn in (("plain", lambda x: ctx.simba(x, **kw)), ("learned", lambda x: mlp.solve(x, **kw)),
                     ("untrained", lambda x: untrained.solve(x, **kw))):
        fn(ev[0])
        t0 = time.perf_counter()
        out = []
        for k, x in enumerate(ev):
            out.append(fn(x))
            if (k + 1) % 1024 == 0:
                log("c5 %s per instance: %d/%d, %.1f s" % (name, k + 1, len(ev),
                                                         time.perf_counter() - t0))
        runs[name] = out
        runs[name + "_seconds"] = time.perf_counter() - t0
    plain, lrn, unt = runs["plain"], runs["learned"], runs["untrained"]
    solved = [i for i, r in enumerate(plain) if r["status"] == 0]
    engaged = [i for i in solved if lrn[i]["ordered_sizes"] > 0 and
               lrn[i]["size"] >= 10]
    same = {n: all(runs[n][i]["status"] == 0 and runs[n][i]["size"] == plain[i]["size"]
                   for i in solved) for n in ("learned", "untrained")}
    verified = all(_verified(lrn[i], ev[i]) and _verified(unt[i], ev[i]) for i in solved)
    pct = lambda name, idx, q: round(1e3 * float(np.percentile(
        [runs[name][i]["seconds"] for i in idx], q)), 4) if idx else None
    t_rec = {q: {n: pct(n, solved, q) for n in ("plain", "learned", "untrained")}
             for q in (50, 95)}
    e_rec = {q: {n: pct(n, engaged, q) for n in ("plain", "learned", "untrained")}
             for q in (50, 95)}
    log("c5 batched warm-up")
    mlp.batch_solve(ev, **bkw)
    batched = {"launch": tuned}
    for name, fn in (("learned", lambda: mlp.batch_solve(ev, **bkw)),
                     ("untrained", lambda: untrained.batch_solve(ev, **bkw)),
                     ("plain", lambda: ctx.batch(ev, **bkw))):
        t0 = time.perf_counter()
        rb = fn()
        dt = time.perf_counter() - t0
        log("c5 batched %s: %.1f s" % (name, dt))
        batched[name] = {"solves_per_s": round(sum(r["status"] == 0 for r in rb) / dt, 1),
                         "solve_rate": round(sum(r["status"] == 0 for r in rb) / len(ev), 4),
                         "same_size": all(rb[i]["status"] == 0 and rb[i]["size"] ==
                                          plain[i]["size"] for i in solved)}
    ctx.close()
    plain_rate = len(solved) / runs["plain_seconds"]
    return {"workload": "learned_train%d_eval%d_sizes9to19_budget%d"
                        % (args.train, args.eval_count, args.learned_budget),
            "train_instances": args.train, "train_rows": int(len(Y)),
            "rows_sizes_10_13": int((Z >= 10).sum()),
            "label_seconds": round(t_label, 2), "train_seconds": round(t_train, 2),
            "final_loss": round(losses[-1], 4), "ranking": ranking,
            "solved": len(solved), "solve_rate": round(len(solved) / len(ev), 4),
            "same_size": same, "verified": verified,
            "p50_ms": t_rec[50], "p95_ms": t_rec[95],
            "engaged_instances": len(engaged),
            "engaged_p50_ms": e_rec[50], "engaged_p95_ms": e_rec[95],
            "batched": batched,
            "plain_per_instance_solves_per_s": round(plain_rate, 1),
            "batched_learned_vs_plain_per_instance":
                round(batched["learned"]["solves_per_s"] / plain_rate, 2) if plain_rate else None}


# ---- single-record workloads ------------------------------------------------
def _line(metric, value, unit, hib, ms, rec, dtype="u64", steps=1, warmup=1):
    return {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps,
            "warmup": warmup, "ms_per_step": ms, "higher_is_better": hib, "scaling": "weak",
            "vs_baseline": None, "dtype": dtype, "data": "synthetic",
            "config": rec}


def run_simba_small(args):
    rec = rec_c1(args)
    return _line("time to first solution", rec["p50_ms"], "ms", False, rec["p50_ms"], rec,
                 steps=100, warmup=3)


def run_vfb(args):
    rec = rec_c2(args)
    return _line("Gcand/s evaluated", rec["Gcand_s"], "Gcand/s", True, rec["ms_per_sweep"],
                 rec, steps=3)


def run_batch(args):
    rec = rec_c4(args)
    return _line("batched solves/s", rec["solves_per_s"], "solves/s", True,
                 rec["size_budget_%d" % args.budget_size]["ms_per_call"], rec, steps=3)


def run_learned(args):
    rec = rec_c5(args)
    p = rec["p50_ms"]
    v = round(p["plain"] / p["learned"], 3) if p["plain"] and p["learned"] else None
    return _line("learned vs plain p50 time-to-solution", v, "x (plain p50 / learned p50)",
                 True, p["learned"], rec, dtype="u64 search, bf16 MFMA MLP")


def run_simba_cap(args):
    rec = rec_c3_cap(args)
    return _line("Gcand/s evaluated", rec["Gcand_s"], "Gcand/s", True, rec["wall_ms"], rec,
                 warmup=0)


def rec_c3_cap(args):
    """The default instance searched to size 20 under a time cap, as the
    reference's runs are (one run): how far the search gets (the completed
    sizes and the splits of the next size checked whole), and its rate over
    the candidates those decided."""
    import simba_amd
    from simba_amd.synth import count_table, make_instance

    inst = make_instance(args.seed, 4, 64, 15, 64, (1,))
    ctx = simba_amd.Context(0)
    t0 = time.perf_counter()
    r = ctx.simba(inst, min_size=1, max_size=20, exhaustive=True, budget_s=args.cap,
                  table_bytes=args.table_gb << 30)
    wall = time.perf_counter() - t0
    T = count_table(5, 20)
    done = sum(T[s][8] for s in range(1, r["size_reached"] + 1))
    # the size the cap stopped: its splits {j, s-1-j} checked whole (every
    # binary operator instance; the unary trees of that size are the windows
    # of the smaller sizes, already checked)
    s1, part = r["size_reached"] + 1, 0
    for j in range(1, r["split_reached"] + 1):
        m = s1 - 1 - j
        part += (5 + (1 if j == m else 2)) * T[j][8] * T[m][8]
    # and of the split after those, the value pairs of the B blocks whose
    # launches had completed (each pair decides the split's 6 or 7 trees)
    j1 = r["split_reached"] + 1
    m1 = s1 - 1 - j1
    part += (5 + (1 if j1 == m1 else 2)) * r["split_pairs"]
    # the unary-rooted trees of that size, decided by the windows of the
    # completed sizes (counted once the size is entered)
    unary = 2 * T[s1 - 1][8] if (r["split_reached"] or r["split_pairs"]) else 0
    part += unary
    # the completed sizes alone, same windows, no cap: their own time
    rs = ctx.simba(inst, min_size=1, max_size=20, exhaustive=True, stop_size=r["size_reached"],
                   table_bytes=args.table_gb << 30)
    rec = {"workload": "simba_4vars_const1_u64_64samples_sizes1to20_cap%ds" % args.cap,
           "status": r["status"], "size_completed": r["size_reached"],
           "candidates_completed": done,
           "completed_sizes_alone": {"seconds": round(rs["seconds"], 3),
                                     "Gcand_s": round(done / rs["seconds"] / 1e9, 3)},
           "furthest": {"size": s1, "splits_checked": r["split_reached"],
                        "of_splits": (s1 - 1) // 2,
                        "next_split_pairs_checked": r["split_pairs"],
                        "next_split_pairs": T[j1][8] * T[m1][8],
                        "unary_trees": unary, "candidates": part},
           "search_seconds": round(r["seconds"], 3),
           "rate_completed_sizes_Gcand_s": round(done / r["seconds"] / 1e9, 3),
           "first_solution": r["expr"], "first_solution_size": r["size"],
           "Gcand_s": round((done + part) / r["seconds"] / 1e9, 3),
           "wall_ms": round(1e3 * wall, 1)}
    ctx.close()
    return rec


def run_copy(args):
    """Device-to-device copy of 1 GiB per launch: the kernel of known traffic
    (1 GiB read, 1 GiB written) that scales the FETCH_SIZE / WRITE_SIZE
    counters of the profiled kernels."""
    import torch
    n = 1 << 30
    src = torch.ones(n, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(args.warmup):
        dst.copy_(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dst.copy_(src)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    return _line("copy GB/s", round(2 * n / dt / 1e9, 1), "GB/s", True, round(1e3 * dt, 3),
                 {"workload": "d2d_copy_1GiB"}, dtype="u8", steps=args.steps,
                 warmup=args.warmup)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="simba",
                    choices=["simba", "simba_small", "simba_cap", "vfb", "batch", "learned",
                             "copy"])
    ap.add_argument("--no-configs", action="store_true",
                    help="simba: leave out the records of the other configurations")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--size", type=int, default=15, help="simba: last size swept")
    ap.add_argument("--window-max", type=int, default=20,
                    help="simba: the search's max size (unary-chain windows of span max - s)")
    ap.add_argument("--cpu-size", type=int, default=11)
    ap.add_argument("--table-gb", type=int, default=40)
    ap.add_argument("--seed", type=int, default=15)
    ap.add_argument("--small-seed", type=int, default=27,
                    help="C1 instance seed (27: no equivalent below size 7)")
    ap.add_argument("--vfb-levels", type=int, default=13)
    ap.add_argument("--vfb-cpu-levels", type=int, default=11)
    ap.add_argument("--cache-gb", type=int, default=96)
    ap.add_argument("--count", type=int, default=4096)
    ap.add_argument("--budget-size", type=int, default=9)
    ap.add_argument("--batch-only", action="store_true",
                    help="batch: time the batched calls alone (no per-instance comparison)")
    ap.add_argument("--time-budget-ms", type=float, default=2.0,
                    help="C4: per-instance time budget of the sizes-to-17 comparison")
    ap.add_argument("--learned-budget", type=int, default=13,
                    help="C5: largest size searched (the learned order applies to sizes "
                         "of at least 2^26 trees)")
    ap.add_argument("--train", type=int, default=200000)
    ap.add_argument("--eval-count", type=int, default=4096)
    ap.add_argument("--label-size", type=int, default=11,
                    help="C5: labels from the search up to this size (else the generating tree)")
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--cap", type=int, default=60, help="simba_cap: time cap in seconds")
    ap.add_argument("--lanes-b", type=int, default=0, help="B values per lane (4 or 8)")
    ap.add_argument("--wg-per-cu", type=int, default=0, help="workgroups per CU")
    ap.add_argument("--autotune", action="store_true",
                    help="tune launch parameters on the workload before timing")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: every box here has a single MI355X; "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    out = {"simba": run_simba, "vfb": run_vfb, "batch": run_batch, "learned": run_learned,
           "copy": run_copy, "simba_cap": run_simba_cap,
           "simba_small": run_simba_small}[args.workload](args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
