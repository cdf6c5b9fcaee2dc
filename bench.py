"""Benchmark of the MI355X SIMBA/VFB search.

Default workload (one JSON line on stdout): the largest single-instance SIMBA
shape -- leaves {a, b, c, d, 1}, uint64 words, 64 sample points, a uniform
random size-15 target -- swept exhaustively over every size 1..S (S = 15 by
default, 3.90e13 candidate trees).  One step = one full sweep; inputs are
resident on the device before the timed region.  `value` is candidates per
second over the K timed steps (Gcand/s).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload simba|vfb]

roofline: the fused pair kernel (k_pair_check) is integer-VALU bound.
Algorithmic work per candidate tree = sum over its internal nodes of c(op)
(64-bit words: & | ^ ~ + - neg = 2 lane-ops, * = 4; 32-bit words: 1), times
the samples evaluated (1 for all but the rare sample-0 hits).  `achieved` sums
that over the candidates the pair kernel decides and divides by the kernel's
HIP-event time on the library stream (every launch of a step).  The kernel
takes most of a tree's value from the per-size value tables, so this
naive-equivalent rate can exceed the lane-op rate the hardware executes; the
counter figures (VALU busy, issued VALU instructions) and the HBM traffic per
launch come from the rocprofv3 summary under profiles/ for this workload.
Peak: 256 CU x 64 lanes x 2.4 GHz = 39.3 T lane-ops/s.

cpu_baseline: the CPU oracle (multithreaded C restatement of the reference
semantics, oracle/) on the same instance for sizes 1..11.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

VALU_PEAK = 256 * 64 * 2.4e9  # integer lane-ops/s
HBM_PEAK_GBS = 8000.0


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


def pair_kernel_ops(T, C, L, max_size, W):
    """Naive-equivalent lane-ops of the trees the pair kernel decides: every
    tree of sizes 3..max_size except the unary chains over a single leaf
    (L * 2^(s-1) trees of s-1 unary nodes each), which the leaf scan covers."""
    c_un = 2 if W == 64 else 1
    tot = 0
    for s in range(3, max_size + 1):
        tot += C[s] - L * (1 << (s - 1)) * (s - 1) * c_un
    return tot


def profile_fields(name):
    """Counter figures recorded by rocprofv3 for the pair kernel on this
    workload: the `bench-fields:` JSON line of profiles/<name>, or {}."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("bench-fields:"):
                    return json.loads(line.split(":", 1)[1])
    except OSError:
        pass
    return {}


def cpu_baseline_simba(inst, max_size, threads):
    from oracle import oracle as orc
    r = orc.simba_search(inst.inputs, inst.outputs, inst.n_variables,
                         [int(c) for c in inst.consts], inst.width, 1, max_size, threads,
                         exhaustive=True)
    return r


def run_simba(args):
    import simba_amd
    from simba_amd.synth import make_instance

    inst = make_instance(args.seed, 4, 64, 15, 64, (1,))
    S = args.size
    L = 5
    T, C = naive_ops_per_size(L, S, 64)
    cands = sum(T[s][8] for s in range(1, S + 1))
    ctx = simba_amd.Context(0)
    kw = dict(min_size=1, max_size=args.window_max, stop_size=S, exhaustive=True,
              table_bytes=args.table_gb << 30)
    tuned = {k: v for k, v in (("lanes_b", args.lanes_b), ("wg_per_cu", args.wg_per_cu)) if v}
    kw.update(tuned)
    if args.autotune:
        from simba_amd import autotune
        tuned, _ = autotune.tune_simba(ctx, inst, S - 1)
        kw.update(tuned)
    for _ in range(args.warmup):
        ctx.simba(inst, **kw)
    times, ktimes = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = ctx.simba(inst, **kw)
        times.append(r["seconds"])
        ktimes.append(r["kernel_seconds"])
        assert r["size_reached"] == S and r["candidates"] == cands, r
    wall = time.perf_counter() - t0
    ms_per_step = 1e3 * wall / args.steps
    value = cands * args.steps / wall / 1e9
    kavg = sum(ktimes) / len(ktimes)
    launches = r["kernel_launches"]
    pk_ops = pair_kernel_ops(T, C, L, S, 64)
    achieved = pk_ops / kavg / 1e12
    prof = profile_fields("k_pair_check_simba_sweep_sizes1to%d.md" % S)
    out = {
        "metric": "Gcand/s evaluated",
        "value": round(value, 3),
        "unit": "Gcand/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": "simba_sweep_4vars_const1_u64_64samples_sizes1to%d" % S,
                   "target": inst.original_formula, "target_size": 15,
                   "candidates_per_step": cands, "first_solution": r["expr"],
                   "first_solution_size": r["size"], "launch": tuned or "default"},
        "roofline": {"bound": "valu", "achieved": round(achieved, 3),
                     "peak": round(VALU_PEAK / 1e12, 3), "unit": "TOP/s",
                     "frac": round(achieved * 1e12 / VALU_PEAK, 4),
                     "traffic": prof.get("traffic_bytes_per_launch"),
                     "kernel": "k_pair_check",
                     "definition": "naive-equivalent 32-bit lane-ops of the candidates the "
                                   "pair kernel decides (sum of c(op) over every internal "
                                   "node; u64: 2 per op, 4 per *) / pair-kernel HIP-event "
                                   "time; the kernel reuses tabulated subtree values, so "
                                   "this can exceed the executed rate",
                     "launches_per_step": launches,
                     "avg_launch_ms": round(1e3 * kavg / max(launches, 1), 4),
                     "kernel_seconds_per_step": round(kavg, 4),
                     "counter_valu_busy": prof.get("valu_busy"),
                     "counter_valu_inst_rate_frac": prof.get("valu_inst_frac"),
                     "counter_source": prof.get("source")},
    }
    ctx.close()
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        cb = cpu_baseline_simba(inst, args.cpu_size, threads)
        out["cpu_baseline"] = {
            "value": round(cb["candidates"] / cb["seconds"] / 1e9, 5), "unit": "Gcand/s",
            "cores": threads, "kind": "port",
            "sample": "oracle SIMBA sweep of the same instance, every size 1..%d "
                      "(%d candidates, %.1f s)" % (args.cpu_size, cb["candidates"], cb["seconds"])}
    return out


def run_simba_small(args):
    """Leaves {x, y, 1}, u64, 16 samples, a uniform size-7 target, sizes 1..9:
    time to the first solution (latency-bound: at most 6.3e6 candidates).
    The default seed draws a target whose first solution has size 7."""
    import numpy as np

    import simba_amd
    from simba_amd.synth import make_instance

    inst = make_instance(args.small_seed, 2, 16, 7, 64, (1,))
    ctx = simba_amd.Context(0)
    kw = dict(min_size=1, max_size=9)
    for _ in range(max(args.warmup, 1)):
        r = ctx.simba(inst, **kw)
    reps = max(args.steps, 1) * 20
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        r = ctx.simba(inst, **kw)
        times.append(time.perf_counter() - t0)
    ctx.close()
    med = float(np.median(times))
    out = {
        "metric": "time to first solution", "value": round(1e3 * med, 4), "unit": "ms",
        "n_gpus": 1, "steps": reps, "warmup": max(args.warmup, 1),
        "ms_per_step": round(1e3 * med, 4), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "u64", "data": "synthetic",
        "config": {"workload": "simba_2vars_const1_u64_16samples_size7target_sizes1to9",
                   "target": inst.original_formula, "first_solution": r["expr"],
                   "size": r["size"], "rank": r["rank"], "candidates": r["candidates"],
                   "p95_ms": round(1e3 * float(np.percentile(times, 95)), 4)},
    }
    if not args.no_cpu:
        from oracle import oracle as orc
        threads = min(16, os.cpu_count() or 1)
        cb = orc.simba_search(inst.inputs, inst.outputs, 2, [1], 64, 1, 9, threads)
        assert (cb["size"], cb["rank"]) == (r["size"], r["rank"])
        out["cpu_baseline"] = {"value": round(1e3 * cb["seconds"], 4), "unit": "ms",
                               "cores": threads, "kind": "port",
                               "sample": "oracle SIMBA search of the same instance, sizes 1..9"}
    return out


def run_simba_cap(args):
    """The same instance swept exhaustively over sizes 1..20 under a time cap
    (one run): how far the search gets, and its rate over completed sizes."""
    import simba_amd
    from simba_amd.synth import count_table, make_instance

    inst = make_instance(args.seed, 4, 64, 15, 64, (1,))
    ctx = simba_amd.Context(0)
    t0 = time.perf_counter()
    r = ctx.simba(inst, min_size=1, max_size=20, exhaustive=True, budget_s=args.cap,
                  table_bytes=args.table_gb << 30)
    wall = time.perf_counter() - t0
    ctx.close()
    T = count_table(5, 20)
    done = sum(T[s][8] for s in range(1, r["size_reached"] + 1))
    return {
        "metric": "Gcand/s evaluated", "value": round(done / r["seconds"] / 1e9, 3),
        "unit": "Gcand/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": round(1e3 * wall, 1), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u64", "data": "synthetic",
        "config": {"workload": "simba_4vars_const1_u64_64samples_sizes1to20_cap%ds" % args.cap,
                   "status": r["status"], "size_completed": r["size_reached"],
                   "candidates_completed": done, "search_seconds": round(r["seconds"], 3),
                   "first_solution": r["expr"], "first_solution_size": r["size"]},
    }


def run_vfb(args):
    import simba_amd
    from simba_amd.synth import make_instance

    inst = make_instance(args.seed, 3, 32, 11, 64, (1, (1 << 64) - 1))
    ctx = simba_amd.Context(0)
    kw = dict(max_size=args.vfb_levels, exhaustive=True, cache_bytes=args.cache_gb << 30)
    for _ in range(args.warmup):
        ctx.vfb(inst, **kw)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = ctx.vfb(inst, **kw)
    wall = time.perf_counter() - t0
    cands = r["candidates"]
    nbytes = sum(r["level_bytes"])
    out = {
        "metric": "Gcand/s evaluated", "value": round(cands * args.steps / wall / 1e9, 3),
        "unit": "Gcand/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u64", "data": "synthetic",
        "config": {"workload": "vfb_3vars_consts1m1_u64_32samples_levels1to%d" % args.vfb_levels,
                   "level_new": r["level_new"][1:args.vfb_levels + 1],
                   "level_cand": r["level_cand"][1:args.vfb_levels + 1]},
        "roofline": {"bound": "hbm", "achieved": round(nbytes / r["kernel_seconds"] / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(nbytes / r["kernel_seconds"] / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": None},
    }
    ctx.close()
    if not args.no_cpu:
        from oracle import oracle as orc
        lv = min(args.vfb_levels, 8)
        t = time.perf_counter()
        w = orc.vfb_search(inst.inputs, inst.outputs, 3, [1, (1 << 64) - 1], 64, lv,
                           exhaustive=True)
        dt = time.perf_counter() - t
        out["cpu_baseline"] = {"value": round(w["candidates"] / dt / 1e9, 5), "unit": "Gcand/s",
                               "cores": 1, "kind": "port",
                               "sample": "oracle VFB, same instance, levels 1..%d" % lv}
    return out


def _key(r):
    return (r["status"], r["size"], r["rank"], r["expr"])


def _verified(r, inst):
    """Re-evaluates a returned SIMBA RPN on every sample (product evaluator)."""
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


def run_batch(args):
    """4096 instances (vars 2..5, target size 5..17, samples 16..64) solved by
    one batched call vs one plain call per instance, same size budget."""
    import simba_amd
    from simba_amd.learn import sample_instances

    insts = sample_instances(args.seed, args.count, (5, 17), (2, 5), (16, 64))
    ctx = simba_amd.Context(0)
    kw = dict(max_size=args.budget_size)
    tuned = {"wg_per_cu": args.wg_per_cu} if args.wg_per_cu else {}
    kw.update(tuned)
    if args.autotune:
        from simba_amd import autotune
        tuned, _ = autotune.tune_batch(ctx, insts[:1024], args.budget_size)
        kw.update(tuned)
    ctx.batch(insts, **kw)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        rb = ctx.batch(insts, verbose=args.verbose, **kw)
    t_batch = (time.perf_counter() - t0) / args.steps
    t0 = time.perf_counter()
    rp = [ctx.simba(x, **kw) for x in insts]
    t_plain = time.perf_counter() - t0
    solved = sum(r["status"] == 0 for r in rb)
    exact = sum(_key(a) == _key(b) for a, b in zip(rb, rp)) / len(insts)
    ctx.close()
    return {
        "metric": "batched solves/s", "value": round(solved / t_batch, 2), "unit": "solves/s",
        "n_gpus": 1, "steps": args.steps, "warmup": 1, "ms_per_step": round(1e3 * t_batch, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": "batch_%d_vars2to5_size5to17_samples16to64_budget%d"
                               % (args.count, args.budget_size),
                   "solved": solved, "exact_match_vs_plain": exact,
                   "plain_per_instance_solves_per_s": round(solved / t_plain, 2),
                   "plain_seconds": round(t_plain, 3),
                   "speedup_vs_plain": round(t_plain / t_batch, 2),
                   "launch": tuned or "default"},
    }


def run_learned(args):
    """Learned block proposal: trained on plain-solved instances (sizes 3..13,
    2..4 vars, 32 samples), evaluated on held-out instances (sizes 9..19)."""
    import numpy as np

    import simba_amd
    from simba_amd import learn

    log = lambda m: print(m, file=sys.stderr, flush=True)
    ctx = simba_amd.Context(0)
    t0 = time.perf_counter()
    X, Y = learn.build_dataset(ctx, args.train, args.seed, max_size=args.label_size, log=log)
    t_label = time.perf_counter() - t0
    mlp = simba_amd.Mlp(ctx, seed=args.seed)
    untrained = simba_amd.Mlp(ctx, seed=args.seed + 1)
    t0 = time.perf_counter()
    losses = learn.train(mlp, X, Y, epochs=args.epochs, log=log)
    t_train = time.perf_counter() - t0
    ev = learn.sample_instances(args.seed + 10**6, args.count, (9, 19), (2, 4), (32, 32))
    kw = dict(max_size=args.learned_budget)
    mlp.solve(ev[0], **kw)
    t0 = time.perf_counter()
    plain = [ctx.simba(x, **kw) for x in ev]
    t_plain_all = time.perf_counter() - t0
    log("plain per-instance done")
    lrn = [mlp.solve(x, **kw) for x in ev]
    log("learned per-instance done")
    unt = [untrained.solve(x, **kw) for x in ev]
    log("untrained per-instance done")
    solved = [i for i, r in enumerate(plain) if r["status"] == 0]
    same_size = all(lrn[i]["status"] == 0 and lrn[i]["size"] == plain[i]["size"] for i in solved)
    verified = all(lrn[i]["status"] == 0 and _verified(lrn[i], ev[i]) for i in solved)
    missed = [i for i in solved if lrn[i]["status"] != 0 or lrn[i]["size"] != plain[i]["size"]]
    tp = np.array([plain[i]["seconds"] for i in solved])
    tl = np.array([lrn[i]["seconds"] for i in solved])
    tu = np.array([unt[i]["seconds"] for i in solved])
    mlp.batch_solve(ev, **kw)  # warm-up: device buffers sized for the set
    ctx.batch(ev, **kw)
    t0 = time.perf_counter()
    bl = mlp.batch_solve(ev, **kw)
    t_bl = time.perf_counter() - t0
    t0 = time.perf_counter()
    bp = ctx.batch(ev, **kw)
    t_bp = time.perf_counter() - t0
    bp_solved = sum(r["status"] == 0 for r in bp)
    b_solved = sum(r["status"] == 0 for r in bl)
    b_same = all(bl[i]["status"] == 0 and bl[i]["size"] == plain[i]["size"] for i in solved)
    ctx.close()
    pct = lambda a, q: round(float(np.percentile(a, q)) * 1e3, 3) if len(a) else None
    # the per-instance path over the whole set, unsolved instances included
    plain_rate = len(solved) / t_plain_all if len(solved) else 0.0
    return {
        "metric": "learned vs plain p50 time-to-solution", "value": round(
            float(np.percentile(tp, 50) / np.percentile(tl, 50)), 3) if len(solved) else None,
        "unit": "x (plain p50 / learned p50)", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": pct(tl, 50), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u64 search, bf16 MFMA MLP", "data": "synthetic",
        "config": {"workload": "learned_train%d_eval%d_sizes9to19_budget%d"
                               % (args.train, args.count, args.learned_budget),
                   "train_examples": int(len(Y)), "label_seconds": round(t_label, 2),
                   "train_seconds": round(t_train, 2), "final_loss": round(losses[-1], 4),
                   "solved_plain": len(solved), "same_size": same_size,
                   "verified": verified, "mismatched": len(missed),
                   "p50_ms": {"plain": pct(tp, 50), "learned": pct(tl, 50),
                              "untrained": pct(tu, 50)},
                   "p95_ms": {"plain": pct(tp, 95), "learned": pct(tl, 95),
                              "untrained": pct(tu, 95)},
                   "batched_learned_solves_per_s": round(b_solved / t_bl, 2),
                   "batched_plain_solves_per_s": round(bp_solved / t_bp, 2),
                   "plain_per_instance_solves_per_s": round(plain_rate, 2),
                   "batched_learned_vs_plain_per_instance": round(
                       (b_solved / t_bl) / plain_rate, 2) if plain_rate else None,
                   "batched_learned_same_size": b_same},
    }


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
    return {"metric": "copy GB/s", "value": round(2 * n / dt / 1e9, 1), "unit": "GB/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(1e3 * dt, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "u8", "data": "synthetic",
            "config": {"workload": "d2d_copy_1GiB"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="simba",
                    choices=["simba", "simba_small", "simba_cap", "vfb", "batch", "learned",
                             "copy"])
    ap.add_argument("--count", type=int, default=4096)
    ap.add_argument("--budget-size", type=int, default=9)
    ap.add_argument("--learned-budget", type=int, default=11,
                    help="learned: largest size searched (the learned block order applies "
                         "to sizes of at least 2^26 trees)")
    ap.add_argument("--train", type=int, default=200000)
    ap.add_argument("--label-size", type=int, default=9)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--size", type=int, default=15)
    ap.add_argument("--window-max", type=int, default=20,
                    help="simba: the search's max size (unary-chain windows of size "
                         "max - s); the sweep ends after --size")
    ap.add_argument("--cpu-size", type=int, default=11)
    ap.add_argument("--table-gb", type=int, default=40)
    ap.add_argument("--vfb-levels", type=int, default=13)
    ap.add_argument("--cache-gb", type=int, default=96)
    ap.add_argument("--seed", type=int, default=15)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cap", type=int, default=60, help="simba_cap: time cap in seconds")
    ap.add_argument("--small-seed", type=int, default=27,
                    help="simba_small: instance seed (27: no equivalent below size 7)")
    ap.add_argument("--lanes-b", type=int, default=0, help="B values per lane (4 or 8)")
    ap.add_argument("--verbose", action="store_true", help="library progress on stderr")
    ap.add_argument("--wg-per-cu", type=int, default=0, help="workgroups per CU")
    ap.add_argument("--autotune", action="store_true",
                    help="tune launch parameters on the workload before timing")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: every box here has a single MI355X; "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    out = {"simba": run_simba, "vfb": run_vfb, "batch": run_batch,
           "learned": run_learned, "copy": run_copy,
           "simba_cap": run_simba_cap, "simba_small": run_simba_small}[args.workload](args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
