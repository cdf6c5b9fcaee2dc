"""Launch autotuner of the SIMBA kernels.

Tunes, by timing real sweeps on the device:
  lanes_b     -- B values (candidates' right operands) held per lane: 4 or 8;
  wg_per_cu   -- workgroups per CU of the grid-stride launches: 8 to 48;
  table_delta -- how many of the largest value tables to leave to per-tile
                 generation instead of HBM (0 or 1): the table-build /
                 on-the-fly-evaluation trade-off.
Every configuration returns the same answer (checked); only time differs.
Results are cached in a JSON file (SIMBA_AMD_TUNE, default
~/.cache/simba_amd/tune.json) keyed by word width, leaf count and size.
"""
import itertools
import json
import os

SIMBA_SPACE = {"lanes_b": (4, 8), "wg_per_cu": (8, 16, 32, 48), "table_delta": (0, 1)}
BATCH_SPACE = {"wg_per_cu": (4, 8, 16, 24, 32)}


def cache_path():
    return os.environ.get("SIMBA_AMD_TUNE",
                          os.path.join(os.path.expanduser("~"), ".cache", "simba_amd", "tune.json"))


def _load():
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(key, value):
    d = _load()
    d[key] = value
    os.makedirs(os.path.dirname(cache_path()), exist_ok=True)
    with open(cache_path(), "w") as f:
        json.dump(d, f, indent=1, sort_keys=True)


def simba_key(inst, size):
    return f"simba:w{inst.width}:l{inst.n_variables + len(inst.consts)}:s{size}"


def tune_simba(ctx, inst, size, space=SIMBA_SPACE, repeats=1, store=True):
    """Times an exhaustive sweep of sizes 1..size for every configuration;
    returns (best config, {config: seconds})."""
    names = list(space)
    timing = {}
    answer = None
    for combo in itertools.product(*(space[n] for n in names)):
        cfg = dict(zip(names, combo))
        best = None
        for _ in range(repeats + 1):  # first run warms up allocations
            r = ctx.simba(inst, min_size=1, max_size=size, exhaustive=True, **cfg)
            best = r["seconds"] if best is None else min(best, r["seconds"])
        key = (r["status"], r["size"], r["rank"], r["expr"])
        if answer is None:
            answer = key
        elif key != answer:
            raise RuntimeError(f"configuration {cfg} changed the answer")
        timing[json.dumps(cfg, sort_keys=True)] = best
    best_cfg = json.loads(min(timing, key=timing.get))
    if store:
        _store(simba_key(inst, size), best_cfg)
    return best_cfg, timing


def tune_batch(ctx, insts, max_size, space=BATCH_SPACE, store=True):
    """Times a batched call on `insts` for every configuration (each must
    give the same answers); returns (best config, {config: seconds})."""
    import time
    names = list(space)
    timing = {}
    answer = None
    for combo in itertools.product(*(space[n] for n in names)):
        cfg = dict(zip(names, combo))
        ctx.batch(insts[: min(64, len(insts))], max_size=max_size, **cfg)
        t = time.perf_counter()
        res = ctx.batch(insts, max_size=max_size, **cfg)
        timing[json.dumps(cfg, sort_keys=True)] = time.perf_counter() - t
        key = [(r["status"], r["size"], r["rank"]) for r in res]
        if answer is None:
            answer = key
        elif key != answer:
            raise RuntimeError(f"configuration {cfg} changed the answers")
    best_cfg = json.loads(min(timing, key=timing.get))
    if store:
        _store(f"batch:w{insts[0].width}:s{max_size}", best_cfg)
    return best_cfg, timing


def tuned(key, default=None):
    """The cached configuration for `key`, or `default`."""
    return _load().get(key, default or {})
