"""Experiment harness: every instance of a folder through one search tool,
one CSV (run_all.sh:6-43).

The reference runs `timeout 60 ./<tool>_exe <instance> 20` once per instance
and appends each process's stdout line `<path>,<formula>,<seconds>` to a CSV
whose header is `instance,formula,time`; a run that times out or fails adds
no line.  Here one process and one device context serve the whole folder:
the per-instance cap is the library's time budget (a capped search returns
MBA_BUDGET and adds no line; a search that ends past the cap adds none
either, as the killed process would not), and with --batch the folder is
solved by the batched search, the cap bounding each instance's share of the
device time (each step's time split over the instances by their work), so a
folder may take longer than the cap while every instance stays within it.
Lines match the
CLIs' (path as globbed, the reference's printing of the formula, seconds with
std::cout's default 6 significant digits).

    python -m simba_amd.harness instances/ --tool simba -o res_simba.csv
    python -m simba_amd.harness instances/ --tool vfb --max-len 20 --timeout 60
    python -m simba_amd.harness instances/ --tool rtid     (SIMBA-RTid order)
"""
import argparse
import glob
import os
import sys

HEADER = "instance,formula,time"


def cout_double(x):
    """std::cout << double with the default stream state (%g, precision 6)."""
    return "%g" % x


def csv_line(path, formula, seconds):
    return "%s,%s,%s" % (path, formula, cout_double(seconds))


def instance_paths(folder):
    """The reference's `for i in ./instances/*.json` (glob order is sorted)."""
    return sorted(glob.glob(os.path.join(folder, "*.json")))


def run(paths, tool="simba", max_len=20, timeout_s=60.0, batch=False, native=False,
        ctx=None, log=None):
    """Solves every instance; returns the CSV lines (header first).  A search
    stopped by the cap contributes no line, as a killed process does."""
    import simba_amd
    own = ctx is None
    ctx = ctx or simba_amd.Context(0)
    lines = [HEADER]
    try:
        insts, kept = [], []
        for p in paths:
            try:
                insts.append(simba_amd.MbaInstance.from_json_file(p))
                kept.append(p)
            except simba_amd.MbaError as e:  # the reference's CLI exits; no line
                if log:
                    log("skip %s: %s" % (p, e))
        kw = dict(max_size=max_len, compat=not native, budget_s=timeout_s)
        if tool == "simba" and batch:
            results = ctx.batch(insts, **kw)
        else:
            results = []
            for p, inst in zip(kept, insts):
                if tool == "simba":
                    r = ctx.simba(inst, **kw)
                elif tool == "rtid":
                    r = ctx.simba_rtid(inst, **kw)
                elif tool == "vfb":
                    if inst.n_examples > 100:  # full-cache.cu:671-675
                        r = {"status": 0, "expr": "see_the_error", "seconds": 0.0}
                    else:
                        r = ctx.vfb(inst, **kw)
                else:
                    raise ValueError("unknown tool %r" % tool)
                results.append(r)
                if log:
                    log("%s: status %d" % (p, r["status"]))
        for p, r in zip(kept, results):
            if r["status"] == simba_amd.BUDGET:
                continue
            if not batch and timeout_s > 0 and r["seconds"] > timeout_s:
                continue  # finished after the cap: `timeout` would have killed it
            lines.append(csv_line(p, r["expr"], r["seconds"]))
    finally:
        if own:
            ctx.close()
    return lines


def main(argv=None):
    ap = argparse.ArgumentParser(description="Run a search tool over a folder of instances.")
    ap.add_argument("folder")
    ap.add_argument("--tool", choices=["simba", "rtid", "vfb"], default="simba")
    ap.add_argument("--max-len", type=int, default=20)
    ap.add_argument("--timeout", type=float, default=60.0, help="per-instance cap, seconds")
    ap.add_argument("--batch", action="store_true",
                    help="SIMBA: solve the folder with the batched persistent kernel")
    ap.add_argument("--native", action="store_true", help="sizes 1..max-len inclusive")
    ap.add_argument("-o", "--output", default=None, help="CSV file (default: stdout)")
    a = ap.parse_args(argv)
    lines = run(instance_paths(a.folder), a.tool, a.max_len, a.timeout, a.batch, a.native,
                log=lambda m: print(m, file=sys.stderr, flush=True))
    text = "\n".join(lines) + "\n"
    if a.output:
        with open(a.output, "w") as f:
            f.write(text)
    else:
        sys.stdout.write(text)
    return 0


if __name__ == "__main__":
    sys.exit(main())
