"""Writes the golden fixtures of tests/golden/.

Instances: seeded synthetic instances in the reference's JSON schema
(scripts/sample.py:43-65), 32-bit words like the reference's data set
(16 samples, 1-3 variables), plus two 64-bit ones.  Expected results: the CPU
oracle's SIMBA answer in the reference's size range (simba_exe <file> 8:
sizes 1..7) and VFB answer (vfb_exe <file> 7: levels 1..7), as the strings
the two CLIs print.

Run from the repository root:  python tests/golden/make_golden.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as orc  # noqa: E402
from simba_amd import synth  # noqa: E402

CASES = [
    # (seed, n_vars, n_samples, target size, width)
    (1, 1, 16, 3, 32), (2, 2, 16, 3, 32), (3, 2, 16, 4, 32), (4, 2, 16, 5, 32),
    (5, 3, 16, 5, 32), (6, 2, 16, 6, 32), (7, 3, 16, 6, 32), (8, 1, 16, 7, 32),
    (9, 2, 16, 7, 32), (10, 3, 16, 4, 32), (11, 2, 16, 1, 32), (12, 3, 16, 9, 32),
    (13, 2, 16, 5, 64), (14, 3, 16, 6, 64),
]
SIMBA_MAXLEN, VFB_MAXLEN = 8, 7


def main():
    expected = {}
    for seed, nv, ns, ts, w in CASES:
        inst = synth.make_instance(1000 + seed, nv, ns, ts, w)
        name = f"mba_{seed}_w{w}.json"
        with open(os.path.join(HERE, name), "w") as f:
            f.write(synth.to_json(inst))
        s = orc.simba_search(inst.inputs, inst.outputs, nv, (), w, 2, SIMBA_MAXLEN - 1, 4)
        v = orc.vfb_search(inst.inputs, inst.outputs, nv, (), w, VFB_MAXLEN)
        expected[name] = {
            "simba": {"expr": s["expr"] if s["status"] == 0 else "Not found !",
                      "size": s["size"], "rank": s["rank"]},
            "vfb": {"expr": v["expr"], "level": v["size"], "index": v["rank"]},
        }
    with open(os.path.join(HERE, "expected.json"), "w") as f:
        json.dump({"simba_maxlen": SIMBA_MAXLEN, "vfb_maxlen": VFB_MAXLEN,
                   "cases": expected}, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
