"""Writes the golden vectors of the instance pipeline and the Xyntia converter
(tests/golden/pipeline/) by running the reference's own scripts.

Nothing else fixes these functions: their output is defined by the exact
sequence of RNG calls (Python's MT19937 `randrange` in scripts/generate.py,
numpy's PCG64 `integers` in scripts/sample.py) and by the converter's
formatting (scripts/convert_to_xyntia.py), and the reference holds no
fixtures for them.  So this script imports those three files from the
reference checkout, writes their outputs for seed 42, and only the outputs
(formulas, instance JSON, Xyntia JSON) are committed.

Run from the repository root (needs the reference at /root/reference):
    python tests/golden/make_pipeline_golden.py
"""
import importlib.util
import json
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "pipeline")
REF_SCRIPTS = "/root/reference/scripts"

SEED = 42
SIZES = (3, 9)      # generate.py --size_range 3 9
VARS = (1, 3)       # --var_range 1 3
REPEATS = 2         # --repeats 2
N_EXAMPLES = 16     # sample.py --n_examples 16
N_INSTANCES = 8     # instances written for the first formulas
N_XYNTIA = 3


def _load(name):
    spec = importlib.util.spec_from_file_location("ref_" + name,
                                                  os.path.join(REF_SCRIPTS, name + ".py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def main():
    gen, smp, xy = _load("generate"), _load("sample"), _load("convert_to_xyntia")
    if os.path.isdir(OUT):
        shutil.rmtree(OUT)
    os.makedirs(os.path.join(OUT, "instances"))
    os.makedirs(os.path.join(OUT, "xyntia"))
    # generate.py's main loop: sizes, then variable counts, then repeats
    sampler = gen.Sampler(SEED)
    formulas = []
    for size in range(SIZES[0], SIZES[1] + 1):
        for nv in range(VARS[0], VARS[1] + 1):
            for _ in range(REPEATS):
                formulas.append(gen.random_formula(size, n_variables=nv, sampler=sampler))
    with open(os.path.join(OUT, "formulas_seed%d.txt" % SEED), "w") as f:
        for x in formulas:
            f.write(x + "\n")
    # sample.py's main loop over the first formulas
    s = smp.Sampler(SEED)
    names = []
    for i, formula in enumerate(formulas[:N_INSTANCES]):
        doc = smp.sample_formula(formula, N_EXAMPLES, s)
        name = "mba_%d_seed=%d.json" % (i, SEED)
        smp.write_to_file(doc, __import__("pathlib").Path(OUT, "instances", name))
        names.append(name)
    # the converter on the first instances
    from pathlib import Path
    for name in names[:N_XYNTIA]:
        xy.convert_to_xyntia(Path(OUT, "instances", name), Path(OUT, "xyntia", name))
    with open(os.path.join(OUT, "params.json"), "w") as f:
        json.dump({"seed": SEED, "size_range": SIZES, "var_range": VARS, "repeats": REPEATS,
                   "n_examples": N_EXAMPLES, "instances": names,
                   "xyntia": names[:N_XYNTIA]}, f, indent=1)
    print("wrote %d formulas, %d instances, %d Xyntia files to %s"
          % (len(formulas), len(names), N_XYNTIA, OUT))


if __name__ == "__main__":
    sys.exit(main())
