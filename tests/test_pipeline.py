"""The callers around the search: the instance pipeline (formulas and
sampled instances), the Xyntia converter and the experiment harness.

The pipeline and converter are pinned by the golden vectors that
tests/golden/make_pipeline_golden.py wrote with the reference's own scripts
(seed 42); the harness by the CLIs' CSV contract (run_all.sh, nocache.cu:340,
full-cache.cu:845) and, on the GPU, by the plain path it calls."""
import json
import os

import numpy as np
import pytest

from simba_amd import harness, instances, xyntia

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "pipeline")
PARAMS = json.load(open(os.path.join(GOLD, "params.json")))


def test_formulas_match_reference_generator():
    want = [ln.rstrip("\n") for ln in open(os.path.join(GOLD, "formulas_seed42.txt"))]
    got = instances.generate(PARAMS["size_range"], PARAMS["var_range"], PARAMS["repeats"],
                             PARAMS["seed"])
    assert got == want


def test_formula_sizes_and_variables():
    fs = instances.generate((1, 12), (1, 4), 3, 7)
    for f in fs:
        tree = instances.parse(f)

        def size(t):
            return 1 if t[0] == "v" else 1 + sum(size(c) for c in t[1:])
        assert size(tree) in range(1, 13)
        assert instances.count_variables(f) <= 4


def test_instances_match_reference_sampler_byte_for_byte(tmp_path):
    formulas = [ln.strip() for ln in open(os.path.join(GOLD, "formulas_seed42.txt"))]
    names = instances.write_instances(formulas[:len(PARAMS["instances"])],
                                      PARAMS["n_examples"], tmp_path, PARAMS["seed"])
    assert names == PARAMS["instances"]
    for n in names:
        assert open(tmp_path / n).read() == open(os.path.join(GOLD, "instances", n)).read(), n


def test_uint32_evaluation_against_python_ints():
    rng = np.random.default_rng(3)
    for f in instances.generate((1, 10), (1, 3), 4, 11):
        tree = instances.parse(f)
        vals = [int(x) for x in rng.integers(0, 2**32, size=3, dtype=np.uint64)]
        env = {"v%d" % i: vals[i] for i in range(3)}
        # an independent reading: Python big ints, reduced mod 2^32 at the end
        # (wrap-around + - * ~ commute with the final reduction)
        py = eval(f.replace("~", "-1-"), {}, env) % (1 << 32)  # noqa: S307 (own formulas)
        assert instances.evaluate(tree, vals) == py, f


def test_xyntia_matches_reference_converter(tmp_path):
    src = tmp_path / "in"
    src.mkdir()
    for n in PARAMS["xyntia"]:
        (src / n).write_text(open(os.path.join(GOLD, "instances", n)).read())
    done = xyntia.convert_folder(src, tmp_path / "out")
    assert done == sorted(PARAMS["xyntia"])
    for n in PARAMS["xyntia"]:
        assert (tmp_path / "out" / n).read_text() == open(os.path.join(GOLD, "xyntia", n)).read()


def test_xyntia_rejects_empty_instance():
    with pytest.raises(ValueError):
        xyntia.convert({"examples": []})


def test_harness_csv_contract():
    assert harness.HEADER == "instance,formula,time"
    # std::cout << double: 6 significant digits, trailing zeros dropped
    assert harness.cout_double(0.5) == "0.5"
    assert harness.cout_double(0.000123456789) == "0.000123457"
    assert harness.cout_double(1234567.0) == "1.23457e+06"
    assert harness.cout_double(12.0) == "12"
    assert harness.csv_line("./instances/a.json", "(v0 + v1)", 0.25) == \
        "./instances/a.json,(v0 + v1),0.25"


def test_harness_instance_order(tmp_path):
    for n in ("b.json", "a.json", "c.txt"):
        (tmp_path / n).write_text("{}")
    assert [os.path.basename(p) for p in harness.instance_paths(str(tmp_path))] == \
        ["a.json", "b.json"]


@pytest.mark.gpu
def test_harness_runs_the_plain_path(tmp_path):
    import simba_amd
    from tests.gpu_util import note, oracle_simba
    note("harness over the golden pipeline instances")
    names = PARAMS["instances"][:4]
    paths = [os.path.join(GOLD, "instances", n) for n in names]
    lines = harness.run(paths, "simba", max_len=8, timeout_s=30)
    assert lines[0] == harness.HEADER and len(lines) == 1 + len(paths)
    for p, ln in zip(paths, lines[1:]):
        path, formula, secs = ln.rsplit(",", 2)
        assert path == p and float(secs) >= 0
        inst = simba_amd.MbaInstance.from_json_file(p)
        want = oracle_simba(inst, 2, 7)
        assert formula == (want["expr"] if want["status"] == 0 else "Not found !")
    vl = harness.run(paths, "vfb", max_len=6, timeout_s=30)
    assert len(vl) == 1 + len(paths)
    bl = harness.run(paths, "simba", max_len=8, timeout_s=30, batch=True)
    assert [x.split(",")[1] for x in bl[1:]] == [x.split(",")[1] for x in lines[1:]]
