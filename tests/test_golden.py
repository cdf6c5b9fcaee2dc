"""Golden fixtures (tests/golden/): instances in the reference schema and the
expected answers of the two reference CLIs' semantics."""
import json
import os
import subprocess

import pytest

from oracle import oracle as orc
from tests.gpu_util import note

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
EXP = json.load(open(os.path.join(GOLD, "expected.json")))
NAMES = sorted(EXP["cases"])


def _inst(name):
    import simba_amd
    return simba_amd.MbaInstance.from_json_file(os.path.join(GOLD, name))


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_golden(name):
    inst = _inst(name)
    e = EXP["cases"][name]
    s = orc.simba_search(inst.inputs, inst.outputs, inst.n_variables, (), inst.width, 2,
                         EXP["simba_maxlen"] - 1, 2)
    assert (s["expr"] if s["status"] == 0 else "Not found !") == e["simba"]["expr"]
    v = orc.vfb_search(inst.inputs, inst.outputs, inst.n_variables, (), inst.width,
                       EXP["vfb_maxlen"])
    assert v["expr"] == e["vfb"]["expr"]


@pytest.mark.parametrize("name", NAMES)
def test_reader_matches_reference_reader_on_fixture(name):
    ref = os.path.join(ROOT, "oracle", "_ref", "simba_instance_dump")
    if not os.path.exists(ref):
        pytest.skip("reference reader not built (needs the reference tree)")
    inst = _inst(name)
    if inst.width != 32:
        pytest.skip("the reference reader is 32-bit")
    out = subprocess.run([ref, os.path.join(GOLD, name)], capture_output=True, text=True,
                         check=True).stdout.split("\n")
    assert [int(x) for x in out[1].split()] == [int(x) for x in inst.inputs]
    assert [int(x) for x in out[2].split()] == [int(x) for x in inst.outputs]


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_cli_reproduces_golden(name):
    e = EXP["cases"][name]
    path = os.path.join(GOLD, name)
    note(f"cli golden {name}")
    r = subprocess.run([os.path.join(ROOT, "simba_amd", "bin", "simba_exe"), path,
                        str(EXP["simba_maxlen"])], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    p, formula, secs = r.stdout.strip().rsplit(",", 2)
    assert p == path and formula == e["simba"]["expr"] and float(secs) >= 0
    r = subprocess.run([os.path.join(ROOT, "simba_amd", "bin", "vfb_exe"), path,
                        str(EXP["vfb_maxlen"])], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    p, formula, secs = r.stdout.strip().rsplit(",", 2)
    assert formula == e["vfb"]["expr"]


@pytest.mark.gpu
def test_python_mirror_reproduces_golden():
    import simba_amd
    for name in NAMES:
        e = EXP["cases"][name]
        inst = _inst(name)
        note(f"mirror golden {name}")
        assert simba_amd.NoCache.solve(inst, EXP["simba_maxlen"]) == e["simba"]["expr"]
        assert simba_amd.MBA(EXP["vfb_maxlen"], inst.n_examples, inst.inputs, inst.outputs,
                             inst.n_variables, inst.width) == e["vfb"]["expr"]
