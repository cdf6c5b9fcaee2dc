"""CPU-side checks of the product library: it loads, exports every symbol the
C header declares, reads instances like the reference's reader, and refuses
to run without a gfx950 device (no CPU fallback)."""
import os
import re
import subprocess

import numpy as np
import pytest

import simba_amd
from simba_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mba_api.h")


def _declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(mba_[a-z_]+)\s*\(", txt)))


def test_header_and_exports_agree():
    names = _declared()
    assert set(names) == set(simba_amd.EXPORTS)
    lib = simba_amd.lib()
    for n in names:
        assert hasattr(lib, n), n


def test_struct_layouts_match_header(tmp_path):
    """The ctypes mirror has the header's sizes and field offsets (gcc)."""
    import ctypes
    structs = {"mba_instance": simba_amd.Instance, "mba_opts": simba_amd.Opts,
               "mba_result": simba_amd.Result, "mba_vfb_stats": simba_amd.VfbStats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "mba_api.h"',
             'int main(void) {']
    for cname, py in structs.items():
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (cname, cname))
        for f in py._fields_:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));'
                         % (cname, f[0], cname, f[0]))
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)],
                   check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                 check=True).stdout.splitlines())
    for cname, py in structs.items():
        assert int(got[cname]) == ctypes.sizeof(py), cname
        for f in py._fields_:
            assert int(got[cname + "." + f[0]]) == getattr(py, f[0]).offset, (cname, f[0])


def test_context_refuses_without_gpu():
    import torch
    if torch.cuda.is_available() or os.path.exists("/dev/kfd"):
        pytest.skip("GPU present")
    with pytest.raises(simba_amd.MbaError):
        simba_amd.Context(0)


def test_json_round_trip(tmp_path):
    inst = synth.make_instance(3, 3, 16, 7, 32)
    p = tmp_path / "x.json"
    p.write_text(synth.to_json(inst))
    back = simba_amd.MbaInstance.from_json_file(str(p))
    assert back.width == 32 and back.n_variables == 3 and back.n_examples == 16
    assert np.array_equal(back.inputs, inst.inputs)
    assert np.array_equal(back.outputs, inst.outputs)


def test_json_reader_matches_reference_reader(tmp_path):
    ref = os.path.join(ROOT, "oracle", "_ref", "simba_instance_dump")
    if not os.path.exists(ref):
        pytest.skip("reference reader not built (needs the reference tree)")
    for seed in range(5):
        inst = synth.make_instance(seed, 1 + seed % 4, 5 + seed, 5, 32)
        p = tmp_path / f"i{seed}.json"
        p.write_text(synth.to_json(inst))
        out = subprocess.run([ref, str(p)], capture_output=True, text=True, check=True).stdout
        lines = out.split("\n")
        nv, ne = map(int, lines[0].split())
        ins = [int(x) for x in lines[1].split()]
        outs = [int(x) for x in lines[2].split()]
        mine = simba_amd.MbaInstance.from_json_file(str(p))
        assert (nv, ne) == (mine.n_variables, mine.n_examples)
        assert ins == [int(x) for x in mine.inputs]
        assert outs == [int(x) for x in mine.outputs]


def test_json_reader_rejects_malformed(tmp_path):
    p = tmp_path / "bad.json"
    p.write_text('{"num_variables": 2, "num_examples": 1, "examples": [')
    with pytest.raises(simba_amd.MbaError):
        simba_amd.MbaInstance.from_json_file(str(p))
    with pytest.raises(simba_amd.MbaError):
        simba_amd.MbaInstance.from_json_file(str(tmp_path / "missing.json"))


def test_rpn_to_infix_reference_format():
    import ctypes
    out = ctypes.create_string_buffer(256)
    simba_amd.lib().mba_rpn_to_infix(b"0~~", 2, None, 0, 32, out, 256)
    assert out.value == b"(~(~v0))"
    simba_amd.lib().mba_rpn_to_infix(b"01+1#*", 2, None, 0, 32, out, 256)
    assert out.value == b"((v0 + v1) * (-v1))"


@pytest.mark.parametrize("name", ["simba_exe", "simba_rtid_exe"])
def test_cli_usage_contract(name):
    exe = os.path.join(ROOT, "simba_amd", "bin", name)
    if not os.path.exists(exe):
        pytest.skip("CLI not built")
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout == "" and "<maxLen>" in r.stderr
    r = subprocess.run([exe, "x.json", "99"], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout == ""


def test_synth_targets_are_uniform_over_trees():
    # size-3 trees with 2 leaves: 44 trees; ranks drawn uniformly
    T = synth.count_table(2, 3)
    n = int(T[3][8])
    counts = np.zeros(n)
    rng = synth.SplitMix64(9)
    for _ in range(20000):
        counts[rng.below(n)] += 1
    chi2 = ((counts - counts.mean()) ** 2 / counts.mean()).sum()
    assert chi2 < 90  # 43 dof; p ~ 1e-5 threshold


def test_synth_instances_reproduce_their_target():
    inst = synth.make_instance(11, 3, 8, 9, 64, (1,))
    for i in range(inst.n_examples):
        lv = [int(inst.inputs[i * 3 + v]) for v in range(3)] + [1]
        assert synth.evaluate(inst.target_rpn, lv, 64) == int(inst.outputs[i])
