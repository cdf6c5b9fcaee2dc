"""GPU parity of the VFB bottom-up cached search against the CPU oracle:
same level, same enumeration index of the solution, same expression, and the
same number of kept rows per level (exhaustive mode)."""
import pytest

from oracle import oracle as orc
from tests.gpu_util import note

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import simba_amd
    c = simba_amd.Context(0)
    yield c
    c.close()


def _oracle(inst, max_len, exhaustive=False):
    return orc.vfb_search(inst.inputs, inst.outputs, inst.n_variables,
                          [int(c) for c in inst.consts], inst.width, max_len,
                          exhaustive=exhaustive)


CASES = [(seed, 64 if seed % 2 else 32, 1 + seed % 3, [(), (1,), (1, -1)][seed % 3],
          2 + seed % 6) for seed in range(24)]


@pytest.mark.parametrize("seed,W,nv,consts,tsize", CASES)
def test_vfb_matches_oracle(ctx, seed, W, nv, consts, tsize):
    from simba_amd.synth import make_instance
    consts = [c & ((1 << W) - 1) for c in consts]
    inst = make_instance(3000 + seed, nv, 8 + seed % 7, tsize, W, consts)
    note(f"vfb parity seed={seed} W={W} nv={nv} consts={consts} tsize={tsize}")
    want = _oracle(inst, 8)
    got = ctx.vfb(inst, max_size=8)
    assert got["status"] == want["status"]
    if want["status"] == 0:
        assert (got["size"], got["rank"], got["expr"]) == (want["size"], want["rank"], want["expr"])


@pytest.mark.parametrize("seed", range(4))
def test_vfb_level_counts_exhaustive(ctx, seed):
    from simba_amd.synth import make_instance
    W = 32 if seed % 2 == 0 else 64
    inst = make_instance(4000 + seed, 2 + seed % 2, 16, 9, W)
    note(f"vfb exhaustive counts seed={seed}")
    want = _oracle(inst, 7, exhaustive=True)
    got = ctx.vfb(inst, max_size=7, exhaustive=True)
    assert got["level_new"][1:7] == want["level_new"][1:7]
    assert got["level_cand"][1:8] == want["level_cand"][1:8]


def test_vfb_leaf_target(ctx):
    from simba_amd.synth import make_instance
    inst = make_instance(9, 3, 16, 1, 32)
    note("vfb leaf")
    got = ctx.vfb(inst, max_size=5)
    assert got["size"] == 1 and got["expr"] == "v%d" % inst.target_rpn[0]


def test_vfb_capacity_overflow_is_reported(ctx):
    import simba_amd
    from simba_amd.synth import make_instance
    inst = make_instance(10, 3, 16, 12, 64)
    note("vfb overflow")
    got = ctx.vfb(inst, max_size=9, cache_bytes=200 * (16 * 8 + 9))
    assert got["status"] == simba_amd.OVERFLOW


def test_vfb_largest_hash_table_keeps_counts():
    # a 96 GB row budget asks for the largest hash table (2^31 slots, 32 GB):
    # slot indices near the top of the u32 range must address their own
    # companion words (regression: 2 * slot in 32-bit arithmetic)
    import simba_amd
    from simba_amd.synth import make_instance
    inst = make_instance(4100, 2, 16, 9, 64)
    note("vfb largest table")
    want = _oracle(inst, 6, exhaustive=True)
    c = simba_amd.Context(0)
    try:
        got = c.vfb(inst, max_size=6, exhaustive=True, cache_bytes=96 << 30)
    finally:
        c.close()
    assert got["table_slots"] == 1 << 31
    assert got["level_new"][1:6] == want["level_new"][1:6]
    assert (got["size"], got["rank"]) == (want["size"], want["rank"])


def _splitmix64(x):
    """SplitMix64 finalizer on a uint64 array (wrapping arithmetic)."""
    import numpy as np
    with np.errstate(over="ignore"):
        x = x + np.uint64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return x ^ (x >> np.uint64(31))


def _ceiling_keys(count, distinct, salt):
    import numpy as np
    k = _splitmix64(_splitmix64(np.arange(count, dtype=np.uint64) + np.uint64(salt))
                    % np.uint64(distinct))
    return np.where(k >= np.uint64(2**64 - 2), k - np.uint64(2), k)


@pytest.mark.parametrize("slots_log2,count,distinct,prefill", [(20, 700001, 300007, 5000),
                                                               (16, 20000, 25000, 30000)])
def test_probe_ceiling_inserts_each_new_key_once(ctx, slots_log2, count, distinct, prefill):
    """The insert pass's table code on the measurement's keys adds exactly the
    keys that are neither duplicates nor prefilled (counted on the CPU from
    the same SplitMix64 draws)."""
    note(f"vfb probe ceiling slots 2^{slots_log2} count={count} distinct={distinct} "
         f"prefill={prefill}")
    r = ctx.vfb_probe_ceiling(slots_log2, count, distinct, prefill)
    keys = set(_ceiling_keys(count, distinct, 1).tolist())
    pre = set(_ceiling_keys(prefill, 2**62 - 1, 2**62).tolist())
    assert r["fresh"] == len(keys - pre)
    assert r["probe_seconds"] > 0 and r["insert_seconds"] > 0
