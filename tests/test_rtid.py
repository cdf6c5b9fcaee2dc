"""SIMBA-RTid order (nocache_rtid.cu:85-95): oracle pinned by a plain Python
walk over the launched threads in index order, GPU parity against the
oracle (same size, rank and thread)."""
import numpy as np
import pytest

from oracle import oracle as orc

C = orc.RTID_C
M64 = (1 << 64) - 1


def _py_rtid(inputs, outputs, n_vars, W, max_size):
    """Walk (size, op) blocks in order and their threads g = 0, 1, ... of the
    ceil(N/1024)*1024 launched; return the first (size, rank, g) whose tree
    reproduces every sample."""
    L = n_vars
    n = len(outputs)
    T = orc.count_table(L, max_size)
    lv = [[int(inputs[i * n_vars + v]) for v in range(n_vars)] for i in range(n)]
    for s in range(2, max_size + 1):
        off = 0
        for op in range(8):
            N = int(T[s][op])
            if N == 0:
                continue
            G = (N + 1023) // 1024 * 1024
            for g in range(G):
                r = off + ((g * C) & M64) % N
                rpn = orc.unrank(T, s, r)
                if all(orc.evaluate(rpn, lv[i], W) == int(outputs[i]) for i in range(n)):
                    return s, r, g
            off += N
    return None


def _instance(seed, n_vars, W, target_size):
    rng = np.random.default_rng(seed)
    n = 6
    inp = rng.integers(0, 2**W - 1, size=n * n_vars, dtype=np.uint64, endpoint=True)
    T = orc.count_table(n_vars, target_size)
    rpn = orc.unrank(T, target_size, int(rng.integers(0, int(T[target_size][8]))))
    out = np.array([orc.evaluate(rpn, [int(inp[i * n_vars + v]) for v in range(n_vars)], W)
                    for i in range(n)], dtype=np.uint64)
    return inp, out


@pytest.mark.parametrize("seed", range(6))
def test_rtid_oracle_matches_thread_walk(seed):
    W = 32 if seed % 2 else 64
    inp, out = _instance(seed, 2, W, 2 + seed % 3)
    want = _py_rtid(inp, out, 2, W, 5)
    got = orc.simba_rtid_search(inp, out, 2, (), W, 1, 5, 4)
    if want is None:
        assert got["status"] == 1
    else:
        assert got["status"] == 0
        assert (got["size"], got["rank"], got["thread"]) == want


def test_rtid_visits_only_multiples_of_the_gcd():
    # C = 15809 * 142123: a block whose size shares a factor with C is only
    # partly visited (the ablation's coverage gap)
    N = 15809 * 3
    G = (N + 1023) // 1024 * 1024
    seen = {((g * C) & M64) % N for g in range(G)}
    assert seen == {0, 15809, 2 * 15809}


def test_rtid_size_never_below_plain():
    for seed in range(6, 12):
        inp, out = _instance(seed, 2, 64, 3 + seed % 3)
        plain = orc.simba_search(inp, out, 2, (), 64, 2, 6, 4)
        rt = orc.simba_rtid_search(inp, out, 2, (), 64, 2, 6, 4)
        assert plain["status"] == 0 and rt["status"] == 0
        # tiny blocks are fully visited (N < 15809), so the size agrees
        assert rt["size"] == plain["size"]


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(10))
def test_rtid_gpu_matches_oracle(seed):
    import simba_amd
    from simba_amd import synth
    from tests.gpu_util import note, verify_rpn_chars
    W = 64 if seed % 2 else 32
    inst = synth.make_instance(300 + seed, 2 + seed % 3, 16, 3 + seed % 5, W)
    note("rtid seed %d" % seed)
    want = orc.simba_rtid_search(inst.inputs, inst.outputs, inst.n_variables, (), W, 2, 8, 8)
    ctx = simba_amd.Context(0)
    try:
        got = ctx.simba_rtid(inst, min_size=2, max_size=8)
    finally:
        ctx.close()
    assert got["status"] == want["status"]
    if want["status"] == 0:
        assert (got["size"], got["rank"], got["expr"]) == (want["size"], want["rank"],
                                                          want["expr"])
        assert verify_rpn_chars(got["rpn"], inst)
