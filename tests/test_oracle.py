"""Pins the CPU oracle (oracle/) before it is trusted as the parity checker.

Each check is independent of the oracle's own arithmetic: a brute-force tree
generator written from the grammar (tests/treegen.py), Python big-integer
evaluation, closed-form known answers, and invariants of the hash mixers.
"""
import random

import numpy as np
import pytest

from oracle import oracle as orc
from tests.treegen import TOK_OP, py_eval, trees


@pytest.mark.parametrize("L", [1, 2, 3, 4])
def test_count_table_matches_brute_force(L):
    T = orc.count_table(L, 7)
    for s in range(1, 8 if L <= 2 else 7):
        ts = list(trees(s, L))
        assert int(T[s][8]) == len(ts)
        if s > 1:
            for op in range(8):
                assert int(T[s][op]) == sum(1 for t in ts if t[-1] == TOK_OP + op)
        else:
            assert all(int(T[1][op]) == 0 for op in range(8))


def test_count_table_overflow_guard():
    orc.count_table(5, 20)  # 3.38e18 < 2^64
    with pytest.raises(OverflowError):
        orc.count_table(5, 21)  # 3.48e19 > 2^64


@pytest.mark.parametrize("L", [2, 3])
def test_unrank_is_the_enumeration_order(L):
    T = orc.count_table(L, 6)
    for s in range(1, 7):
        for r, t in enumerate(trees(s, L)):
            got = orc.unrank(T, s, r)
            assert tuple(got) == t, (s, r)
            assert orc.rank(T, got) == r


def test_unrank_round_trip_large_sizes():
    rng = random.Random(7)
    for L in (3, 5):
        T = orc.count_table(L, 20)
        for _ in range(300):
            s = rng.randint(1, 20)
            r = rng.randrange(int(T[s][8]))
            assert orc.rank(T, orc.unrank(T, s, r)) == r


@pytest.mark.parametrize("W", [32, 64])
def test_eval_matches_python_integers(W):
    rng = random.Random(W)
    L = 4
    T = orc.count_table(L, 12)
    for _ in range(400):
        s = rng.randint(1, 12)
        rpn = orc.unrank(T, s, rng.randrange(int(T[s][8])))
        vals = [rng.getrandbits(64) for _ in range(L)]
        assert orc.evaluate(rpn, vals, W) == py_eval(rpn, vals, W)


def test_eval_textbook_identities():
    # (x ^ y) + 2*(x & y) == x + y ;  ~x == -x - 1  (mod 2^W)
    rng = random.Random(3)
    for W in (32, 64):
        m = (1 << W) - 1
        for _ in range(50):
            x, y = rng.getrandbits(W), rng.getrandbits(W)
            assert orc.lib().orc_apply(5, x ^ y, (2 * (x & y)) & m, W) == (x + y) & m
            assert orc.lib().orc_apply(0, x, 0, W) == (-x - 1) & m
            assert orc.lib().orc_apply(7, x, y, W) == (x * y) & m


def _brute_first(inputs, outputs, n_vars, consts, W, min_size, max_size):
    L = n_vars + len(consts)
    n = len(outputs)
    leafvals = [[inputs[i * n_vars + v] for v in range(n_vars)] + list(consts)
                for i in range(n)]
    for s in range(min_size, max_size + 1):
        for r, t in enumerate(trees(s, L)):
            if all(py_eval(t, leafvals[i], W) == outputs[i] & ((1 << W) - 1)
                   for i in range(n)):
                return s, r, t
    return None


@pytest.mark.parametrize("seed", range(12))
def test_simba_search_is_min_rank_brute_force(seed):
    rng = random.Random(seed)
    W = 32 if seed % 2 else 64
    n_vars = rng.randint(1, 3)
    consts = [1] if seed % 3 == 0 else []
    L = n_vars + len(consts)
    n = 6
    inputs = [rng.getrandbits(W) for _ in range(n * n_vars)]
    T = orc.count_table(L, 6)
    ts = rng.randint(1, 5)
    target = orc.unrank(T, ts, rng.randrange(int(T[ts][8])))
    leafvals = [[inputs[i * n_vars + v] for v in range(n_vars)] + consts for i in range(n)]
    outputs = [py_eval(target, leafvals[i], W) for i in range(n)]
    want = _brute_first(inputs, outputs, n_vars, consts, W, 1, 5)
    for threads in (1, 4):
        got = orc.simba_search(inputs, outputs, n_vars, consts, W, 1, 5, threads)
        assert got["status"] == 0
        assert (got["size"], got["rank"], tuple(got["rpn"])) == want


def test_simba_exhaustive_sweep_keeps_first_solution():
    rng = np.random.default_rng(9)
    x = rng.integers(0, 2**64, size=8, dtype=np.uint64)
    y = rng.integers(0, 2**64, size=8, dtype=np.uint64)
    inp = np.stack([x, y], axis=1).reshape(-1)
    first = orc.simba_search(inp, x + y, 2, (), 64, 1, 6, 4)
    full = orc.simba_search(inp, x + y, 2, (), 64, 1, 6, 4, exhaustive=True)
    T = orc.count_table(2, 6)
    assert (full["size"], full["rank"], full["expr"]) == (first["size"], first["rank"],
                                                          first["expr"]) == (3, first["rank"],
                                                                             "(v0 + v1)")
    assert full["candidates"] == sum(int(T[s][8]) for s in range(1, 7))
    assert first["candidates"] < full["candidates"]


def test_simba_known_answers():
    rng = np.random.default_rng(5)
    x = rng.integers(0, 2**32, size=16, dtype=np.uint64)
    y = rng.integers(0, 2**32, size=16, dtype=np.uint64)
    inp = np.stack([x, y], axis=1).reshape(-1)
    m = np.uint64(0xFFFFFFFF)
    # target v0 with the reference size range (compat: sizes 2..maxLen-1):
    r = orc.simba_search(inp, x, 2, (), 32, 2, 19)
    assert r["expr"] == "(~(~v0))" and r["size"] == 3 and r["rank"] == 0
    # native range finds the leaf itself
    r = orc.simba_search(inp, x, 2, (), 32, 1, 9)
    assert r["expr"] == "v0" and r["size"] == 1
    r = orc.simba_search(inp, (x + y) & m, 2, (), 32, 1, 9)
    assert r["expr"] == "(v0 + v1)"
    r = orc.simba_search(inp, (~y) & m, 2, (), 32, 1, 9)
    assert r["expr"] == "(~v1)"
    r = orc.simba_search(inp, (x * y) & m, 2, (), 32, 1, 9)
    assert r["expr"] == "(v0 * v1)"
    r = orc.simba_search(inp, (x - y) & m, 2, (), 32, 1, 9)
    assert r["expr"] == "(v0 - v1)"
    r = orc.simba_search(inp, (y - x) & m, 2, (), 32, 1, 9)
    assert r["expr"] == "(v1 - v0)"
    # a constant leaf (extension): x + 1 is (v0 + 1) in SIMBA order? -(~x) is
    # size 3 too and comes first (neg block precedes '+').
    r = orc.simba_search(x, (x + np.uint64(1)) & m, 1, (1,), 32, 1, 9)
    assert r["expr"] == "(-(~v0))"


def test_simba_infix_format():
    T = orc.count_table(2, 5)
    # rank 0 of size 3 is ~(~v0) (nocache.cu:162-195 prints unary as (~X))
    assert orc.infix(orc.unrank(T, 3, 0), 2, W=32) == "(~(~v0))"
    assert orc.to_reference_rpn(orc.unrank(T, 3, 0), 2) == "0~~"


def test_fmix_are_bijections_and_fix_zero():
    # MurmurHash3 finalizers: xorshift and odd multiplications are invertible.
    lib = orc.lib()
    assert lib.orc_fmix32(0) == 0 and lib.orc_fmix64(0) == 0
    rng = random.Random(11)
    xs32 = [rng.getrandbits(32) for _ in range(20000)]
    assert len({lib.orc_fmix32(x) for x in xs32}) == len(set(xs32))
    xs = [rng.getrandbits(64) for _ in range(20000)]
    assert len({lib.orc_fmix64(x) for x in xs}) == len(set(xs))


def _murmur3_x86_32(data, seed, fmix32):
    """MurmurHash3_x86_32 (Appleby's published algorithm) with the oracle's
    finalizer plugged in, so the published hash vectors check the finalizer."""
    rotl = lambda x, r: ((x << r) | (x >> (32 - r))) & 0xffffffff
    c1, c2, h, n = 0xcc9e2d51, 0x1b873593, seed, len(data)
    for i in range(0, n - n % 4, 4):
        k = int.from_bytes(data[i:i + 4], "little")
        k = rotl((k * c1) & 0xffffffff, 15) * c2 & 0xffffffff
        h = (rotl(h ^ k, 13) * 5 + 0xe6546b64) & 0xffffffff
    tail = data[n - n % 4:]
    if tail:
        k = int.from_bytes(tail, "little")
        h ^= rotl((k * c1) & 0xffffffff, 15) * c2 & 0xffffffff
    return fmix32(h ^ n)


def test_fmix32_matches_published_murmur3_vectors():
    # fmix32 is the finalizer of simpleHash (full-cache.cu:113-120); the
    # published MurmurHash3_x86_32 vectors (empty input: the hash is
    # fmix32(seed)) pin its shifts and constants
    f = orc.lib().orc_fmix32
    vectors = [(b"", 0, 0), (b"", 1, 0x514E28B7), (b"", 0xffffffff, 0x81F16F39),
               (b"\0\0\0\0", 0, 0x2362F9DE), (b"aaaa", 0x9747b28c, 0x5A97808A),
               (b"Hello, world!", 0x9747b28c, 0x24884CBA),
               (b"The quick brown fox jumps over the lazy dog", 0x9747b28c, 0x2FA826CD)]
    for data, seed, want in vectors:
        assert _murmur3_x86_32(data, seed, f) == want, (data, seed)


def _murmur3_x64_128(data, seed, fmix64):
    """MurmurHash3_x64_128 (Appleby's published algorithm) with the oracle's
    64-bit finalizer plugged in; returns (h1, h2)."""
    M = (1 << 64) - 1
    rotl = lambda x, r: ((x << r) | (x >> (64 - r))) & M
    c1, c2 = 0x87c37b91114253d5, 0x4cf5ad432745937f
    h1 = h2 = seed
    n = len(data)
    for i in range(0, n - n % 16, 16):
        k1 = int.from_bytes(data[i:i + 8], "little")
        k2 = int.from_bytes(data[i + 8:i + 16], "little")
        h1 ^= rotl(k1 * c1 & M, 31) * c2 & M
        h1 = (rotl(h1, 27) + h2) * 5 + 0x52dce729 & M
        h2 ^= rotl(k2 * c2 & M, 33) * c1 & M
        h2 = (rotl(h2, 31) + h1) * 5 + 0x38495ab5 & M
    tail = data[n - n % 16:]
    if len(tail) > 8:
        h2 ^= rotl(int.from_bytes(tail[8:], "little") * c2 & M, 33) * c1 & M
    if tail:
        h1 ^= rotl(int.from_bytes(tail[:8], "little") * c1 & M, 31) * c2 & M
    h1 ^= n
    h2 ^= n
    h1 = (h1 + h2) & M
    h2 = (h2 + h1) & M
    h1, h2 = fmix64(h1), fmix64(h2)
    h1 = (h1 + h2) & M
    h2 = (h2 + h1) & M
    return h1, h2


def test_fmix64_matches_published_murmur3_x64_128_vectors():
    # fmix64 is the warpcore MurmurHash finalizer (hashers.cuh:130-139) and
    # the W=64 fingerprint's hash; MurmurHash3_x64_128 ends with it, so the
    # published 128-bit vectors pin its shifts and constants: the mmh3
    # package's documented hash64("foo") / hash128("foo", 42), and Guava's
    # murmur3_128 digest of the quick-brown-fox sentence
    f = orc.lib().orc_fmix64
    assert _murmur3_x64_128(b"", 0, f) == (0, 0)
    assert _murmur3_x64_128(b"foo", 0, f) == (16316970633193145697, 9128664383759220103)
    h1, h2 = _murmur3_x64_128(b"foo", 42, f)
    assert (h2 << 64) | h1 == 215966891540331383248189432718888555506
    h1, h2 = _murmur3_x64_128(b"The quick brown fox jumps over the lazy dog", 0, f)
    assert (h1.to_bytes(8, "little") + h2.to_bytes(8, "little")).hex() == \
        "6c1b07bc7bbc4be347939ac4a93c437a"


def test_make_unq_chk_one_sample_wraps_mod_2_32():
    # makeUnqChk (full-cache.cu:126-134) with one sample: seed 0, so the
    # fingerprint is simpleHash(x) + 0x9e3779b9 computed in 32-bit unsigned
    # arithmetic (both operands are unsigned int), not rotated (index 0)
    f32 = orc.lib().orc_fmix32
    rng = random.Random(5)
    xs = [0, 1, 0xffffffff] + [rng.getrandbits(32) for _ in range(200)]
    wrapped = [x for x in xs if f32(x) + 0x9e3779b9 >= 1 << 32]
    assert len(wrapped) > 10  # the sum overflows for most inputs
    for x in xs:
        fp = orc.fingerprint([x], 32)
        assert fp == (f32(x) + 0x9e3779b9) & 0xffffffff
        assert fp < 1 << 32


def test_make_unq_chk_rotates_odd_indices_only():
    # Index 1 is rotated by 32 bits: with s1 = fp([x0]) < 2^32 the second
    # term v = simpleHash(x1) + 0x9e3779b9 (mod 2^32) + (s1 << 6) + (s1 >> 2)
    # is below 2^39, so s1 ^ fp([x0, x1]) = rotl64(v, 32) has a low half
    # below 2^7 and the low half of v on top.  Index 2 is not rotated.
    f32 = orc.lib().orc_fmix32
    M = (1 << 64) - 1
    rng = random.Random(6)
    for _ in range(300):
        x0, x1, x2 = (rng.getrandbits(32) for _ in range(3))
        s1 = orc.fingerprint([x0], 32)
        s2 = orc.fingerprint([x0, x1], 32)
        v = ((f32(x1) + 0x9e3779b9) & 0xffffffff) + (s1 << 6) + (s1 >> 2)
        assert v < 1 << 39
        d = s1 ^ s2
        assert d & 0xffffffff == v >> 32 and d >> 32 == v & 0xffffffff
        s3 = orc.fingerprint([x0, x1, x2], 32)
        w = (((f32(x2) + 0x9e3779b9) & 0xffffffff) + (s2 << 6) + (s2 >> 2)) & M
        assert s2 ^ s3 == w


def test_fingerprint_w64_one_sample():
    # the 64-bit extension: fmix64(x) + 0x9e3779b9 in 64-bit arithmetic
    f64 = orc.lib().orc_fmix64
    rng = random.Random(7)
    for x in [0, 1, (1 << 64) - 1] + [rng.getrandbits(64) for _ in range(100)]:
        assert orc.fingerprint([x], 64) == (f64(x) + 0x9e3779b9) & ((1 << 64) - 1)


def test_fingerprint_depends_on_order_and_values():
    a = [1, 2, 3, 4]
    assert orc.fingerprint(a, 32) == orc.fingerprint(list(a), 32)
    assert orc.fingerprint(a, 32) != orc.fingerprint([2, 1, 3, 4], 32)
    assert orc.fingerprint(a, 64) != orc.fingerprint(a, 32)


def _vfb_exact(inputs, outputs, n_vars, consts, W, max_len):
    """VFB with exact-signature dedup (no hashing): independent restatement
    of the intended keep-iff-new semantics, full-cache.cu:568-663."""
    m = (1 << W) - 1
    n = len(outputs)
    L = n_vars + len(consts)
    rows, levels = [], {1: []}
    seen = set()
    target = tuple(o & m for o in outputs)
    for v in range(L):
        sig = tuple((inputs[i * n_vars + v] if v < n_vars else consts[v - n_vars]) & m
                    for i in range(n))
        rows.append((sig, None))
        levels[1].append(len(rows) - 1)
        if sig == target:
            return 1, [L], f"v{v}"
        seen.add(sig)
    counts = [0, L]
    ops = [lambda a, b: ~a, lambda a, b: a & b, lambda a, b: a | b,
           lambda a, b: a ^ b, lambda a, b: -a, lambda a, b: a + b,
           lambda a, b: a - b, lambda a, b: a * b]
    for lev in range(2, max_len + 1):
        levels[lev] = []
        for op in range(8):
            pairs = []
            if op in (0, 4):
                pairs = [(l, None) for l in levels[lev - 1]]
            else:
                for i in range(1, (lev - 1) // 2 + 1):
                    for l in levels[i]:
                        for r in levels[lev - 1 - i]:
                            pairs.append((l, r))
                            if op == 6:
                                pairs.append((r, l))
            for l, r in pairs:
                A = rows[l][0]
                B = rows[r][0] if r is not None else A
                sig = tuple(ops[op](a, b) & m for a, b in zip(A, B))
                if sig in seen:
                    continue
                seen.add(sig)
                rows.append((sig, (op, l, r)))
                levels[lev].append(len(rows) - 1)
                if sig == target:
                    counts.append(len(levels[lev]))
                    return lev, counts, _vfb_str(rows, len(rows) - 1, n_vars)
        counts.append(len(levels[lev]))
    return None, counts, "Not found !"


def _vfb_str(rows, idx, n_vars):
    d = rows[idx][1]
    if d is None:
        return f"v{idx}"
    op, l, r = d
    ch = "~&|^-+-*"[op]
    if op in (0, 4):
        return f"{ch}({_vfb_str(rows, l, n_vars)})"
    return f"({_vfb_str(rows, l, n_vars)}){ch}({_vfb_str(rows, r, n_vars)})"


@pytest.mark.parametrize("seed", range(6))
def test_vfb_matches_exact_signature_dedup(seed):
    rng = random.Random(100 + seed)
    W = 32 if seed % 2 == 0 else 64
    n_vars = 2 + seed % 2
    n = 8
    inputs = [rng.getrandbits(W) for _ in range(n * n_vars)]
    T = orc.count_table(n_vars, 8)
    ts = rng.randint(3, 6)
    target = orc.unrank(T, ts, rng.randrange(int(T[ts][8])))
    lv = [[inputs[i * n_vars + v] for v in range(n_vars)] for i in range(n)]
    outputs = [py_eval(target, lv[i], W) for i in range(n)]
    lev, counts, expr = _vfb_exact(inputs, outputs, n_vars, [], W, 7)
    got = orc.vfb_search(inputs, outputs, n_vars, (), W, 7)
    if lev is None:
        assert got["status"] == 1
    else:
        assert got["status"] == 0 and got["size"] == lev
        assert got["expr"] == expr
    assert got["level_new"][1:len(counts)] == counts[1:]


def test_vfb_exhaustive_level_counts_match_exact_dedup():
    rng = random.Random(9)
    n_vars, n, W = 2, 8, 32
    inputs = [rng.getrandbits(W) for _ in range(n * n_vars)]
    outputs = [0xDEADBEEF ^ i for i in range(n)]  # unreachable target
    _, counts, _ = _vfb_exact(inputs, outputs, n_vars, [], W, 6)
    got = orc.vfb_search(inputs, outputs, n_vars, (), W, 6, exhaustive=True)
    assert got["level_new"][1:7] == counts[1:7]


def test_vfb_literal_mode_keeps_second_occurrences():
    # In the literal reading the first occurrence of a signature is dropped
    # and later duplicates are kept: v0+v1 is dropped at its first (v0,v1)
    # occurrence and found at the swapped pair (v1,v0).
    rng = np.random.default_rng(2)
    x = rng.integers(0, 2**32, size=8, dtype=np.uint64)
    y = rng.integers(0, 2**32, size=8, dtype=np.uint64)
    inp = np.stack([x, y], axis=1).reshape(-1)
    t = (x + y) & np.uint64(0xFFFFFFFF)
    a = orc.vfb_search(inp, t, 2, (), 32, 5, dedup="intended")
    b = orc.vfb_search(inp, t, 2, (), 32, 5, dedup="literal")
    assert a["expr"] == "(v0)+(v1)"
    assert b["expr"] == "(v1)+(v0)"


@pytest.mark.parametrize("seed", range(6))
def test_vfb_multithreaded_equals_single_threaded(seed):
    # the CPU baseline's level-synchronous VFB gives the checker's answers
    from simba_amd.synth import make_instance
    W = 64 if seed % 2 else 32
    m = (1 << W) - 1
    consts = [(), (1,), (1, m)][seed % 3]
    inst = make_instance(700 + seed, 2 + seed % 2, 8 + 8 * (seed % 3), 4 + seed, W, consts)
    a = orc.vfb_search(inst.inputs, inst.outputs, inst.n_variables, list(consts), W, 7,
                       exhaustive=True)
    b = orc.vfb_search_mt(inst.inputs, inst.outputs, inst.n_variables, list(consts), W, 7,
                          threads=4)
    assert a["level_new"] == b["level_new"] and a["level_cand"] == b["level_cand"]
    assert (a["status"], a["size"], a["rank"], a["expr"]) == \
        (b["status"], b["size"], b["rank"], b["expr"])
