"""GPU parity of the SIMBA search against the CPU oracle (bit-exact: same
size, same rank in the reference's enumeration order, same expression)."""
import numpy as np
import pytest

from tests.gpu_util import note, oracle_simba, verify_rpn_chars

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import simba_amd
    c = simba_amd.Context(0)
    yield c
    c.close()


def _check(ctx, inst, min_size, max_size, **kw):
    want = oracle_simba(inst, min_size, max_size)
    got = ctx.simba(inst, min_size=min_size, max_size=max_size, **kw)
    assert got["status"] == want["status"], (got, want)
    if want["status"] == 0:
        assert (got["size"], got["rank"], got["expr"]) == (want["size"], want["rank"], want["expr"])
        assert verify_rpn_chars(got["rpn"], inst)
    return got


CASES = []
for seed in range(40):
    W = 64 if seed % 2 == 0 else 32
    nv = 1 + seed % 3
    consts = [(), (1,), (1, (1 << W) - 1)][seed % 3]
    CASES.append((seed, W, nv, consts, 1 + seed % 6))


@pytest.mark.parametrize("seed,W,nv,consts,tsize", CASES)
def test_simba_matches_oracle(ctx, seed, W, nv, consts, tsize):
    from simba_amd.synth import make_instance
    inst = make_instance(1000 + seed, nv, 8 + seed % 9, tsize, W, consts)
    note(f"simba parity seed={seed} W={W} nv={nv} consts={consts} tsize={tsize}")
    _check(ctx, inst, 1, 7)


@pytest.mark.parametrize("seed", range(8))
def test_simba_generated_sides_match_oracle(ctx, seed):
    # A tiny table budget forces Kt=1..3 so the per-lane table-assisted
    # evaluation of both pair sides is exercised.
    from simba_amd.synth import make_instance
    inst = make_instance(2000 + seed, 2, 12, 5 + seed % 3, 64 if seed % 2 else 32)
    note(f"simba gen-sides seed={seed}")
    for tb in (8, 200, 4000):
        _check(ctx, inst, 1, 8, table_bytes=tb)


@pytest.mark.parametrize("seed", range(4))
def test_simba_deep_generated_blocks_match_oracle(ctx, seed):
    # Targets at the largest size with tables up to Kt = 3..5: B sides of up to
    # Kt + 6 nodes are formed from table blocks under fixed-sibling and unary
    # path steps (and generic blocks past them).
    from simba_amd.synth import make_instance
    inst = make_instance(7000 + seed, 2, 16, 9, 64 if seed % 2 else 32, (1,))
    for tb, rb in ((2000, 4), (2000, 8), (60000, 8)):
        note(f"simba deep blocks seed={seed} table_bytes={tb} lanes_b={rb}")
        _check(ctx, inst, 1, 9, table_bytes=tb, lanes_b=rb)


def test_simba_compat_range_matches_reference_semantics(ctx):
    import simba_amd
    from simba_amd.synth import make_instance
    inst = make_instance(77, 2, 16, 1, 32)  # target is a bare variable
    note("simba compat")
    s = simba_amd.NoCache.solve(inst, 20)
    assert s == "(~(~v%d))" % inst.target_rpn[0]
    want = oracle_simba(inst, 2, 19)
    assert want["expr"] == s


def test_simba_not_found(ctx):
    from simba_amd.synth import make_instance
    inst = make_instance(5, 3, 16, 9, 64)
    note("simba not found")
    got = ctx.simba(inst, min_size=1, max_size=4)
    want = oracle_simba(inst, 1, 4)
    assert want["status"] == 1 and got["status"] == 1
    assert got["expr"] == "Not found !"
    assert got["candidates"] == want["candidates"]


@pytest.mark.parametrize("seed", range(6))
def test_simba_larger_sizes(ctx, seed):
    # sizes up to 9 with 3 leaves (the small single-instance configuration)
    from simba_amd.synth import make_instance
    inst = make_instance(300 + seed, 2, 16, 7, 64, (1,))
    note(f"simba size<=9 seed={seed}")
    _check(ctx, inst, 1, 9)


def test_simba_degenerate_first_sample(ctx):
    # sample 0 has all-zero inputs: the filter sample moves elsewhere and the
    # answer is unchanged.
    from simba_amd import MbaInstance
    from simba_amd.synth import make_instance
    base = make_instance(42, 2, 10, 6, 32)
    inp = base.inputs.copy()
    inp[0:2] = 0
    from simba_amd.synth import evaluate
    out = base.outputs.copy()
    out[0] = evaluate(base.target_rpn, [0, 0], 32)
    inst = MbaInstance(inp, out, 2, 32)
    note("simba degenerate sample 0")
    _check(ctx, inst, 1, 7)


@pytest.mark.parametrize("cfg", [dict(lanes_b=4), dict(wg_per_cu=4), dict(wg_per_cu=16),
                                 dict(table_delta=1), dict(lanes_b=8, table_delta=1)])
def test_simba_launch_configs_keep_answers(ctx, cfg):
    from simba_amd.synth import make_instance
    for seed in range(4):
        inst = make_instance(5000 + seed, 2 + seed % 2, 16, 6 + seed % 2, 64 if seed % 2 else 32,
                             (1,) if seed == 3 else ())
        note(f"simba cfg {cfg} seed={seed}")
        _check(ctx, inst, 1, 8, **cfg)


# (instance index in the held-out learned evaluation set, max size, hit
# buffer): launch groups whose window hits overflow the small buffer are
# replayed in parts; the answers stay those of the default buffer
REPLAY_CASES = [(19, 9, 256), (15, 11, 256), (31, 9, 1024), (34, 9, 4096), (46, 11, 64)]


@pytest.mark.parametrize("idx,max_size,cap", REPLAY_CASES)
def test_hit_buffer_overflow_replay_keeps_answers(ctx, idx, max_size, cap):
    from simba_amd.learn import sample_instances
    inst = sample_instances(10**6 + 15, 64, (9, 19), (2, 4), (32, 32))[idx]
    note(f"simba hit replay idx={idx} max={max_size} cap={cap}")
    base = ctx.simba(inst, max_size=max_size)
    got = ctx.simba(inst, max_size=max_size, hit_cap=cap)
    assert got["hit_replays"] >= 1
    key = lambda r: (r["status"], r["size"], r["rank"], r["expr"])
    assert key(got) == key(base)
    if base["status"] == 0:
        assert verify_rpn_chars(got["rpn"], inst)


def test_leaf_hits_above_the_buffer_are_an_error(ctx):
    # v1 = t and v0 = t - 1 on every sample: both leaves hit the size-1
    # windows.  A one-record buffer overflows in the fused step; the leaf
    # scan runs again on its own and its two hits overflow it: an error, not
    # an answer missing the leaf solution.  The default buffer finds v1.
    import simba_amd
    rng = np.random.default_rng(5)
    v0 = rng.integers(0, 2**63, 16, dtype=np.uint64)
    v1 = v0 + np.uint64(1)
    inst = simba_amd.MbaInstance(np.stack([v0, v1], 1), v1, 2, width=64)
    note("simba leaf hits above the buffer")
    base = ctx.simba(inst, max_size=7)
    assert (base["status"], base["size"], base["expr"]) == (0, 1, "v1")
    with pytest.raises(simba_amd.MbaError):
        ctx.simba(inst, max_size=7, hit_cap=1)


def test_autotune_picks_a_config_with_identical_answers(ctx, tmp_path, monkeypatch):
    from simba_amd import autotune
    from simba_amd.synth import make_instance
    monkeypatch.setenv("SIMBA_AMD_TUNE", str(tmp_path / "tune.json"))
    inst = make_instance(77, 3, 16, 9, 64)
    note("autotune")
    best, timing = autotune.tune_simba(ctx, inst, 8)
    assert set(best) == {"lanes_b", "wg_per_cu", "table_delta"}
    n = 1
    for v in autotune.SIMBA_SPACE.values():
        n *= len(v)
    assert len(timing) == n
    assert autotune.tuned(autotune.simba_key(inst, 8)) == best


@pytest.mark.parametrize("seed", range(10))
def test_simba_block_mode_matches_oracle(ctx, seed):
    # one (op, split) block per launch with the rank-order early exit
    from simba_amd.synth import make_instance
    inst = make_instance(6000 + seed, 1 + seed % 3, 12, 3 + seed % 5, 64 if seed % 2 else 32,
                         (1,) if seed % 4 == 0 else ())
    note(f"simba block mode seed={seed}")
    _check(ctx, inst, 1, 8, block_log2=1)
    _check(ctx, inst, 1, 8, block_log2=6)


@pytest.mark.parametrize("seed", range(10))
def test_learned_block_mode_is_minimal_and_verified(ctx, seed):
    import simba_amd
    from simba_amd.synth import make_instance
    m = simba_amd.Mlp(ctx, seed=seed)
    inst = make_instance(6100 + seed, 2 + seed % 3, 16, 4 + seed % 5, 64)
    note(f"learned block mode seed={seed}")
    plain = ctx.simba(inst, max_size=8)
    got = m.solve(inst, max_size=8, block_log2=1)
    assert got["status"] == plain["status"], (got, plain)
    if plain["status"] == 0:
        assert got["size"] == plain["size"]
        assert verify_rpn_chars(got["rpn"], inst)
    m.close()
