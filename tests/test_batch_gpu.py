"""Batched search: every instance's answer equals the per-instance plain
search's (and the CPU oracle's) on the same size budget."""
import random

import pytest

from tests.gpu_util import note, oracle_simba

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import simba_amd
    c = simba_amd.Context(0)
    yield c
    c.close()


def _instances(seed, count, W, max_target):
    from simba_amd.synth import make_instance
    rng = random.Random(seed)
    out = []
    for i in range(count):
        nv = rng.randint(2, 5)
        n = rng.randint(16, 64)
        ts = rng.randint(3, max_target)
        out.append(make_instance(seed * 100000 + i, nv, n, ts, W))
    return out


@pytest.mark.parametrize("W", [64, 32])
def test_batch_matches_plain_and_oracle(ctx, W):
    insts = _instances(7 + W, 48, W, 7)
    note(f"batch vs plain W={W}")
    got = ctx.batch(insts, min_size=1, max_size=6)
    for inst, g in zip(insts, got):
        p = ctx.simba(inst, min_size=1, max_size=6)
        assert (g["status"], g["size"], g["rank"], g["expr"]) == \
            (p["status"], p["size"], p["rank"], p["expr"])
    # every instance against the oracle: at max size 6 the leaves' B side
    # (size 4) is above the tables, in both paths formed per B block
    for inst, g in zip(insts, got):
        w = oracle_simba(inst, 1, 6)
        assert g["status"] == w["status"]
        if w["status"] == 0:
            assert (g["size"], g["rank"], g["expr"]) == (w["size"], w["rank"], w["expr"])


@pytest.mark.parametrize("table_kib", [256, 1024])
def test_batch_cohorts_settle_only_swept_sizes(ctx, table_kib):
    # a table pool far below the instances' tables: from the size where they
    # no longer fit, the instances are swept in cohorts, the first of a
    # cohort with whatever of its tables fits; a waiting instance may already
    # hold a window hit of a larger size (a unary chain over a smaller tree)
    # and must not settle there before its own sizes are swept
    from simba_amd.learn import sample_instances
    insts = sample_instances(77, 256, (5, 13), (2, 5), (16, 64))
    note(f"batch cohorts, max size 10, {table_kib} KiB of tables")
    got = ctx.batch(insts, max_size=10, table_bytes=table_kib << 10)
    for inst, g in zip(insts, got):
        p = ctx.simba(inst, max_size=10)
        assert (g["status"], g["size"], g["rank"], g["expr"]) == \
            (p["status"], p["size"], p["rank"], p["expr"])


def test_batch_small_table_budget_and_compat(ctx):
    # tiny table budget: most B sides are formed per block from the few
    # tables (or evaluated per value)
    insts = _instances(99, 24, 64, 6)
    note("batch small tables + compat")
    got = ctx.batch(insts, max_size=8, compat=True, table_bytes=24 * 64)
    for inst, g in zip(insts, got):
        p = ctx.simba(inst, max_size=8, compat=True)
        assert (g["status"], g["size"], g["rank"], g["expr"]) == \
            (p["status"], p["size"], p["rank"], p["expr"])


def test_batch_single_and_repeated(ctx):
    insts = _instances(5, 1, 64, 5)
    note("batch single")
    a = ctx.batch(insts, max_size=7)
    b = ctx.batch(insts, max_size=7)
    key = lambda r: (r["status"], r["size"], r["rank"], r["expr"], r["candidates"])
    assert [key(x) for x in a] == [key(x) for x in b]


def test_batch_b_sides_above_the_tables(ctx):
    # two leaves, sizes up to 9: the B sides above the tables (size 7) are
    # formed per (op, split) block, many of them smaller than a tile; the
    # answers stay the per-instance search's
    from simba_amd.synth import make_instance
    insts = [make_instance(4200 + i, 2, 24, 7 + i % 3, 64) for i in range(24)]
    note("batch B blocks")
    got = ctx.batch(insts, max_size=9)
    for inst, g in zip(insts, got):
        p = ctx.simba(inst, max_size=9)
        assert (g["status"], g["size"], g["rank"], g["expr"]) == \
            (p["status"], p["size"], p["rank"], p["expr"])


@pytest.mark.parametrize("table_mb", [0, 48])
def test_batch_tables_packed_again_keep_answers(ctx, table_mb):
    # sizes to 11 on instances that settle at many sizes: as the active
    # instances halve the table pool is packed again and the tables of the
    # remaining ones grow (a small pool: most B sides stay above the tables
    # for a while); every answer stays the per-instance search's
    from simba_amd.learn import sample_instances
    insts = sample_instances(10**6 + 15, 160, (9, 19), (2, 4), (32, 32))
    note(f"batch tables packed again table_mb={table_mb}")
    got = ctx.batch(insts, max_size=11, table_bytes=table_mb << 20)
    key = lambda r: (r["status"], r["size"], r["rank"], r["expr"])
    for inst, g in zip(insts, got):
        assert key(g) == key(ctx.simba(inst, max_size=11))


# (max size, hit buffer): a step whose window hits overflow the small buffer
# is run again with its jobs (and then one job's tiles) in parts; the answers
# stay those of the default buffer
BATCH_REPLAY_CASES = [(9, 4096), (11, 16384)]


@pytest.mark.parametrize("max_size,cap", BATCH_REPLAY_CASES)
def test_batch_hit_buffer_overflow_split_keeps_answers(ctx, max_size, cap):
    from simba_amd.learn import sample_instances
    insts = sample_instances(10**6 + 15, 64, (9, 19), (2, 4), (32, 32))
    note(f"batch hit split max={max_size} cap={cap}")
    base = ctx.batch(insts, max_size=max_size)
    got = ctx.batch(insts, max_size=max_size, hit_cap=cap)
    assert got[0]["hit_replays"] >= 1 and base[0]["hit_replays"] == 0
    key = lambda r: (r["status"], r["size"], r["rank"], r["expr"])
    assert [key(r) for r in got] == [key(r) for r in base]


def test_batch_plain_search_is_one_launch_sequence(ctx):
    # without a budget or cancel flag every size step of a plain batch runs in
    # one captured launch sequence (one sync): sizes settle on the device
    insts = _instances(31, 64, 64, 7)
    note("batch one launch sequence")
    got = ctx.batch(insts, max_size=8)
    assert got[0]["kernel_launches"] == 1
    for inst, g in list(zip(insts, got))[:24]:
        p = ctx.simba(inst, max_size=8)
        assert (g["status"], g["size"], g["rank"]) == (p["status"], p["size"], p["rank"])


def test_batch_budget_is_per_instance(ctx):
    # the budget bounds each instance's share of the device time, not the
    # batch's: a batch longer than the budget still gives every instance its
    # unbudgeted answer when each share stays below it
    import simba_amd
    from simba_amd.learn import sample_instances
    insts = sample_instances(77, 2048, (5, 11), (2, 5), (16, 64))
    note("batch per-instance budget")
    base = ctx.batch(insts, max_size=9)
    budget = 0.002
    got = ctx.batch(insts, max_size=9, budget_s=budget)
    assert max(r["seconds"] for r in got) > budget  # the batch outlasts the budget
    key = lambda r: (r["status"], r["size"], r["rank"], r["expr"])
    assert [key(r) for r in got] == [key(r) for r in base]
    # a budget every instance exceeds after its first launch sequence (sizes
    # up to 11 take several): answers found before it are the unbudgeted ones
    base11 = ctx.batch(insts[:512], max_size=11)
    tiny = ctx.batch(insts[:512], max_size=11, budget_s=1e-9)
    assert any(r["status"] == simba_amd.BUDGET for r in tiny)
    for t, b in zip(tiny, base11):
        assert t["status"] == simba_amd.BUDGET or key(t) == key(b)
