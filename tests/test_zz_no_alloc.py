"""mba_simba_solve_noalloc (every device buffer in the library's static
scratch array) gives mba_simba_solve's answers, call after call on the same
array, and rejects requests beyond the array's regions."""
import pytest

from tests.gpu_util import note, verify_rpn_chars

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import simba_amd
    c = simba_amd.Context(0)
    yield c
    c.close()


def _key(r):
    return (r["status"], r["size"], r["rank"], r["expr"], r["candidates"], r["size_reached"])


# (seed, vars, samples, target size, width, consts, max size, options):
# each pair of calls is a larger input followed by a smaller one
PAIRS = [
    ((11, 4, 64, 8, 64, (1,), 9, {}), (12, 2, 16, 6, 64, (1,), 8, {})),
    ((13, 3, 32, 9, 32, (), 10, {"block_log2": 16}), (14, 1, 8, 5, 32, (), 7, {})),
    # tables of size 9 for 5 leaves (407 MB) do not fit the scratch: the
    # search forms deeper B sides instead
    ((15, 4, 64, 11, 64, (1,), 11, {}), (16, 2, 16, 7, 64, (1,), 9, {})),
]


@pytest.mark.parametrize("first,second", PAIRS)
def test_noalloc_equals_plain_two_calls(ctx, first, second):
    from simba_amd.synth import make_instance
    for seed, nv, n, ts, W, consts, mx, kw in (first, second):
        inst = make_instance(seed, nv, n, ts, W, consts)
        note(f"noalloc seed={seed} vars={nv} samples={n} W={W} max={mx} {kw}")
        want = ctx.simba(inst, min_size=1, max_size=mx, **kw)
        got = ctx.simba_noalloc(inst, min_size=1, max_size=mx, **kw)
        assert _key(got) == _key(want)
        if got["status"] == 0:
            assert verify_rpn_chars(got["rpn"], inst)


def test_noalloc_exhaustive_sweep(ctx):
    from simba_amd.synth import make_instance
    inst = make_instance(17, 4, 64, 15, 64, (1,))
    note("noalloc exhaustive sweep windows of 20, sizes 1..10")
    kw = dict(min_size=1, max_size=20, stop_size=10, exhaustive=True)
    assert _key(ctx.simba_noalloc(inst, **kw)) == _key(ctx.simba(inst, **kw))


def test_noalloc_rejects_requests_beyond_scratch(ctx):
    import simba_amd
    from simba_amd.synth import make_instance
    inst = make_instance(18, 2, 16, 5, 64, (1,))
    note("noalloc limits")
    with pytest.raises(simba_amd.MbaError, match="hit_cap above"):
        ctx.simba_noalloc(inst, max_size=7, hit_cap=(1 << 17) + 1)
    with pytest.raises(simba_amd.MbaError, match="table_bytes above"):
        ctx.simba_noalloc(inst, max_size=7, table_bytes=1 << 30)
    # a small hit buffer within the limit still gives the plain answer
    want = ctx.simba(inst, max_size=7)
    assert _key(ctx.simba_noalloc(inst, max_size=7, hit_cap=4096)) == _key(want)
