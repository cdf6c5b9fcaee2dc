"""The block-proposal MLP on the GPU: forward and gradients (BF16 MFMA, FP32
accumulation) against the CPU reference with the same bf16 operand
rounding; training lowers the loss; the learned search returns verified
solutions of the plain search's size."""
import numpy as np
import pytest

from oracle import mlp_ref
from tests.gpu_util import note, verify_rpn_chars

pytestmark = pytest.mark.gpu

# Tolerance: operands are rounded identically on both sides; only the FP32
# accumulation order differs (K <= 512 terms), so 1e-3 relative to the
# largest magnitude of each tensor bounds the difference.
RTOL = 1e-3


@pytest.fixture(scope="module")
def ctx():
    import simba_amd
    c = simba_amd.Context(0)
    yield c
    c.close()


def _close(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    scale = max(np.abs(b).max(), 1e-6)
    return np.abs(a - b).max() / scale


def test_forward_matches_reference(ctx):
    import simba_amd
    note("mlp forward")
    m = simba_amd.Mlp(ctx, seed=3)
    rng = np.random.default_rng(1)
    for batch in (1, 16, 37, 256):
        x = rng.random((batch, 256), dtype=np.float32)
        z = m.forward(x)
        _, zr = mlp_ref.forward(x, *m.get_weights())
        assert _close(z, zr) < RTOL, batch
    m.close()


def test_gradients_match_reference(ctx):
    import simba_amd
    note("mlp grads")
    m = simba_amd.Mlp(ctx, seed=4)
    rng = np.random.default_rng(2)
    x = rng.random((48, 256), dtype=np.float32)
    y = rng.integers(0, 128, size=48).astype(np.int32)
    y[5] = -1  # unlabelled row
    loss, g = m.grads(x, y)
    lr, gr = mlp_ref.loss_and_grads(x, y, *m.get_weights())
    assert abs(loss - lr) < 1e-4 * max(1, abs(lr))
    for a, b in zip(g, gr):
        assert _close(a, b) < RTOL
    m.close()


def test_training_reduces_loss(ctx):
    import simba_amd
    note("mlp train")
    m = simba_amd.Mlp(ctx, seed=5)
    rng = np.random.default_rng(3)
    x = rng.random((64, 256), dtype=np.float32)
    y = (np.arange(64) % 10).astype(np.int32)
    l0 = m.train_step(x, y, 1e-3)
    for _ in range(60):
        l1 = m.train_step(x, y, 1e-3)
    assert l1 < 0.5 * l0
    m.close()


def test_learned_solve_is_minimal_size_and_verified(ctx):
    import simba_amd
    from simba_amd.synth import make_instance
    m = simba_amd.Mlp(ctx, seed=6)
    for seed in range(12):
        inst = make_instance(7000 + seed, 2 + seed % 3, 32, 4 + seed % 4, 64)
        note(f"learned solve seed={seed}")
        plain = ctx.simba(inst, max_size=8)
        got = m.solve(inst, max_size=8)
        assert got["status"] == plain["status"]
        if plain["status"] == 0:
            assert got["size"] == plain["size"]
            assert verify_rpn_chars(got["rpn"], inst)
    insts = [make_instance(7100 + s, 3, 32, 6, 64) for s in range(16)]
    note("learned batch solve")
    got = m.batch_solve(insts, max_size=8)
    for inst, g in zip(insts, got):
        p = ctx.simba(inst, max_size=8)
        assert g["status"] == p["status"]
        if p["status"] == 0:
            assert g["size"] == p["size"] and verify_rpn_chars(g["rpn"], inst)
    m.close()


def test_trained_network_drives_learned_order(ctx):
    """A network trained on plain-solved labels orders the blocks of every
    size of at least 2^12 trees (block_log2 lowers the threshold from 2^26),
    per instance and batched: each solution has the plain search's size and
    holds on every sample, and the order was engaged (ordered_sizes)."""
    import simba_amd
    from simba_amd import learn
    from simba_amd.synth import make_instance
    note("learned order engaged: labelling and training")
    X, Y, _ = learn.build_dataset(ctx, 8000, seed=31, max_size=8)
    m = simba_amd.Mlp(ctx, seed=3)
    learn.train(m, X, Y, epochs=2)
    insts = [make_instance(7300 + s, 2 + s % 3, 32, 6 + s % 4, 64) for s in range(16)]
    kw = dict(max_size=9, block_log2=12)
    plain = [ctx.simba(x, max_size=9) for x in insts]
    ordered = 0
    for seed, (inst, p) in enumerate(zip(insts, plain)):
        note(f"learned order engaged: instance {seed}")
        got = m.solve(inst, **kw)
        assert got["status"] == p["status"]
        ordered += got["ordered_sizes"]
        if p["status"] == 0:
            assert got["size"] == p["size"] and verify_rpn_chars(got["rpn"], inst)
    assert ordered > 0
    note("learned order engaged: batched")
    got = m.batch_solve(insts, **kw)
    assert sum(g["ordered_sizes"] for g in got) > 0
    for inst, g, p in zip(insts, got, plain):
        assert g["status"] == p["status"]
        if p["status"] == 0:
            assert g["size"] == p["size"] and verify_rpn_chars(g["rpn"], inst)
    m.close()


def test_trained_network_proposes_solution_blocks_earlier(ctx):
    """What training gains: on held-out instances the trained network puts
    the block that holds the plain search's solution earlier in its order
    (fewer blocks visited before it) than the same network untrained."""
    import simba_amd
    from simba_amd import learn
    note("learned gain: labelling")
    X, Y, _ = learn.build_dataset(ctx, 24000, seed=11, max_size=8)
    Xe, Ye, _ = learn.build_dataset(ctx, 4000, seed=10**6 + 11, max_size=8)
    trained = simba_amd.Mlp(ctx, seed=1)
    untrained = simba_amd.Mlp(ctx, seed=1)
    note("learned gain: training")
    learn.train(trained, X, Y, epochs=4)
    pos_t = learn.mean_label_position(trained, Xe, Ye)
    pos_u = learn.mean_label_position(untrained, Xe, Ye)
    acc_t = learn.top1_accuracy(trained, Xe, Ye)
    acc_u = learn.top1_accuracy(untrained, Xe, Ye)
    note(f"held-out: position {pos_t:.2f} vs {pos_u:.2f} untrained, "
         f"top-1 {acc_t:.3f} vs {acc_u:.3f}")
    assert pos_t < 0.6 * pos_u
    assert acc_t > acc_u + 0.1
    trained.close()
    untrained.close()


def test_device_features_equal_host_features(ctx):
    """mba_features_batch (one workgroup per instance on the device) gives the
    host mba_features bit for bit: widths 32/64, 1..6 variables (the first
    five are hashed), 1..128 samples."""
    import simba_amd
    from simba_amd import synth
    note("device features")
    insts = [synth.make_instance(700 + i, 1 + i % 6, (1, 5, 32, 64, 100, 128)[i % 6], 3 + i % 5,
                                 64 if i % 3 else 32) for i in range(24)]
    dev = simba_amd.features_batch(ctx, insts)
    host = np.stack([simba_amd.features(x) for x in insts])
    assert dev.shape == host.shape
    assert np.array_equal(dev, host)
