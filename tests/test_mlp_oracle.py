"""Pins the MLP CPU reference: its analytic gradients (exact math, float64)
agree with central finite differences of its own loss, and its bf16 rounding
matches the IEEE definition on hand-checked values."""
import numpy as np

from oracle import mlp_ref


def _small_net(rng, din=6, dh=5, dout=4):
    return (rng.normal(size=(din, dh)), rng.normal(size=dh) * 0.1,
            rng.normal(size=(dh, dout)), rng.normal(size=dout) * 0.1)


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(7, 6))
    labels = np.array([0, 3, 1, -1, 2, 2, 0])
    params = list(_small_net(rng))
    loss, grads = mlp_ref.loss_and_grads(x, labels, *params, bf16=False)
    eps = 1e-6
    for pi, (p, g) in enumerate(zip(params, grads)):
        flat = p.reshape(-1)
        for idx in range(0, flat.size, max(1, flat.size // 7)):
            old = flat[idx]
            flat[idx] = old + eps
            lp, _ = mlp_ref.loss_and_grads(x, labels, *params, bf16=False)
            flat[idx] = old - eps
            lm, _ = mlp_ref.loss_and_grads(x, labels, *params, bf16=False)
            flat[idx] = old
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - g.reshape(-1)[idx]) < 1e-6 * max(1.0, abs(fd)), (pi, idx)


def test_loss_of_uniform_logits_is_log_classes():
    x = np.zeros((3, 6))
    params = (np.zeros((6, 5)), np.zeros(5), np.zeros((5, 4)), np.zeros(4))
    loss, _ = mlp_ref.loss_and_grads(x, np.array([0, 1, 2]), *params, bf16=False)
    assert abs(loss - np.log(4)) < 1e-12


def test_bf16_rounding():
    vals = np.array([1.0, 1.00390625, 1.005859375, 1.0078125, -2.5, 3.0e38, 1e-40],
                    dtype=np.float32)
    r = mlp_ref.to_bf16(vals)
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to even -> 1
    assert r[1] == 1.0
    # 1 + 3*2^-9 rounds up to 1 + 2^-7
    assert r[2] == np.float32(1.0078125)
    assert r[3] == np.float32(1.0078125)
    assert r[4] == np.float32(-2.5)
    assert np.isfinite(r[5])
    assert (r.view(np.uint32) & 0xFFFF).max() == 0
