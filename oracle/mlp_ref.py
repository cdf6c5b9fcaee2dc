"""CPU reference of the block-proposal MLP (numpy).

TEST INFRASTRUCTURE ONLY (the product never imports it).  Restates the
network of simba_amd/csrc/mlp_kernels.hip / mlp_host.cpp:

  H = relu(X W1 + b1)       X [B x 256], W1 [256 x 512]
  Z = H W2 + b2             W2 [512 x 128]
  loss = mean over labelled rows of logsumexp(Z) - Z[label]

with every GEMM operand rounded to bfloat16 (round to nearest even) and
products accumulated in higher precision, and the gradients of that loss.
With bf16=False and float64 inputs it is the exact math, which the tests pin
with central finite differences.
"""
import numpy as np


def to_bf16(x):
    """Round float32 values to the nearest bfloat16 (ties to even), kept as float32."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    special = (u & 0x7F800000) == 0x7F800000
    r = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    r = np.where(special, u & 0xFFFF0000, r).astype(np.uint32)
    return r.view(np.float32).reshape(x.shape)


def _q(a, bf16):
    return to_bf16(a).astype(np.float64) if bf16 else np.asarray(a, dtype=np.float64)


def forward(x, w1, b1, w2, b2, bf16=True):
    h = np.maximum(_q(x, bf16) @ _q(w1, bf16) + b1, 0.0)
    if bf16:
        h = h.astype(np.float32)
    z = _q(h, bf16) @ _q(w2, bf16) + b2
    return h, (z.astype(np.float32) if bf16 else z)


def loss_and_grads(x, labels, w1, b1, w2, b2, bf16=True):
    h, z = forward(x, w1, b1, w2, b2, bf16)
    z64 = np.asarray(z, dtype=np.float64)
    labels = np.asarray(labels)
    mask = labels >= 0
    n = max(int(mask.sum()), 1)
    m = z64.max(axis=1, keepdims=True)
    lse = m[:, 0] + np.log(np.exp(z64 - m).sum(axis=1))
    p = np.exp(z64 - lse[:, None])
    onehot = np.zeros_like(p)
    onehot[np.arange(len(labels))[mask], labels[mask]] = 1.0
    dz = (p - onehot) * mask[:, None] / n
    loss = float(((lse - z64[np.arange(len(labels)), np.where(mask, labels, 0)]) * mask).sum() / n)
    if bf16:
        dz = dz.astype(np.float32)
    gw2 = _q(h, bf16).T @ _q(dz, bf16)
    gb2 = np.asarray(dz, dtype=np.float64).sum(axis=0)
    dh = (_q(dz, bf16) @ _q(w2, bf16).T) * (np.asarray(h) > 0)
    if bf16:
        dh = dh.astype(np.float32)
    gw1 = _q(x, bf16).T @ _q(dh, bf16)
    gb1 = np.asarray(dh, dtype=np.float64).sum(axis=0)
    return loss, (gw1, gb1, gw2, gb2)
