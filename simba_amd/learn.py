"""Training and evaluation pipeline of the learned block proposal.

Labels come from the plain enumerator itself: seeded synthetic instances are
solved by the (batched) plain search, and the label of an instance is the
(root operator, left child size) block of its solution, as the class
op * 16 + min(left size, 15).  The network is trained on the GPU with the
library's own forward/backward kernels (Adam).  Weights are stored as a
plain numpy .npz archive (no pickled objects).
"""
import time

import numpy as np

from . import Mlp, block_class, features, features_batch
from .synth import SplitMix64, make_instance


def sample_instances(seed, count, sizes, n_vars, n_samples, width=64):
    """`count` seeded instances with target size, variable count and sample
    count drawn uniformly from the inclusive ranges given."""
    rng = SplitMix64(seed)
    out = []
    for i in range(count):
        ts = sizes[0] + rng.below(sizes[1] - sizes[0] + 1)
        nv = n_vars[0] + rng.below(n_vars[1] - n_vars[0] + 1)
        ns = n_samples[0] + rng.below(n_samples[1] - n_samples[0] + 1)
        out.append(make_instance(rng.next(), nv, ns, ts, width))
    return out


def label_results(results):
    """Class label of each solved result (-1 when unsolved)."""
    labels = np.full(len(results), -1, dtype=np.int32)
    for i, r in enumerate(results):
        if r["status"] == 0 and r["size"] > 1:
            labels[i] = block_class(r["op"], r["split"])
    return labels


def build_dataset(ctx, count, seed, sizes=(3, 13), n_vars=(2, 4), n_samples=(32, 32),
                  max_size=9, chunk=4096, log=None):
    """Features and plain-search labels of `count` instances (unsolved ones
    within the size budget are dropped)."""
    X, Y = [], []
    done = 0
    t0 = time.time()
    while done < count:
        k = min(chunk, count - done)
        insts = sample_instances(seed + done, k, sizes, n_vars, n_samples)
        res = ctx.batch(insts, max_size=max_size)
        lab = label_results(res)
        keep = [i for i, y in enumerate(lab) if y >= 0]
        if keep:
            X.extend(features_batch(ctx, [insts[i] for i in keep]))
            Y.extend(lab[i] for i in keep)
        done += k
        if log:
            log(f"labelled {done}/{count} ({len(Y)} solved) {time.time() - t0:.1f}s")
    return np.stack(X).astype(np.float32), np.asarray(Y, dtype=np.int32)


def train(mlp, X, Y, epochs=4, batch=256, lr=1e-3, seed=0, log=None):
    rng = np.random.default_rng(seed)
    n = X.shape[0]
    losses = []
    for ep in range(epochs):
        perm = rng.permutation(n)
        tot, nb = 0.0, 0
        for b in range(0, n, batch):
            idx = perm[b:b + batch]
            tot += mlp.train_step(X[idx], Y[idx], lr)
            nb += 1
        losses.append(tot / max(nb, 1))
        if log:
            log(f"epoch {ep}: loss {losses[-1]:.4f}")
    return losses


def top1_accuracy(mlp, X, Y):
    z = mlp.forward(X)
    return float((z.argmax(axis=1) == Y).mean())


def mean_label_position(mlp, X, Y):
    """Mean position (0 = first) of each instance's solution block in the
    order the network proposes: how many blocks the learned search visits
    before the one that holds the solution."""
    z = mlp.forward(X)
    true = z[np.arange(len(Y)), Y][:, None]
    return float((z > true).sum(axis=1).mean())


def save(mlp, path):
    w1, b1, w2, b2 = mlp.get_weights()
    np.savez(path, w1=w1, b1=b1, w2=w2, b2=b2)


def load(mlp, path):
    with np.load(path, allow_pickle=False) as d:
        mlp.set_weights(d["w1"], d["b1"], d["w2"], d["b2"])
