"""Training and evaluation pipeline of the learned block proposal.

A training row is an instance's features with the one-hot of a size, and
the (root operator, left child size) block, as the class
op * 16 + min(left size, 15), of a solution of that size produced by the
enumerator: the plain (batched) search's first solution when it finds one
within the labelling size budget, else the instance's generating tree -- a
uniform rank unranked by the enumerator's own order -- at its own size (a
solution of that size by construction; the learned order only matters at
sizes the plain search has not settled).  The network is trained on the GPU
with the library's own forward/backward kernels (Adam).  Weights are stored
as a plain numpy .npz archive (no pickled objects).
"""
import time

import numpy as np

from . import Mlp, block_class, features, features_batch, set_size_slot
from .synth import TOK_OP, SplitMix64, make_instance


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


def tree_block(rpn):
    """(root op, left child size) of an RPN token list (left size 0 for a
    unary root)."""
    n = len(rpn)
    op = rpn[-1] - TOK_OP
    if op in (0, 4):
        return op, 0
    need, i = 1, n - 2  # walk the right subtree back from the root
    while need > 0:
        t = rpn[i]
        need += -1 if t < TOK_OP else (0 if t - TOK_OP in (0, 4) else 1)
        i -= 1
    return op, i + 1


def build_dataset(ctx, count, seed, sizes=(3, 13), n_vars=(2, 4), n_samples=(32, 32),
                  max_size=9, chunk=4096, log=None):
    """Features (size one-hot set) and labels of `count` instances: the plain
    search's solution when it finds one within `max_size`, else the
    generating tree at its own size; returns X, Y and the label sizes."""
    X, Y, Z = [], [], []
    done, from_search = 0, 0
    t0 = time.time()
    while done < count:
        k = min(chunk, count - done)
        insts = sample_instances(seed + done, k, sizes, n_vars, n_samples)
        res = ctx.batch(insts, max_size=max_size)
        feats = features_batch(ctx, insts)
        for i, (inst, r) in enumerate(zip(insts, res)):
            if r["status"] == 0:
                if r["size"] < 2:
                    continue
                y, z = block_class(r["op"], r["split"]), r["size"]
                from_search += 1
            else:
                op, split = tree_block(inst.target_rpn)
                y, z = block_class(op, split), len(inst.target_rpn)
            X.append(feats[i])
            Y.append(y)
            Z.append(z)
        done += k
        if log:
            log(f"labelled {done}/{count} ({from_search} by the search) {time.time() - t0:.1f}s")
    X = set_size_slot(np.stack(X).astype(np.float32), np.asarray(Z))
    return X, np.asarray(Y, dtype=np.int32), np.asarray(Z, dtype=np.int32)


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
