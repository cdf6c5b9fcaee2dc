"""Seeded synthetic instances for the SIMBA/VFB workloads.

Targets are drawn uniformly over all expression trees of a given size: a
rank is drawn uniformly in [0, T[size]) and decoded in the reference's
enumeration order (nocache.cu:23-83).  Sample values come from SplitMix64
(Steele, Lea, Flood, "Fast splittable pseudorandom number generators",
OOPSLA 2014), uniform over [0, 2^W).  The reference's own generator
(scripts/generate.py) draws operators and splits uniformly instead, which is
not uniform over trees; both produce valid instances of the same schema.
"""
import json

import numpy as np

from . import MbaInstance

NOT, AND, OR, XOR, NEG, ADD, SUB, MUL = range(8)
TOK_OP = 32
M64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.state = seed & M64

    def next(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & M64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
        return z ^ (z >> 31)

    def below(self, n):
        """Uniform integer in [0, n) by rejection (n <= 2^64)."""
        if n <= 0:
            raise ValueError("empty range")
        lim = ((1 << 64) // n) * n
        while True:
            x = self.next()
            if x < lim:
                return x % n


def count_table(L, max_size):
    """T[s][0..8] (nocache.cu:117-160); Python integers, no overflow."""
    T = [[0] * 9 for _ in range(max_size + 1)]
    T[1][8] = L
    for s in range(2, max_size + 1):
        U = T[s - 1][8]
        C = sum(T[j][8] * T[s - 1 - j][8] for j in range(1, (s - 1) // 2 + 1))
        NC = sum(T[j][8] * T[s - 1 - j][8] for j in range(1, s - 1))
        T[s] = [U, C, C, C, U, C, NC, C, 2 * U + 5 * C + NC]
    return T


def unrank(T, size, rank):
    """RPN tokens of tree `rank` of `size` (leaf < 32, op = 32 + o)."""
    out = [0] * size

    def rec(s, n, pos):
        if s == 1:
            out[pos] = n
            return
        op = 0
        while n >= T[s][op]:
            n -= T[s][op]
            op += 1
        cs = s - 1
        out[pos + cs] = TOK_OP + op
        if op in (NOT, NEG):
            rec(cs, n, pos)
            return
        j = 1
        while n >= T[j][8] * T[cs - j][8]:
            n -= T[j][8] * T[cs - j][8]
            j += 1
        rs = T[cs - j][8]
        rec(j, n // rs, pos)
        rec(cs - j, n % rs, pos + j)

    rec(size, rank, 0)
    return out


def evaluate(rpn, leafvals, W):
    m = (1 << W) - 1
    st = []
    for t in rpn:
        if t < TOK_OP:
            st.append(int(leafvals[t]) & m)
            continue
        op = t - TOK_OP
        if op == NOT:
            st.append(~st.pop() & m)
        elif op == NEG:
            st.append(-st.pop() & m)
        else:
            r, l = st.pop(), st.pop()
            st.append((l & r if op == AND else l | r if op == OR else l ^ r if op == XOR
                       else l + r if op == ADD else l - r if op == SUB else l * r) & m)
    return st[0]


def infix(rpn, n_vars, consts=(), W=64):
    ch = "~&|^-+-*"
    st = []
    m = (1 << W) - 1
    for t in rpn:
        if t < TOK_OP:
            if t < n_vars:
                st.append(f"v{t}")
            else:
                c = int(consts[t - n_vars]) & m
                st.append(f"-{(-c) & m}" if c >> (W - 1) else str(c))
            continue
        op = t - TOK_OP
        if op in (NOT, NEG):
            st.append(f"({ch[op]}{st.pop()})")
        else:
            r, l = st.pop(), st.pop()
            st.append(f"({l} {ch[op]} {r})")
    return st[0]


def make_instance(seed, n_vars, n_samples, target_size, W=64, consts=()):
    """Uniform random target tree of `target_size` over the leaves
    [v0..v{n_vars-1}] + consts, sampled at `n_samples` uniform points."""
    rng = SplitMix64(seed)
    L = n_vars + len(consts)
    T = count_table(L, target_size)
    rank = rng.below(T[target_size][8])
    rpn = unrank(T, target_size, rank)
    m = (1 << W) - 1
    inputs = np.zeros((n_samples, n_vars), dtype=np.uint64)
    outputs = np.zeros(n_samples, dtype=np.uint64)
    for i in range(n_samples):
        row = [rng.next() & m for _ in range(n_vars)]
        inputs[i] = row
        outputs[i] = evaluate(rpn, row + [int(c) & m for c in consts], W)
    inst = MbaInstance(inputs, outputs, n_vars, W, consts)
    inst.target_rpn = rpn
    inst.target_size = target_size
    inst.target_rank = rank
    inst.original_formula = infix(rpn, n_vars, consts, W)
    return inst


def to_json(inst):
    """Instance JSON in the reference schema (scripts/sample.py:43-65)."""
    n, v, W = inst.n_examples, inst.n_variables, inst.width
    ex = []
    for i in range(n):
        ex.append({"inputs": [{"value": int(inst.inputs[i * v + k]), "size": W} for k in range(v)],
                   "output": [{"value": int(inst.outputs[i]), "size": W}]})
    return json.dumps({"original_formula": getattr(inst, "original_formula", ""),
                       "num_variables": v, "num_examples": n, "examples": ex})
