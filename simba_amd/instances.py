"""Instance pipeline: random target formulas and their sampled instances
(scripts/generate.py:1-116 and scripts/sample.py:1-129).

Formulas: a Python `random.Random(seed)` drives the recursive draw of
generate.py:28-49 -- a leaf draws its variable with randrange(n_vars); an
inner node of size s draws its operator uniformly from the unary list
[-, ~] (s >= 2) followed by the binary list [-, +, *, &, |, ^] (s >= 3), and
a binary node then draws its left size with randrange(1, s - 1); the left
subtree is drawn before the right.  Printing: `v<i>`, `(<op><x>)`,
`(<x> <op> <y>)`.  Formulas are produced for every size, then every
variable count, then every repeat (generate.py:103-116).

Instances: a numpy `default_rng(seed)` draws each example's inputs as one
`integers(2^32 - 1, size=n_vars, dtype=uint32, endpoint=True)` call
(sample.py:21-31); the output is the formula evaluated in uint32 arithmetic
(wrap-around + - * and unary -, bitwise ~ & | ^).  `num_variables` is the
largest variable index + 1 (sample.py:13-15).  The document layout and file
names (`<prefix>_<i>_seed=<seed>.json`, json.dump defaults) follow
sample.py:34-52,119-129.  The formula is evaluated by this module's own
parser; no text is passed to `eval`.

    python -m simba_amd.instances generate -s 3 8 -v 1 3 -r 2 --seed 42 -o formulas.txt
    python -m simba_amd.instances sample -i formulas.txt -n 16 --seed 42 -o instances/
"""
import argparse
import json
import random
import re
import sys
from pathlib import Path

import numpy as np

UNARY_OPS = ("-", "~")
BINARY_OPS = ("-", "+", "*", "&", "|", "^")
_VAR = re.compile(r"v(\d+)")
_M32 = 0xFFFFFFFF


# ---------------------------------------------------------------- formulas
def random_formula(size, n_vars, rng):
    """One formula of `size` nodes over v0..v{n_vars-1} (generate.py:37-49)."""
    if size == 1:
        return "v%d" % rng.randrange(n_vars)
    ops = [(o, 1) for o in UNARY_OPS]
    if size >= 3:
        ops += [(o, 2) for o in BINARY_OPS]
    op, arity = ops[rng.randrange(len(ops))]
    if arity == 1:
        return "(%s%s)" % (op, random_formula(size - 1, n_vars, rng))
    left = rng.randrange(1, size - 1)
    lf = random_formula(left, n_vars, rng)
    rf = random_formula(size - 1 - left, n_vars, rng)
    return "(%s %s %s)" % (lf, op, rf)


def generate(size_range, var_range, repeats=1, seed=None):
    """All formulas of generate.py's main loop, in its order."""
    rng = random.Random(seed)
    out = []
    for size in range(size_range[0], size_range[1] + 1):
        for nv in range(var_range[0], var_range[1] + 1):
            for _ in range(repeats):
                out.append(random_formula(size, nv, rng))
    return out


# ------------------------------------------------------------- evaluation
def count_variables(formula):
    idx = [int(m) for m in _VAR.findall(formula)]
    return (max(idx) if idx else 0) + 1


def parse(formula):
    """Fully parenthesised infix -> nested tuples ('v', i) | (op, x) | (op, x, y)."""
    toks = re.findall(r"v\d+|[()+\-*&|^~]", formula)
    if "".join(toks) != re.sub(r"\s+", "", formula):
        raise ValueError("unsupported token in %r" % formula)
    pos = 0

    def expr():
        nonlocal pos
        t = toks[pos]
        if t.startswith("v"):
            pos += 1
            return ("v", int(t[1:]))
        if t != "(":
            raise ValueError("unexpected %r in %r" % (t, formula))
        pos += 1
        if toks[pos] in UNARY_OPS and toks[pos + 1] != ")" and _starts_operand(toks[pos + 1]):
            op = toks[pos]
            pos += 1
            x = expr()
            _expect(")")
            return (op, x)
        x = expr()
        op = toks[pos]
        if op not in BINARY_OPS:
            raise ValueError("expected a binary operator in %r" % formula)
        pos += 1
        y = expr()
        _expect(")")
        return (op, x, y)

    def _expect(c):
        nonlocal pos
        if toks[pos] != c:
            raise ValueError("expected %r in %r" % (c, formula))
        pos += 1

    tree = expr()
    if pos != len(toks):
        raise ValueError("trailing tokens in %r" % formula)
    return tree


def _starts_operand(t):
    return t.startswith("v") or t == "("


def evaluate(tree, values):
    """uint32 value of a parsed formula (values: v0, v1, ...)."""
    if tree[0] == "v":
        return int(values[tree[1]]) & _M32
    if len(tree) == 2:
        x = evaluate(tree[1], values)
        return (-x) & _M32 if tree[0] == "-" else (~x) & _M32
    x, y = evaluate(tree[1], values), evaluate(tree[2], values)
    op = tree[0]
    if op == "+":
        return (x + y) & _M32
    if op == "-":
        return (x - y) & _M32
    if op == "*":
        return (x * y) & _M32
    if op == "&":
        return x & y
    if op == "|":
        return x | y
    return x ^ y


# --------------------------------------------------------------- instances
def sample_formula(formula, n_examples, rng):
    """Instance document of one formula (sample.py:55-77)."""
    nv = count_variables(formula)
    tree = parse(formula)
    examples = []
    for _ in range(n_examples):
        vals = rng.integers(np.iinfo(np.uint32).max, size=nv, dtype=np.uint32, endpoint=True)
        out = evaluate(tree, vals)
        examples.append({"inputs": [{"value": int(x), "size": 32} for x in vals],
                         "output": [{"value": int(out), "size": 32}]})
    return {"original_formula": formula, "num_variables": nv, "num_examples": n_examples,
            "examples": examples}


def sample_all(formulas, n_examples, seed=None):
    rng = np.random.default_rng(seed=seed)
    return [sample_formula(f, n_examples, rng) for f in formulas]


def instance_name(prefix, i, seed):
    return "%s_%d_seed=%s.json" % (prefix, i, seed)


def write_instances(formulas, n_examples, out_dir, seed=None, prefix="mba"):
    out_dir = Path(out_dir)
    out_dir.mkdir(exist_ok=True)
    names = []
    for i, doc in enumerate(sample_all(formulas, n_examples, seed)):
        name = instance_name(prefix, i, seed)
        with open(out_dir / name, "w") as f:
            json.dump(doc, f)
        names.append(name)
    return names


def read_formulas(path):
    with open(path) as f:
        return [ln.strip() for ln in f if ln.strip()]


def main(argv=None):
    ap = argparse.ArgumentParser(description="Generate formulas or sample instances.")
    sub = ap.add_subparsers(dest="cmd", required=True)
    g = sub.add_parser("generate")
    g.add_argument("--size_range", "-s", type=int, nargs=2, required=True)
    g.add_argument("--var_range", "-v", type=int, nargs=2, required=True)
    g.add_argument("--repeats", "-r", type=int, default=1)
    g.add_argument("--output_file", "-o", type=Path, required=True)
    g.add_argument("--seed", type=int, default=None)
    s = sub.add_parser("sample")
    s.add_argument("--input_file", "-i", type=Path, required=True)
    s.add_argument("--n_examples", "-n", type=int, required=True)
    s.add_argument("--output_folder", "-o", type=Path, required=True)
    s.add_argument("--seed", "-s", type=int, default=None)
    s.add_argument("--prefix", "-p", default="mba")
    a = ap.parse_args(argv)
    if a.cmd == "generate":
        fs = generate(a.size_range, a.var_range, a.repeats, a.seed)
        with open(a.output_file, "w") as f:
            for x in fs:
                f.write(x + "\n")
    else:
        write_instances(read_formulas(a.input_file), a.n_examples, a.output_folder, a.seed,
                        a.prefix)
    return 0


if __name__ == "__main__":
    sys.exit(main())
