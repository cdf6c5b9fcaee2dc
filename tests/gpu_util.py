"""Shared helpers for the GPU parity tests."""
import sys

import numpy as np

from oracle import oracle as orc


def note(msg):
    """Names the case before a GPU step, flushed, so a failing run shows it."""
    print(msg, flush=True)
    sys.stdout.flush()


def oracle_simba(inst, min_size, max_size, threads=8):
    return orc.simba_search(inst.inputs, inst.outputs, inst.n_variables,
                            [int(c) for c in inst.consts], inst.width, min_size,
                            max_size, threads)


def verify_rpn_chars(rpn_chars, inst):
    """Evaluates a reference-format RPN string on every sample."""
    L = inst.n_variables + len(inst.consts)
    toks = []
    for ch in rpn_chars:
        if "0" <= ch <= chr(ord("0") + L - 1):
            toks.append(ord(ch) - ord("0"))
        else:
            toks.append(orc.TOK_OP + "~&|^#+-*".index(ch))
    n = inst.n_examples
    for i in range(n):
        lv = [int(inst.inputs[i * inst.n_variables + v]) for v in range(inst.n_variables)]
        lv += [int(c) for c in inst.consts]
        if orc.evaluate(toks, lv, inst.width) != int(inst.outputs[i]):
            return False
    return True


def eval_vfb_infix(expr, leafvals, W):
    """Value of a VFB-format expression (full-cache.cu:483-509: leaves `v<i>`
    or decimal constants, unary `~(X)` / `-(X)`, binary `(L)op(R)`) on one
    sample's leaf values, mod 2^W.  Constant leaves are read from the text."""
    M = (1 << W) - 1
    ops = {"&": lambda a, b: a & b, "|": lambda a, b: a | b, "^": lambda a, b: a ^ b,
           "+": lambda a, b: a + b, "-": lambda a, b: a - b, "*": lambda a, b: a * b}

    def expect(i, ch):
        assert expr[i] == ch, (expr, i, ch)
        return i + 1

    def parse(i):
        c = expr[i]
        if c in "~-" and expr[i + 1] == "(":
            x, i = parse(i + 2)
            i = expect(i, ")")
            return ((~x) if c == "~" else -x) & M, i
        if c == "(":
            a, i = parse(i + 1)
            i = expect(i, ")")
            op = expr[i]
            i = expect(i + 1, "(")
            b, i = parse(i)
            i = expect(i, ")")
            return ops[op](a, b) & M, i
        j = i + 1
        while j < len(expr) and (expr[j].isdigit()):
            j += 1
        tok = expr[i:j]
        if tok[0] == "v":
            return leafvals[int(tok[1:])] & M, j
        return int(tok) & M, j

    v, end = parse(0)
    assert end == len(expr), expr
    return v


def verify_vfb_expr(expr, inst):
    """Evaluates a VFB result expression on every sample of the instance."""
    for i in range(inst.n_examples):
        lv = [int(inst.inputs[i * inst.n_variables + v]) for v in range(inst.n_variables)]
        if eval_vfb_infix(expr, lv, inst.width) != int(inst.outputs[i]):
            return False
    return True
