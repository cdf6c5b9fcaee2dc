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
