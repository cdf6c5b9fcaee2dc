"""ctypes binding of the CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg.  The product package (simba_amd) never imports
this module.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
MAX_RPN = 128
TOK_OP = 32
OPS = "~&|^#+-*"  # reference RPN characters, globals.cuh:13-15


class OrcResult(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int32),
        ("size", ctypes.c_int32),
        ("rank", ctypes.c_uint64),
        ("op", ctypes.c_int32),
        ("split", ctypes.c_int32),
        ("rpn", ctypes.c_uint8 * MAX_RPN),
        ("expr", ctypes.c_char * 1024),
        ("seconds", ctypes.c_double),
        ("candidates", ctypes.c_uint64),
        ("rows", ctypes.c_uint64),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(path)
        u64p = ctypes.POINTER(ctypes.c_uint64)
        L.orc_count_table.argtypes = [ctypes.c_int, ctypes.c_int, u64p]
        L.orc_unrank.argtypes = [u64p, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                 ctypes.POINTER(ctypes.c_uint8)]
        L.orc_rank.argtypes = [u64p, ctypes.POINTER(ctypes.c_uint8), ctypes.c_int]
        L.orc_rank.restype = ctypes.c_uint64
        L.orc_apply.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int]
        L.orc_apply.restype = ctypes.c_uint64
        L.orc_eval.argtypes = [ctypes.POINTER(ctypes.c_uint8), ctypes.c_int, u64p, ctypes.c_int]
        L.orc_eval.restype = ctypes.c_uint64
        L.orc_infix.argtypes = [ctypes.POINTER(ctypes.c_uint8), ctypes.c_int, ctypes.c_int,
                                u64p, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
        L.orc_simba_search.argtypes = [ctypes.c_int, ctypes.c_int, u64p, ctypes.c_int,
                                       ctypes.c_int, u64p, u64p, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                       ctypes.POINTER(OrcResult)]
        L.orc_simba_rtid_search.argtypes = [ctypes.c_int, ctypes.c_int, u64p, ctypes.c_int,
                                            ctypes.c_int, u64p, u64p, ctypes.c_int,
                                            ctypes.c_int, ctypes.c_int,
                                            ctypes.POINTER(OrcResult), u64p]
        L.orc_fmix32.argtypes = [ctypes.c_uint32]
        L.orc_fmix32.restype = ctypes.c_uint32
        L.orc_fmix64.argtypes = [ctypes.c_uint64]
        L.orc_fmix64.restype = ctypes.c_uint64
        L.orc_fingerprint.argtypes = [u64p, ctypes.c_int, ctypes.c_int]
        L.orc_fingerprint.restype = ctypes.c_uint64
        L.orc_vfb_search.argtypes = [ctypes.c_int, ctypes.c_int, u64p, ctypes.c_int,
                                     ctypes.c_int, u64p, u64p, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_uint64,
                                     ctypes.POINTER(OrcResult), u64p, u64p]
        L.orc_vfb_search_mt.argtypes = [ctypes.c_int, ctypes.c_int, u64p, ctypes.c_int,
                                        ctypes.c_int, u64p, u64p, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, ctypes.POINTER(OrcResult), u64p, u64p]
        _lib = L
    return _lib


def _u64(a):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.uint64))
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


def count_table(L, max_size):
    T = np.zeros((max_size + 1) * 9, dtype=np.uint64)
    rc = lib().orc_count_table(L, max_size, T.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
    if rc != 0:
        raise OverflowError("count table overflows uint64")
    return T.reshape(max_size + 1, 9)


def unrank(T, size, rank):
    Tf = np.ascontiguousarray(T.reshape(-1))
    buf = (ctypes.c_uint8 * MAX_RPN)()
    rc = lib().orc_unrank(Tf.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                          T.shape[0] - 1, size, rank, buf)
    if rc != 0:
        raise ValueError("rank out of range")
    return list(buf[:size])


def rank(T, rpn):
    Tf = np.ascontiguousarray(T.reshape(-1))
    buf = (ctypes.c_uint8 * MAX_RPN)(*rpn)
    return int(lib().orc_rank(Tf.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), buf, len(rpn)))


def evaluate(rpn, leafvals, W):
    buf = (ctypes.c_uint8 * MAX_RPN)(*rpn)
    _, p = _u64(leafvals)
    keep = _
    return int(lib().orc_eval(buf, len(rpn), p, W))


def infix(rpn, n_vars, consts=(), W=64, vfb_style=False):
    buf = (ctypes.c_uint8 * MAX_RPN)(*rpn)
    c, cp = _u64(list(consts) if len(consts) else [0])
    out = ctypes.create_string_buffer(1024)
    lib().orc_infix(buf, len(rpn), n_vars, cp, W, int(vfb_style), out, 1024)
    return out.value.decode()


def to_reference_rpn(rpn, n_vars):
    """RPN in the reference's character encoding ('0'+var, op chars)."""
    s = []
    for t in rpn:
        s.append(chr(ord("0") + t) if t < TOK_OP else OPS[t - TOK_OP])
    return "".join(s)


def _result(r):
    return {
        "status": r.status, "size": r.size, "rank": int(r.rank), "op": r.op,
        "split": r.split, "rpn": list(r.rpn[: max(r.size, 0)]),
        "expr": r.expr.decode(), "seconds": r.seconds,
        "candidates": int(r.candidates), "rows": int(r.rows),
    }


def simba_search(inputs, outputs, n_vars, consts=(), W=64, min_size=1, max_size=9,
                 threads=1, budget_s=0.0, exhaustive=False):
    """Minimum-rank first solution; `exhaustive` evaluates every size up to
    max_size (the answer is unchanged)."""
    inputs = np.asarray(inputs, dtype=np.uint64).reshape(-1)
    outputs = np.asarray(outputs, dtype=np.uint64)
    n = outputs.shape[0]
    ia, ip = _u64(inputs)
    oa, op = _u64(outputs)
    ca, cp = _u64(list(consts) if len(consts) else [0])
    r = OrcResult()
    lib().orc_simba_search(n_vars, len(consts), cp, W, n, ip, op, min_size, max_size,
                           threads, budget_s, int(exhaustive), ctypes.byref(r))
    return _result(r)


RTID_C = 2246822507


def simba_rtid_search(inputs, outputs, n_vars, consts=(), W=64, min_size=1, max_size=9,
                      threads=1):
    """SIMBA-RTid order (nocache_rtid.cu:85-95); `thread` is the smallest
    thread index that checks the returned tree."""
    inputs = np.asarray(inputs, dtype=np.uint64).reshape(-1)
    outputs = np.asarray(outputs, dtype=np.uint64)
    ia, ip = _u64(inputs)
    oa, op = _u64(outputs)
    ca, cp = _u64(list(consts) if len(consts) else [0])
    g = np.zeros(1, dtype=np.uint64)
    r = OrcResult()
    lib().orc_simba_rtid_search(n_vars, len(consts), cp, W, outputs.shape[0], ip, op,
                                min_size, max_size, threads, ctypes.byref(r),
                                g.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
    d = _result(r)
    d["thread"] = int(g[0]) if d["status"] == 0 else None
    return d


def fingerprint(values, W):
    a, p = _u64(values)
    return int(lib().orc_fingerprint(p, len(a), W))


def vfb_search_mt(inputs, outputs, n_vars, consts=(), W=32, max_len=8, threads=1,
                  exhaustive=True):
    """Multithreaded VFB (intended dedup; every level swept fully): the CPU
    baseline.  Same solution and per-level counts as vfb_search."""
    inputs = np.asarray(inputs, dtype=np.uint64).reshape(-1)
    outputs = np.asarray(outputs, dtype=np.uint64)
    n = outputs.shape[0]
    ia, ip = _u64(inputs)
    oa, op = _u64(outputs)
    ca, cp = _u64(list(consts) if len(consts) else [0])
    new = np.zeros(max_len + 1, dtype=np.uint64)
    cand = np.zeros(max_len + 1, dtype=np.uint64)
    r = OrcResult()
    lib().orc_vfb_search_mt(n_vars, len(consts), cp, W, n, ip, op, max_len, int(exhaustive),
                            threads, ctypes.byref(r),
                            new.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                            cand.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
    out = _result(r)
    out["level_new"] = [int(x) for x in new]
    out["level_cand"] = [int(x) for x in cand]
    return out


def vfb_search(inputs, outputs, n_vars, consts=(), W=32, max_len=8, dedup="intended",
               exhaustive=False, max_rows=0):
    inputs = np.asarray(inputs, dtype=np.uint64).reshape(-1)
    outputs = np.asarray(outputs, dtype=np.uint64)
    n = outputs.shape[0]
    ia, ip = _u64(inputs)
    oa, op = _u64(outputs)
    ca, cp = _u64(list(consts) if len(consts) else [0])
    new = np.zeros(max_len + 1, dtype=np.uint64)
    cand = np.zeros(max_len + 1, dtype=np.uint64)
    r = OrcResult()
    mode = {"intended": 0, "literal": 1}[dedup]
    lib().orc_vfb_search(n_vars, len(consts), cp, W, n, ip, op, max_len, mode,
                         int(exhaustive), max_rows, ctypes.byref(r),
                         new.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                         cand.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
    out = _result(r)
    out["level_new"] = [int(x) for x in new]
    out["level_cand"] = [int(x) for x in cand]
    return out
