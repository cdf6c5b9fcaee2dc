"""simba_amd -- MI355X (gfx950) MBA synthesis: SIMBA and VFB search.

Python mirror of the reference's search interface over the C ABI of
libsimba_amd.so (include/mba_api.h):

  MbaInstance.from_json_file(path)     simba/src/instance.h:19-29
  NoCache.solve(mba, max_size) -> str  simba/src/nocache.cu:197-289
  MBA(maxLen, numExamples, inputData, outputData, numVar) -> str
                                       vfb/src/full-cache.cu:665-774

plus richer entry points (simba_solve, vfb_solve, batch_solve) that return
the full result record.  There is no CPU fallback: without the built library
or without a gfx950 device every entry point raises.
"""
import ctypes
import os
import weakref
from collections.abc import Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsimba_amd.so")

MAX_SIZE = 32
EXPR_CAP = 1024
FOUND, NOT_FOUND, BUDGET, OVERFLOW = 0, 1, 2, 3


class Instance(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("n_vars", ctypes.c_uint32),
                ("n_samples", ctypes.c_uint32), ("n_consts", ctypes.c_uint32),
                ("inputs", ctypes.POINTER(ctypes.c_uint64)),
                ("outputs", ctypes.POINTER(ctypes.c_uint64)),
                ("consts", ctypes.POINTER(ctypes.c_uint64))]


class Opts(ctypes.Structure):
    _fields_ = [("min_size", ctypes.c_int32), ("max_size", ctypes.c_int32),
                ("compat", ctypes.c_int32), ("exhaustive", ctypes.c_int32),
                ("time_budget_s", ctypes.c_double),
                ("cancel", ctypes.POINTER(ctypes.c_int32)),
                ("table_bytes", ctypes.c_uint64), ("cache_bytes", ctypes.c_uint64),
                ("verbose", ctypes.c_int32), ("wg_per_cu", ctypes.c_int32),
                ("lanes_b", ctypes.c_int32), ("table_delta", ctypes.c_int32),
                ("block_log2", ctypes.c_int32), ("hit_cap", ctypes.c_int32),
                ("stop_size", ctypes.c_int32), ("reserved", ctypes.c_int32 * 1)]


class Result(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("size", ctypes.c_int32),
                ("rank", ctypes.c_uint64), ("op", ctypes.c_int32),
                ("split", ctypes.c_int32), ("expr", ctypes.c_char * EXPR_CAP),
                ("rpn", ctypes.c_char * (MAX_SIZE + 1)), ("seconds", ctypes.c_double),
                ("kernel_seconds", ctypes.c_double), ("candidates", ctypes.c_uint64),
                ("size_reached", ctypes.c_int32), ("kernel_launches", ctypes.c_uint32),
                ("hit_replays", ctypes.c_int32), ("ordered_sizes", ctypes.c_int32),
                ("split_reached", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("split_pairs", ctypes.c_uint64)]

    def as_dict(self):
        return {"status": self.status, "size": self.size, "rank": int(self.rank),
                "op": self.op, "split": self.split, "expr": self.expr.decode(),
                "rpn": self.rpn.decode(), "seconds": self.seconds,
                "kernel_seconds": self.kernel_seconds,
                "kernel_launches": self.kernel_launches,
                "candidates": int(self.candidates), "size_reached": self.size_reached,
                "hit_replays": self.hit_replays, "ordered_sizes": self.ordered_sizes,
                "split_reached": self.split_reached, "split_pairs": int(self.split_pairs)}


def _result_array(n):
    """n result structs in an uninitialised buffer (the library writes every
    field of every result)."""
    return (Result * n).from_buffer(np.empty(n * ctypes.sizeof(Result), np.uint8))


class BatchResults(Sequence):
    """Results of a batched call, one per instance in input order: the C
    result structs, each turned into a dict (Result.as_dict) when read."""

    def __init__(self, res):
        self._res = res
        self._d = [None] * len(res)

    def __len__(self):
        return len(self._res)

    def __getitem__(self, i):
        if isinstance(i, slice):
            return [self[k] for k in range(*i.indices(len(self)))]
        d = self._d[i]
        if d is None:
            d = self._d[i] = self._res[i].as_dict()
        return d


class VfbStats(ctypes.Structure):
    _fields_ = [("level_new", ctypes.c_uint64 * (MAX_SIZE + 1)),
                ("level_cand", ctypes.c_uint64 * (MAX_SIZE + 1)),
                ("level_seconds", ctypes.c_double * (MAX_SIZE + 1)),
                ("level_bytes", ctypes.c_uint64 * (MAX_SIZE + 1)),
                ("table_slots", ctypes.c_uint64), ("rows_capacity", ctypes.c_uint64),
                ("level_cas", ctypes.c_uint64 * (MAX_SIZE + 1)),
                ("level_atomics", ctypes.c_uint64 * (MAX_SIZE + 1))]


# Every symbol the C ABI exports (include/mba_api.h).
EXPORTS = ["mba_ctx_create", "mba_ctx_destroy", "mba_strerror", "mba_ctx_last_error",
           "mba_default_opts", "mba_simba_solve", "mba_simba_solve_noalloc",
           "mba_simba_rtid_solve", "mba_vfb_solve",
           "mba_batch_solve", "mba_vfb_fingerprints", "mba_vfb_probe_ceiling",
           "mba_instance_read_json", "mba_instance_free", "mba_rpn_to_infix",
           "mba_features", "mba_features_batch", "mba_mlp_create", "mba_mlp_destroy",
           "mba_mlp_get_weights", "mba_mlp_set_weights", "mba_mlp_forward", "mba_mlp_grads",
           "mba_mlp_train_step",
           "mba_learned_solve", "mba_learned_batch_solve"]
MLP_IN, MLP_HIDDEN, MLP_OUT = 256, 512, 128
FEAT_HASHED = 240  # hashed sample-bit buckets; the last 16 inputs: size one-hot

_lib = None


def lib():
    """Loads libsimba_amd.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: build with `python -m simba_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        L.mba_ctx_create.argtypes = [ctypes.c_int, P(ctypes.c_void_p)]
        L.mba_ctx_destroy.argtypes = [ctypes.c_void_p]
        L.mba_strerror.argtypes = [ctypes.c_int]
        L.mba_strerror.restype = ctypes.c_char_p
        L.mba_ctx_last_error.argtypes = [ctypes.c_void_p]
        L.mba_ctx_last_error.restype = ctypes.c_char_p
        L.mba_default_opts.argtypes = [P(Opts)]
        L.mba_simba_solve.argtypes = [ctypes.c_void_p, P(Instance), P(Opts), P(Result)]
        L.mba_simba_rtid_solve.argtypes = [ctypes.c_void_p, P(Instance), P(Opts), P(Result)]
        L.mba_simba_solve_noalloc.argtypes = [ctypes.c_void_p, P(Instance), P(Opts),
                                              P(Result)]
        L.mba_vfb_solve.argtypes = [ctypes.c_void_p, P(Instance), P(Opts), P(Result),
                                    P(VfbStats)]
        L.mba_batch_solve.argtypes = [ctypes.c_void_p, P(Instance), ctypes.c_size_t,
                                      P(Opts), P(Result)]
        L.mba_vfb_fingerprints.argtypes = [ctypes.c_void_p, ctypes.c_uint32,
                                           P(ctypes.c_uint64), ctypes.c_uint32,
                                           ctypes.c_size_t, P(ctypes.c_uint64)]
        L.mba_vfb_probe_ceiling.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64,
                                            ctypes.c_uint64, ctypes.c_uint64,
                                            P(ctypes.c_double), P(ctypes.c_double),
                                            P(ctypes.c_uint64)]
        L.mba_instance_read_json.argtypes = [ctypes.c_char_p, P(Instance)]
        L.mba_instance_free.argtypes = [P(Instance)]
        L.mba_rpn_to_infix.argtypes = [ctypes.c_char_p, ctypes.c_uint32, P(ctypes.c_uint64),
                                       ctypes.c_uint32, ctypes.c_uint32, ctypes.c_char_p,
                                       ctypes.c_size_t]
        fp = P(ctypes.c_float)
        L.mba_features.argtypes = [P(Instance), fp]
        L.mba_features_batch.argtypes = [ctypes.c_void_p, P(Instance), ctypes.c_size_t, fp]
        L.mba_mlp_create.argtypes = [ctypes.c_void_p, ctypes.c_uint64, P(ctypes.c_void_p)]
        L.mba_mlp_destroy.argtypes = [ctypes.c_void_p]
        L.mba_mlp_get_weights.argtypes = [ctypes.c_void_p, fp, fp, fp, fp]
        L.mba_mlp_set_weights.argtypes = [ctypes.c_void_p, fp, fp, fp, fp]
        L.mba_mlp_forward.argtypes = [ctypes.c_void_p, fp, ctypes.c_int, fp]
        L.mba_mlp_grads.argtypes = [ctypes.c_void_p, fp, P(ctypes.c_int32), ctypes.c_int, fp,
                                    fp, fp, fp, fp]
        L.mba_mlp_train_step.argtypes = [ctypes.c_void_p, fp, P(ctypes.c_int32), ctypes.c_int,
                                         ctypes.c_float, fp]
        L.mba_learned_solve.argtypes = [ctypes.c_void_p, ctypes.c_void_p, P(Instance), P(Opts),
                                        P(Result)]
        L.mba_learned_batch_solve.argtypes = [ctypes.c_void_p, ctypes.c_void_p, P(Instance),
                                              ctypes.c_size_t, P(Opts), P(Result)]
        _lib = L
    return _lib


class MbaError(RuntimeError):
    pass


class Context:
    """One device context (stream + reusable device buffers)."""

    def __init__(self, device=0):
        self._h = ctypes.c_void_p()
        self._mlps = weakref.WeakSet()  # networks created on this context
        rc = lib().mba_ctx_create(device, ctypes.byref(self._h))
        if rc != 0:
            raise MbaError(f"mba_ctx_create failed: {lib().mba_strerror(rc).decode()}")

    def close(self):
        """Destroys the context and, first, every network created on it."""
        if self._h:
            for m in list(self._mlps):
                m.close()
            lib().mba_ctx_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc < 0:
            msg = lib().mba_ctx_last_error(self._h).decode()
            raise MbaError(f"{lib().mba_strerror(rc).decode()}: {msg}")
        return rc

    def simba(self, inst, **kw):
        ci, keep = inst.to_c()
        o = make_opts(**kw)
        r = Result()
        self._check(lib().mba_simba_solve(self._h, ctypes.byref(ci), ctypes.byref(o),
                                          ctypes.byref(r)))
        return r.as_dict()

    def simba_noalloc(self, inst, **kw):
        """SIMBA search without device allocation (mba_simba_solve_noalloc)."""
        ci, keep = inst.to_c()
        o = make_opts(**kw)
        r = Result()
        self._check(lib().mba_simba_solve_noalloc(self._h, ctypes.byref(ci), ctypes.byref(o),
                                                  ctypes.byref(r)))
        return r.as_dict()

    def simba_rtid(self, inst, **kw):
        """SIMBA-RTid order (nocache_rtid.cu:85-95)."""
        ci, keep = inst.to_c()
        o = make_opts(**kw)
        r = Result()
        self._check(lib().mba_simba_rtid_solve(self._h, ctypes.byref(ci), ctypes.byref(o),
                                               ctypes.byref(r)))
        return r.as_dict()

    def vfb(self, inst, **kw):
        ci, keep = inst.to_c()
        o = make_opts(**kw)
        r = Result()
        st = VfbStats()
        self._check(lib().mba_vfb_solve(self._h, ctypes.byref(ci), ctypes.byref(o),
                                        ctypes.byref(r), ctypes.byref(st)))
        d = r.as_dict()
        d["level_new"] = list(st.level_new)
        d["level_cand"] = list(st.level_cand)
        d["level_seconds"] = list(st.level_seconds)
        d["level_bytes"] = list(st.level_bytes)
        d["table_slots"] = int(st.table_slots)
        d["level_cas"] = [int(x) for x in st.level_cas]
        d["level_atomics"] = [int(x) for x in st.level_atomics]
        return d

    def vfb_fingerprints(self, values, width):
        """Device fingerprints (the VFB dedup key) of the rows of `values`
        ([count][n] words), as numpy uint64 (mba_vfb_fingerprints)."""
        v = np.ascontiguousarray(np.asarray(values, dtype=np.uint64))
        if v.ndim != 2:
            raise ValueError("values must be [count][n_words]")
        out = np.zeros(v.shape[0], np.uint64)
        P = ctypes.POINTER(ctypes.c_uint64)
        self._check(lib().mba_vfb_fingerprints(self._h, width, v.ctypes.data_as(P), v.shape[1],
                                               v.shape[0], out.ctypes.data_as(P)))
        return out

    def vfb_probe_ceiling(self, slots_log2, count, distinct, prefill=0):
        """Random-access ceiling of the VFB insert pass (mba_vfb_probe_ceiling):
        seconds of one probe per key and of the insert pass's table code on
        `count` keys with `distinct` values, after `prefill` other keys."""
        ps, ins, fresh = ctypes.c_double(), ctypes.c_double(), ctypes.c_uint64()
        self._check(lib().mba_vfb_probe_ceiling(self._h, slots_log2, count, distinct, prefill,
                                                ctypes.byref(ps), ctypes.byref(ins),
                                                ctypes.byref(fresh)))
        return {"probe_seconds": ps.value, "insert_seconds": ins.value, "fresh": fresh.value}

    def batch(self, insts, **kw):
        """Solves the instances in one batched call; returns a BatchResults
        sequence of result dicts (mba_batch_solve)."""
        n = len(insts)
        arr = pack_instances(insts)
        o = make_opts(**kw)
        res = _result_array(n)
        self._check(lib().mba_batch_solve(self._h, arr, n, ctypes.byref(o), res))
        return BatchResults(res)


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def features(inst):
    """256 features of an instance (float32): 240 hashed input/output-bit
    buckets and a size one-hot left at zero (set_size_slot)."""
    ci, keep = inst.to_c()
    out = np.zeros(MLP_IN, dtype=np.float32)
    rc = lib().mba_features(ctypes.byref(ci), _fptr(out))
    if rc != 0:
        raise MbaError(lib().mba_strerror(rc).decode())
    return out


def set_size_slot(x, size):
    """Sets the size one-hot of feature rows x (n x 256, in place): the size
    whose blocks the network orders (sizes above 16 share the last slot)."""
    x = np.asarray(x)
    x[..., FEAT_HASHED:] = 0.0
    sizes = np.broadcast_to(np.asarray(size), x.shape[:-1])
    slot = FEAT_HASHED + np.clip(sizes, 1, MLP_IN - FEAT_HASHED) - 1
    np.put_along_axis(x, slot[..., None].astype(np.int64), 1.0, axis=-1)
    return x


def features_batch(ctx, insts):
    """Features of many instances computed on the device (count x 256 float32,
    equal to features() row by row)."""
    out = np.zeros((len(insts), MLP_IN), dtype=np.float32)
    if len(insts):
        ctx._check(lib().mba_features_batch(ctx._h, pack_instances(insts), len(insts),
                                            _fptr(out)))
    return out


def block_class(op, split):
    """Output class of a root block: op * 16 + min(left size, 15)."""
    return op * 16 + min(max(split, 0), 15)


class Mlp:
    """The learned block-proposal network on one context (device weights)."""

    def __init__(self, ctx, seed=0):
        self.ctx = ctx
        self._h = ctypes.c_void_p()
        rc = lib().mba_mlp_create(ctx._h, seed, ctypes.byref(self._h))
        if rc != 0:
            raise MbaError(lib().mba_strerror(rc).decode())
        ctx._mlps.add(self)

    def close(self):
        if self._h:
            lib().mba_mlp_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_weights(self):
        w1 = np.zeros((MLP_IN, MLP_HIDDEN), np.float32)
        b1 = np.zeros(MLP_HIDDEN, np.float32)
        w2 = np.zeros((MLP_HIDDEN, MLP_OUT), np.float32)
        b2 = np.zeros(MLP_OUT, np.float32)
        self.ctx._check(lib().mba_mlp_get_weights(self._h, _fptr(w1), _fptr(b1), _fptr(w2),
                                                  _fptr(b2)))
        return w1, b1, w2, b2

    def set_weights(self, w1, b1, w2, b2):
        arrs = [np.ascontiguousarray(x, dtype=np.float32) for x in (w1, b1, w2, b2)]
        self.ctx._check(lib().mba_mlp_set_weights(self._h, *[_fptr(x) for x in arrs]))

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, MLP_IN)
        out = np.zeros((x.shape[0], MLP_OUT), np.float32)
        self.ctx._check(lib().mba_mlp_forward(self._h, _fptr(x), x.shape[0], _fptr(out)))
        return out

    def grads(self, x, labels):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, MLP_IN)
        y = np.ascontiguousarray(labels, dtype=np.int32)
        loss = np.zeros(1, np.float32)
        g = [np.zeros((MLP_IN, MLP_HIDDEN), np.float32), np.zeros(MLP_HIDDEN, np.float32),
             np.zeros((MLP_HIDDEN, MLP_OUT), np.float32), np.zeros(MLP_OUT, np.float32)]
        self.ctx._check(lib().mba_mlp_grads(
            self._h, _fptr(x), y.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), x.shape[0],
            _fptr(loss), *[_fptr(a) for a in g]))
        return float(loss[0]), g

    def train_step(self, x, labels, lr):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, MLP_IN)
        y = np.ascontiguousarray(labels, dtype=np.int32)
        loss = np.zeros(1, np.float32)
        self.ctx._check(lib().mba_mlp_train_step(
            self._h, _fptr(x), y.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), x.shape[0],
            ctypes.c_float(lr), _fptr(loss)))
        return float(loss[0])

    def solve(self, inst, **kw):
        ci, keep = inst.to_c()
        o = make_opts(**kw)
        r = Result()
        self.ctx._check(lib().mba_learned_solve(self.ctx._h, self._h, ctypes.byref(ci),
                                                ctypes.byref(o), ctypes.byref(r)))
        return r.as_dict()

    def batch_solve(self, insts, **kw):
        n = len(insts)
        arr = pack_instances(insts)
        o = make_opts(**kw)
        res = _result_array(n)
        self.ctx._check(lib().mba_learned_batch_solve(self.ctx._h, self._h, arr, n,
                                                      ctypes.byref(o), res))
        return BatchResults(res)


def make_opts(min_size=1, max_size=9, compat=False, exhaustive=False, budget_s=0.0,
              table_bytes=0, cache_bytes=0, verbose=False, cancel=None, wg_per_cu=0,
              lanes_b=0, table_delta=0, block_log2=0, hit_cap=0, stop_size=0):
    o = Opts()
    lib().mba_default_opts(ctypes.byref(o))
    o.wg_per_cu, o.lanes_b, o.table_delta = wg_per_cu, lanes_b, table_delta
    o.block_log2, o.hit_cap, o.stop_size = block_log2, hit_cap, stop_size
    o.min_size, o.max_size = min_size, max_size
    o.compat, o.exhaustive = int(compat), int(exhaustive)
    o.time_budget_s = budget_s
    o.table_bytes, o.cache_bytes = table_bytes, cache_bytes
    o.verbose = int(verbose)
    if cancel is not None:
        o.cancel = ctypes.cast(ctypes.addressof(cancel), ctypes.POINTER(ctypes.c_int32))
    return o


class MbaInstance:
    """A synthesis instance: inputs [sample][var], outputs [sample] (mirror of
    the reference's MbaInstance, simba/src/instance.h:10-30)."""

    def __init__(self, inputs, outputs, n_variables, width=32, consts=()):
        self.width = int(width)
        m = np.uint64((1 << self.width) - 1) if self.width < 64 else np.uint64(~0 & (2**64 - 1))
        self.inputs = (np.asarray(inputs, dtype=np.uint64).reshape(-1) & m).copy()
        self.outputs = (np.asarray(outputs, dtype=np.uint64).reshape(-1) & m).copy()
        self.n_variables = int(n_variables)
        self.n_examples = int(self.outputs.shape[0])
        self.consts = np.asarray(list(consts), dtype=np.uint64) & m
        if self.inputs.shape[0] != self.n_variables * self.n_examples:
            raise ValueError("inputs must hold n_examples * n_variables values")

    @staticmethod
    def from_json_file(path):
        ci = Instance()
        rc = lib().mba_instance_read_json(os.fsencode(path), ctypes.byref(ci))
        if rc != 0:
            raise MbaError(f"Error: Could not read {path}: {lib().mba_strerror(rc).decode()}")
        try:
            n, v = ci.n_samples, ci.n_vars
            inp = np.ctypeslib.as_array(ci.inputs, shape=(max(n * v, 1),))[: n * v].copy()
            out = np.ctypeslib.as_array(ci.outputs, shape=(n,)).copy()
            return MbaInstance(inp, out, v, ci.width)
        finally:
            lib().mba_instance_free(ctypes.byref(ci))

    def to_c(self):
        """The C view of the instance (cached: an MbaInstance is not changed
        after construction; build a new one for new data)."""
        c = getattr(self, "_c", None)
        if c is None:
            c = self._c = self._make_c()
            self._cb = bytes(c[0])
        return c

    def to_c_bytes(self):
        """The bytes of the C view (its arrays stay referenced by to_c)."""
        self.to_c()
        return self._cb

    def _make_c(self):
        ci = Instance()
        ci.width = self.width
        ci.n_vars = self.n_variables
        ci.n_samples = self.n_examples
        ci.n_consts = len(self.consts)
        inp = np.ascontiguousarray(self.inputs) if self.inputs.size else np.zeros(1, np.uint64)
        out = np.ascontiguousarray(self.outputs)
        cst = np.ascontiguousarray(self.consts) if len(self.consts) else np.zeros(1, np.uint64)
        P = ctypes.POINTER(ctypes.c_uint64)
        ci.inputs = inp.ctypes.data_as(P)
        ci.outputs = out.ctypes.data_as(P)
        ci.consts = cst.ctypes.data_as(P)
        return ci, (inp, out, cst)


def pack_instances(insts):
    """ctypes array of the instances' C views, for the batched calls."""
    return (Instance * len(insts)).from_buffer_copy(b"".join(x.to_c_bytes() for x in insts))


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class NoCache:
    """Mirror of the reference's SIMBA entry point (nocache.cu:197-289)."""

    @staticmethod
    def solve(mba, max_size):
        """Reference semantics: sizes 1..max_size-1 (exclusive, size 1 holds no
        candidates); returns the infix string or "Not found !"."""
        if max_size > MAX_SIZE:
            raise MbaError(f"This version supports formulas of size at most {MAX_SIZE}")
        r = default_context().simba(mba, max_size=max_size, compat=True)
        return r["expr"]


def MBA(maxLen, numExamples, inputData, outputData, numVar, width=32):
    """Mirror of VFB's MBA() (full-cache.cu:665-774): levels 1..maxLen."""
    if numExamples > 100:
        return "see_the_error"
    inst = MbaInstance(inputData, outputData, numVar, width)
    r = default_context().vfb(inst, max_size=maxLen, compat=True)
    return r["expr"]


def simba_solve(inst, **kw):
    return default_context().simba(inst, **kw)


def vfb_solve(inst, **kw):
    return default_context().vfb(inst, **kw)


def batch_solve(insts, **kw):
    return default_context().batch(insts, **kw)
