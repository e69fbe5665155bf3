"""ctypes binding of libsc_amd.so (the C ABI in include/sc_hip.h).

Host-side plumbing for the tests and the benchmark: device buffers, the
context and thin wrappers that raise on a non-zero status.  All compute runs
in the HIP library; there is no CPU fallback here -- a missing library or
device raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libsc_amd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "sc_hip.h")

SC_OK = 0
SC_ERR_ARG, SC_ERR_NOMEM, SC_ERR_CONFIG, SC_ERR_HIP = -1, -2, -3, -4
SC_ERR_TOO_LARGE, SC_ERR_NO_DEVICE, SC_ERR_HEAP_EMPTY = -5, -6, -7
KIND_IVFADC, KIND_IMI, KIND_SC, KIND_SC3, KIND_SCIMI = 0, 1, 2, 3, 4
PREC_FP32, PREC_BF16 = 0, 1
FORM_PLAIN, FORM_MIXED, FORM_REAL = 0, 1, 2


class ScError(RuntimeError):
    def __init__(self, fn, rc, msg):
        super().__init__("%s failed (rc=%d): %s" % (fn, rc, msg))
        self.rc = rc


class Mix(C.Structure):
    _fields_ = [("seed", C.c_uint32), ("ncomp", C.c_int32),
                ("frac_bits", C.c_int32), ("out_bits", C.c_int32),
                ("mu_lo", C.c_int64), ("mu_span", C.c_int64),
                ("sig_lo", C.c_int64), ("sig_span", C.c_int64),
                ("maxq", C.c_int64)]


class EncodeStats(C.Structure):
    _fields_ = [("rechecked", C.c_int64), ("fallback", C.c_int64)]


class SearchParams(C.Structure):
    _fields_ = [("kind", C.c_int32), ("form", C.c_int32), ("R", C.c_int32),
                ("w", C.c_int32), ("T", C.c_int32), ("M", C.c_int32)]


class SearchStats(C.Structure):
    _fields_ = [("queries", C.c_int64), ("candidates", C.c_int64),
                ("cells", C.c_int64), ("visited", C.c_int64),
                ("reranked", C.c_int64), ("heap_fallback", C.c_int64),
                ("exhausted", C.c_int64), ("ms_lut", C.c_double),
                ("ms_rank", C.c_double), ("ms_traverse", C.c_double),
                ("ms_scan", C.c_double), ("fused_phase_ms", C.c_double * 5),
                ("ms_scan_kernel", C.c_double)]


def mix_d128(seed=20151001, ncomp=64):
    """D=128 integer data in [0,255]: mu ~ U[24,232], sigma ~ U[8,32]."""
    return Mix(seed, ncomp, 8, 0, 24 * 256, 208 * 256, 8 * 256, 24 * 256, 255)


def mix_d960(seed=20151001, ncomp=32):
    """D=960 data on the 2^-16 grid in [0,1]: mu ~ U[0.2,0.8], sigma 0.1."""
    return Mix(seed, ncomp, 16, 16, 13107, 39321, 6554, 0, 65536)


STREAM_TRAIN, STREAM_BASE, STREAM_QUERY = 1, 2, 3

_P = C.c_void_p
_SIGS = {
    "sc_last_error": (C.c_char_p, []),
    "sc_version": (C.c_char_p, []),
    "sc_ctx_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "sc_ctx_destroy": (C.c_int, [_P]),
    "sc_ctx_sync": (C.c_int, [_P]),
    "sc_ctx_stream": (_P, [_P]),
    "sc_malloc": (C.c_int, [_P, C.c_size_t, C.POINTER(_P)]),
    "sc_free": (C.c_int, [_P, _P]),
    "sc_memcpy_h2d": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "sc_memcpy_d2h": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "sc_memset": (C.c_int, [_P, _P, C.c_int, C.c_size_t]),
    "sc_gen_rows": (C.c_int, [_P, C.POINTER(Mix), C.c_uint32, C.c_int64, C.c_int64,
                              C.c_int32, _P]),
    "sc_encode": (C.c_int, [_P, _P, C.c_int64, C.c_int32, _P, C.c_int32, C.c_int32, _P,
                            C.c_int32, C.c_int32, C.c_int32, C.c_int32, _P, _P,
                            C.POINTER(EncodeStats)]),
    "sc_index_build": (C.c_int, [_P, _P, _P, C.c_int64, C.c_int32, _P, C.c_int32,
                                 C.c_int32, _P, C.c_int32, C.c_int32, C.c_int32,
                                 C.POINTER(_P)]),
    "sc_index_load": (C.c_int, [_P, _P, C.c_int64, _P, _P, C.c_int64, C.c_int32, _P,
                                C.c_int32, C.c_int32, _P, C.c_int32, C.c_int32, C.c_int32,
                                C.POINTER(_P)]),
    "sc_index_load_sparse": (C.c_int, [_P, _P, _P, C.c_int64, _P, _P, C.c_int64, C.c_int32, _P,
                                       C.c_int32, _P, C.c_int32, C.c_int32, C.POINTER(_P)]),
    "sc_index_cell_keys": (C.c_int, [_P, _P, _P]),
    "sc_index_destroy": (C.c_int, [_P]),
    "sc_index_attach_raw": (C.c_int, [_P, _P, C.c_int64]),
    "sc_encode_general": (C.c_int, [_P, _P, C.c_int64, C.c_int32, _P, C.c_int32, C.c_int32, _P,
                                    C.c_int32, C.c_int32, C.c_int32, _P, _P]),
    "sc_encode_given": (C.c_int, [_P, _P, C.c_int64, C.c_int32, _P, C.c_int64, _P, _P,
                                  C.c_int32, C.c_int32, _P, C.c_int32, C.c_int32, _P]),
    "sc_groundtruth_nn1": (C.c_int, [_P, _P, C.c_int64, _P, C.c_int64, C.c_int32, _P, _P]),
    "sc_groundtruth_nn1_noalloc": (C.c_int, [_P, _P, C.c_int64, _P, C.c_int64, C.c_int32, _P,
                                             _P]),
    "sc_index_shape": (C.c_int, [_P, _P]),
    "sc_index_export": (C.c_int, [_P, _P, _P, _P, _P, _P, _P, _P]),
    "sc_search": (C.c_int, [_P, _P, _P, C.c_int64, C.POINTER(SearchParams), _P, _P, _P,
                            C.POINTER(SearchStats)]),
    "sc_search_budgets": (C.c_int, [_P, _P, _P, C.c_int64, C.POINTER(SearchParams), _P, _P, _P,
                                    _P, C.POINTER(SearchStats)]),
    "sc_plan_create": (C.c_int, [_P, _P, C.POINTER(SearchParams), _P, C.c_int64, _P, _P, _P,
                                 C.POINTER(_P)]),
    "sc_plan_create_budgets": (C.c_int, [_P, _P, C.POINTER(SearchParams), _P, _P, C.c_int64,
                                         _P, _P, _P, C.POINTER(_P)]),
    "sc_plan_execute": (C.c_int, [_P, C.POINTER(SearchStats)]),
    "sc_plan_autotune": (C.c_int, [_P, _P]),
    "sc_plan_destroy": (C.c_int, [_P]),
    "sc_kmeans": (C.c_int, [_P, _P, C.c_int64, C.c_int32, C.c_int32, C.c_int32,
                            C.c_uint32, C.c_int32, C.c_int32, _P, _P]),
    "sc_residual": (C.c_int, [_P, _P, C.c_int64, C.c_int32, _P, C.c_int32, C.c_int32,
                              C.c_int32]),
}

_LIB = None


def lib():
    """Loads libsc_amd.so; raises when it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libsc_amd.so is missing (%s): run __graft_entry__.build()"
                               % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def header_symbols():
    """Names of the functions include/sc_hip.h declares."""
    import re
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?(?:int|void|char)\s*\*?\s*(sc_\w+)\s*\(",
                                 text, re.M)))


def _check(fn, rc):
    if rc != SC_OK:
        raise ScError(fn, rc, lib().sc_last_error().decode())


class Ctx:
    def __init__(self, device=0):
        self.h = _P()
        _check("sc_ctx_create", lib().sc_ctx_create(device, C.byref(self.h)))

    def sync(self):
        _check("sc_ctx_sync", lib().sc_ctx_sync(self.h))

    def close(self):
        if self.h:
            lib().sc_ctx_destroy(self.h)
            self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Dev:
    """A device buffer with a numpy dtype and shape."""

    def __init__(self, ctx, shape, dtype):
        self.ctx = ctx
        self.shape = tuple(int(s) for s in np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = _P()
        _check("sc_malloc", lib().sc_malloc(ctx.h, max(self.nbytes, 1), C.byref(self.ptr)))

    @classmethod
    def from_numpy(cls, ctx, arr):
        arr = np.ascontiguousarray(arr)
        d = cls(ctx, arr.shape, arr.dtype)
        if d.nbytes:
            _check("sc_memcpy_h2d", lib().sc_memcpy_h2d(ctx.h, d.ptr, arr.ctypes.data, d.nbytes))
        return d

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            _check("sc_memcpy_d2h", lib().sc_memcpy_d2h(self.ctx.h, out.ctypes.data, self.ptr,
                                                        self.nbytes))
        return out

    def copy_from(self, arr):
        """Overwrites the buffer with a host array of the same size (synchronous)."""
        arr = np.ascontiguousarray(arr, dtype=self.dtype)
        if arr.nbytes != self.nbytes:
            raise ValueError("copy_from: %d bytes into a %d-byte buffer" % (arr.nbytes, self.nbytes))
        if self.nbytes:
            _check("sc_memcpy_h2d", lib().sc_memcpy_h2d(self.ctx.h, self.ptr, arr.ctypes.data,
                                                        self.nbytes))
        return self

    def offset(self, elems):
        return _P(self.ptr.value + int(elems) * self.dtype.itemsize)

    def rows(self, r0, n):
        """Rows [r0, r0 + n) of the buffer as a Dev that does not own its
        memory (no copy; it keeps this buffer alive)."""
        inner = self.shape[1:]
        per = int(np.prod(inner)) if inner else 1
        if r0 < 0 or n < 0 or r0 + n > self.shape[0]:
            raise ValueError("rows: [%d, %d) outside %d rows" % (r0, r0 + n, self.shape[0]))
        v = Dev.__new__(Dev)
        v.ctx, v.dtype, v.shape = self.ctx, self.dtype, (int(n),) + tuple(inner)
        v.nbytes = int(n) * per * self.dtype.itemsize
        v.ptr = self.offset(int(r0) * per)
        v._owner = self
        return v

    def free(self):
        if self.ptr and getattr(self, "_owner", None) is None:
            lib().sc_free(self.ctx.h, self.ptr)
        self.ptr = _P()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def gen_rows(ctx, mix, stream, row0, n, dim, out=None):
    out = out if out is not None else Dev(ctx, (n, dim), np.float32)
    _check("sc_gen_rows", lib().sc_gen_rows(ctx.h, C.byref(mix), stream, row0, n, dim, out.ptr))
    return out


def encode(ctx, x, cq, pq, mc, kc, mp, kp, nc, prec=PREC_FP32):
    """x, cq, pq: Dev buffers.  Returns (cid Dev, codes Dev, EncodeStats)."""
    n, dim = x.shape
    cid = Dev(ctx, (n, mc * nc), np.uint16)
    codes = Dev(ctx, (n, mp), np.uint8)
    st = EncodeStats()
    _check("sc_encode", lib().sc_encode(ctx.h, x.ptr, n, dim, cq.ptr, mc, kc, pq.ptr, mp, kp,
                                        nc, prec, cid.ptr, codes.ptr, C.byref(st)))
    return cid, codes, st


def encode_stream(ctx, fill, n, dim, cq, pq, mc, kc, mp, kp, nc, prec=PREC_FP32,
                  chunk=1_000_000):
    """Encodes n rows produced a chunk at a time: fill(row0, m, buf) writes
    rows [row0, row0 + m) into buf (Dev [chunk][dim] float32); only one chunk
    of rows is ever resident.  Returns (cid Dev [n][mc*nc], codes Dev
    [n][mp], codewords re-checked exactly, (row, block) pairs evaluated over
    the whole codebook)."""
    cid = Dev(ctx, (n, mc * nc), np.uint16)
    codes = Dev(ctx, (n, mp), np.uint8)
    buf = Dev(ctx, (max(1, min(chunk, n)), dim), np.float32)
    st = EncodeStats()
    rech = exact = 0
    for r0 in range(0, n, chunk):
        m = min(chunk, n - r0)
        fill(r0, m, buf)
        _check("sc_encode", lib().sc_encode(ctx.h, buf.ptr, m, dim, cq.ptr, mc, kc, pq.ptr, mp,
                                            kp, nc, prec, cid.offset(r0 * mc * nc),
                                            codes.offset(r0 * mp), C.byref(st)))
        rech += st.rechecked
        exact += st.fallback
    buf.free()
    return cid, codes, rech, exact


class Index:
    def __init__(self, ctx, handle, dim, mc, kc, mp, kp, nc):
        self.ctx, self.h = ctx, handle
        self.dim, self.mc, self.kc, self.mp, self.kp, self.nc = dim, mc, kc, mp, kp, nc

    @classmethod
    def build(cls, ctx, cid, codes, dim, cq, mc, kc, pq, mp, kp, nc):
        h = _P()
        n = codes.shape[0]
        _check("sc_index_build", lib().sc_index_build(ctx.h, cid.ptr, codes.ptr, n, dim, cq.ptr,
                                                      mc, kc, pq.ptr, mp, kp, nc, C.byref(h)))
        return cls(ctx, h, dim, mc, kc, mp, kp, nc)

    @classmethod
    def load(cls, ctx, L, pid, codes, dim, cq, mc, kc, pq, mp, kp, nc):
        h = _P()
        L = np.ascontiguousarray(L, np.int32)
        pid = np.ascontiguousarray(pid, np.int32)
        codes = np.ascontiguousarray(codes, np.uint8)
        cq = np.ascontiguousarray(cq, np.float32)
        pq = np.ascontiguousarray(pq, np.float32)
        _check("sc_index_load", lib().sc_index_load(
            ctx.h, L.ctypes.data, L.size, pid.ctypes.data, codes.ctypes.data, pid.size, dim,
            cq.ctypes.data, mc, kc, pq.ctypes.data, mp, kp, nc, C.byref(h)))
        return cls(ctx, h, dim, mc, kc, mp, kp, nc)

    @classmethod
    def load_sparse(cls, ctx, ckeys, L, pid, codes, dim, cq, kc, pq, mp, kp):
        """A two-halves index from host arrays (non-empty cells only)."""
        h = _P()
        ckeys = np.ascontiguousarray(ckeys, np.uint64)
        L = np.ascontiguousarray(L, np.int32)
        pid = np.ascontiguousarray(pid, np.int32)
        codes = np.ascontiguousarray(codes, np.uint8)
        cq = np.ascontiguousarray(cq, np.float32)
        pq = np.ascontiguousarray(pq, np.float32)
        _check("sc_index_load_sparse", lib().sc_index_load_sparse(
            ctx.h, ckeys.ctypes.data, L.ctypes.data, L.size, pid.ctypes.data, codes.ctypes.data,
            pid.size, dim, cq.ctypes.data, kc, pq.ctypes.data, mp, kp, C.byref(h)))
        return cls(ctx, h, dim, 2, kc, mp, kp, 2)

    def cell_keys(self):
        """Keys of the sparse directory (two-halves index), ascending."""
        ncells = int(self.shape()[1])
        out = np.zeros(max(ncells, 1), np.uint64)
        _check("sc_index_cell_keys", lib().sc_index_cell_keys(self.ctx.h, self.h,
                                                              out.ctypes.data))
        return out[:ncells]

    def attach_raw(self, raw):
        """Raw base vectors (Dev [n][dim], by id) for FORM_REAL searches; the
        index keeps a reference to `raw`."""
        _check("sc_index_attach_raw", lib().sc_index_attach_raw(
            self.h, raw.ptr if raw is not None else None, raw.shape[0] if raw is not None else 0))
        self.raw = raw

    def shape(self):
        out = np.zeros(6, np.int64)
        _check("sc_index_shape", lib().sc_index_shape(self.h, out.ctypes.data))
        return out

    def export(self):
        n, ncells, mc, kc, mp, kp = (int(v) for v in self.shape())
        L = np.zeros(ncells, np.int32)
        pid = np.zeros(n, np.int32)
        codes = np.zeros((n, mp), np.uint8)
        nc_ = np.zeros(mc * kc, np.float32)
        nr = np.zeros(mp * kp, np.float32)
        dcr = np.zeros(kc * mp * kp, np.float32)
        _check("sc_index_export", lib().sc_index_export(
            self.ctx.h, self.h, L.ctypes.data, pid.ctypes.data, codes.ctypes.data,
            nc_.ctypes.data, nr.ctypes.data, dcr.ctypes.data))
        return dict(L=L, pid=pid, codes=codes, norm_c=nc_, norm_r=nr, dot_cr=dcr)

    def search(self, q, kind, R, w, T, M=0, form=FORM_PLAIN, ids=None, dists=None, sums=None,
               stats=True, budgets=None):
        """q: Dev [nq][dim]; budgets: Dev [nq] int32 of per-query candidate
        budgets (T is then the largest) or None.  Returns (ids Dev, dists Dev,
        sums Dev, SearchStats|None)."""
        nq = q.shape[0]
        ids = ids if ids is not None else Dev(self.ctx, (nq, R), np.int32)
        dists = dists if dists is not None else Dev(self.ctx, (nq, R), np.float32)
        sums = sums if sums is not None else Dev(self.ctx, (nq,), np.int32)
        p = SearchParams(kind, form, R, w, T, M or self.kc)
        st = SearchStats()
        if budgets is None:
            _check("sc_search", lib().sc_search(self.ctx.h, self.h, q.ptr, nq, C.byref(p), ids.ptr,
                                                dists.ptr, sums.ptr,
                                                C.byref(st) if stats else None))
        else:
            _check("sc_search_budgets", lib().sc_search_budgets(
                self.ctx.h, self.h, q.ptr, nq, C.byref(p), budgets.ptr, ids.ptr, dists.ptr,
                sums.ptr, C.byref(st) if stats else None))
        return ids, dists, sums, (st if stats else None)

    def close(self):
        if self.h:
            lib().sc_index_destroy(self.h)
            self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Plan:
    """A batched search captured in a hipGraph (sc_plan_*): fixed queries and
    output buffers, replayed by execute()."""

    def __init__(self, index, q, kind, R, w, T, M=0, form=FORM_PLAIN, budgets=None):
        """budgets: Dev [nq] int32 of per-query candidate budgets (T is then the
        largest), or None."""
        self.index, self.q, self.budgets = index, q, budgets
        nq = q.shape[0]
        ctx = index.ctx
        self.ids = Dev(ctx, (nq, R), np.int32)
        self.dists = Dev(ctx, (nq, R), np.float32)
        self.sums = Dev(ctx, (nq,), np.int32)
        self.params = SearchParams(kind, form, R, w, T, M or index.kc)
        self.h = _P()
        if budgets is None:
            _check("sc_plan_create", lib().sc_plan_create(
                ctx.h, index.h, C.byref(self.params), q.ptr, nq, self.ids.ptr, self.dists.ptr,
                self.sums.ptr, C.byref(self.h)))
        else:
            _check("sc_plan_create_budgets", lib().sc_plan_create_budgets(
                ctx.h, index.h, C.byref(self.params), q.ptr, budgets.ptr, nq, self.ids.ptr,
                self.dists.ptr, self.sums.ptr, C.byref(self.h)))

    def execute(self, stats=True):
        st = SearchStats()
        _check("sc_plan_execute", lib().sc_plan_execute(self.h, C.byref(st) if stats else None))
        return st if stats else None

    def autotune(self):
        out = np.zeros(4, np.int32)
        _check("sc_plan_autotune", lib().sc_plan_autotune(self.h, out.ctypes.data))
        return {"threads": int(out[0]), "grid": int(out[1]), "us": int(out[2]),
                "variant": int(out[3])}

    def close(self):
        if self.h:
            lib().sc_plan_destroy(self.h)
            self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def encode_general(ctx, x, cq, pq, mc, kc, mp, kp, nc):
    """SCEncoder::encode with general nc (1..3): cid Dev [n][mc*nc]
    (block-major), codes Dev [n][mp]."""
    n, dim = x.shape
    cid = Dev(ctx, (n, mc * nc), np.uint16)
    codes = Dev(ctx, (n, mp), np.uint8)
    _check("sc_encode_general", lib().sc_encode_general(ctx.h, x.ptr, n, dim, cq.ptr, mc, kc,
                                                        pq.ptr, mp, kp, nc, cid.ptr, codes.ptr))
    return cid, codes


def encode_given(ctx, x, cid, cq, pq, mc, kc, mp, kp, rows=None):
    """Encoder::rencode: PQ codes (Dev [n][mp]) of raw rows `rows` (Dev int32
    [n], or None: every row of x) against the given cells cid (Dev [n][mc])."""
    nx, dim = x.shape
    n = rows.shape[0] if rows is not None else nx
    codes = Dev(ctx, (n, mp), np.uint8)
    _check("sc_encode_given", lib().sc_encode_given(ctx.h, x.ptr, nx, dim,
                                                    rows.ptr if rows is not None else None, n,
                                                    cid.ptr, cq.ptr, mc, kc, pq.ptr, mp, kp,
                                                    codes.ptr))
    return codes


def groundtruth_nn1(ctx, base, q):
    """Nearest base vector of each query (Dev [nq] int32) by the exact S1
    distance, lowest id on ties, and that distance (Dev [nq] float64)."""
    nq, dim = q.shape
    ids = Dev(ctx, (nq,), np.int32)
    dists = Dev(ctx, (nq,), np.float64)
    _check("sc_groundtruth_nn1", lib().sc_groundtruth_nn1(ctx.h, base.ptr, base.shape[0], q.ptr,
                                                          nq, dim, ids.ptr, dists.ptr))
    return ids, dists


def kmeans(ctx, x, k, iters, seed, lagged=0, prec=PREC_FP32, want_labels=True):
    n, d = x.shape
    centers = Dev(ctx, (k, d), np.float32)
    labels = Dev(ctx, (n,), np.int32) if want_labels else None
    _check("sc_kmeans", lib().sc_kmeans(ctx.h, x.ptr, n, d, k, iters, seed, lagged, prec,
                                        centers.ptr, labels.ptr if labels else None))
    return centers, labels


def residual(ctx, x, cq, mc, kc, which):
    n, dim = x.shape
    _check("sc_residual", lib().sc_residual(ctx.h, x.ptr, n, dim, cq.ptr, mc, kc, which))
    return x
