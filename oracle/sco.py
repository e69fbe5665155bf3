"""ctypes binding of the C oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this module.  It loads oracle/liboracle.so (built by oracle/Makefile) and, when
present, oracle/_ref/libscref.so (the reference's own encoder/heap sources).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libscref.so")

IVFADC, IMI, SC, SC3, SCIMI = 0, 1, 2, 3, 4

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u16p = np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")


class Mix(C.Structure):
    _fields_ = [("seed", C.c_uint32), ("ncomp", C.c_int32),
                ("frac_bits", C.c_int32), ("out_bits", C.c_int32),
                ("mu_lo", C.c_int64), ("mu_span", C.c_int64),
                ("sig_lo", C.c_int64), ("sig_span", C.c_int64),
                ("maxq", C.c_int64)]


class Index(C.Structure):
    _fields_ = [("dim", C.c_int32), ("mc", C.c_int32), ("kc", C.c_int32),
                ("mp", C.c_int32), ("kp", C.c_int32),
                ("cq", C.c_void_p), ("pq", C.c_void_p), ("norm_c", C.c_void_p),
                ("norm_r", C.c_void_p), ("dot_cr", C.c_void_p),
                ("ncells", C.c_int64), ("L", C.c_void_p), ("pid", C.c_void_p),
                ("codes", C.c_void_p), ("ckeys", C.c_void_p)]


class Stats(C.Structure):
    _fields_ = [("sum", C.c_int32), ("ncells", C.c_int32),
                ("nvisited", C.c_int32), ("empty", C.c_int32),
                ("exhausted", C.c_int32)]


class Heap4(C.Structure):
    _fields_ = [("v", C.c_void_p), ("h1", C.c_void_p), ("h2", C.c_void_p),
                ("h3", C.c_void_p), ("h4", C.c_void_p), ("pos", C.c_int32),
                ("cap", C.c_int32)]


class Heap1(C.Structure):
    _fields_ = [("v", C.c_void_p), ("id", C.c_void_p), ("pos", C.c_int32),
                ("cap", C.c_int32)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("oracle/liboracle.so missing: run `make -C oracle`")
    lib = C.CDLL(LIB_PATH)
    lib.sco_philox.argtypes = [_u32p, _u32p, _u32p]
    lib.sco_gen_rows.argtypes = [C.POINTER(Mix), C.c_uint32, C.c_int64, C.c_int64,
                                 C.c_int32, _f32p, C.c_int32]
    lib.sco_dist_l2sq.argtypes = [_f32p, _f32p, C.c_int32]
    lib.sco_dist_l2sq.restype = C.c_double
    lib.sco_encode.argtypes = [_f32p, C.c_int64, C.c_int32, _f32p, C.c_int32,
                               C.c_int32, _f32p, C.c_int32, C.c_int32, C.c_int32,
                               _u16p, _u8p, C.c_int32]
    lib.sco_encode_general.argtypes = lib.sco_encode.argtypes
    lib.sco_cell_keys.argtypes = [_u16p, C.c_int64, C.c_int32, C.c_int32,
                                  C.c_int32, _u64p]
    lib.sco_build_dense.argtypes = [_u64p, C.c_int64, C.c_int64, _u8p, C.c_int32,
                                    _i32p, _i32p, _u8p]
    lib.sco_build_sparse.argtypes = [_u64p, C.c_int64, _u8p, C.c_int32, _u64p, _i32p, _i32p,
                                     _u8p]
    lib.sco_build_sparse.restype = C.c_int64
    lib.sco_tables.argtypes = [_f32p, _f32p] + [C.c_int32] * 5 + [_f32p] * 3
    lib.sco_lut.argtypes = [_f32p] * 5 + [C.c_int32] * 5 + [_f32p] * 2
    lib.sco_search.argtypes = [C.POINTER(Index), C.c_int32, _f32p, C.c_int32,
                               C.c_int32, C.c_int32, C.c_int32, _i32p, _f32p,
                               C.POINTER(Stats), C.c_void_p, C.c_void_p,
                               C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]
    lib.sco_search_batch.argtypes = [C.POINTER(Index), C.c_int32, _f32p, C.c_int64,
                                     C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                     _i32p, _f32p, C.c_void_p, C.c_int32]
    lib.sco_scimi_set_select_min.argtypes = [C.c_int64]
    lib.sco_search_real.argtypes = [C.POINTER(Index), C.c_int32, _f32p, _f32p, C.c_int32,
                                    C.c_int32, C.c_int32, _i32p, _f32p, C.c_void_p]
    lib.sco_heap4_insert.argtypes = [C.POINTER(Heap4), C.c_float] + [C.c_int32] * 4
    lib.sco_heap4_pop.argtypes = [C.POINTER(Heap4)] + [C.POINTER(C.c_int32)] * 4
    lib.sco_heap1_insert.argtypes = [C.POINTER(Heap1), C.c_float, C.c_int32]
    lib.sco_heap1_pop.argtypes = [C.POINTER(Heap1)]
    lib.sco_heap1_pop.restype = C.c_int32
    lib.sco_topr.argtypes = [_f32p, C.c_int32, C.c_int32, _i32p]
    lib.sco_topr.restype = C.c_int32
    lib.sco_bruteforce_nn1.argtypes = [_f32p, C.c_int64, _f32p, C.c_int64,
                                       C.c_int32, _i32p, C.c_int32]
    lib.sco_recall.argtypes = [_i32p, _i32p, C.c_int64, C.c_int32, C.c_int32]
    lib.sco_recall.restype = C.c_double
    lib.sco_kmeans.argtypes = [_f32p, C.c_int64, C.c_int32, C.c_int32, C.c_int32,
                               C.c_uint32, C.c_int32, _f32p, _i32p, C.c_int32]
    lib.sco_kmeans_pp.argtypes = [_f32p, C.c_int64, C.c_int32, C.c_int32,
                                  C.c_uint32, C.c_int32, _i64p, C.c_int32]
    lib.sco_residual.argtypes = [_f32p, C.c_int64, C.c_int32, _f32p, C.c_int32,
                                 C.c_int32, C.c_int32, C.c_int32]
    return lib


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = _load()
    return _LIB


def nthreads():
    return max(1, min(16, os.cpu_count() or 1))


# ---------------------------------------------------------------- data
# Seeded mixture specs (fixed here; the GPU generator shares the layout).
SEED = 20151001


def mix_d128(seed=SEED, ncomp=64):
    """D=128 integer data in [0,255]: mu ~ U[24,232], sigma ~ U[8,32]."""
    return Mix(seed, ncomp, 8, 0, 24 * 256, 208 * 256, 8 * 256, 24 * 256, 255)


def mix_d960(seed=SEED, ncomp=32):
    """D=960 data on the 2^-16 grid in [0,1]: mu ~ U[0.2,0.8], sigma 0.1."""
    return Mix(seed, ncomp, 16, 16, 13107, 39321, 6554, 0, 65536)


STREAM_TRAIN, STREAM_BASE, STREAM_QUERY = 1, 2, 3


def gen(mix, stream, row0, n, dim):
    out = np.empty((n, dim), np.float32)
    lib().sco_gen_rows(C.byref(mix), stream, row0, n, dim, out, nthreads())
    return out


def philox(ctr, key):
    out = np.zeros(4, np.uint32)
    lib().sco_philox(np.asarray(ctr, np.uint32), np.asarray(key, np.uint32), out)
    return out


def dist(x, y):
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    return lib().sco_dist_l2sq(x, y, x.size)


def encode(x, cq, pq, mc, kc, mp, kp, nc):
    x = np.ascontiguousarray(x, np.float32)
    n, dim = x.shape
    cid = np.zeros((n, mc * nc), np.uint16)
    codes = np.zeros((n, mp), np.uint8)
    rc = lib().sco_encode(x, n, dim, np.ascontiguousarray(cq, np.float32), mc, kc,
                          np.ascontiguousarray(pq, np.float32), mp, kp, nc, cid,
                          codes, nthreads())
    if rc:
        raise ValueError("sco_encode rc=%d" % rc)
    return cid, codes


def encode_given(x, rows, cid, cq, pq, mc, kc, mp, kp):
    """Encoder::rencode<T> (include/encoder.h:202-309): position i takes raw
    row rows[i] of x, subtracts the centroid of its given cell cid[i][j] in
    each coarse block (float subtraction, encoder.h:271-285) and is coded by
    Encoder::encode's PQ argmin (encoder.h:287-303: S1 distance, first strict
    minimum) -- here sco_encode over the residuals with one all-zero coarse
    centroid (mc=1, kc=1), which leaves them unchanged (r - 0 = r)."""
    x = np.ascontiguousarray(x, np.float32)
    dim = x.shape[1]
    bsc = dim // mc
    rows = np.asarray(rows, np.int64)
    cid = np.asarray(cid, np.int64).reshape(len(rows), mc)
    cq = np.ascontiguousarray(cq, np.float32).reshape(mc, kc, bsc)
    res = x[rows].copy()
    for j in range(mc):
        res[:, j * bsc:(j + 1) * bsc] -= cq[j][cid[:, j]]
    _, codes = encode(res, np.zeros((1, 1, dim), np.float32), pq, 1, 1, mp, kp, 1)
    return codes


def encode_general(x, cq, pq, mc, kc, mp, kp, nc):
    """SCEncoder::encode (general nc): cid [n][mc*nc] (block-major)."""
    x = np.ascontiguousarray(x, np.float32)
    n, dim = x.shape
    cid = np.zeros((n, mc * nc), np.uint16)
    codes = np.zeros((n, mp), np.uint8)
    rc = lib().sco_encode_general(x, n, dim, np.ascontiguousarray(cq, np.float32), mc, kc,
                                  np.ascontiguousarray(pq, np.float32), mp, kp, nc, cid,
                                  codes, nthreads())
    if rc:
        raise ValueError("sco_encode_general rc=%d" % rc)
    return cid, codes


def cell_keys(cid, mc, kc, nc):
    cid = np.ascontiguousarray(cid, np.uint16)
    keys = np.zeros(cid.shape[0], np.uint64)
    lib().sco_cell_keys(cid, cid.shape[0], mc, kc, nc, keys)
    return keys


def build_dense(keys, ncells, codes):
    n, mp = codes.shape
    L = np.zeros(ncells, np.int32)
    pid = np.zeros(n, np.int32)
    cs = np.zeros((n, mp), np.uint8)
    rc = lib().sco_build_dense(np.ascontiguousarray(keys, np.uint64), n, ncells,
                               np.ascontiguousarray(codes), mp, L, pid, cs)
    if rc:
        raise ValueError("sco_build_dense rc=%d" % rc)
    return L, pid, cs


def build_sparse(keys, codes):
    """Non-empty cells only: (ckeys ascending, L inclusive prefix, pid, codes
    in cell order)."""
    n, mp = codes.shape
    ck = np.zeros(max(n, 1), np.uint64)
    L = np.zeros(max(n, 1), np.int32)
    pid = np.zeros(n, np.int32)
    cs = np.zeros((n, mp), np.uint8)
    nc = lib().sco_build_sparse(np.ascontiguousarray(keys, np.uint64), n,
                                np.ascontiguousarray(codes), mp, ck, L, pid, cs)
    if nc < 0:
        raise ValueError("sco_build_sparse rc=%d" % nc)
    return ck[:nc].copy(), L[:nc].copy(), pid, cs


def tables(cq, pq, dim, mc, kc, mp, kp):
    nc_ = np.zeros(mc * kc, np.float32)
    nr = np.zeros(mp * kp, np.float32)
    dcr = np.zeros(kc * mp * kp, np.float32)
    lib().sco_tables(np.ascontiguousarray(cq, np.float32),
                     np.ascontiguousarray(pq, np.float32), dim, mc, kc, mp, kp,
                     nc_, nr, dcr)
    return nc_, nr, dcr


def lut(q, cq, pq, norm_c, norm_r, dim, mc, kc, mp, kp):
    dqc = np.zeros(mc * kc, np.float32)
    dqr = np.zeros(mp * kp, np.float32)
    lib().sco_lut(np.ascontiguousarray(q, np.float32), cq, pq, norm_c, norm_r,
                  dim, mc, kc, mp, kp, dqc, dqr)
    return dqc, dqr


class OracleIndex:
    """Dense-cell index held in numpy arrays (kept alive with the struct)."""

    def __init__(self, dim, mc, kc, mp, kp, cq, pq, L, pid, codes_sorted, ckeys=None):
        """ckeys: the sparse directory's keys (one per entry of L), or None for
        dense cells."""
        self.dim, self.mc, self.kc, self.mp, self.kp = dim, mc, kc, mp, kp
        self.ckeys = None if ckeys is None else np.ascontiguousarray(ckeys, np.uint64)
        self.cq = np.ascontiguousarray(cq, np.float32)
        self.pq = np.ascontiguousarray(pq, np.float32)
        self.norm_c, self.norm_r, self.dot_cr = tables(self.cq, self.pq, dim, mc, kc, mp, kp)
        self.L = np.ascontiguousarray(L, np.int32)
        self.pid = np.ascontiguousarray(pid, np.int32)
        self.codes = np.ascontiguousarray(codes_sorted, np.uint8)
        p = lambda a: a.ctypes.data
        self.s = Index(dim, mc, kc, mp, kp, p(self.cq), p(self.pq), p(self.norm_c),
                       p(self.norm_r), p(self.dot_cr), self.L.size, p(self.L),
                       p(self.pid), p(self.codes),
                       None if self.ckeys is None else p(self.ckeys))

    def search(self, kind, q, R, w, T, M=None, trace=False):
        M = self.kc if M is None else M
        ids = np.zeros(R, np.int32)
        dists = np.zeros(R, np.float32)
        st = Stats()
        q = np.ascontiguousarray(q, np.float32)
        extra = {}
        if trace:
            cells = np.full(w, -1, np.int32)
            s1 = np.full(w, -1, np.int32)
            s2 = np.full(w, -1, np.int32)
            cap = int(self.pid.size)
            aid = np.full(cap, -1, np.int32)
            ad = np.zeros(cap, np.float32)
            rc = lib().sco_search(C.byref(self.s), kind, q, R, w, T, M, ids, dists,
                                  C.byref(st), cells.ctypes.data, s1.ctypes.data,
                                  s2.ctypes.data, aid.ctypes.data, ad.ctypes.data, cap)
            n = st.ncells
            extra = dict(cells=cells[:n], s1=s1[:n], s2=s2[:n],
                         all_ids=aid[:min(st.sum, cap)], all_dists=ad[:min(st.sum, cap)])
        else:
            rc = lib().sco_search(C.byref(self.s), kind, q, R, w, T, M, ids, dists,
                                  C.byref(st), None, None, None, None, None, 0)
        if rc:
            raise ValueError("sco_search rc=%d" % rc)
        return ids, dists, st, extra

    def search_real(self, kind, q, raw, R, w, T):
        """The search with real_dist = true: candidates ranked by their S1
        distance to the raw vectors raw [n][dim] (by id)."""
        ids = np.zeros(R, np.int32)
        dists = np.zeros(R, np.float32)
        st = Stats()
        rc = lib().sco_search_real(C.byref(self.s), kind, np.ascontiguousarray(q, np.float32),
                                   np.ascontiguousarray(raw, np.float32), R, w, T, ids, dists,
                                   C.cast(C.pointer(st), C.c_void_p))
        if rc:
            raise ValueError("sco_search_real rc=%d" % rc)
        return ids, dists, st

    def search_batch(self, kind, Q, R, w, T, M=None, threads=None):
        M = self.kc if M is None else M
        Q = np.ascontiguousarray(Q, np.float32)
        nq = Q.shape[0]
        ids = np.zeros((nq, R), np.int32)
        dists = np.zeros((nq, R), np.float32)
        st = (Stats * nq)()
        rc = lib().sco_search_batch(C.byref(self.s), kind, Q, nq, R, w, T, M, ids,
                                    dists, C.cast(st, C.c_void_p),
                                    threads or nthreads())
        if rc:
            raise ValueError("sco_search_batch rc=%d" % rc)
        sums = np.array([s.sum for s in st], np.int32)
        return ids, dists, sums


def topr(d, R):
    d = np.ascontiguousarray(d, np.float32)
    out = np.zeros(R, np.int32)
    m = lib().sco_topr(d, d.size, R, out)
    return out[:m]


def bruteforce_nn1(base, q):
    base = np.ascontiguousarray(base, np.float32)
    q = np.ascontiguousarray(q, np.float32)
    gt = np.zeros(q.shape[0], np.int32)
    lib().sco_bruteforce_nn1(base, base.shape[0], q, q.shape[0], base.shape[1], gt,
                             nthreads())
    return gt


def recall(gt, ids, R):
    ids = np.ascontiguousarray(ids, np.int32)
    return lib().sco_recall(np.ascontiguousarray(gt, np.int32), ids, ids.shape[0], R,
                            ids.shape[1])


def kmeans(x, k, iters, seed, lagged=0):
    x = np.ascontiguousarray(x, np.float32)
    n, d = x.shape
    centers = np.zeros((k, d), np.float32)
    labels = np.zeros(max(n, k), np.int32)
    rc = lib().sco_kmeans(x, n, d, k, iters, seed, lagged, centers, labels, nthreads())
    if rc:
        raise ValueError("sco_kmeans rc=%d" % rc)
    return centers, labels[:n]


def kmeans_pp(x, k, seed, lagged=0):
    x = np.ascontiguousarray(x, np.float32)
    idx = np.zeros(k, np.int64)
    rc = lib().sco_kmeans_pp(x, x.shape[0], x.shape[1], k, seed, lagged, idx, nthreads())
    if rc:
        raise ValueError("sco_kmeans_pp rc=%d" % rc)
    return idx


def residual(x, cq, mc, kc, which):
    x = np.array(x, np.float32, copy=True, order="C")
    lib().sco_residual(x, x.shape[0], x.shape[1], np.ascontiguousarray(cq, np.float32),
                       mc, kc, which, nthreads())
    return x


# ---------------------------------------------------------------- reference
_REF = None


def ref():
    """The reference's own encoder/heap code (oracle/_ref), or None."""
    global _REF
    if _REF is None and os.path.exists(REF_PATH):
        r = C.CDLL(REF_PATH)
        r.scref_dist.argtypes = [_f32p, _f32p, C.c_int]
        r.scref_dist.restype = C.c_double
        r.scref_vector_base.argtypes = [_u16p, C.c_int, C.c_int]
        r.scref_encode_imi.argtypes = [C.c_char_p] * 3 + [_u16p, _u8p, _i32p, _i32p,
                                                          _u8p, C.c_int64, C.c_int64]
        r.scref_heap4_insert.argtypes = [_f32p, _i32p, _i32p, _i32p, _i32p, C.c_float,
                                         C.c_int, C.c_int, C.c_int, C.c_int,
                                         C.POINTER(C.c_int32), C.c_int]
        r.scref_heap4_pop.argtypes = [_f32p, _i32p, _i32p, _i32p, _i32p, _i32p,
                                      C.POINTER(C.c_int32)]
        r.scref_heap1_insert.argtypes = [_f32p, _i32p, C.c_float, C.c_int,
                                         C.POINTER(C.c_int32), C.c_int]
        r.scref_heap1_pop.argtypes = [_f32p, _i32p, C.POINTER(C.c_int32)]
        r.scref_heap1_pop.restype = C.c_int
        r.scref_heap6_insert.argtypes = [_f32p] + [_i32p] * 6 + [C.c_float] + [C.c_int] * 6 + \
            [C.POINTER(C.c_int32), C.c_int]
        r.scref_heap6_pop.argtypes = [_f32p] + [_i32p] * 6 + [_i32p, C.POINTER(C.c_int32)]
        r.scref_residual.argtypes = [_f32p, C.c_int, C.c_int, C.c_int, C.c_char_p, C.c_int]
        _REF = r
    return _REF


def write_fvecs(path, x):
    x = np.ascontiguousarray(x, np.float32)
    n, d = x.shape
    rows = np.empty((n, d + 1), np.float32)
    rows[:, 0] = np.frombuffer(np.int32(d).tobytes(), np.float32)[0]
    rows[:, 1:] = x
    rows.tofile(path)


def write_codebook(path, cb, k, m, dim):
    """`.ctr_` layout: float header [k, m, dim] then k*dim floats
    (include/quantizer.h:370-431 writer, include/sc_utilities.h:896-980 reader)."""
    hdr = np.array([k, m, dim], np.float32)
    np.concatenate([hdr, np.ascontiguousarray(cb, np.float32).ravel()]).tofile(path)
