"""ctypes access to the CPU oracle (oracle/liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and the benchmark's CPU-baseline leg import this.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
_lib = None

_vp = ctypes.c_void_p
_i = ctypes.c_int
_i64 = ctypes.c_int64
_u64 = ctypes.c_uint64

_SIGS = {
    "orc_zipf_thresholds": (None, [_u64, _i, ctypes.c_double, _vp]),
    "orc_gen_centres": (None, [_u64, _i, _i, _vp]),
    "orc_gen_rows": (None, [_u64, _u64, _i64, _i64, _i, _vp, _i, _vp, ctypes.c_float, _vp]),
    "orc_cross_dist": (None, [_i, _i, _i, _vp, _i, _vp, _i, _vp, _i]),
    "orc_topk_keys": (None, [_vp, _i64, _i, _vp]),
    "orc_knn": (None, [_i64, _i, _i, _i, _vp, _i, _vp, _i, _vp, _vp, _i]),
    "orc_residual": (None, [_i, _vp, _vp, _vp]),
    "orc_tables": (None, [_i, _i, _i, _vp, _vp, _vp]),
    "orc_adc": (None, [_i, _i, _vp, _vp, _i64, _vp]),
    "orc_search": (_i, [_i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _vp, _vp, _vp, _vp, _i]),
    "orc_search_phases": (_i, [_i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _vp, _vp, _vp, _i]),
    "orc_search_var": (_i, [_i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _i, _vp, _vp, _i, _vp, _vp, _i]),
    "orc_search_residual": (_i, [_i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _vp, _vp, _i]),
    "orc_search_rerank": (_i, [_i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _i, _i, _i, _i,
                               _vp, _vp, _i]),
    "orc_kmeans": (_i, [_vp, _i64, _i, _i, _i, _i, _u64, _vp, _i]),
    "orc_ivfpq_train": (_i, [_vp, _i64, _i, _i, _i, _i, _i, _i, _u64, _vp, _vp, _i]),
    "orc_ivfpq_encode": (_i, [_i, _i, _i, _i, _vp, _vp, _vp, _i64, _vp, _vp, _i]),
    "orc_ivfpq_bucket": (_i, [_i, _i, _vp, _vp, _i64, ctypes.c_int32, _vp, _vp, _vp]),
    "orc_bruteforce_f64": (None, [_vp, _i64, _vp, _i, _i, _i, _vp, _vp, _i]),
    "orc_recall_at": (ctypes.c_double, [_vp, _i, _i, _vp, _i]),
}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.run(["make", "-C", HERE, "liboracle.so"], check=True, capture_output=True)
        L = ctypes.CDLL(LIB)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def P(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


# -- keys (same definition as the C side) ------------------------------------
def fkey(d):
    d = np.asarray(d, dtype=np.float32) + np.float32(0.0)
    b = d.view(np.uint32)
    return np.where(b & np.uint32(0x80000000), ~b, b | np.uint32(0x80000000)).astype(np.uint32)


def ckey(d, ids):
    return (fkey(d).astype(np.uint64) << np.uint64(32)) | np.asarray(ids, dtype=np.int64).astype(np.uint32).astype(np.uint64)


# -- generator ----------------------------------------------------------------
def gen_centres(seed, ncomp, d):
    c = np.empty((ncomp, d), np.float32)
    lib().orc_gen_centres(seed, ncomp, d, P(c))
    return c


def zipf_thresholds(seed, ncomp, s):
    t = np.empty(ncomp, np.uint64)
    lib().orc_zipf_thresholds(seed, ncomp, s, P(t))
    return t


def gen_rows(seed, stream, row0, n, centres, thr=None, sigma=20.0):
    centres = f32(centres)
    ncomp, d = centres.shape
    out = np.empty((n, d), np.float32)
    thr = None if thr is None else np.ascontiguousarray(thr, dtype=np.uint64)
    lib().orc_gen_rows(seed, stream, row0, n, d, P(centres), ncomp, P(thr), sigma, P(out))
    return out


# -- primitives ------------------------------------------------------------------
def cross_dist(a, b):
    """out[j, i] = dist(a_i, b_j) (a: f32-accumulated norm, b: f64)."""
    a, b = f32(a), f32(b)
    na, d = a.shape
    nb = b.shape[0]
    out = np.empty((nb, na), np.float32)
    lib().orc_cross_dist(d, na, nb, P(a), d, P(b), d, P(out), na)
    return out


def topk_keys(keys, k):
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    out = np.empty(k, np.uint64)
    lib().orc_topk_keys(P(keys), keys.size, k, P(out))
    return out


def knn(mat2, mat1, k, nthreads=0):
    mat2, mat1 = f32(mat2), f32(mat1)
    n2, d = mat2.shape
    n1 = mat1.shape[0]
    idx = np.empty((n1, k), np.int32)
    dis = np.empty((n1, k), np.float32)
    lib().orc_knn(n1, n2, d, k, P(mat2), d, P(mat1), d, P(idx), P(dis), nthreads)
    return idx, dis


def tables(r, pq):
    r, pq = f32(r), f32(pq)
    m, ks, ds = pq.shape
    tab = np.empty((m, ks), np.float32)
    lib().orc_tables(m * ds, m, ks, P(r), P(pq), P(tab))
    return tab


def adc(tab, codes):
    tab = f32(tab)
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    m, ks = tab.shape
    out = np.empty(codes.shape[0], np.float32)
    lib().orc_adc(m, ks, P(tab), P(codes), codes.shape[0], P(out))
    return out


# -- search ------------------------------------------------------------------------
def _ix(index):
    coarse, pq, off, ids, codes = index
    coarse, pq = f32(coarse), f32(pq)
    off = np.ascontiguousarray(off, dtype=np.int64)
    ids = np.ascontiguousarray(ids, dtype=np.int32)
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    nlist, d = coarse.shape
    m, ks, _ = pq.shape
    return d, nlist, m, ks, coarse, pq, off, ids, codes


def search(index, q, w, k, nthreads=0, want_coarse=False):
    d, nlist, m, ks, coarse, pq, off, ids, codes = _ix(index)
    q = f32(q)
    nq = q.shape[0]
    oi = np.empty((nq, k), np.int32)
    od = np.empty((nq, k), np.float32)
    ci = np.empty((nq, w), np.int32) if want_coarse else None
    cd = np.empty((nq, w), np.float32) if want_coarse else None
    rc = lib().orc_search(d, nlist, m, ks, P(coarse), P(pq), P(off), P(ids), P(codes), P(q), nq, w, k,
                          P(oi), P(od), P(ci), P(cd), nthreads)
    assert rc == 0, rc
    return (oi, od, ci, cd) if want_coarse else (oi, od)


def search_phases(index, q, w, k, nthreads=0):
    """orc_search staged as the reference's parallel/ loop; returns
    (ids, dis, seconds per phase {coarse, tables, adc, select} summed over threads)."""
    d, nlist, m, ks, coarse, pq, off, ids, codes = _ix(index)
    q = f32(q)
    nq = q.shape[0]
    oi = np.empty((nq, k), np.int32)
    od = np.empty((nq, k), np.float32)
    ph = np.zeros(4, np.float64)
    rc = lib().orc_search_phases(d, nlist, m, ks, P(coarse), P(pq), P(off), P(ids), P(codes), P(q), nq, w, k,
                                 P(oi), P(od), P(ph), nthreads)
    assert rc == 0, rc
    return oi, od, dict(zip(("coarse", "tables", "adc", "select"), ph.tolist()))


def search_var(index, q, wq, kq, kmax, nthreads=0):
    d, nlist, m, ks, coarse, pq, off, ids, codes = _ix(index)
    q = f32(q)
    nq = q.shape[0]
    wq = np.ascontiguousarray(wq, dtype=np.int32)
    kq = np.ascontiguousarray(kq, dtype=np.int32)
    oi = np.empty((nq, kmax), np.int32)
    od = np.empty((nq, kmax), np.float32)
    rc = lib().orc_search_var(d, nlist, m, ks, P(coarse), P(pq), P(off), P(ids), P(codes), P(q), nq, P(wq),
                              P(kq), kmax, P(oi), P(od), nthreads)
    assert rc == 0, rc
    return oi, od


def search_residual(index, residual, list_of, w, k, nthreads=0):
    d, nlist, m, ks, coarse, pq, off, ids, codes = _ix(index)
    residual = f32(residual)
    list_of = np.ascontiguousarray(list_of, dtype=np.int32)
    nq = residual.shape[0] // w
    oi = np.empty((nq, k), np.int32)
    od = np.empty((nq, k), np.float32)
    rc = lib().orc_search_residual(d, m, ks, P(pq), P(off), P(ids), P(codes), P(residual), P(list_of), nq, w,
                                   k, P(oi), P(od), nthreads)
    assert rc == 0, rc
    return oi, od


def search_rerank(index, base, q, w, kl, k, nthreads=0):
    d, nlist, m, ks, coarse, pq, off, ids, codes = _ix(index)
    q, base = f32(q), f32(base)
    nq = q.shape[0]
    oi = np.empty((nq, k), np.int32)
    od = np.empty((nq, k), np.float32)
    rc = lib().orc_search_rerank(d, nlist, m, ks, P(coarse), P(pq), P(off), P(ids), P(codes), P(base),
                                 base.shape[0], P(q), nq, w, kl, k, P(oi), P(od), nthreads)
    assert rc == 0, rc
    return oi, od


# -- build --------------------------------------------------------------------------
def kmeans(x, k, niter, seed, nthreads=0):
    x = f32(x)
    n, d = x.shape
    cent = np.empty((k, d), np.float32)
    rc = lib().orc_kmeans(P(x), n, d, d, k, niter, seed, P(cent), nthreads)
    assert rc == 0, rc
    return cent


def train(x, nlist, m, ks=256, niter_coarse=20, niter_pq=25, seed=1234, nthreads=0):
    x = f32(x)
    n, d = x.shape
    coarse = np.empty((nlist, d), np.float32)
    pq = np.empty((m, ks, d // m), np.float32)
    rc = lib().orc_ivfpq_train(P(x), n, d, nlist, m, ks, niter_coarse, niter_pq, seed, P(coarse), P(pq),
                               nthreads)
    assert rc == 0, rc
    return coarse, pq


def encode(coarse, pq, x, nthreads=0):
    coarse, pq, x = f32(coarse), f32(pq), f32(x)
    nlist, d = coarse.shape
    m, ks, _ = pq.shape
    n = x.shape[0]
    assign = np.empty(n, np.int32)
    codes = np.empty((n, m), np.uint8)
    rc = lib().orc_ivfpq_encode(d, nlist, m, ks, P(coarse), P(pq), P(x), n, P(assign), P(codes), nthreads)
    assert rc == 0, rc
    return assign, codes


def bucket(nlist, assign, codes, id0=0):
    assign = np.ascontiguousarray(assign, dtype=np.int32)
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    n, m = codes.shape
    off = np.empty(nlist + 1, np.int64)
    ids = np.empty(n, np.int32)
    co = np.empty((n, m), np.uint8)
    rc = lib().orc_ivfpq_bucket(nlist, m, P(assign), P(codes), n, id0, P(off), P(ids), P(co))
    assert rc == 0, rc
    return off, ids, co


def build_index(coarse, pq, base, id0=0, nthreads=0):
    a, c = encode(coarse, pq, base, nthreads)
    off, ids, co = bucket(coarse.shape[0], a, c, id0)
    return (coarse, pq, off, ids, co)


def bruteforce_f64(base, q, k=1, nthreads=0):
    base, q = f32(base), f32(q)
    nb, d = base.shape
    nq = q.shape[0]
    ids = np.empty((nq, k), np.int32)
    dis = np.empty((nq, k), np.float64)
    lib().orc_bruteforce_f64(P(base), nb, P(q), nq, d, k, P(ids), P(dis), nthreads)
    return ids, dis


def recall_at(ids, gt, R):
    ids = np.ascontiguousarray(ids, dtype=np.int32)
    gt = np.ascontiguousarray(gt, dtype=np.int32)
    nq, k = ids.shape
    return lib().orc_recall_at(P(ids), nq, k, P(gt), R)
