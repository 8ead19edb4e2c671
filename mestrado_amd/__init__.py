"""mestrado_amd -- MI355X (gfx950) IVF-PQ search engine.

Python access to libmivf.so, the HIP/C++ library behind the C ABI in
include/mivf.h.  The library is the product; this module only loads it with
ctypes and moves numpy arrays in and out of device memory.  Nothing here
computes search results: when the library is missing, every call raises.
"""
import ctypes
import os

import numpy as np

__all__ = ["lib", "MivfError", "DeviceArray", "Index", "Plan", "Server", "serve_bench", "train", "kmeans",
           "synth_centres", "zipf_thresholds", "synth_rows", "bruteforce_nn",
           "device_count", "synchronize", "LIB_PATH"]

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmivf.so")
_lib = None

_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_f32p = ctypes.POINTER(ctypes.c_float)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_vp = ctypes.c_void_p
_i = ctypes.c_int
_i64 = ctypes.c_int64
_u64 = ctypes.c_uint64

# name -> argtypes (all return int status)
_SIGS = {
    "mivf_version": [],
    "mivf_device_count": [_i32p],
    "mivf_malloc": [ctypes.POINTER(_vp), ctypes.c_size_t],
    "mivf_free": [_vp],
    "mivf_memcpy_htod": [_vp, _vp, ctypes.c_size_t],
    "mivf_memcpy_dtoh": [_vp, _vp, ctypes.c_size_t],
    "mivf_memcpy_dtod": [_vp, _vp, ctypes.c_size_t],
    "mivf_synchronize": [],
    "mivf_train": [_vp, _i64, _i, _i, _i, _i, _i, _i, _u64, _vp, _vp],
    "mivf_train_device": [_vp, _i64, _i, _i, _i, _i, _i, _i, _u64, _vp, _vp],
    "mivf_kmeans": [_vp, _i64, _i, _i, _i, _i, _u64, _vp],
    "mivf_index_create": [_i, _i, _i, _i, _vp, _vp, ctypes.POINTER(_vp)],
    "mivf_index_free": [_vp],
    "mivf_index_info": [_vp, _i32p, _i32p, _i32p, _i32p, _i64p],
    "mivf_index_set_lists": [_vp, _vp, _vp, _vp],
    "mivf_index_set_lists_i32": [_vp, _vp, _vp, _vp],
    "mivf_index_get_lists": [_vp, _vp, _vp, _vp],
    "mivf_index_add": [_vp, _vp, _i64, ctypes.c_int32],
    "mivf_index_add_device": [_vp, _vp, _i64, ctypes.c_int32],
    "mivf_index_encode": [_vp, _vp, _i64, _vp, _vp, _vp],
    "mivf_search": [_vp, _vp, _i, _i, _i, _vp, _vp, _vp],
    "mivf_search_noalloc": [_vp, _vp, _i, _i, _i, _vp, _vp, _vp],
    "mivf_search_var": [_vp, _vp, _i, _vp, _vp, _vp, _vp, _i, _vp],
    "mivf_coarse": [_vp, _vp, _i, _i, _vp, _vp, _vp],
    "mivf_search_residual": [_vp, _vp, _vp, _i, _i, _i, _vp, _vp, _vp],
    "mivf_search_rerank": [_vp, _vp, _i, _i, _i, _i, _vp, _i64, _vp, _vp, _vp],
    "mivf_adc": [_vp, _vp, _vp, _i64, _vp, _vp],
    "mivf_kmin": [_vp, _vp, _i64, _i, _vp, _vp, _vp],
    "mivf_search_fast": [_vp, _vp, _i, _i, _i, _i, _vp, _i64, _vp, _vp, _vp],
    "mivf_fast_stats": [_vp, _i64p, _i64p],
    "mivf_index_attach_base": [_vp, _vp, _i64],
    "mivf_fast_tables": [_vp, _vp, _i, _i, _vp, _vp],
    "mivf_fast_var_adc": [_vp, _vp, _i, _i, _i, _vp],
    "mivf_fast_seed_tau": [_vp, _vp, _i, _i, _i, _vp],
    "mivf_fast_select": [_vp, _vp, _i, _i, _i, _vp],
    "mivf_search_batched": [_vp, _vp, _i, _vp, _vp, _i, _vp, _vp, _vp],
    "mivf_batched_stats": [_vp, _i64p],
    "mivf_plan_create": [_vp, _i, _i, _i, _i, ctypes.POINTER(_vp)],
    "mivf_plan_search": [_vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "mivf_plan_info": [_vp, ctypes.POINTER(_i), ctypes.POINTER(_i), _f32p],
    "mivf_plan_free": [_vp],
    "mivf_exact_tables": [_vp, _vp, _i, _i, _vp],
    "mivf_exact_adc": [_vp, _vp, _i, _i, _vp],
    "mivf_server_create": [_vp, _i, _i, _i64, ctypes.POINTER(_vp)],
    "mivf_server_free": [_vp],
    "mivf_server_submit": [_vp, _vp, _i, _i, _i, _vp, _vp, _i64p],
    "mivf_server_progress": [_vp, ctypes.POINTER(_i)],
    "mivf_server_wait": [_vp, _i64],
    "mivf_server_drain": [_vp],
    "mivf_server_latency": [_vp, _i64, ctypes.POINTER(ctypes.c_double)],
    "mivf_server_stats": [_vp, _i64p, _i64p, _i64p],
    "mivf_server_warm": [_vp, _vp, _i64, _i, _i],
    "mivf_serve_bench": [_vp, _i, _i, _i64, _vp, _i64, ctypes.c_double, _i, _u64, _i, _vp, _i, _vp, _i,
                         ctypes.POINTER(ctypes.c_double)],
    "mivf_set_option": [_vp, ctypes.c_char_p, _i64],
    "mivf_get_timings": [_vp, _f32p, _i],
    "mivf_synth_rows": [_u64, _u64, _i64, _i64, _i, _vp, _i, _vp, ctypes.c_float, _vp, _vp],
    "mivf_synth_centres": [_u64, _i, _i, _vp],
    "mivf_zipf_thresholds": [_u64, _i, ctypes.c_double, _vp],
    "mivf_bruteforce_nn": [_vp, _i64, _vp, _i, _i, _vp, _vp],
    "mivf_lds_read_rate": [_i, _i, ctypes.POINTER(ctypes.c_double)],
    "mivf_write_cent": [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, _i, _i, _i, _i, _i, _vp, _vp],
    "mivf_read_cent_header": [ctypes.c_char_p, _i32p, _i32p, _i32p, _i32p, _i32p],
    "mivf_read_cent": [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, _vp, _vp],
    "mivf_write_ivf_dir": [ctypes.c_char_p, _i, _i, _vp, _vp, _vp],
    "mivf_read_ivf_dir": [ctypes.c_char_p, _i, _i, _vp, _vp, _vp, _i64],
    "mivf_write_ivf_file": [ctypes.c_char_p, _i, _i, _vp, _vp, _vp],
    "mivf_read_ivf_file": [ctypes.c_char_p, _i, _i, _vp, _vp, _vp, _i64],
    "mivf_vecs_info": [ctypes.c_char_p, _i, _i32p, _i64p],
    "mivf_read_fvecs": [ctypes.c_char_p, _i64, _i64, _i, _vp],
    "mivf_read_bvecs": [ctypes.c_char_p, _i64, _i64, _i, _vp],
    "mivf_read_ivecs": [ctypes.c_char_p, _i64, _i64, _i, _vp],
    "mivf_index_add_file": [_vp, ctypes.c_char_p, _i, _i64, _i64, ctypes.c_int32, _i64],
}


class MivfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libmivf error {code}: {msg}")
        self.code = code


def lib():
    """Load libmivf.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m mestrado_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ctypes.c_int
        L.mivf_last_error.argtypes = []
        L.mivf_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise MivfError(rc, lib().mivf_last_error().decode(errors="replace"))


def _ptr(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def device_count():
    n = ctypes.c_int32(0)
    _check(lib().mivf_device_count(ctypes.byref(n)))
    return n.value


def synchronize():
    _check(lib().mivf_synchronize())


class DeviceArray:
    """A device (HBM) buffer with a numpy dtype and shape."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        _check(lib().mivf_malloc(ctypes.byref(p), max(self.nbytes, 1)))
        self.ptr = p

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        if d.nbytes:
            _check(lib().mivf_memcpy_htod(d.ptr, _ptr(a), d.nbytes))
        return d

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            _check(lib().mivf_memcpy_dtoh(_ptr(out), self.ptr, self.nbytes))
        return out

    def head(self, n):
        """Host copy of the first n rows (along the first axis)."""
        n = min(int(n), self.shape[0])
        out = np.empty((n,) + self.shape[1:], self.dtype)
        if out.nbytes:
            _check(lib().mivf_memcpy_dtoh(_ptr(out), self.ptr, out.nbytes))
        return out

    def offset(self, nbytes):
        return ctypes.c_void_p(self.ptr.value + int(nbytes))

    def free(self):
        if self.ptr is not None and _lib is not None:
            lib().mivf_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _dev(x, dtype):
    """Device pointer for x (DeviceArray or numpy array copied to the device)."""
    if isinstance(x, DeviceArray):
        return x, None
    d = DeviceArray.from_numpy(np.ascontiguousarray(x, dtype=dtype))
    return d, d


def train(x, nlist, m, ks=256, niter_coarse=20, niter_pq=25, seed=1234):
    """ivfpq_new: returns (coarse [nlist][d], pq [m][ks][d/m]) float32."""
    if isinstance(x, DeviceArray):
        n, d = x.shape
        fn, ptr = lib().mivf_train_device, x.ptr
    else:
        x = np.ascontiguousarray(x, dtype=np.float32)
        n, d = x.shape
        fn, ptr = lib().mivf_train, _ptr(x)
    coarse = np.empty((nlist, d), np.float32)
    pq = np.empty((m, ks, d // m), np.float32)
    _check(fn(ptr, n, d, nlist, m, ks, niter_coarse, niter_pq, seed, _ptr(coarse), _ptr(pq)))
    return coarse, pq


def kmeans(x, k, niter, seed, ldx=None):
    x = np.ascontiguousarray(x, dtype=np.float32)
    n, d = x.shape
    cent = np.empty((k, d), np.float32)
    _check(lib().mivf_kmeans(_ptr(x), n, d, ldx or d, k, niter, seed, _ptr(cent)))
    return cent


def synth_centres(seed, ncomp, d):
    c = np.empty((ncomp, d), np.float32)
    _check(lib().mivf_synth_centres(seed, ncomp, d, _ptr(c)))
    return c


def zipf_thresholds(seed, ncomp, s):
    t = np.empty(ncomp, np.uint64)
    _check(lib().mivf_zipf_thresholds(seed, ncomp, s, _ptr(t)))
    return t


def synth_rows(seed, stream, row0, n, centres, thr=None, sigma=20.0, out=None):
    """Mixture rows generated on the device; returns a DeviceArray [n][d]."""
    cd, _ = _dev(centres, np.float32)
    ncomp, d = cd.shape
    td = None
    if thr is not None:
        td, _ = _dev(thr, np.uint64)
    if out is None:
        out = DeviceArray((n, d), np.float32)
    _check(lib().mivf_synth_rows(seed, stream, row0, n, d, cd.ptr, ncomp,
                                 td.ptr if td is not None else None, sigma, out.ptr, None))
    synchronize()
    return out


def lds_read_rate(nbytes, random_rows):
    """Measured LDS read rate (GB/s) of 4- or 16-byte reads at random or
    conflict-free rows (the table-gather ceiling of the scan kernels)."""
    v = ctypes.c_double(0)
    _check(lib().mivf_lds_read_rate(int(nbytes), 1 if random_rows else 0, ctypes.byref(v)))
    return v.value


_VECS_ELEM = {"fvecs": 4, "ivecs": 4, "bvecs": 1}


def vecs_info(path, kind):
    """(d, n) of an fvecs / ivecs / bvecs file."""
    d = ctypes.c_int32(0)
    n = ctypes.c_int64(0)
    _check(lib().mivf_vecs_info(str(path).encode(), _VECS_ELEM[kind], ctypes.byref(d), ctypes.byref(n)))
    return d.value, n.value


def read_vecs(path, kind, row0=0, n=None):
    """Records [row0, row0 + n) of a vector file as an [n, d] array: float32
    for fvecs and bvecs (bytes widened), int32 for ivecs."""
    d, total = vecs_info(path, kind)
    if n is None:
        n = max(total - row0, 0)
    out = np.empty((n, d), np.int32 if kind == "ivecs" else np.float32)
    fn = {"fvecs": lib().mivf_read_fvecs, "bvecs": lib().mivf_read_bvecs, "ivecs": lib().mivf_read_ivecs}[kind]
    if n:
        _check(fn(str(path).encode(), row0, n, d, _ptr(out)))
    return out


def bruteforce_nn(base, q):
    """Exact FP64 nearest neighbour: returns (ids [nq] int32, dist [nq] float64)."""
    bd, _ = _dev(base, np.float32)
    qd, _ = _dev(q, np.float32)
    nb, d = bd.shape
    nq = qd.shape[0]
    ids = DeviceArray((nq,), np.int32)
    dis = DeviceArray((nq,), np.float64)
    _check(lib().mivf_bruteforce_nn(bd.ptr, nb, qd.ptr, nq, d, ids.ptr, dis.ptr))
    return ids.numpy(), dis.numpy()


def fast_select(cand, cnt, kl):
    """Shortlist step of the reduced-precision form: for each row of cand
    [nq][capq] uint64 keys, the ids (low 32 bits) of the kl smallest of its
    first cnt[q] keys, in any order, padded with -1."""
    cd, _ = _dev(cand, np.uint64)
    nd, _ = _dev(cnt, np.int32)
    nq, capq = cd.shape
    out = DeviceArray((nq, kl), np.int32)
    _check(lib().mivf_fast_select(cd.ptr, nd.ptr, capq, nq, kl, out.ptr))
    return out.numpy()


def kmin(dis, ids, k):
    """Exact k smallest (distance, id) pairs of one candidate array, ascending
    (my_k_min done exactly); padded with (-1, +inf) past len(dis)."""
    dd, _ = _dev(dis, np.float32)
    di, _ = _dev(ids, np.int32)
    oi = DeviceArray((k,), np.int32)
    od = DeviceArray((k,), np.float32)
    _check(lib().mivf_kmin(dd.ptr, di.ptr, dd.shape[0], k, oi.ptr, od.ptr, None))
    return oi.numpy(), od.numpy()


class Index:
    """Device-resident IVF-PQ index (coarse + PQ quantizers and inverted lists)."""

    def __init__(self, coarse, pq):
        coarse = np.ascontiguousarray(coarse, dtype=np.float32)
        pq = np.ascontiguousarray(pq, dtype=np.float32)
        self.nlist, self.d = coarse.shape
        self.m, self.ks, self.ds = pq.shape
        h = ctypes.c_void_p()
        _check(lib().mivf_index_create(self.d, self.nlist, self.m, self.ks, _ptr(coarse), _ptr(pq),
                                       ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            lib().mivf_index_free(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def ntotal(self):
        n = ctypes.c_int64(0)
        _check(lib().mivf_index_info(self.h, None, None, None, None, ctypes.byref(n)))
        return n.value

    def add(self, x, id0=0):
        if isinstance(x, DeviceArray):
            _check(lib().mivf_index_add_device(self.h, x.ptr, x.shape[0], id0))
        else:
            x = np.ascontiguousarray(x, dtype=np.float32)
            _check(lib().mivf_index_add(self.h, _ptr(x), x.shape[0], id0))

    def add_file(self, path, kind="fvecs", row0=0, n=None, id0=0, chunk=0):
        """Add records [row0, row0 + n) of an fvecs or bvecs file, `chunk`
        records at a time (0: 2^20)."""
        if kind not in ("fvecs", "bvecs"):
            raise ValueError("add_file reads fvecs or bvecs")
        if n is None:
            n = max(vecs_info(path, kind)[1] - row0, 0)
        _check(lib().mivf_index_add_file(self.h, str(path).encode(), _VECS_ELEM[kind], row0, n, id0, chunk))

    def encode(self, x):
        xd, _ = _dev(x, np.float32)
        n = xd.shape[0]
        a = DeviceArray((n,), np.int32)
        c = DeviceArray((n, self.m), np.uint8)
        _check(lib().mivf_index_encode(self.h, xd.ptr, n, a.ptr, c.ptr, None))
        return a.numpy(), c.numpy()

    def set_lists(self, offsets, ids, codes):
        offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        if np.asarray(codes).dtype == np.int32:
            codes = np.ascontiguousarray(codes, dtype=np.int32)
            _check(lib().mivf_index_set_lists_i32(self.h, _ptr(offsets), _ptr(ids), _ptr(codes)))
        else:
            codes = np.ascontiguousarray(codes, dtype=np.uint8)
            _check(lib().mivf_index_set_lists(self.h, _ptr(offsets), _ptr(ids), _ptr(codes)))

    def get_lists(self):
        n = self.ntotal
        off = np.empty(self.nlist + 1, np.int64)
        ids = np.empty(max(n, 1), np.int32)
        codes = np.empty((max(n, 1), self.m), np.uint8)
        _check(lib().mivf_index_get_lists(self.h, _ptr(off), _ptr(ids), _ptr(codes)))
        return off, ids[:n], codes[:n]

    def set_option(self, name, value):
        _check(lib().mivf_set_option(self.h, name.encode(), int(value)))

    def timings(self):
        t = np.zeros(8, np.float32)
        _check(lib().mivf_get_timings(self.h, t.ctypes.data_as(_f32p), 8))
        return t

    # -- search ------------------------------------------------------------
    def search(self, q, w, k, out=None, stream=None):
        qd, _ = _dev(q, np.float32)
        nq = qd.shape[0]
        ids, dis = out if out is not None else (DeviceArray((nq, k), np.int32), DeviceArray((nq, k), np.float32))
        _check(lib().mivf_search(self.h, qd.ptr, nq, w, k, ids.ptr, dis.ptr, stream))
        if out is not None:
            return ids, dis
        return ids.numpy(), dis.numpy()

    def search_var(self, q, w, k, kmax=None, out=None):
        qd, _ = _dev(q, np.float32)
        nq = qd.shape[0]
        w = np.ascontiguousarray(w, dtype=np.int32)
        k = np.ascontiguousarray(k, dtype=np.int32)
        kmax = int(kmax or k.max())
        ids, dis = out if out is not None else (DeviceArray((nq, kmax), np.int32), DeviceArray((nq, kmax), np.float32))
        _check(lib().mivf_search_var(self.h, qd.ptr, nq, _ptr(w), _ptr(k), ids.ptr, dis.ptr, kmax, None))
        if out is not None:
            return ids, dis
        return ids.numpy(), dis.numpy()

    def coarse(self, q, w):
        qd, _ = _dev(q, np.float32)
        nq = qd.shape[0]
        idx = DeviceArray((nq, w), np.int32)
        dis = DeviceArray((nq, w), np.float32)
        _check(lib().mivf_coarse(self.h, qd.ptr, nq, w, idx.ptr, dis.ptr, None))
        return idx.numpy(), dis.numpy()

    def search_residual(self, residual, list_of, w, k):
        rd, _ = _dev(residual, np.float32)
        ld, _ = _dev(list_of, np.int32)
        nq = rd.shape[0] // w
        ids = DeviceArray((nq, k), np.int32)
        dis = DeviceArray((nq, k), np.float32)
        _check(lib().mivf_search_residual(self.h, rd.ptr, ld.ptr, nq, w, k, ids.ptr, dis.ptr, None))
        return ids.numpy(), dis.numpy()

    def adc(self, residual, codes):
        """ADC distances of one residual over codes [n][m] (uint8), in code order."""
        rd, _ = _dev(residual, np.float32)
        cd, _ = _dev(codes, np.uint8)
        n = cd.shape[0]
        out = DeviceArray((max(n, 1),), np.float32)
        _check(lib().mivf_adc(self.h, rd.ptr, cd.ptr, n, out.ptr, None))
        return out.numpy()[:n]

    def attach_base(self, base):
        """Keep a list-ordered copy of the raw vectors for search_fast(base=None)."""
        bd, _ = _dev(base, np.float32)
        _check(lib().mivf_index_attach_base(self.h, bd.ptr, bd.shape[0]))

    def batched_stats(self):
        """Overflowed (recomputed) queries of the last batched search."""
        o = ctypes.c_int64(0)
        _check(lib().mivf_batched_stats(self.h, ctypes.byref(o)))
        return o.value

    def search_batched(self, q, w, k, kmax=None, out=None):
        """Batched exact search with per-query w and k (same results as search_var)."""
        qd, _ = _dev(q, np.float32)
        nq = qd.shape[0]
        w = np.ascontiguousarray(w, dtype=np.int32)
        k = np.ascontiguousarray(k, dtype=np.int32)
        kmax = int(kmax or k.max())
        ids, dis = out if out is not None else (DeviceArray((nq, kmax), np.int32), DeviceArray((nq, kmax), np.float32))
        _check(lib().mivf_search_batched(self.h, qd.ptr, nq, _ptr(w), _ptr(k), kmax, ids.ptr, dis.ptr, None))
        if out is not None:
            return ids, dis
        return ids.numpy(), dis.numpy()

    def exact_tables(self, q, list_no):
        """FP32 tables [m][256][8] of the batched exact kernel (queries q[:n], n <= 8)."""
        qd, _ = _dev(q, np.float32)
        out = DeviceArray((self.m, 256, 8), np.float32)
        _check(lib().mivf_exact_tables(self.h, qd.ptr, qd.shape[0], int(list_no), out.ptr))
        return out.numpy()

    def exact_adc(self, q, list_no, list_len):
        """ADC distances [len][8] of list `list_no` from the batched exact scan loop."""
        qd, _ = _dev(q, np.float32)
        out = DeviceArray((max(list_len, 1), 8), np.float32)
        _check(lib().mivf_exact_adc(self.h, qd.ptr, qd.shape[0], int(list_no), out.ptr))
        return out.numpy()[:list_len]

    def search_fast(self, q, w, kl, k, base=None, out=None):
        """Batched reduced-precision search with exact FP32 rerank (against
        `base`, or the attached copy when base is None)."""
        qd, _ = _dev(q, np.float32)
        nq = qd.shape[0]
        bptr, nb = None, 0
        if base is not None:
            bd, _ = _dev(base, np.float32)
            bptr, nb = bd.ptr, bd.shape[0]
        ids, dis = out if out is not None else (DeviceArray((nq, k), np.int32), DeviceArray((nq, k), np.float32))
        _check(lib().mivf_search_fast(self.h, qd.ptr, nq, w, kl, k, bptr, nb, ids.ptr, dis.ptr, None))
        if out is not None:
            return ids, dis
        return ids.numpy(), dis.numpy()

    def fast_tables(self, q, list_no):
        """8-bit tables [m][256][16] and deltas [16] of one reduced-precision work item."""
        qd, _ = _dev(q, np.float32)
        t8 = DeviceArray((self.m, 256, 16), np.uint8)
        dl = DeviceArray((16,), np.float32)
        _check(lib().mivf_fast_tables(self.h, qd.ptr, qd.shape[0], int(list_no), t8.ptr, dl.ptr))
        return t8.numpy(), dl.numpy()

    def fast_var_adc(self, q, list_no, list_len, tables):
        """Approximate ADC [len][16] of list `list_no` for queries q[:n]
        (n <= 16) with the FP32-built table type `tables` (1 FP16, 2 8-bit
        per-subquantizer offset/scale)."""
        qd, _ = _dev(q, np.float32)
        out = DeviceArray((max(list_len, 1), 16), np.float32)
        _check(lib().mivf_fast_var_adc(self.h, qd.ptr, qd.shape[0], int(list_no), int(tables), out.ptr))
        return out.numpy()[:list_len]

    def fast_seed_tau(self, q, w, kl):
        """Seed thresholds [nq] of the reduced-precision form."""
        qd, _ = _dev(q, np.float32)
        t = DeviceArray((qd.shape[0],), np.float32)
        _check(lib().mivf_fast_seed_tau(self.h, qd.ptr, qd.shape[0], w, kl, t.ptr))
        return t.numpy()

    def fast_stats(self):
        c = ctypes.c_int64(0)
        o = ctypes.c_int64(0)
        _check(lib().mivf_fast_stats(self.h, ctypes.byref(c), ctypes.byref(o)))
        return c.value, o.value

    def search_rerank(self, q, w, kl, k, base):
        qd, _ = _dev(q, np.float32)
        bd, _ = _dev(base, np.float32)
        nq = qd.shape[0]
        ids = DeviceArray((nq, k), np.int32)
        dis = DeviceArray((nq, k), np.float32)
        _check(lib().mivf_search_rerank(self.h, qd.ptr, nq, w, kl, k, bd.ptr, bd.shape[0], ids.ptr,
                                        dis.ptr, None))
        return ids.numpy(), dis.numpy()


class Server:
    """Continuous-batching request loop over one index (exact search, per-request w and k)."""

    def __init__(self, index, slots=4, max_batch=2048, mem_budget=0):
        h = ctypes.c_void_p()
        _check(lib().mivf_server_create(index.h, slots, max_batch, mem_budget, ctypes.byref(h)))
        self.h, self.index = h, index
        self._keep = {}

    def submit(self, q, w, k):
        """Queue a request; returns a ticket.  q: DeviceArray [nq][d]."""
        nq = q.shape[0]
        ids = DeviceArray((nq, k), np.int32)
        dis = DeviceArray((nq, k), np.float32)
        t = ctypes.c_int64(0)
        _check(lib().mivf_server_submit(self.h, q.ptr, nq, w, k, ids.ptr, dis.ptr, ctypes.byref(t)))
        self._keep[t.value] = (q, ids, dis)
        return t.value

    def progress(self):
        c = ctypes.c_int(0)
        _check(lib().mivf_server_progress(self.h, ctypes.byref(c)))
        return c.value

    def drain(self):
        _check(lib().mivf_server_drain(self.h))

    def result(self, ticket):
        _check(lib().mivf_server_wait(self.h, ticket))
        _, ids, dis = self._keep[ticket]
        return ids.numpy(), dis.numpy()

    def latency_ms(self, ticket):
        v = ctypes.c_double(0)
        _check(lib().mivf_server_latency(self.h, ticket, ctypes.byref(v)))
        return v.value

    def warm(self, rows, w, k):
        """Grow every slot's workspaces for micro-batches of w and k (untimed
        searches of `rows`); only while nothing waits or runs."""
        rd, _ = _dev(rows, np.float32)
        self._warm_rows = rd
        _check(lib().mivf_server_warm(self.h, rd.ptr, rd.shape[0], w, k))
        self._keep = {}

    def stats(self):
        """{batches, list_major_batches, rerun_queries} since creation or warm()."""
        v = [ctypes.c_int64(0) for _ in range(3)]
        _check(lib().mivf_server_stats(self.h, *[ctypes.byref(x) for x in v]))
        return dict(zip(("batches", "list_major_batches", "rerun_queries"), (x.value for x in v)))

    def close(self):
        if self.h is not None:
            lib().mivf_server_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def serve_bench(index, pool, rate, nreq, seed=1, bmax=1024, w_choices=(32,), k_choices=(100,), slots=4,
                max_batch=2048, mem_budget=0):
    """Open-loop Poisson load on a Server; returns a dict of QPS and latency percentiles."""
    pd, _ = _dev(pool, np.float32)
    wc = np.ascontiguousarray(w_choices, dtype=np.int32)
    kc = np.ascontiguousarray(k_choices, dtype=np.int32)
    st = (ctypes.c_double * 10)()
    _check(lib().mivf_serve_bench(index.h, slots, max_batch, mem_budget, pd.ptr, pd.shape[0], float(rate), nreq,
                                  seed, bmax, _ptr(wc), len(wc), _ptr(kc), len(kc), st))
    keys = ("qps", "p50_ms", "p99_ms", "mean_ms", "queries", "makespan_ms", "batches", "offered_qps",
            "list_major_batches", "rerun_queries")
    return {k: float(v) for k, v in zip(keys, st)}


class Plan:
    """Graph-captured, optionally autotuned batched exact search for nq queries."""

    def __init__(self, index, nq, wmax, kmax, autotune=True):
        h = ctypes.c_void_p()
        _check(lib().mivf_plan_create(index.h, nq, wmax, kmax, 1 if autotune else 0, ctypes.byref(h)))
        self.h, self.index, self.nq, self.kmax = h, index, nq, kmax

    def search(self, q, w, k, out=None):
        qd, _ = _dev(q, np.float32)
        w = np.ascontiguousarray(w, dtype=np.int32)
        k = np.ascontiguousarray(k, dtype=np.int32)
        ids, dis = out if out is not None else (DeviceArray((self.nq, self.kmax), np.int32),
                                                DeviceArray((self.nq, self.kmax), np.float32))
        _check(lib().mivf_plan_search(self.h, qd.ptr, _ptr(w), _ptr(k), ids.ptr, dis.ptr, None))
        if out is not None:
            return ids, dis
        return ids.numpy(), dis.numpy()

    def info(self):
        g, t, ms = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_float(0)
        _check(lib().mivf_plan_info(self.h, ctypes.byref(g), ctypes.byref(t), ctypes.byref(ms)))
        return {"group": g.value, "threads": t.value, "tuned_ms": ms.value}

    def close(self):
        if self.h is not None:
            lib().mivf_plan_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

