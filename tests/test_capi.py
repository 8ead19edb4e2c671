"""CPU tests of the C ABI: exported symbols, reference on-disk formats,
argument validation, and a C++ caller compiled against the drop-in headers.
Nothing here launches a kernel."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


@pytest.fixture(scope="module")
def L():
    import mestrado_amd as M
    return M.lib()


def declared_c_functions():
    names = []
    txt = open(os.path.join(ROOT, "include", "mivf.h")).read()
    for m in re.finditer(r"^(?:int|const char \*)\s*(mivf_\w+)\s*\(", txt, re.M):
        names.append(m.group(1))
    return names


def test_every_declared_function_is_exported(L):
    names = declared_c_functions()
    assert len(names) >= 35
    for n in names:
        assert hasattr(L, n), n


def test_compat_symbols_have_reference_mangling(L):
    # C++ linkage with the reference's struct tags, as the reference declares
    # them (parallel2/ivf_pq/mycuda.h:7-9, ivf_training.h, sequential ivf_search.h)
    for sym in ["_Z8core_gpu6pqtipo3matP3ivfiPiS3_4matIS0_ii",
                "_Z19preallocate_gpu_mem6pqtipoP3ivfi",
                "_Z18deallocate_gpu_memv",
                "_Z9ivfpq_newii3mati",
                "_Z12ivfpq_assign5ivfpq3matP3ivf",
                "_Z10write_centPcS_S_5ivfpq",
                "_Z9read_centPcS_S_P5ivfpq",
                "_Z8read_ivf5ivfpqiiPc",
                "_Z12ivfpq_search5ivfpqP3ivf3matS2_iiiPiPf",
                # parallel/ fork: per-list search and my_k_min (parallel/ivf_pq/ivf_search.h:39, k_min.h:15)
                "_Z12ivfpq_searchP3ivfPf6pqtipoiPdS3_",
                "_Z8my_k_min5dis_tiPfPi"]:
        assert hasattr(L, sym), sym


@pytest.mark.parametrize("fork", ["parallel2", "sequential", "parallel"])
def test_compat_callers_load(L, fork):
    """The C++ callers built against each fork's header (tests/compat/) were
    linked with --no-undefined against libmivf and load here; their entry
    points exist.  (They run on the GPU in test_gpu_compat.py.)"""
    path = os.path.join(ROOT, "tests", "compat", f"libcaller_{fork}.so")
    assert os.path.exists(path), "build() compiles the callers"
    C = ctypes.CDLL(path)
    for fn in {"parallel2": ["run_core_gpu", "run_build", "run_read_ivf"], "sequential": ["run_seq_search"],
               "parallel": ["run_par_search", "run_par_list", "run_kmin"]}[fork]:
        assert hasattr(C, fn), fn


def test_read_reference_centroid_files(L):
    """The reference's own trained 256-list quantizer files (parallel/bin)."""
    h = os.path.join(GOLD, "ref", "file_ivfpq_256.bin").encode()
    c = os.path.join(GOLD, "ref", "cent_ivfpq_256.bin").encode()
    a = os.path.join(GOLD, "ref", "coa_ivfpq_256.bin").encode()
    v = [ctypes.c_int() for _ in range(5)]
    assert L.mivf_read_cent_header(h, *[ctypes.byref(x) for x in v]) == 0
    nsq, ks, ds, coarsek, d = [x.value for x in v]
    assert (nsq, ks, ds, coarsek, d) == (8, 256, 16, 256, 128)
    pq = np.empty((nsq, ks, ds), np.float32)
    co = np.empty((coarsek, d), np.float32)
    assert L.mivf_read_cent(h, c, a, pq.ctypes.data, co.ctypes.data) == 0
    assert np.array_equal(pq.ravel(), np.fromfile(c.decode(), np.float32))
    assert np.array_equal(co.ravel(), np.fromfile(a.decode(), np.float32))


def test_cent_roundtrip_and_corruption(L, tmp_path):
    rng = np.random.default_rng(0)
    pq = rng.standard_normal((8, 256, 16)).astype(np.float32)
    co = rng.standard_normal((64, 128)).astype(np.float32)
    h, c, a = [str(tmp_path / n).encode() for n in ("h", "c", "a")]
    assert L.mivf_write_cent(h, c, a, 8, 256, 16, 64, 128, pq.ctypes.data, co.ctypes.data) == 0
    hdr = np.fromfile(h.decode(), np.int32)
    assert list(hdr) == [8, 256, 16, 4096, 8, 64, 64, 128]  # ivf_training.cpp:104-111 order
    pq2 = np.empty_like(pq)
    co2 = np.empty_like(co)
    assert L.mivf_read_cent(h, c, a, pq2.ctypes.data, co2.ctypes.data) == 0
    assert np.array_equal(pq, pq2) and np.array_equal(co, co2)
    # truncated centroid file
    open(c.decode(), "r+b").truncate(1000)
    assert L.mivf_read_cent(h, c, a, pq2.ctypes.data, co2.ctypes.data) == 4
    assert b"truncated" in L.mivf_last_error()
    # inconsistent header
    bad = hdr.copy()
    bad[3] = 7
    bad.tofile(h.decode())
    assert L.mivf_read_cent_header(h, None, None, None, None, None) == 4
    assert L.mivf_read_cent(b"/nonexistent/h", c, a, pq2.ctypes.data, co2.ctypes.data) == 4


def random_lists(nlist, m, seed=1):
    rng = np.random.default_rng(seed)
    sizes = rng.integers(0, 30, nlist)
    sizes[::5] = 0
    off = np.zeros(nlist + 1, np.int64)
    off[1:] = np.cumsum(sizes)
    n = int(off[-1])
    ids = rng.permutation(n).astype(np.int32)
    codes = rng.integers(0, 256, (n, m)).astype(np.uint8)
    return off, ids, codes


@pytest.mark.parametrize("kind", ["dir", "file"])
def test_ivf_roundtrip(L, tmp_path, kind):
    nlist, m = 40, 8
    off, ids, codes = random_lists(nlist, m)
    n = int(off[-1])
    path = str(tmp_path / "ivf").encode()
    if kind == "dir":
        os.makedirs(path)
        w, r = L.mivf_write_ivf_dir, L.mivf_read_ivf_dir
    else:
        w, r = L.mivf_write_ivf_file, L.mivf_read_ivf_file
    assert w(path, nlist, m, off.ctypes.data, ids.ctypes.data, codes.ctypes.data) == 0
    off2 = np.empty_like(off)
    assert r(path, nlist, m, off2.ctypes.data, None, None, ctypes.c_int64(0)) == 0
    assert np.array_equal(off, off2)
    ids2 = np.empty(n, np.int32)
    codes2 = np.empty((n, m), np.uint8)
    assert r(path, nlist, m, off2.ctypes.data, ids2.ctypes.data, codes2.ctypes.data, ctypes.c_int64(n)) == 0
    assert np.array_equal(ids, ids2) and np.array_equal(codes, codes2)
    # too small output arrays are rejected, not overrun
    if n > 1:
        assert r(path, nlist, m, off2.ctypes.data, ids2.ctypes.data, codes2.ctypes.data,
                 ctypes.c_int64(n - 1)) == 1


def test_ivf_file_oversized_list_length(L, tmp_path):
    """A list length far beyond the file's size is reported as a corrupt file
    (status 4) before anything is sized by it."""
    path = tmp_path / "ivf"
    np.array([2, 1 << 30, 5, 6], np.int32).tofile(str(path))  # coarsek 2, list 0 claims 2^30 vectors
    off = np.empty(3, np.int64)
    assert L.mivf_read_ivf_file(str(path).encode(), 2, 8, off.ctypes.data, None, None, ctypes.c_int64(0)) == 4
    assert b"larger than the file" in L.mivf_last_error()
    d = tmp_path / "dir"
    d.mkdir()
    for j in range(2):
        np.array([2, 1 << 29, 1, 2], np.int32).tofile(str(d / str(j)))
    assert L.mivf_read_ivf_dir(str(d).encode(), 2, 8, off.ctypes.data, None, None, ctypes.c_int64(0)) == 4


def test_ivf_dir_layout_matches_reference(L, tmp_path):
    """<dir>/<j>: int coarsek, int n, ids[n], int n, int m, int32 codes[n][m]
    (parallel2/genivf.cpp:72-78)."""
    nlist, m = 3, 4
    off = np.array([0, 2, 2, 3], np.int64)
    ids = np.array([7, 9, 4], np.int32)
    codes = np.array([[1, 2, 3, 4], [5, 6, 7, 8], [255, 0, 1, 2]], np.uint8)
    d = str(tmp_path)
    assert L.mivf_write_ivf_dir(d.encode(), nlist, m, off.ctypes.data, ids.ctypes.data, codes.ctypes.data) == 0
    f0 = np.fromfile(os.path.join(d, "0"), np.int32)
    assert list(f0) == [3, 2, 7, 9, 2, 4, 1, 2, 3, 4, 5, 6, 7, 8]
    f1 = np.fromfile(os.path.join(d, "1"), np.int32)
    assert list(f1) == [3, 0, 0, 4]
    # corrupt: code outside 0..255, id/code count mismatch, wrong coarsek
    bad = f0.copy()
    bad[6] = 300
    bad.tofile(os.path.join(d, "0"))
    off2 = np.empty(4, np.int64)
    assert L.mivf_read_ivf_dir(d.encode(), nlist, m, off2.ctypes.data, None, None, ctypes.c_int64(0)) == 4
    bad = f0.copy()
    bad[4] = 1
    bad.tofile(os.path.join(d, "0"))
    assert L.mivf_read_ivf_dir(d.encode(), nlist, m, off2.ctypes.data, None, None, ctypes.c_int64(0)) == 4
    f0.tofile(os.path.join(d, "0"))
    assert L.mivf_read_ivf_dir(d.encode(), 4, m, off2.ctypes.data, None, None, ctypes.c_int64(0)) == 4


def _write_vecs(path, a, dtype):
    """Vector-file writer of this test: int32 d then d values per record."""
    a = np.ascontiguousarray(a)
    n, d = a.shape
    with open(path, "wb") as f:
        for r in a:
            f.write(np.int32(d).tobytes())
            f.write(r.astype(dtype).tobytes())


def test_read_fvecs_reference_queries():
    """The reference's own siftsmall query file: 100 records of 128 floats
    (the file the reference's tests load with fvecs_read)."""
    import mestrado_amd as M
    sys_path = os.path.join(GOLD, "ref", "siftsmall_query.fvecs")
    assert M.vecs_info(sys_path, "fvecs") == (128, 100)
    raw = np.fromfile(sys_path, np.int32).reshape(100, 129)
    assert (raw[:, 0] == 128).all()
    want = raw[:, 1:].copy().view(np.float32)
    assert np.array_equal(M.read_vecs(sys_path, "fvecs"), want)
    assert np.array_equal(M.read_vecs(sys_path, "fvecs", 37, 11), want[37:48])
    assert M.read_vecs(sys_path, "fvecs", 100).shape == (0, 128)


def test_read_bvecs_ivecs_roundtrip(tmp_path):
    import mestrado_amd as M
    rng = np.random.default_rng(5)
    b = rng.integers(0, 256, (1000, 128)).astype(np.uint8)
    iv = rng.integers(-2 ** 31, 2 ** 31 - 1, (300, 100)).astype(np.int32)
    _write_vecs(tmp_path / "x.bvecs", b, np.uint8)
    _write_vecs(tmp_path / "gt.ivecs", iv, np.int32)
    assert M.vecs_info(tmp_path / "x.bvecs", "bvecs") == (128, 1000)
    got = M.read_vecs(tmp_path / "x.bvecs", "bvecs", 123, 500)
    assert got.dtype == np.float32 and np.array_equal(got, b[123:623].astype(np.float32))
    # the record boundary of a 4097-row read (two internal chunks)
    big = rng.integers(0, 256, (5000, 16)).astype(np.uint8)
    _write_vecs(tmp_path / "big.bvecs", big, np.uint8)
    assert np.array_equal(M.read_vecs(tmp_path / "big.bvecs", "bvecs", 3, 4097), big[3:4100].astype(np.float32))
    assert np.array_equal(M.read_vecs(tmp_path / "gt.ivecs", "ivecs"), iv)
    (tmp_path / "empty.fvecs").write_bytes(b"")
    assert M.vecs_info(tmp_path / "empty.fvecs", "fvecs") == (0, 0)


def test_read_vecs_errors(L, tmp_path):
    """Past the end, a dimension change inside the file, a ragged tail and a
    missing file are MIVF_EIO; bad arguments are MIVF_EINVAL."""
    rng = np.random.default_rng(6)
    x = rng.standard_normal((10, 8)).astype(np.float32)
    p = tmp_path / "x.fvecs"
    _write_vecs(p, x, np.float32)
    out = np.empty((11, 8), np.float32)
    assert L.mivf_read_fvecs(str(p).encode(), 0, 11, 8, out.ctypes.data) == 4
    assert L.mivf_read_fvecs(str(p).encode(), 5, 6, 8, out.ctypes.data) == 4
    assert L.mivf_read_fvecs(str(p).encode(), 5, 5, 8, out.ctypes.data) == 0
    assert np.array_equal(out[:5], x[5:])
    # dimension 7 asked of an 8-dim file: records misalign, the d field differs
    assert L.mivf_read_fvecs(str(p).encode(), 0, 3, 7, out.ctypes.data) == 4
    raw = bytearray(p.read_bytes())
    raw[36 * 4:36 * 4 + 4] = np.int32(9).tobytes()  # record 4's d field
    q = tmp_path / "bad.fvecs"
    q.write_bytes(bytes(raw))
    assert L.mivf_read_fvecs(str(q).encode(), 0, 4, 8, out.ctypes.data) == 0
    assert L.mivf_read_fvecs(str(q).encode(), 0, 5, 8, out.ctypes.data) == 4
    assert b"dimension" in L.mivf_last_error()
    (tmp_path / "ragged.fvecs").write_bytes(p.read_bytes()[:-3])
    d = ctypes.c_int32(0)
    n = ctypes.c_int64(0)
    assert L.mivf_vecs_info(str(tmp_path / "ragged.fvecs").encode(), 4, ctypes.byref(d), ctypes.byref(n)) == 4
    assert L.mivf_vecs_info(b"/nonexistent/x.fvecs", 4, ctypes.byref(d), ctypes.byref(n)) == 4
    assert L.mivf_vecs_info(str(p).encode(), 2, ctypes.byref(d), ctypes.byref(n)) == 1
    assert L.mivf_read_fvecs(str(p).encode(), -1, 1, 8, out.ctypes.data) == 1
    assert L.mivf_read_fvecs(str(p).encode(), 0, 1, 0, out.ctypes.data) == 1


def test_argument_validation_without_gpu(L):
    """Bad shapes are rejected before any device work."""
    co = np.zeros((16, 128), np.float32)
    pq = np.zeros((8, 256, 16), np.float32)
    h = ctypes.c_void_p()
    assert L.mivf_index_create(128, 16, 7, 256, co.ctypes.data, pq.ctypes.data, ctypes.byref(h)) == 1
    assert L.mivf_index_create(128, 16, 8, 300, co.ctypes.data, pq.ctypes.data, ctypes.byref(h)) == 1
    assert L.mivf_index_create(96, 16, 8, 256, co.ctypes.data, pq.ctypes.data, ctypes.byref(h)) == 5
    bad = co.copy()
    bad[3, 3] = np.nan
    assert L.mivf_index_create(128, 16, 8, 256, bad.ctypes.data, pq.ctypes.data, ctypes.byref(h)) == 1
    assert b"non-finite" in L.mivf_last_error()
    assert L.mivf_search(None, None, 1, 1, 1, None, None, None) == 1
    assert L.mivf_set_option(None, b"profile", 1) == 1
    assert L.mivf_index_free(None) == 0


def test_cpp_caller_against_compat_header(tmp_path):
    """A C++ program written against the reference's signatures compiles and
    links against libmivf; its host-only calls (write_cent/read_cent) run."""
    src = tmp_path / "caller.cpp"
    src.write_text(r'''
#include "mestrado_compat.h"
#include <cstdio>
#include <cstdlib>
#include <cstring>
int main(int argc, char** argv) {
    ivfpq_t q;
    q.pq.nsq = 2; q.pq.ks = 4; q.pq.ds = 2; q.pq.centroidsn = 8; q.pq.centroidsd = 2;
    q.coarsek = q.coa_centroidsn = 3; q.coa_centroidsd = 4;
    float pq[16], co[12];
    for (int i = 0; i < 16; ++i) pq[i] = i * 0.5f;
    for (int i = 0; i < 12; ++i) co[i] = -i;
    q.pq.centroids = pq; q.coa_centroids = co;
    write_cent(argv[1], argv[2], argv[3], q);
    ivfpq_t r;
    read_cent(argv[1], argv[2], argv[3], &r);
    if (r.pq.nsq != 2 || r.pq.ks != 4 || r.coarsek != 3 || r.coa_centroidsd != 4) return 2;
    if (memcmp(r.pq.centroids, pq, sizeof pq) || memcmp(r.coa_centroids, co, sizeof co)) return 3;
    // the device entry points link (not called: no GPU here)
    void (*f)(pqtipo, mat, ivf_t*, int, int*, int*, matI, mat, int, int) = &core_gpu;
    void (*g)(pqtipo, ivf_t*, int) = &preallocate_gpu_mem;
    std::printf("ok %p %p\n", (void*)f, (void*)g);
    return 0;
}
''')
    exe = tmp_path / "caller"
    lib = os.path.join(ROOT, "mestrado_amd")
    r = subprocess.run(["g++", "-std=c++11", str(src), "-I", os.path.join(ROOT, "include"), "-L", lib,
                        "-lmivf", "-Wl,-rpath," + lib, "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    args = [str(tmp_path / n) for n in ("h", "c", "a")]
    r = subprocess.run([str(exe), *args], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.startswith("ok"), (r.returncode, r.stdout, r.stderr[-2000:])
