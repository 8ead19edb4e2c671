// Readers and writers for the reference's on-disk index formats (host code).
//
//  header file  8 x int32: nsq, ks, ds, centroidsn (= ks*ds), centroidsd (= nsq),
//               coarsek, coa_centroidsn (= coarsek), coa_centroidsd (= d)
//  cent file    float32 [nsq][ks][ds]          (parallel/ivf_pq/ivf_training.cpp:92-143)
//  coa file     float32 [coarsek][d]
//  list file    one per list, <dir>/<j>: int32 coarsek, int32 n, int32 ids[n],
//               int32 n, int32 m, int32 codes[n][m] (parallel2/genivf.cpp:62-86,
//               parallel2/ivf_pq/ivf_search.cpp:293-321)
//  single file  int32 coarsek, then per list: int32 n, int32 ids[n], int32 n,
//               int32 m, int32 codes[n][m] (parallel/ivf_pq/ivf_search.cpp:396-441)
//  vector files fvecs / ivecs / bvecs: records of int32 d followed by d float32
//               / int32 / uint8 values (parallel/yael_needs/vector.c:296-350,
//               parallel/pq-utils/pq_test_load_vectors.cpp:236-325); bvecs values
//               widen to float32, and a read may start at any record (as
//               my_bvecs_read's offset does)
// Every read is size-checked: a truncated or inconsistent file gives MIVF_EIO.
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/mivf.h"
#include "mivf_internal.h"

using namespace mivf;

namespace {

struct File {
    FILE *f = nullptr;
    explicit File(const char *path, const char *mode) : f(path ? std::fopen(path, mode) : nullptr) {}
    ~File() {
        if (f) std::fclose(f);
    }
};

int read_exact(FILE *f, void *dst, size_t elem, size_t n, const char *what) {
    if (n && std::fread(dst, elem, n, f) != n) {
        set_error(std::string("truncated file while reading ") + what);
        return kIo;
    }
    return kOk;
}

int write_exact(FILE *f, const void *src, size_t elem, size_t n, const char *what) {
    if (n && std::fwrite(src, elem, n, f) != n) {
        set_error(std::string("write failed: ") + what);
        return kIo;
    }
    return kOk;
}

int open_or_fail(File &fh, const char *path) {
    if (!fh.f) {
        set_error(std::string("cannot open ") + (path ? path : "(null)"));
        return kIo;
    }
    return kOk;
}

// Bytes from the current position to the end of the file (-1 if unknown).
int64_t bytes_left(FILE *f) {
    const long pos = std::ftell(f);
    if (pos < 0 || std::fseek(f, 0, SEEK_END) != 0) return -1;
    const long end = std::ftell(f);
    if (std::fseek(f, pos, SEEK_SET) != 0 || end < pos) return -1;
    return int64_t(end) - int64_t(pos);
}

// Read one list record (ids + int32 codes) and append it in uint8 form.
int read_list_record(FILE *f, int m, std::vector<int32_t> &ids, std::vector<uint8_t> &codes,
                     const char *what) {
    int32_t n = 0, n2 = 0, mm = 0;
    MIVF_TRY(read_exact(f, &n, 4, 1, what));
    MIVF_REQUIRE(n >= 0, kIo, std::string("negative list length in ") + what);
    // a list of n vectors takes 4n (ids) + 8 + 4nm (codes) more bytes: check
    // the file holds them before sizing anything by n
    MIVF_REQUIRE(int64_t(n) * 4 * (1 + int64_t(m)) + 8 <= bytes_left(f), kIo,
                 std::string("list length larger than the file in ") + what);
    const size_t i0 = ids.size();
    ids.resize(i0 + size_t(n));
    MIVF_TRY(read_exact(f, ids.data() + i0, 4, size_t(n), what));
    MIVF_TRY(read_exact(f, &n2, 4, 1, what));
    MIVF_TRY(read_exact(f, &mm, 4, 1, what));
    MIVF_REQUIRE(n2 == n, kIo, std::string("code count differs from id count in ") + what);
    MIVF_REQUIRE(mm == m || n == 0, kIo, std::string("code width differs from nsq in ") + what);
    std::vector<int32_t> c32(size_t(n) * m);
    MIVF_TRY(read_exact(f, c32.data(), 4, c32.size(), what));
    for (int32_t v : c32) {
        MIVF_REQUIRE(v >= 0 && v < 256, kIo, std::string("code outside 0..255 in ") + what);
        codes.push_back(uint8_t(v));
    }
    return kOk;
}

int write_list_record(FILE *f, int m, const int32_t *ids, const uint8_t *codes, int64_t n,
                      const char *what) {
    const int32_t n32 = int32_t(n), m32 = m;
    MIVF_TRY(write_exact(f, &n32, 4, 1, what));
    MIVF_TRY(write_exact(f, ids, 4, size_t(n), what));
    MIVF_TRY(write_exact(f, &n32, 4, 1, what));
    MIVF_TRY(write_exact(f, &m32, 4, 1, what));
    std::vector<int32_t> c32(size_t(n) * m);
    for (size_t i = 0; i < c32.size(); ++i) c32[i] = codes[i];
    MIVF_TRY(write_exact(f, c32.data(), 4, c32.size(), what));
    return kOk;
}

int finish_lists(int nlist, int m, std::vector<std::vector<int32_t>> &lid,
                 std::vector<std::vector<uint8_t>> &lcodes, int64_t *offsets, int32_t *ids,
                 uint8_t *codes, int64_t cap) {
    int64_t o = 0;
    offsets[0] = 0;
    for (int l = 0; l < nlist; ++l) {
        const int64_t n = int64_t(lid[l].size());
        if (ids && codes) {
            MIVF_REQUIRE(o + n <= cap, kInvalid, "read_ivf: output arrays too small");
            std::copy(lid[l].begin(), lid[l].end(), ids + o);
            std::copy(lcodes[l].begin(), lcodes[l].end(), codes + o * m);
        }
        o += n;
        offsets[l + 1] = o;
    }
    return kOk;
}

}  // namespace

namespace mivf {

// Records [row0, row0 + n) of a vector file with elem-byte values of
// dimension d, converted to T (float from float32 or uint8, int32 from int32).
template <typename T>
int read_vecs_rows(const char *path, int elem, int64_t row0, int64_t n, int d, T *out) {
    MIVF_REQUIRE(path && d > 0 && row0 >= 0 && n >= 0 && (n == 0 || out), kInvalid, "read_vecs: bad arguments");
    MIVF_REQUIRE(elem == 1 || elem == 4, kInvalid, "read_vecs: elem must be 1 (bvecs) or 4 (fvecs, ivecs)");
    File fh(path, "rb");
    MIVF_TRY(open_or_fail(fh, path));
    const int64_t rec = 4 + int64_t(d) * elem;
    const int64_t left = bytes_left(fh.f);
    MIVF_REQUIRE(left >= 0 && (row0 + n) * rec <= left, kIo,
                 std::string("vector file shorter than the requested records: ") + path);
    MIVF_REQUIRE(std::fseek(fh.f, long(row0 * rec), SEEK_SET) == 0, kIo, std::string("seek failed in ") + path);
    const int64_t CH = 4096;  // records per read
    std::vector<unsigned char> buf(size_t(std::min(n, CH) * rec));
    for (int64_t r0 = 0; r0 < n; r0 += CH) {
        const int64_t nn = std::min(CH, n - r0);
        MIVF_TRY(read_exact(fh.f, buf.data(), size_t(rec), size_t(nn), path));
        for (int64_t i = 0; i < nn; ++i) {
            const unsigned char *p = buf.data() + i * rec;
            int32_t dd;
            std::memcpy(&dd, p, 4);
            MIVF_REQUIRE(dd == d, kIo, std::string("unexpected vector dimension in ") + path);
            T *o = out + (r0 + i) * d;
            if (elem == 1) {
                for (int t = 0; t < d; ++t) o[t] = T(p[4 + t]);
            } else {
                std::memcpy(o, p + 4, size_t(d) * 4);
            }
        }
    }
    return kOk;
}
template int read_vecs_rows<float>(const char *, int, int64_t, int64_t, int, float *);

}  // namespace mivf

extern "C" {

int mivf_vecs_info(const char *path, int elem_bytes, int *d, int64_t *n) try {
    MIVF_REQUIRE(path && (elem_bytes == 1 || elem_bytes == 4), kInvalid,
                 "vecs_info: elem_bytes must be 1 (bvecs) or 4 (fvecs, ivecs)");
    File fh(path, "rb");
    MIVF_TRY(open_or_fail(fh, path));
    const int64_t size = bytes_left(fh.f);
    int32_t dd = 0;
    if (size == 0) {
        if (d) *d = 0;
        if (n) *n = 0;
        return kOk;
    }
    MIVF_TRY(read_exact(fh.f, &dd, 4, 1, path));
    MIVF_REQUIRE(dd > 0, kIo, std::string("non-positive vector dimension in ") + path);
    const int64_t rec = 4 + int64_t(dd) * elem_bytes;
    MIVF_REQUIRE(size % rec == 0, kIo, std::string("vector file size is not a whole number of records: ") + path);
    if (d) *d = dd;
    if (n) *n = size / rec;
    return kOk;
} MIVF_CATCH_ALL

int mivf_read_fvecs(const char *path, int64_t row0, int64_t n, int d, float *out) try {
    return read_vecs_rows<float>(path, 4, row0, n, d, out);
} MIVF_CATCH_ALL

int mivf_read_bvecs(const char *path, int64_t row0, int64_t n, int d, float *out) try {
    return read_vecs_rows<float>(path, 1, row0, n, d, out);
} MIVF_CATCH_ALL

int mivf_read_ivecs(const char *path, int64_t row0, int64_t n, int d, int32_t *out) try {
    return read_vecs_rows<int32_t>(path, 4, row0, n, d, out);
} MIVF_CATCH_ALL

int mivf_write_cent(const char *header, const char *cent, const char *coa, int nsq, int ks, int ds,
                    int coarsek, int d, const float *pq, const float *coarse) try {
    MIVF_REQUIRE(nsq > 0 && ks > 0 && ds > 0 && coarsek > 0 && d > 0 && pq && coarse, kInvalid,
                 "write_cent: bad arguments");
    File h(header, "wb"), c(cent, "wb"), a(coa, "wb");
    MIVF_TRY(open_or_fail(h, header));
    MIVF_TRY(open_or_fail(c, cent));
    MIVF_TRY(open_or_fail(a, coa));
    const int32_t hdr[8] = {nsq, ks, ds, ks * ds, nsq, coarsek, coarsek, d};
    MIVF_TRY(write_exact(h.f, hdr, 4, 8, header));
    MIVF_TRY(write_exact(c.f, pq, 4, size_t(nsq) * ks * ds, cent));
    MIVF_TRY(write_exact(a.f, coarse, 4, size_t(coarsek) * d, coa));
    return kOk;
} MIVF_CATCH_ALL

int mivf_read_cent_header(const char *header, int *nsq, int *ks, int *ds, int *coarsek, int *d) try {
    File h(header, "rb");
    MIVF_TRY(open_or_fail(h, header));
    int32_t hdr[8];
    MIVF_TRY(read_exact(h.f, hdr, 4, 8, header));
    MIVF_REQUIRE(hdr[0] > 0 && hdr[1] > 0 && hdr[2] > 0 && hdr[5] > 0 && hdr[7] > 0, kIo,
                 "read_cent: non-positive size in header");
    MIVF_REQUIRE(hdr[3] == hdr[1] * hdr[2] && hdr[4] == hdr[0] && hdr[6] == hdr[5], kIo,
                 "read_cent: inconsistent header");
    MIVF_REQUIRE(hdr[0] * hdr[2] == hdr[7], kIo, "read_cent: nsq * ds differs from d");
    if (nsq) *nsq = hdr[0];
    if (ks) *ks = hdr[1];
    if (ds) *ds = hdr[2];
    if (coarsek) *coarsek = hdr[5];
    if (d) *d = hdr[7];
    return kOk;
} MIVF_CATCH_ALL

int mivf_read_cent(const char *header, const char *cent, const char *coa, float *pq, float *coarse) try {
    int nsq, ks, ds, coarsek, d;
    MIVF_TRY(mivf_read_cent_header(header, &nsq, &ks, &ds, &coarsek, &d));
    MIVF_REQUIRE(pq && coarse, kInvalid, "read_cent: null output");
    File c(cent, "rb"), a(coa, "rb");
    MIVF_TRY(open_or_fail(c, cent));
    MIVF_TRY(open_or_fail(a, coa));
    MIVF_TRY(read_exact(c.f, pq, 4, size_t(nsq) * ks * ds, cent));
    MIVF_TRY(read_exact(a.f, coarse, 4, size_t(coarsek) * d, coa));
    return kOk;
} MIVF_CATCH_ALL

int mivf_write_ivf_dir(const char *dir, int coarsek, int m, const int64_t *offsets, const int32_t *ids,
                       const uint8_t *codes) try {
    MIVF_REQUIRE(dir && coarsek > 0 && m > 0 && offsets, kInvalid, "write_ivf_dir: bad arguments");
    for (int j = 0; j < coarsek; ++j) {
        const std::string path = std::string(dir) + "/" + std::to_string(j);
        File f(path.c_str(), "wb");
        MIVF_TRY(open_or_fail(f, path.c_str()));
        const int32_t ck = coarsek;
        MIVF_TRY(write_exact(f.f, &ck, 4, 1, path.c_str()));
        const int64_t b = offsets[j], n = offsets[j + 1] - offsets[j];
        MIVF_TRY(write_list_record(f.f, m, ids + b, codes + b * m, n, path.c_str()));
    }
    return kOk;
} MIVF_CATCH_ALL

// offsets [coarsek+1] always written; ids/codes may be NULL to query sizes.
int mivf_read_ivf_dir(const char *dir, int coarsek, int m, int64_t *offsets, int32_t *ids, uint8_t *codes,
                      int64_t cap) try {
    MIVF_REQUIRE(dir && coarsek > 0 && m > 0 && offsets, kInvalid, "read_ivf_dir: bad arguments");
    std::vector<std::vector<int32_t>> lid(coarsek);
    std::vector<std::vector<uint8_t>> lcodes(coarsek);
    for (int j = 0; j < coarsek; ++j) {
        const std::string path = std::string(dir) + "/" + std::to_string(j);
        File f(path.c_str(), "rb");
        MIVF_TRY(open_or_fail(f, path.c_str()));
        int32_t ck = 0;
        MIVF_TRY(read_exact(f.f, &ck, 4, 1, path.c_str()));
        MIVF_REQUIRE(ck == coarsek, kIo, "read_ivf_dir: list file from another coarsek: " + path);
        MIVF_TRY(read_list_record(f.f, m, lid[j], lcodes[j], path.c_str()));
    }
    return finish_lists(coarsek, m, lid, lcodes, offsets, ids, codes, cap);
} MIVF_CATCH_ALL

int mivf_write_ivf_file(const char *path, int coarsek, int m, const int64_t *offsets, const int32_t *ids,
                        const uint8_t *codes) try {
    MIVF_REQUIRE(path && coarsek > 0 && m > 0 && offsets, kInvalid, "write_ivf_file: bad arguments");
    File f(path, "wb");
    MIVF_TRY(open_or_fail(f, path));
    const int32_t ck = coarsek;
    MIVF_TRY(write_exact(f.f, &ck, 4, 1, path));
    for (int j = 0; j < coarsek; ++j) {
        const int64_t b = offsets[j], n = offsets[j + 1] - offsets[j];
        MIVF_TRY(write_list_record(f.f, m, ids + b, codes + b * m, n, path));
    }
    return kOk;
} MIVF_CATCH_ALL

int mivf_read_ivf_file(const char *path, int coarsek, int m, int64_t *offsets, int32_t *ids, uint8_t *codes,
                       int64_t cap) try {
    MIVF_REQUIRE(path && coarsek > 0 && m > 0 && offsets, kInvalid, "read_ivf_file: bad arguments");
    File f(path, "rb");
    MIVF_TRY(open_or_fail(f, path));
    int32_t ck = 0;
    MIVF_TRY(read_exact(f.f, &ck, 4, 1, path));
    MIVF_REQUIRE(ck == coarsek, kIo, "read_ivf_file: file holds another coarsek");
    std::vector<std::vector<int32_t>> lid(coarsek);
    std::vector<std::vector<uint8_t>> lcodes(coarsek);
    for (int j = 0; j < coarsek; ++j) MIVF_TRY(read_list_record(f.f, m, lid[j], lcodes[j], path));
    return finish_lists(coarsek, m, lid, lcodes, offsets, ids, codes, cap);
} MIVF_CATCH_ALL

}  // extern "C"
