// Fused IVF-PQ search kernel: for one query and a group of its probed lists,
//   residual r = fl32(q - c_list)                 (parallel/ivf_pq/ivf_assign.cpp:72-78)
//   table T[s][j] = dist(a = r_s, b = P[s][j])     (parallel/ivf_pq/ivf_search.cpp:206-209)
//   ADC  d = 0f; d = fl32(d + T[s][code_s]), s ascending   (ivf_search.cpp:451-468)
//   running exact top-k by (distance, id)          (parallel2/ivf_pq/topk.cu:26-32)
//
// Workgroup = 512 threads (8 wavefronts).  A prologue kernel writes the
// residual row and the f64 sub-vector norms of every (query, probe) pair;
// the scan reads them with scalar loads, so a table entry costs one vector
// load of P[s][j] (shared by the TB tables of a phase), DS multiplies and
// adds with a scalar operand, one f64 add and one LDS store.  The m x 256
// FP32 tables live in LDS; every lane scans U vectors per step with one
// 8/16-byte coalesced code load each and M LDS gathers.  Candidates that
// beat the current threshold are appended to an LDS buffer with one LDS
// atomic per wavefront; when the buffer fills, a radix select on the distance
// keys shrinks it to the candidates within the k-th distance (positions are
// resolved to ids only for the final exact selection by (distance, id)).
#include "select_dev.h"

namespace mivf {

namespace {

constexpr int D = 128;
constexpr int NT = 512;

template <int M>
struct CodeVec;
template <>
struct CodeVec<4> {
    static constexpr int W = 1;
    __device__ static void load(const uint8_t *p, uint32_t *w) { w[0] = *reinterpret_cast<const uint32_t *>(p); }
};
template <>
struct CodeVec<8> {
    static constexpr int W = 2;
    __device__ static void load(const uint8_t *p, uint32_t *w) {
        uint2 v = *reinterpret_cast<const uint2 *>(p);
        w[0] = v.x; w[1] = v.y;
    }
};
template <>
struct CodeVec<16> {
    static constexpr int W = 4;
    __device__ static void load(const uint8_t *p, uint32_t *w) {
        uint4 v = *reinterpret_cast<const uint4 *>(p);
        w[0] = v.x; w[1] = v.y; w[2] = v.z; w[3] = v.w;
    }
};
template <>
struct CodeVec<32> {
    static constexpr int W = 8;
    __device__ static void load(const uint8_t *p, uint32_t *w) {
        uint4 v0 = *reinterpret_cast<const uint4 *>(p);
        uint4 v1 = *reinterpret_cast<const uint4 *>(p + 16);
        w[0] = v0.x; w[1] = v0.y; w[2] = v0.z; w[3] = v0.w;
        w[4] = v1.x; w[5] = v1.y; w[6] = v1.z; w[7] = v1.w;
    }
};

// Control words of one workgroup.
struct ScanCtl {
    SelectShared sel;
    int inc[3];  // per-step append counters, rotated so one barrier per step suffices
    int newcnt;
    int eqt;
    uint32_t maxhi;
    uint32_t seed;  // first-step threshold key (max over the wavefronts)
};

constexpr int TB = 2;  // tables built per phase (probes scanned per phase)
constexpr int PD = 1;  // subquantizers of P loaded ahead during a table build

// Dynamic LDS layout.  The tables sit at LDS address 0 so that a gather is
// one ds_read_b32 at code*4 with the table row as immediate offset.
template <int M, int CAP>
struct Lay {
    static constexpr size_t tab = 0;  // float [TB][M][256]
    static constexpr size_t cand = tab + size_t(TB) * M * 256 * 4;
    static constexpr size_t ctl = cand + size_t(CAP) * 8;
    static constexpr size_t total = ctl + sizeof(ScanCtl);
};

// Float whose order key is `kh` (+inf for keys at or past +inf).
__device__ inline float key_hi_dist(uint32_t kh) {
    return kh >= 0xFF800000u ? __builtin_huge_valf() : key_dist(uint64_t(kh) << 32);
}

// Candidates carry (distance key, list position) while scanning, so the hot
// loop issues no id loads.  Two ways to shrink the buffer (all threads call
// with the same arguments; only LDS is exchanged, so the barriers are
// LDS-only and code loads of later steps stay in flight):
//
// compact_dist: keep every candidate whose distance is <= the kk-th smallest
// distance.  Positions stay unresolved: ids only order keys of equal
// distance, and all of those are kept.  Returns the new count and sets
// *thr_f to the largest kept distance.
template <int CAP>
__device__ int compact_dist(ScanCtl &C, uint64_t *cand, int c, int kk, float *thr_f) {
    constexpr int PER = CAP / NT;
    const int tid = threadIdx.x;
    if (c <= kk) return c;
    const SelectResultT<uint32_t> r =
        block_select_lds<NT, uint32_t>([&](int i) { return uint32_t(cand[i] >> 32); }, c, kk, C.sel);
    uint64_t mine[PER];
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const int i = tid + e * NT;
        mine[e] = i < c ? cand[i] : kNoKey;
    }
    if (tid == 0) {
        C.newcnt = 0;
        C.maxhi = 0;
    }
    lds_barrier();
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const uint64_t key = mine[e];
        if (key == kNoKey) continue;
        const uint32_t h = uint32_t(key >> 32);
        if (r.eq > 0 ? h <= r.thr : h < r.thr) {  // eq > 0: thr is the kk-th distance itself
            cand[atomicAdd(&C.newcnt, 1)] = key;
            atomicMax(&C.maxhi, h);
        }
    }
    lds_barrier();
    *thr_f = key_hi_dist(C.maxhi);
    return C.newcnt;
}

// compact: cand[conv..c) become (distance, id) keys, then cand[0..c) shrinks
// to exactly the kk smallest keys.  Returns the new count; *thr receives the
// selection threshold.
template <int CAP>
__device__ int compact(ScanCtl &C, uint64_t *cand, int c, int conv, int kk,
                       const int32_t *__restrict__ list_ids, uint64_t *thr) {
    constexpr int PER = CAP / NT;
    const int tid = threadIdx.x;
    for (int i = conv + tid; i < c; i += NT) {
        const uint64_t k = cand[i];
        cand[i] = (k & 0xFFFFFFFF00000000ull) | uint32_t(list_ids[uint32_t(k)]);
    }
    lds_barrier();
    if (c <= kk) return c;
    const SelectResultT<uint64_t> r = block_select_lds<NT, uint64_t>([&](int i) { return cand[i]; }, c, kk, C.sel);
    uint64_t mine[PER];
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const int i = tid + e * NT;
        mine[e] = i < c ? cand[i] : kNoKey;
    }
    if (tid == 0) {
        C.newcnt = 0;
        C.eqt = 0;
    }
    lds_barrier();
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const uint64_t key = mine[e];
        if (key == kNoKey) continue;
        bool take = key < r.thr;
        if (!take && r.eq > 0 && key == r.thr) take = atomicAdd(&C.eqt, 1) < r.eq;
        if (take) {
            const int pos = atomicAdd(&C.newcnt, 1);
            if (pos < CAP) cand[pos] = key;
        }
    }
    lds_barrier();
    *thr = r.thr;
    return min(C.newcnt, CAP);
}

// Branch-free code loads: lanes past the end re-read the last vector (the
// device code buffer carries tail padding for empty lists).  lc is
// wave-uniform; the lane offset is 32-bit.
template <int M, int U>
__device__ inline void load_step(const uint8_t *lc, int base, int len, uint32_t (&cw)[U][CodeVec<M>::W]) {
    const int last = len > 0 ? len - 1 : 0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const uint32_t v = uint32_t(min(base + u * NT + int(threadIdx.x), last));
        CodeVec<M>::load(lc + v * uint32_t(M), cw[u]);
    }
}

typedef __attribute__((address_space(3))) const float lds_cfloat;

// LDS read at an integer LDS byte address; with a compile-time row offset
// the byte extraction becomes one SDWA shift and the row an immediate.
__device__ inline float lds_at(uint32_t addr) {
    return *reinterpret_cast<lds_cfloat *>(uintptr_t(addr));
}

// ADC of U vectors against the LDS table at byte address tbase:
// acc = 0f; acc = fl32(acc + T[s][c_s]), s ascending.  No table entry is -0
// (an entry is fl32(-2 dot + c) with c >= +0, and an exact zero sum rounds to
// +0), so 0f + T[0][c_0] = T[0][c_0] and the sum starts from the first entry.
// Wavefronts whose vectors all lie past the list end (wave0 + u*NT >= len)
// skip the gathers.
template <int M, int U>
__device__ inline void adc_step(const uint32_t (&cw)[U][CodeVec<M>::W], uint32_t tbase, int wave_v0,
                                int len, float (&acc)[U]) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
        acc[u] = __builtin_huge_valf();
        if (wave_v0 + u * NT >= len) continue;  // wave-uniform
        float s = 0.0f;
#pragma unroll
        for (int w = 0; w < CodeVec<M>::W; ++w) {
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const uint32_t off4 = ((cw[u][w] >> (8 * b)) & 255u) << 2;
                const float t = lds_at(tbase + off4 + uint32_t((w * 4 + b) * 1024));
                s = (w == 0 && b == 0) ? t : s + t;
            }
        }
        acc[u] = s;
    }
}

// The probed lists of one workgroup, in coarse order, skipping invalid ids.
struct ProbeSeq {
    const ScanArgs &a;
    int q, p, p1;
    __device__ int next() {  // next valid probe index >= p, or p1
        while (p < p1) {
            const int l = a.probes[int64_t(q) * a.wstride + p];
            if (l >= 0 && l < a.nlist) return p;
            ++p;
        }
        return p1;
    }
};

// rres: residual rows, row (q * wstride + p) * D; rnd: f64 sub-vector norms
// of the same rows, row (q * wstride + p) * M.  Both are read at uniform
// addresses (scalar loads).
template <int M, int KM>
__global__ __launch_bounds__(NT, 6) void scan_kernel(ScanArgs a, const float *__restrict__ rres,
                                                     const double *__restrict__ rnd) {
    constexpr int DS = D / M;
    constexpr int HM = M / 2;
    constexpr int CW = CodeVec<M>::W;
    constexpr int U = 3;                 // vectors per thread per step
    constexpr int STEP = NT * U;
    constexpr int CAP = KM <= 128 ? 2048 : 4096;
    static_assert(CAP - STEP >= KM, "candidate buffer too small");
    using L = Lay<M, CAP>;
    extern __shared__ __align__(16) unsigned char smem[];
    float *tab = reinterpret_cast<float *>(smem + L::tab);  // written here, read through lds_at()
    uint64_t *cand = reinterpret_cast<uint64_t *>(smem + L::cand);
    ScanCtl &C = *reinterpret_cast<ScanCtl *>(smem + L::ctl);

    const int tid = threadIdx.x;
    // The kernel declares no static LDS, so its dynamic block starts at LDS
    // address 0; the gathers below address the tables absolutely.  If that
    // ever stops holding, report it (the host checks) instead of reading
    // wrong tables.
    if (uint32_t(uintptr_t((__attribute__((address_space(3))) unsigned char *)(smem))) != 0) {
        if (tid == 0) atomicOr(a.err, 1);
        return;
    }
    // Workgroups go to the XCDs round robin (block b to XCD b mod 8).  With
    // a query order, XCD x takes a contiguous run of it, so the workgroups
    // resident on one XCD scan neighbouring queries, whose probed lists
    // overlap and stay in that XCD's L2.
    int b = blockIdx.x;
    if (a.qorder) {
        const int per = int(gridDim.x) >> 3;
        if (b < per * 8) b = (b & 7) * per + (b >> 3);
    }
    const int qs = b / a.ngroups;
    const int g = b - qs * a.ngroups;
    const int q = a.qorder ? a.qorder[qs] : qs;
    const int wq = a.wq ? a.wq[q] : a.w;
    const int kq = a.kq ? a.kq[q] : a.k;
    const int p0 = g * a.ppg;
    const int p1 = min(wq, p0 + a.ppg);
    const int j = tid & 255;
    const int h = __builtin_amdgcn_readfirstlane(tid >> 8);  // wave-uniform half index
    const bool jok = j < a.ks;  // table rows past ks are never indexed by valid codes

    if (tid < 3) C.inc[tid] = 0;
    if (tid == 0) C.seed = 0u;
    int fill = 0;      // candidates held at the start of a step (uniform)
    int conv = 0;      // cand[0..conv) already carry ids
    int step = 0;      // global step counter (uniform)
    uint64_t thr = kNoKey;
    float thr_f = __builtin_huge_valf();  // distances <= thr_f may enter the top-k
    __syncthreads();

    ProbeSeq seq{a, q, p0, p1};
    constexpr int RD = 2;  // register ring depth: loads run RD-1 steps ahead
    uint32_t ring[RD][U][CW];
    for (;;) {
        // ---- pick up to TB probes
        int pp[TB], ll[TB], len[TB];
        int64_t beg[TB];
        int nb = 0;
#pragma unroll
        for (int t = 0; t < TB; ++t) {
            pp[t] = seq.next();
            ll[t] = 0;
            beg[t] = 0;
            len[t] = 0;
            if (pp[t] < p1) {
                ll[t] = a.probes[int64_t(q) * a.wstride + pp[t]];
                beg[t] = a.list_off[ll[t]];
                len[t] = int(a.list_off[ll[t] + 1] - beg[t]);
                seq.p = pp[t] + 1;
                nb = t + 1;
            }
        }
        if (nb == 0) break;

        // residual rows of the batch (a missing second probe repeats the first)
        int64_t row[TB];
#pragma unroll
        for (int t = 0; t < TB; ++t) row[t] = int64_t(q) * a.wstride + pp[t < nb ? t : 0];
        // flat step sequence over the batch's lists: step k covers STEP
        // vectors of list 0 (k < ns0) or list 1; loads run RD-1 steps ahead
        const int ns0 = (len[0] + STEP - 1) / STEP;
        const int S = ns0 + (nb > 1 ? (len[1] + STEP - 1) / STEP : 0);
        auto issue = [&](int k, uint32_t (&dst)[U][CW]) {
            // unconditional (clamped) so the compiler tracks the loads exactly
            const int kc = min(k, S > 0 ? S - 1 : 0);
            const int t1 = kc >= ns0;
            load_step<M, U>(a.codes + beg[t1] * M, (kc - (t1 ? ns0 : 0)) * STEP, len[t1], dst);
        };

        // ---- TB tables: thread (j, h) streams P[s][j] of its subquantizers
        //      from L2 once for the whole batch, PD subquantizers ahead
        // lane offset kept opaque so the per-subquantizer addresses are not
        // hoisted out of the batch loop into long-lived 64-bit registers
        unsigned jj = jok ? unsigned(j) : 0u;
        asm volatile("" : "+v"(jj));
        float pr[PD + 1][DS];
        double pn[PD + 1];
        auto pload = [&](int s2) {
            const int s = h * HM + s2;
            const float *src = a.pq + size_t(s) * a.ks * DS;  // wave-uniform row base
#pragma unroll
            for (int i = 0; i < DS; ++i) pr[s2 % (PD + 1)][i] = src[jj * DS + i];
            pn[s2 % (PD + 1)] = a.pq_norm[size_t(s) * a.ks + jj];
        };
#pragma unroll
        for (int s2 = 0; s2 < PD && s2 < HM; ++s2) pload(s2);
#pragma unroll
        for (int s2 = 0; s2 < HM; ++s2) {
            if (s2 + PD < HM) pload(s2 + PD);
            const int s = h * HM + s2;
            const float *p = pr[s2 % (PD + 1)];
#pragma unroll
            for (int t = 0; t < TB; ++t) {
                // residual sub-vector and its norm: uniform (scalar) loads
                const float *rv = rres + row[t] * D + s * DS;
                // dot = 0f; dot = fl32(dot + fl32(r_i p_i)).  Starting from
                // the first product instead of 0f can only change the sign
                // of an all-zero dot, which the entry below does not see:
                // fl32(-2 (+-0) + c) = fl32(c) for c >= +0.
                float dot = rv[0] * p[0];
#pragma unroll
                for (int i = 1; i < DS; ++i) {
                    const float prod = rv[i] * p[i];
                    dot = dot + prod;
                }
                const float c = float(pn[s2 % (PD + 1)] + rnd[row[t] * M + s]);
                // fl32(fl32(-2*dot) + c): the product by -2 is exact, so one
                // fused rounding gives the same bits
                const float e = __builtin_fmaf(-2.0f, dot, c);
                tab[(t * M + s) * 256 + j] = jok ? e : __builtin_huge_valf();
            }
            __builtin_amdgcn_sched_barrier(0);  // PD + 1 subquantizers' operands live at a time
        }
        // the first code steps are issued after the P loads: loads complete in
        // order, so issued earlier they would hold up every table wait
#pragma unroll
        for (int r = 0; r + 1 < RD; ++r) issue(r, ring[r]);
        lds_barrier();  // tables complete; the code loads stay in flight

        // ---- scan the batch's lists back to back; the prefetch runs one
        //      step ahead, across list boundaries
        for (int k0 = 0; k0 < S; k0 += RD) {
#pragma unroll
            for (int r = 0; r < RD; ++r) {
                const int k = k0 + r;
                if (k >= S) break;
                issue(k + RD - 1, ring[(r + RD - 1) % RD]);
                const int t1 = k >= ns0;
                const int base = (k - (t1 ? ns0 : 0)) * STEP;
                const int ln = t1 ? len[1] : len[0];
                const int64_t bg = t1 ? beg[1] : beg[0];
                const int rot = step % 3;
                if (tid == 0) C.inc[(step + 1) % 3] = 0;
                // the dynamic LDS block starts at address 0 (checked at entry),
                // so every table address is a compile-time constant and a
                // gather is one SDWA byte shift plus a ds_read
                float accs[U];
                // The two calls differ only in the table base.  The distinct
                // empty asm statements stop the optimizer from merging their
                // common tails into one block with a run-time base, which
                // would cost a second VALU instruction per gather there.
                if (t1) {
                    adc_step<M, U>(ring[r], uint32_t(L::tab) + uint32_t(M * 256 * 4), base + (tid & ~63), ln, accs);
                    asm volatile("; second table" ::);
                } else {
                    adc_step<M, U>(ring[r], uint32_t(L::tab), base + (tid & ~63), ln, accs);
                    asm volatile("; first table" ::);
                }
                // First step of the workgroup, threshold still +inf: every
                // wavefront finds the ceil(kq / 8)-th smallest of its own
                // distances by a bitwise search over ballot counts (no LDS,
                // no selection pass); the largest of the eight bounds the
                // kq-th smallest distance of the step, so only the distances
                // within it are appended, instead of the whole step.
                if (step == 0) {  // workgroup-uniform
                    const int need = (kq + NT / 64 - 1) / (NT / 64);
                    uint32_t key[U];
#pragma unroll
                    for (int u = 0; u < U; ++u)
                        key[u] = base + u * NT + tid < ln ? fkey(accs[u]) : 0xFFFFFFFFu;
                    uint32_t ans = 0u;  // need-th smallest key of the wavefront (all ones: fewer keys)
                    for (int bit = 31; bit >= 0; --bit) {
                        const uint32_t t = ans | ((1u << bit) - 1u);
                        int cnt = 0;
#pragma unroll
                        for (int u = 0; u < U; ++u) cnt += __popcll(__ballot(key[u] <= t));
                        if (cnt < need) ans |= 1u << bit;
                    }
                    if (__lane_id() == 0) atomicMax(&C.seed, ans);
                    lds_barrier();
                    thr_f = key_hi_dist(C.seed);
                }
                // ties with the threshold distance pass too; ids are compared
                // exactly at the final selection.  One LDS atomic per
                // wavefront and step reserves the slots of all U vectors.
                uint64_t mask[U];
                int tot = 0;
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int v = base + u * NT + tid;
                    mask[u] = __ballot((v < ln) && accs[u] <= thr_f);
                    tot += __popcll(mask[u]);
                }
                if (tot) {  // wave-uniform
                    const int lane = __lane_id();
                    const uint64_t below = (uint64_t(1) << lane) - 1;
                    int pos = 0;
                    if (lane == 0) pos = atomicAdd(&C.inc[rot], tot);
                    pos = fill + __builtin_amdgcn_readfirstlane(pos);  // lane 0's reservation
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const int at = pos + __popcll(mask[u] & below);
                        if (((mask[u] >> lane) & 1) && at < CAP) {
                            const int v = base + u * NT + tid;
                            cand[at] = (uint64_t(fkey(accs[u])) << 32) | uint32_t(bg + v);
                        }
                        pos += __popcll(mask[u]);
                    }
                }
                // LDS-only: the code loads of the next steps stay in flight
                lds_barrier();
                fill += C.inc[rot];
                ++step;
                if (fill > CAP - STEP) {
                    // distance-only while no id is resolved; the exact form
                    // when the kept set leaves no room for a step (many equal
                    // distances) or ids already are
                    const int nf = conv == 0 ? compact_dist<CAP>(C, cand, fill, kq, &thr_f) : CAP;
                    if (nf > CAP - STEP) {
                        fill = compact<CAP>(C, cand, conv == 0 ? nf : fill, conv, kq, a.list_ids, &thr);
                        conv = fill;
                        thr_f = key_hi_dist(uint32_t(thr >> 32));
                    } else {
                        fill = nf;
                    }
                }
            }
        }
        lds_barrier();  // the next batch rewrites the tables
    }

    // final selection: a distance-only shrink first (no id loads), then the
    // remaining positions become ids and, if equal distances leave more than
    // kq, the exact (distance, id) selection
    if (conv == 0 && fill > kq) {
        const int nf = compact_dist<CAP>(C, cand, fill, kq, &thr_f);
        fill = nf;
    }
    fill = compact<CAP>(C, cand, fill, conv, kq, a.list_ids, &thr);
    const int c = fill;
    if (a.ngroups > 1) {
        uint64_t *dst = a.partial + (int64_t(q) * a.ngroups + g) * a.kstride;
        for (int i = tid; i < a.kstride; i += NT) dst[i] = i < c ? cand[i] : kNoKey;
        return;
    }
    int32_t *oi = a.out_ids + int64_t(q) * a.kstride;
    float *od = a.out_dis + int64_t(q) * a.kstride;
    if (c <= NT) {
        // key i goes to its rank: the count of smaller keys, plus equal keys
        // before it (equal keys only come from repeated ids); every thread
        // reads the keys in the same order
        if (tid < c) {
            const uint64_t key = cand[tid];
            int rank = 0;
            for (int j = 0; j < c; ++j) {
                const uint64_t o = cand[j];
                rank += o < key || (o == key && j < tid);
            }
            oi[rank] = key_id(key);
            od[rank] = key_dist(key);
        }
        for (int i = c + tid; i < a.kstride; i += NT) {
            oi[i] = -1;
            od[i] = __builtin_huge_valf();
        }
        return;
    }
    const int np = next_pow2(c > 0 ? c : 1);
    for (int i = c + tid; i < np; i += NT) cand[i] = kNoKey;
    lds_barrier();
    block_bitonic_sort<NT, true>(cand, np);
    for (int i = tid; i < a.kstride; i += NT) {
        if (i < c) {
            oi[i] = key_id(cand[i]);
            od[i] = key_dist(cand[i]);
        } else {
            oi[i] = -1;
            od[i] = __builtin_huge_valf();
        }
    }
}

constexpr int QO_NT = 1024;
constexpr int QO_MAXL = 16384;  // LDS histogram bins

__global__ __launch_bounds__(QO_NT) void query_order_kernel(const int32_t *__restrict__ probes, int wstride,
                                                            int nq, int nlist, int32_t *__restrict__ order) {
    extern __shared__ int hist[];  // [nlist]
    __shared__ int wsum[QO_NT / 64];
    const int tid = threadIdx.x;
    for (int i = tid; i < nlist; i += QO_NT) hist[i] = 0;
    __syncthreads();
    auto bucket = [&](int q) { return min(max(probes[int64_t(q) * wstride], 0), nlist - 1); };
    for (int q = tid; q < nq; q += QO_NT) atomicAdd(&hist[bucket(q)], 1);
    __syncthreads();
    // exclusive scan: thread t owns bins [t*per, (t+1)*per)
    const int per = (nlist + QO_NT - 1) / QO_NT;
    const int b0 = min(tid * per, nlist), b1 = min(b0 + per, nlist);
    int loc = 0;
    for (int i = b0; i < b1; ++i) loc += hist[i];
    int incl = loc;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int o = __shfl_up(incl, off, 64);
        if ((tid & 63) >= off) incl += o;
    }
    if ((tid & 63) == 63) wsum[tid >> 6] = incl;
    __syncthreads();
    int base = 0;
    for (int w = 0; w < (tid >> 6); ++w) base += wsum[w];
    int run = base + incl - loc;
    for (int i = b0; i < b1; ++i) {
        const int c = hist[i];
        hist[i] = run;
        run += c;
    }
    __syncthreads();
    for (int q = tid; q < nq; q += QO_NT) order[atomicAdd(&hist[bucket(q)], 1)] = q;
}

// Residual rows r = fl32(q - c_l) (unless the caller passed residuals) and
// their f32 sub-vector norms, summed in dimension order and stored as f64,
// for every valid (query, probe) pair.  One 256-thread block per query: each
// half takes every second probe, one dimension per thread, eight probes'
// loads at a time.
template <int M>
__global__ __launch_bounds__(256) void resid_kernel(ScanArgs a, float *__restrict__ rres, double *__restrict__ rnd) {
    constexpr int DS = D / M;
    constexpr int RB = 8;  // probes per half and batch: their loads are issued together
    static_assert(DS <= 64, "a sub-vector must fit in one wavefront");
    const int q = blockIdx.x;
    const int i = threadIdx.x & (D - 1);
    const int half = threadIdx.x >> 7;
    const int wq = a.wq ? a.wq[q] : a.w;
    const float qi = a.resid_mode ? 0.0f : a.queries[int64_t(q) * D + i];
    for (int p0 = half; p0 < wq; p0 += 2 * RB) {  // whole wavefronts
        int lid[RB];
        float val[RB];
#pragma unroll
        for (int b = 0; b < RB; ++b) {
            const int p = p0 + 2 * b;
            lid[b] = p < wq ? a.probes[int64_t(q) * a.wstride + p] : -1;
        }
#pragma unroll
        for (int b = 0; b < RB; ++b) {
            const int p = p0 + 2 * b;
            const bool ok = lid[b] >= 0 && lid[b] < a.nlist;
            val[b] = !ok ? 0.0f
                         : (a.resid_mode ? a.queries[(int64_t(q) * a.wstride + p) * D + i]
                                         : a.coarse[int64_t(lid[b]) * D + i]);
        }
#pragma unroll
        for (int b = 0; b < RB; ++b) {
            if (lid[b] < 0 || lid[b] >= a.nlist) continue;  // wave-uniform
            const int64_t pair = int64_t(q) * a.wstride + p0 + 2 * b;
            float r;
            if (a.resid_mode) {
                r = val[b];
            } else {
                r = qi - val[b];
                rres[pair * D + i] = r;
            }
            // f32 norm of each sub-vector, summed in dimension order
            const float sq = r * r;
            const int g0 = (i & 63) & ~(DS - 1);
            float acc = 0.0f;
#pragma unroll
            for (int t = 0; t < DS; ++t) {
                const float v = __shfl(sq, g0 + t, 64);
                acc = acc + v;
            }
            if ((i & (DS - 1)) == 0) rnd[pair * M + i / DS] = double(acc);
        }
    }
}

template <int M, int KM>
hipError_t launch_mk(const ScanArgs &a, hipStream_t s) {
    constexpr int CAP = KM <= 128 ? 2048 : 4096;
    constexpr size_t lds = Lay<M, CAP>::total;
    static bool attr_set = false;  // raise the dynamic-LDS limit once per instantiation
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&scan_kernel<M, KM>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    const unsigned grid = unsigned(int64_t(a.nq) * a.ngroups);
    scan_kernel<M, KM><<<grid, NT, lds, s>>>(a, a.resid_mode ? a.queries : a.rres, a.rnd);
    return hipGetLastError();
}

template <int M>
hipError_t launch_m(const ScanArgs &a, hipStream_t s) {
    return a.kstride <= 128 ? launch_mk<M, 128>(a, s) : launch_mk<M, 1024>(a, s);
}

}  // namespace

bool query_order_ok(int nlist) { return nlist >= 1 && nlist <= QO_MAXL; }

hipError_t launch_query_order(const int32_t *probes, int wstride, int nq, int nlist, int32_t *order,
                              hipStream_t s) {
    if (nq <= 0) return hipSuccess;
    if (!query_order_ok(nlist)) return hipErrorInvalidValue;
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&query_order_kernel),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, int(QO_MAXL * sizeof(int)));
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    query_order_kernel<<<1, QO_NT, size_t(nlist) * sizeof(int), s>>>(probes, wstride, nq, nlist, order);
    return hipGetLastError();
}

hipError_t launch_scan_prologue(const ScanArgs &a, int m, hipStream_t s) {
    if (a.nq <= 0) return hipSuccess;
    if (!a.rnd || (!a.resid_mode && !a.rres)) return hipErrorInvalidValue;
    const unsigned rgrid = unsigned(a.nq);
    switch (m) {
        case 4: resid_kernel<4><<<rgrid, 256, 0, s>>>(a, a.rres, a.rnd); break;
        case 8: resid_kernel<8><<<rgrid, 256, 0, s>>>(a, a.rres, a.rnd); break;
        case 16: resid_kernel<16><<<rgrid, 256, 0, s>>>(a, a.rres, a.rnd); break;
        case 32: resid_kernel<32><<<rgrid, 256, 0, s>>>(a, a.rres, a.rnd); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_scan(const ScanArgs &a, int m, hipStream_t s) {
    if (a.nq <= 0) return hipSuccess;
    if (a.kstride > 1024) return hipErrorInvalidValue;
    if (!a.rnd || (!a.resid_mode && !a.rres)) return hipErrorInvalidValue;
    switch (m) {
        case 4: return launch_m<4>(a, s);
        case 8: return launch_m<8>(a, s);
        case 16: return launch_m<16>(a, s);
        case 32: return launch_m<32>(a, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace mivf
