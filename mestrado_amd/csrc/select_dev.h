// Block-wide exact k-selection over 64-bit order keys (distance, id).
//
// block_select(): radix select on 8-bit digits from the most significant end.
// Each pass builds a 256-bin LDS histogram of the keys that share the prefix
// found so far, one wavefront scans it, and the pass stops as soon as the
// bucket holding the k-th key is taken whole.  The result {thr, eq} selects
// exactly k keys: every key < thr, plus `eq` keys equal to thr (eq > 0 only
// when the input holds duplicate keys).  kNoKey entries are ignored.
#pragma once
#include "mivf_internal.h"

namespace mivf {

// Workgroup barrier that orders LDS only.  __syncthreads() also fences
// global memory, which makes every wave wait for its outstanding global
// loads (here: the code prefetch of later steps) at each barrier.  Global
// data is never exchanged between the threads of a workgroup at these
// barriers, so the LDS-scope fence is all they need.
__device__ inline void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

struct SelectResult {
    uint64_t thr;
    int eq;
};

struct SelectShared {
    uint32_t hist[256];
    uint64_t prefix;
    int need;
    int digit;
    int below;
    int inbucket;
    int count;
    int eqticket;
};

// keyfn(i) -> key of element i (i < n).  All NT threads must call.
template <int NT, typename KeyFn>
__device__ SelectResult block_select(KeyFn keyfn, int n, int k, SelectShared &sh) {
    const int tid = threadIdx.x;
    // count real keys
    if (tid == 0) sh.count = 0;
    __syncthreads();
    int local = 0;
    for (int i = tid; i < n; i += NT) local += keyfn(i) != kNoKey;
    if (local) atomicAdd(&sh.count, local);
    __syncthreads();
    const int nreal = sh.count;
    if (nreal <= k) return SelectResult{kNoKey, 0};

    uint64_t prefix = 0;
    int need = k;
    for (int shift = 56; shift >= 0; shift -= 8) {
        for (int i = tid; i < 256; i += NT) sh.hist[i] = 0;
        __syncthreads();
        for (int i = tid; i < n; i += NT) {
            uint64_t key = keyfn(i);
            if (key == kNoKey) continue;
            bool match = (shift == 56) ? true : ((key >> (shift + 8)) == (prefix >> (shift + 8)));
            if (match) atomicAdd(&sh.hist[(key >> shift) & 255], 1u);
        }
        __syncthreads();
        if (tid < 64) {
            // lane l owns bins 4l..4l+3
            uint32_t h0 = sh.hist[4 * tid], h1 = sh.hist[4 * tid + 1];
            uint32_t h2 = sh.hist[4 * tid + 2], h3 = sh.hist[4 * tid + 3];
            uint32_t tot = h0 + h1 + h2 + h3;
            uint32_t incl = tot;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                uint32_t o = __shfl_up(incl, off, 64);
                if (tid >= off) incl += o;
            }
            uint32_t excl = incl - tot;
            bool crosses = excl < uint32_t(need) && incl >= uint32_t(need);
            if (crosses) {
                uint32_t c = excl;
                int dgt;
                uint32_t hb;
                if (c + h0 >= uint32_t(need)) { dgt = 0; hb = h0; }
                else if (c + h0 + h1 >= uint32_t(need)) { c += h0; dgt = 1; hb = h1; }
                else if (c + h0 + h1 + h2 >= uint32_t(need)) { c += h0 + h1; dgt = 2; hb = h2; }
                else { c += h0 + h1 + h2; dgt = 3; hb = h3; }
                sh.digit = 4 * tid + dgt;
                sh.below = int(c);
                sh.inbucket = int(hb);
            }
        }
        __syncthreads();
        const int dgt = sh.digit;
        const int below = sh.below;
        const int inb = sh.inbucket;
        __syncthreads();
        need -= below;
        prefix |= uint64_t(dgt) << shift;
        if (inb == need) {
            uint64_t step = uint64_t(1) << shift;
            uint64_t thr = (prefix > kNoKey - step) ? kNoKey : prefix + step;
            return SelectResult{thr, 0};
        }
        if (shift == 0) return SelectResult{prefix, need};
    }
    return SelectResult{kNoKey, 0};  // not reached
}

template <typename K>
struct SelKey;
template <>
struct SelKey<uint32_t> {
    static constexpr int TOP = 24;
};
template <>
struct SelKey<uint64_t> {
    static constexpr int TOP = 56;
};

template <typename K>
struct SelectResultT {
    K thr;
    int eq;
};

// The same radix select over n keys that are all real (no kNoKey), with
// LDS-only barriers and three barriers per digit: for callers whose keys sit
// in LDS or registers and that exchange no global data around the call.
// Returns {thr, eq}: every key < thr is among the k smallest, and so are eq
// of the keys equal to thr (eq > 0 only when keys equal to thr straddle the
// k-th position).  n <= k returns {all ones, 0}.
template <int NT, typename K, typename KeyFn>
__device__ SelectResultT<K> block_select_lds(KeyFn keyfn, int n, int k, SelectShared &sh) {
    constexpr int TOP = SelKey<K>::TOP;
    const int tid = threadIdx.x;
    if (n <= k) return SelectResultT<K>{K(~K(0)), 0};
    K prefix = 0;
    int need = k;
    for (int shift = TOP; shift >= 0; shift -= 8) {
        for (int i = tid; i < 256; i += NT) sh.hist[i] = 0;
        lds_barrier();
        for (int i = tid; i < n; i += NT) {
            const K key = keyfn(i);
            const bool match = shift == TOP || (key >> (shift + 8)) == (prefix >> (shift + 8));
            if (match) atomicAdd(&sh.hist[(key >> shift) & 255], 1u);
        }
        lds_barrier();
        if (tid < 64) {
            // lane l owns bins 4l..4l+3
            const uint32_t h0 = sh.hist[4 * tid], h1 = sh.hist[4 * tid + 1];
            const uint32_t h2 = sh.hist[4 * tid + 2], h3 = sh.hist[4 * tid + 3];
            const uint32_t tot = h0 + h1 + h2 + h3;
            uint32_t incl = tot;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const uint32_t o = __shfl_up(incl, off, 64);
                if (tid >= off) incl += o;
            }
            const uint32_t excl = incl - tot;
            if (excl < uint32_t(need) && incl >= uint32_t(need)) {
                uint32_t c = excl;
                int dgt;
                uint32_t hb;
                if (c + h0 >= uint32_t(need)) { dgt = 0; hb = h0; }
                else if (c + h0 + h1 >= uint32_t(need)) { c += h0; dgt = 1; hb = h1; }
                else if (c + h0 + h1 + h2 >= uint32_t(need)) { c += h0 + h1; dgt = 2; hb = h2; }
                else { c += h0 + h1 + h2; dgt = 3; hb = h3; }
                sh.digit = 4 * tid + dgt;
                sh.below = int(c);
                sh.inbucket = int(hb);
            }
        }
        // the next digit's clear and histogram come after two more barriers,
        // so these reads need no barrier of their own behind them
        lds_barrier();
        const int dgt = sh.digit;
        const int below = sh.below;
        const int inb = sh.inbucket;
        need -= below;
        prefix |= K(dgt) << shift;
        if (inb == need) {
            const K step = K(1) << shift;
            const K thr = (prefix > K(~K(0)) - step) ? K(~K(0)) : prefix + step;
            return SelectResultT<K>{thr, 0};
        }
        if (shift == 0) return SelectResultT<K>{prefix, need};
    }
    return SelectResultT<K>{K(~K(0)), 0};  // not reached
}

// In-place ascending bitonic sort of s[0..n), n a power of two, in LDS.
// LdsOnly: LDS-only barriers (the caller exchanges no global data around it).
template <int NT, bool LdsOnly = false>
__device__ void block_bitonic_sort(uint64_t *s, int n) {
    const int tid = threadIdx.x;
    for (int size = 2; size <= n; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = tid; i < (n >> 1); i += NT) {
                int lo = 2 * i - (i & (stride - 1));
                int hi = lo + stride;
                bool up = (lo & size) == 0;
                uint64_t x = s[lo], y = s[hi];
                if ((x > y) == up) {
                    s[lo] = y;
                    s[hi] = x;
                }
            }
            if (LdsOnly)
                lds_barrier();
            else
                __syncthreads();
        }
    }
}

__device__ inline int next_pow2(int x) {
    int p = 1;
    while (p < x) p <<= 1;
    return p;
}

}  // namespace mivf
