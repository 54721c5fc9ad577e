// Per-particle device arithmetic shared by the fused single-bunch turn and
// the batched ensemble turn: RF kick, drift, and the workgroup-private LDS
// histogram (wavefront-aggregated increments, flushed to global counts).
#pragma once

#include <hip/hip_runtime.h>

#include "blmi_common.h"
#include "particles.h"

namespace blmi {

// ---- shared per-particle arithmetic (same order as the single kernels) ----
template <typename T>
__device__ __forceinline__ T e_rf_kick(T t, int n_rf, T charge, const T *__restrict__ V,
                                       const T *__restrict__ w, const T *__restrict__ phi, T acc) {
    T s = T(0);
    for (int j = 0; j < n_rf; ++j) s = s + V[j] * Trig<T>::sin(w[j] * t + phi[j]);
    return charge * s + acc;
}

template <typename T>
__device__ __forceinline__ T e_drift(const DriftCoeffs<T> &c, T e, int solver) {
    if (solver == 0) return c.coeff * e;
    if (solver == 1) {
        T den;
        if (c.alpha_order == 0)
            den = T(1) - c.e0 * e;
        else if (c.alpha_order == 1)
            den = T(1) - c.e0 * e - c.e1 * e * e;
        else
            den = T(1) - c.e0 * e - c.e1 * e * e - c.e2 * e * e * e;
        return c.tl * (T(1) / den - T(1));
    }
    const T delta = sqrt(T(1) + c.ib2 * (e * e * c.ie2 + T(2) * e / c.energy)) - T(1);
    return c.tl * ((T(1) + c.a0 * delta + c.a1 * (delta * delta) + c.a2 * (delta * delta * delta)) *
                       (T(1) + e / c.energy) / (T(1) + delta) -
                   T(1));
}

// Count `bin` (>= 0; < 0 = no count) into the LDS histogram.  Lanes of the
// wavefront holding the same bin are merged first: up to max_peel times the
// first pending lane's bin is broadcast, the lanes sharing it are counted
// with one LDS atomic, then the remaining lanes add one each.  A narrow
// bunch (few distinct bins per wavefront) thus costs a handful of atomics
// instead of 64 conflicting ones; a wide one costs at most max_peel extra
// scalar steps.  PACK16: bin b is the (b & 1) half of word b >> 1.
template <bool PACK16>
__device__ __forceinline__ void lds_count(unsigned *__restrict__ h, int bin, int max_peel) {
    unsigned long long pending = __ballot(bin >= 0);
    const int lane = (int)__lane_id();
    for (int it = 0; it < max_peel && pending != 0ull; ++it) {
        const int leader = __builtin_ctzll(pending);
        const int lb = __builtin_amdgcn_readlane(bin, leader);
        const unsigned long long same = __ballot(bin == lb) & pending;
        if (lane == leader) {
            const unsigned c = (unsigned)__popcll(same);
            if (PACK16)
                atomicAdd(&h[lb >> 1], c << (16 * (lb & 1)));
            else
                atomicAdd(&h[lb], c);
        }
        pending &= ~same;
    }
    if ((pending >> lane) & 1ull) {
        if (PACK16)
            atomicAdd(&h[bin >> 1], 1u << (16 * (bin & 1)));
        else
            atomicAdd(&h[bin], 1u);
    }
}

// LDS words of a histogram of n_slices bins in mode 1 (u32) or 2 (u16
// pairs, padded to a multiple of 4 words for 16-byte copies)
__host__ __device__ __forceinline__ int hist_words(int mode, int n_slices) {
    return mode == 2 ? (((n_slices + 1) / 2 + 3) & ~3) : n_slices;
}

// Add the workgroup's LDS histogram into the global counts (non-zero bins).
template <int LDSH>
__device__ __forceinline__ void hist_flush(const unsigned *__restrict__ h,
                                           unsigned *__restrict__ counts, int n_slices) {
    const int words = hist_words(LDSH, n_slices);
    for (int w = threadIdx.x; w < words; w += blockDim.x) {
        const unsigned c = h[w];
        if (!c) continue;
        if (LDSH == 1) {
            atomic_add_agent(&counts[w], c);
        } else {
            const unsigned lo = c & 0xffffu, hi = c >> 16;
            if (lo) atomic_add_agent(&counts[2 * w], lo);
            if (hi) atomic_add_agent(&counts[2 * w + 1], hi);
        }
    }
}

// profile[b] = sum over groups of c[g][b], and the counts cleared, for b in [b0, b1)
__device__ __forceinline__ void publish_range(unsigned *__restrict__ c, double *__restrict__ profile,
                                              int n_slices, int groups, int b0, int b1) {
    for (int b = b0 + (int)threadIdx.x; b < b1; b += blockDim.x) {
        unsigned s = 0;
        for (int g = 0; g < groups; ++g) {
            s += c[(size_t)g * n_slices + b];
            c[(size_t)g * n_slices + b] = 0u;
        }
        profile[b] = (double)s;
    }
}

}  // namespace blmi
