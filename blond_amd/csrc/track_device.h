// Per-particle device arithmetic shared by the fused single-bunch turn and
// the batched ensemble turn: RF kick, drift, and the workgroup-private LDS
// histogram (wavefront-aggregated increments, flushed to global counts).
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

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

// The RF programs of one turn, held in registers for a whole launch: the
// first kRfRegs systems are loaded once and the loop over systems is
// unrolled (kick_n: for a count known at compile time, no guard at all), so
// no load or wait precedes a particle's sines; further systems are read
// from the table.  The same sum in the same order as e_rf_kick
// (kick.cpp:27-33).
constexpr int kRfRegs = 4;
template <typename T>
struct RfProg {
    T V[kRfRegs], w[kRfRegs], phi[kRfRegs];
    const T *Vt, *wt, *pt;
    int n;
    __device__ __forceinline__ RfProg(const T *V_, const T *w_, const T *phi_, int n_)
        : Vt(V_), wt(w_), pt(phi_), n(n_) {
#pragma unroll
        for (int j = 0; j < kRfRegs; ++j) {
            V[j] = w[j] = phi[j] = T(0);
            if (j < n) {
                V[j] = V_[j];
                w[j] = w_[j];
                phi[j] = phi_[j];
            }
        }
    }
    __device__ __forceinline__ T kick(T t, T charge, T acc) const {
        T s = T(0);
#pragma unroll
        for (int j = 0; j < kRfRegs; ++j)
            if (j < n) s = s + V[j] * Trig<T>::sin(w[j] * t + phi[j]);
        for (int j = kRfRegs; j < n; ++j) s = s + Vt[j] * Trig<T>::sin(wt[j] * t + pt[j]);
        return charge * s + acc;
    }
    // Moves the first NRF systems into vector registers for the loop that
    // follows: they are uniform, but the scalar file is taken by the sine's
    // constants and the loop state, and spilled scalars cost a v_readlane
    // per use inside the particle loop.
    template <int NRF>
    __device__ __forceinline__ void pin() {
#pragma unroll
        for (int j = 0; j < (NRF < 0 ? 0 : NRF); ++j) {
            asm volatile("" : "+v"(V[j]));
            asm volatile("" : "+v"(w[j]));
            asm volatile("" : "+v"(phi[j]));
        }
    }
    // NRF >= 0: the caller has checked that exactly NRF (<= kRfRegs) systems
    // are present, so the sum is unrolled with no guard at all; NRF < 0: the
    // count is read at run time as in kick().
    template <int NRF>
    __device__ __forceinline__ T kick_n(T t, T charge, T acc) const {
        static_assert(NRF <= kRfRegs, "NRF exceeds the systems held in registers");
        if constexpr (NRF < 0) {
            return kick(t, charge, acc);
        } else {
            T s = T(0);
#pragma unroll
            for (int j = 0; j < NRF; ++j) s = s + V[j] * Trig<T>::sin(w[j] * t + phi[j]);
            return charge * s + acc;
        }
    }
    // kick_n with the sines of cephes_sin_tab (coefficients in LDS):
    // bit-identical, about 13 fewer FP64 operations a sine
    template <int NRF>
    __device__ __forceinline__ T kick_tab(T t, T charge, T acc, LdsConstDoublePtr tab) const {
        static_assert(NRF <= kRfRegs, "NRF exceeds the systems held in registers");
        if constexpr (NRF < 0) {
            T s = T(0);
#pragma unroll
            for (int j = 0; j < kRfRegs; ++j)
                if (j < n) s = s + V[j] * TrigTab<T>::sin(w[j] * t + phi[j], tab);
            for (int j = kRfRegs; j < n; ++j) s = s + Vt[j] * TrigTab<T>::sin(wt[j] * t + pt[j], tab);
            return charge * s + acc;
        } else {
            T s = T(0);
#pragma unroll
            for (int j = 0; j < NRF; ++j) s = s + V[j] * TrigTab<T>::sin(w[j] * t + phi[j], tab);
            return charge * s + acc;
        }
    }
};

// Calls f(std::integral_constant<int, NRF>) with NRF = n_rf for the counts
// that have a specialised particle loop (0, 1, 2) and NRF = -1 otherwise.
// The dispatch is uniform and happens once per launch, outside the loop.
template <typename F>
__device__ __forceinline__ void with_rf_count(int n_rf, F &&f) {
    switch (n_rf) {
        case 0: f(std::integral_constant<int, 0>{}); break;
        case 1: f(std::integral_constant<int, 1>{}); break;
        case 2: f(std::integral_constant<int, 2>{}); break;
        default: f(std::integral_constant<int, -1>{}); break;
    }
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
// The LDS word of u16-pair word w: with pad_shift < 31 one pad word follows
// every 2^pad_shift words (the transposed partial-row store then reads at an
// odd stride, on distinct banks); pad_shift = 31 is the plain layout.
__host__ __device__ __forceinline__ int lds_word(int w, int pad_shift) { return w + (w >> pad_shift); }

// One count of bin (>= 0) into the LDS histogram, no merging.
template <bool PACK16>
__device__ __forceinline__ void lds_add1(unsigned *__restrict__ h, int bin, int pad_shift) {
    if (PACK16)
        atomicAdd(&h[lds_word(bin >> 1, pad_shift)], 1u << (16 * (bin & 1)));
    else
        atomicAdd(&h[bin], 1u);
}

template <bool PACK16>
__device__ __forceinline__ void lds_count(unsigned *__restrict__ h, int bin, int max_peel,
                                          int pad_shift = 31) {
    if (max_peel <= 0) {
        // no merging (a uniform branch): each counted lane adds its one,
        // without the ballot and the pending-mask test
        if (bin >= 0) lds_add1<PACK16>(h, bin, pad_shift);
        return;
    }
    unsigned long long pending = __ballot(bin >= 0);
    const int lane = (int)__lane_id();
    for (int it = 0; it < max_peel && pending != 0ull; ++it) {
        const int leader = __builtin_ctzll(pending);
        const int lb = __builtin_amdgcn_readlane(bin, leader);
        const unsigned long long same = __ballot(bin == lb) & pending;
        if (lane == leader) {
            const unsigned c = (unsigned)__popcll(same);
            if (PACK16)
                atomicAdd(&h[lds_word(lb >> 1, pad_shift)], c << (16 * (lb & 1)));
            else
                atomicAdd(&h[lb], c);
        }
        pending &= ~same;
    }
    if ((pending >> lane) & 1ull) lds_add1<PACK16>(h, bin, pad_shift);
}

// A workgroup's u16-pair row stored transposed for the first pass of the
// four-step transform (whose column m2 takes the words m = M2 m1 + m2):
// word m of row `row` goes to P[(m2 rows + row) L + m1], M2 = 2^shift,
// L = 2^lshift words per column, L * M2 words in all.  Lanes take
// consecutive m1: the stores are contiguous, and the LDS reads (padded
// layout, stride M2 + 1 words) fall on distinct banks.
__device__ __forceinline__ void store_partial_perm(const unsigned *__restrict__ h,
                                                   unsigned *__restrict__ P, int row, int rows,
                                                   int shift, int lshift) {
    const int L = 1 << lshift, W = L << shift;
    for (int idx = threadIdx.x; idx < W; idx += blockDim.x) {
        const int m2 = idx >> lshift, m1 = idx & (L - 1);
        const int m = (m1 << shift) + m2;
        P[((size_t)m2 * rows + row) * L + m1] = h[lds_word(m, shift)];
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
