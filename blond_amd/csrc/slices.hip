// Slicing kernels: nearest-bin histogram and the reference's two-bin
// "smooth" histogram.
//
// Histogram: each workgroup counts into a private LDS copy of the profile
// (u32 per bin), then adds its non-zero bins to the output.  Counts are
// integers, so the result is exact and independent of launch shape and
// atomic order.  Profiles too large for LDS count straight into the output
// with device-scope atomics.
//
// Smooth histogram: weights are accumulated in 64-bit fixed point (2^-S
// resolution, S chosen from N so no bin can overflow), which makes the result
// deterministic -- independent of workgroup count and atomic order -- and
// within a few 1e-13 of the exact weight sum per particle.
#include "blmi_common.h"
#include "particles.h"

namespace blmi {

// histogram.cpp:41-57 : floor((t - cut_left) * inv_bw), bins outside
// [0, n_slices) skipped.  The reference stages the floored value through a
// float before the int cast; for every in-range bin (|bin| < 2^24) that
// staging is exact, so the range test on the floored value is equivalent.
template <typename T>
__device__ __forceinline__ int hist_bin(T t, T cut_left, T inv_bw, int n_slices) {
    const T f = floor((t - cut_left) * inv_bw);
    return (f >= T(0) && f < T(n_slices)) ? (int)f : -1;
}

template <typename T, int VEC, bool LDS>
__global__ void __launch_bounds__(256) histogram_kernel(const T *__restrict__ dt, int64_t n,
                                                        T *__restrict__ out, T cut_left,
                                                        T inv_bw, int n_slices) {
    extern __shared__ unsigned lds_hist[];
    if (LDS) {
        for (int b = threadIdx.x; b < n_slices; b += blockDim.x) lds_hist[b] = 0u;
        __syncthreads();
    }
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    auto count = [&](T t) {
        const int b = hist_bin(t, cut_left, inv_bw, n_slices);
        if (b >= 0) {
            if (LDS)
                atomicAdd(&lds_hist[b], 1u);
            else
                atomic_add_agent(&out[b], T(1));
        }
    };
    if (VEC == 2) {
        using V2 = typename Vec2<T>::type;
        const int64_t n2 = n / 2;
        for (; i < n2; i += stride) {
            const V2 t = reinterpret_cast<const V2 *>(dt)[i];
            count(t.x);
            count(t.y);
        }
        if (blockIdx.x == 0 && threadIdx.x == 0 && (n & 1)) count(dt[n - 1]);
    } else {
        for (; i < n; i += stride) count(dt[i]);
    }
    if (LDS) {
        __syncthreads();
        for (int b = threadIdx.x; b < n_slices; b += blockDim.x) {
            const unsigned c = lds_hist[b];
            if (c) atomic_add_agent(&out[b], T(c));
        }
    }
}

template <typename T>
__global__ void fill_kernel(T *__restrict__ p, int64_t n, T v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

template <typename T>
int launch_fill(T *p, int64_t n, T v, hipStream_t s) {
    if (n <= 0) return BLMI_OK;
    hipLaunchKernelGGL((fill_kernel<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n,
                       v);
    return blmi_check_launch("fill");
}

static inline bool aligned16(const void *p) { return (((uintptr_t)p) & 15u) == 0; }

template <typename T>
int launch_histogram(const T *dt, int64_t n, T *profile, T cut_left, T cut_right, int n_slices,
                     hipStream_t s) {
    if (n_slices <= 0) return BLMI_OK;
    // histogram.cpp:29
    const T inv_bw = T(n_slices) / (cut_right - cut_left);
    int rc = launch_fill<T>(profile, n_slices, T(0), s);
    if (rc) return rc;
    if (n <= 0) return BLMI_OK;
    const size_t lds = (size_t)n_slices * sizeof(unsigned);
    const bool use_lds = lds <= 65536;
    const bool vec = aligned16(dt);
    int grid = stream_grid(vec ? (n + 1) / 2 : n, 256);
    const int cap = device_cu_count() * 4;
    if (grid > cap) grid = cap;
    if (use_lds) {
        if (vec)
            hipLaunchKernelGGL((histogram_kernel<T, 2, true>), dim3(grid), dim3(256), lds, s, dt, n,
                               profile, cut_left, inv_bw, n_slices);
        else
            hipLaunchKernelGGL((histogram_kernel<T, 1, true>), dim3(grid), dim3(256), lds, s, dt, n,
                               profile, cut_left, inv_bw, n_slices);
    } else {
        if (vec)
            hipLaunchKernelGGL((histogram_kernel<T, 2, false>), dim3(grid), dim3(256), 0, s, dt, n,
                               profile, cut_left, inv_bw, n_slices);
        else
            hipLaunchKernelGGL((histogram_kernel<T, 1, false>), dim3(grid), dim3(256), 0, s, dt, n,
                               profile, cut_left, inv_bw, n_slices);
    }
    return blmi_check_launch("histogram");
}

// ---------------------------------------------------------------------------
// sparse histogram (blond/beam/sparse_histogram.cpp:22-87): one profile of
// nsb slices per filled bucket of a filling pattern.  The bucket is found
// relative to the first filled bucket's left cut and looked up in
// bunch_indexes (the bucket's profile row, -1 when empty), the bin by
// truncation toward zero from that row's left cut.  Where the reference
// would read past bunch_indexes (a particle exactly at the last right cut,
// or NaN) or write outside the profile rows, the particle is skipped.
// ---------------------------------------------------------------------------
template <typename T, bool LDS>
__global__ void __launch_bounds__(256)
sparse_hist_kernel(const T *__restrict__ dt, int64_t n, T *__restrict__ out,
                   const T *__restrict__ cl, const T *__restrict__ cr,
                   const T *__restrict__ bidx, int nsb, int nfb, int n_pattern) {
    extern __shared__ unsigned lds_hist[];
    const int total = nfb * nsb;
    if (LDS) {
        for (int b = threadIdx.x; b < total; b += blockDim.x) lds_hist[b] = 0u;
        __syncthreads();
    }
    const T lo = cl[0], hi = cr[nfb - 1];
    // sparse_histogram.cpp:40-41 (the 1.0 literal makes the reciprocal a double)
    const T inv_bucket = (T)(1.0 / (double)(cr[0] - cl[0]));
    const T inv_bw = inv_bucket * (T)nsb;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
        const T a = dt[j];
        if (a < lo || a > hi || isnan(a)) continue;
        const int fb = (int)((a - lo) * inv_bucket);
        if (fb >= n_pattern) continue;
        const int ib = (int)bidx[fb];
        if (ib == -1) continue;
        const int bin = ib * nsb + (int)((a - cl[ib]) * inv_bw);
        if (bin < 0 || bin >= total) continue;
        if (LDS)
            atomicAdd(&lds_hist[bin], 1u);
        else
            atomic_add_agent(&out[bin], T(1));
    }
    if (LDS) {
        __syncthreads();
        for (int b = threadIdx.x; b < total; b += blockDim.x) {
            const unsigned c = lds_hist[b];
            if (c) atomic_add_agent(&out[b], T(c));
        }
    }
}

template <typename T>
int launch_sparse_histogram(const T *dt, int64_t n, T *profile, const T *cut_left,
                            const T *cut_right, const T *bunch_indexes, int n_slices_bucket,
                            int n_filled, int n_pattern, hipStream_t s) {
    const int64_t total = (int64_t)n_slices_bucket * n_filled;
    if (total <= 0) return BLMI_OK;
    int rc = launch_fill<T>(profile, total, T(0), s);
    if (rc || n <= 0) return rc;
    const size_t lds = (size_t)total * sizeof(unsigned);
    int grid = stream_grid(n, 256);
    const int cap = device_cu_count() * 4;
    if (grid > cap) grid = cap;
    if (lds <= 65536)
        hipLaunchKernelGGL((sparse_hist_kernel<T, true>), dim3(grid), dim3(256), lds, s, dt, n,
                           profile, cut_left, cut_right, bunch_indexes, n_slices_bucket, n_filled,
                           n_pattern);
    else
        hipLaunchKernelGGL((sparse_hist_kernel<T, false>), dim3(grid), dim3(256), 0, s, dt, n,
                           profile, cut_left, cut_right, bunch_indexes, n_slices_bucket, n_filled,
                           n_pattern);
    return blmi_check_launch("sparse_histogram");
}

template int launch_sparse_histogram<double>(const double *, int64_t, double *, const double *,
                                             const double *, const double *, int, int, int,
                                             hipStream_t);
template int launch_sparse_histogram<float>(const float *, int64_t, float *, const float *,
                                            const float *, const float *, int, int, int,
                                            hipStream_t);

// ---------------------------------------------------------------------------
// smooth histogram (histogram.cpp:75-132, including its weight quirk)
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ long long to_fixed(T w, T scale) {
    return (long long)__double2ll_rn((double)w * (double)scale);
}

template <typename T, bool LDS>
__global__ void __launch_bounds__(256) smooth_kernel(const T *__restrict__ dt, int64_t n,
                                                     unsigned long long *__restrict__ acc,
                                                     T cut_left, T inv_bw, T lo, T hi,
                                                     int n_slices, T scale) {
    extern __shared__ unsigned long long lds_acc[];
    if (LDS) {
        for (int b = threadIdx.x; b < n_slices; b += blockDim.x) lds_acc[b] = 0ull;
        __syncthreads();
    }
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const T a = dt[i];
        if (a < lo || a > hi) continue;
        const T f = (a - cut_left) * inv_bw;
        const int k = (int)f;
        const T d = f - (T)k;
        const int k2 = (d > T(0.5)) ? (int)(f + T(1)) : (int)(f - T(1));
        // the same two-bin rule as the reference: own bin 0.5-d, partner 0.5+d
        const unsigned long long w1 = (unsigned long long)to_fixed(T(0.5) - d, scale);
        const unsigned long long w2 = (unsigned long long)to_fixed(T(0.5) + d, scale);
        // the range test above keeps k and k2 inside [0, n_slices) for any
        // finite dt; guard anyway so a rounding corner cannot write outside
        if (k >= 0 && k < n_slices) {
            if (LDS) atomicAdd(&lds_acc[k], w1); else atomic_add_agent(&acc[k], w1);
        }
        if (k2 >= 0 && k2 < n_slices) {
            if (LDS) atomicAdd(&lds_acc[k2], w2); else atomic_add_agent(&acc[k2], w2);
        }
    }
    if (LDS) {
        __syncthreads();
        for (int b = threadIdx.x; b < n_slices; b += blockDim.x) {
            const unsigned long long c = lds_acc[b];
            if (c) atomic_add_agent(&acc[b], c);
        }
    }
}

template <typename T>
__global__ void fixed_to_real_kernel(const unsigned long long *__restrict__ acc, T *__restrict__ out,
                                     int n_slices, double inv_scale) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < n_slices) out[b] = (T)((double)(long long)acc[b] * inv_scale);
}

// Fixed-point exponent: the largest S <= 44 with 1.5 * N * 2^S < 2^62.
int smooth_scale_bits(int64_t n) {
    int bits = 0;
    int64_t m = n > 0 ? n : 1;
    while ((1ll << bits) < m) ++bits;
    int S = 61 - 1 - bits;  // 1.5*N*2^S <= 2^(bits+1+S) < 2^62
    if (S > 44) S = 44;
    if (S < 8) S = 8;
    return S;
}

template <typename T>
int launch_smooth_histogram(const T *dt, int64_t n, T *profile, T cut_left, T cut_right,
                            int n_slices, long long *scratch, hipStream_t s) {
    if (n_slices <= 0) return BLMI_OK;
    // histogram.cpp:82-85
    const T inv_bw = T(n_slices) / (cut_right - cut_left);
    const T bw = (cut_right - cut_left) / T(n_slices);
    const T lo = cut_left + bw * T(0.5);
    const T hi = cut_right - bw * T(0.5);
    const int S = smooth_scale_bits(n);
    const T scale = (T)ldexp(1.0, S);
    unsigned long long *acc = reinterpret_cast<unsigned long long *>(scratch);
    int rc = launch_fill<unsigned long long>(acc, n_slices, 0ull, s);
    if (rc) return rc;
    if (n > 0) {
        const size_t lds = (size_t)n_slices * sizeof(unsigned long long);
        int grid = stream_grid(n, 256);
        const int cap = device_cu_count() * 4;
        if (grid > cap) grid = cap;
        if (lds <= 65536)
            hipLaunchKernelGGL((smooth_kernel<T, true>), dim3(grid), dim3(256), lds, s, dt, n, acc,
                               cut_left, inv_bw, lo, hi, n_slices, scale);
        else
            hipLaunchKernelGGL((smooth_kernel<T, false>), dim3(grid), dim3(256), 0, s, dt, n, acc,
                               cut_left, inv_bw, lo, hi, n_slices, scale);
        rc = blmi_check_launch("smooth_histogram");
        if (rc) return rc;
    }
    hipLaunchKernelGGL((fixed_to_real_kernel<T>), dim3((n_slices + 255) / 256), dim3(256), 0, s,
                       acc, profile, n_slices, ldexp(1.0, -S));
    return blmi_check_launch("smooth_histogram_convert");
}

template int launch_fill<double>(double *, int64_t, double, hipStream_t);
template int launch_fill<float>(float *, int64_t, float, hipStream_t);
template int launch_fill<unsigned>(unsigned *, int64_t, unsigned, hipStream_t);
template int launch_fill<unsigned long long>(unsigned long long *, int64_t, unsigned long long,
                                             hipStream_t);
template int launch_histogram<double>(const double *, int64_t, double *, double, double, int,
                                      hipStream_t);
template int launch_histogram<float>(const float *, int64_t, float *, float, float, int,
                                     hipStream_t);
template int launch_smooth_histogram<double>(const double *, int64_t, double *, double, double,
                                             int, long long *, hipStream_t);
template int launch_smooth_histogram<float>(const float *, int64_t, float *, float, float, int,
                                            long long *, hipStream_t);

}  // namespace blmi
