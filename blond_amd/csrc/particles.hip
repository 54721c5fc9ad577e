// Per-particle kernels of the tracking path: RF kick, drift (3 solvers),
// linear-interpolation kick and the RF voltage on bin centres.
//
// Layout: structure-of-arrays dt[N], dE[N] in HBM.  The streaming kernels move
// two particles per lane per iteration with 16-byte loads and stores when both
// arrays are 16-byte aligned (the common case for whole torch tensors), so a
// wavefront reads 1 KiB of dt and 1 KiB of dE per step, fully coalesced.
#include "blmi_common.h"
#include "particles.h"

namespace blmi {

// ---------------------------------------------------------------------------
// Per-particle arithmetic, shared by the single-purpose kernels and the fused
// turn kernel.  Operation order follows the reference's strict build.
// ---------------------------------------------------------------------------

// kick.cpp:26-34 : dE += charge * sum_j V_j sin(w_j t + phi_j) + acc
template <typename T>
__device__ __forceinline__ T rf_kick_delta(T t, int n_rf, T charge, const T *__restrict__ V,
                                           const T *__restrict__ w, const T *__restrict__ phi,
                                           T acc) {
    T s = T(0);
    for (int j = 0; j < n_rf; ++j) s = s + V[j] * Trig<T>::sin(w[j] * t + phi[j]);
    return charge * s + acc;
}

// drift.cpp:33-75
template <typename T>
__device__ __forceinline__ T drift_delta(const DriftCoeffs<T> &c, T e) {
    if (c.solver == 0) {
        return c.coeff * e;
    } else if (c.solver == 1) {
        T den;
        if (c.alpha_order == 0)
            den = T(1) - c.e0 * e;
        else if (c.alpha_order == 1)
            den = T(1) - c.e0 * e - c.e1 * e * e;
        else
            den = T(1) - c.e0 * e - c.e1 * e * e - c.e2 * e * e * e;
        return c.tl * (T(1) / den - T(1));
    } else {
        const T delta = sqrt(T(1) + c.ib2 * (e * e * c.ie2 + T(2) * e / c.energy)) - T(1);
        return c.tl * ((T(1) + c.a0 * delta + c.a1 * (delta * delta) +
                       c.a2 * (delta * delta * delta)) *
                          (T(1) + e / c.energy) / (T(1) + delta) -
                      T(1));
    }
}

// linear_interp_kick.cpp:56-64 with an explicit signed range test in place of
// the reference's unsigned cast (same set of kicked particles: negative,
// NaN and last-bin positions get no kick).
template <typename T>
__device__ __forceinline__ T interp_kick_delta(T t, const T *__restrict__ vk,
                                               const T *__restrict__ fac, T bc0, T inv_bw,
                                               int n_tab, bool &hit) {
    const T f = floor((t - bc0) * inv_bw);
    hit = (f >= T(0)) && (f < T(n_tab));
    if (!hit) return T(0);
    const int k = (int)f;
    return t * vk[k] + fac[k];
}

// ---------------------------------------------------------------------------
// kick
// ---------------------------------------------------------------------------
template <typename T, int VEC>
__global__ void __launch_bounds__(256) kick_kernel(const T *__restrict__ dt, T *__restrict__ dE,
                                                   int64_t n, int n_rf, T charge,
                                                   const T *__restrict__ V,
                                                   const T *__restrict__ w,
                                                   const T *__restrict__ phi, T acc) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (VEC == 2) {
        using V2 = typename Vec2<T>::type;
        const int64_t n2 = n / 2;
        for (; i < n2; i += stride) {
            const V2 t = reinterpret_cast<const V2 *>(dt)[i];
            V2 e = reinterpret_cast<V2 *>(dE)[i];
            e.x = e.x + rf_kick_delta(t.x, n_rf, charge, V, w, phi, acc);
            e.y = e.y + rf_kick_delta(t.y, n_rf, charge, V, w, phi, acc);
            reinterpret_cast<V2 *>(dE)[i] = e;
        }
        if (blockIdx.x == 0 && threadIdx.x == 0 && (n & 1)) {
            const int64_t k = n - 1;
            dE[k] = dE[k] + rf_kick_delta(dt[k], n_rf, charge, V, w, phi, acc);
        }
    } else {
        for (; i < n; i += stride) dE[i] = dE[i] + rf_kick_delta(dt[i], n_rf, charge, V, w, phi, acc);
    }
}

// ---------------------------------------------------------------------------
// drift
// ---------------------------------------------------------------------------
template <typename T, int VEC>
__global__ void __launch_bounds__(256) drift_kernel(T *__restrict__ dt, const T *__restrict__ dE,
                                                    int64_t n, DriftCoeffs<T> c) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (VEC == 2) {
        using V2 = typename Vec2<T>::type;
        const int64_t n2 = n / 2;
        for (; i < n2; i += stride) {
            V2 t = reinterpret_cast<const V2 *>(dt)[i];
            const V2 e = reinterpret_cast<const V2 *>(dE)[i];
            t.x = t.x + drift_delta(c, e.x);
            t.y = t.y + drift_delta(c, e.y);
            reinterpret_cast<V2 *>(dt)[i] = t;
        }
        if (blockIdx.x == 0 && threadIdx.x == 0 && (n & 1)) {
            const int64_t k = n - 1;
            dt[k] = dt[k] + drift_delta(c, dE[k]);
        }
    } else {
        for (; i < n; i += stride) dt[i] = dt[i] + drift_delta(c, dE[i]);
    }
}

// ---------------------------------------------------------------------------
// linear_interp_kick: table build then gather
// ---------------------------------------------------------------------------
// linear_interp_kick.cpp:30-48 : inv_bw from the bin-centre span; per bin
// slope vk = charge*(V[i+1]-V[i])*inv_bw and intercept (charge*V[i] - bc[i]*vk) + acc
template <typename T>
__global__ void interp_table_kernel(const T *__restrict__ volt, const T *__restrict__ bc,
                                    int n_slices, T charge, T acc, T *__restrict__ vk,
                                    T *__restrict__ fac, T *__restrict__ params) {
    const T inv_bw = T(n_slices - 1) / (bc[n_slices - 1] - bc[0]);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        params[0] = bc[0];
        params[1] = inv_bw;
    }
    if (i < n_slices - 1) {
        const T k = charge * (volt[i + 1] - volt[i]) * inv_bw;
        vk[i] = k;
        fac[i] = (charge * volt[i] - bc[i] * k) + acc;
    }
}

template <typename T, int VEC>
__global__ void __launch_bounds__(256) interp_kick_kernel(const T *__restrict__ dt, T *__restrict__ dE,
                                                          int64_t n, const T *__restrict__ vk,
                                                          const T *__restrict__ fac,
                                                          const T *__restrict__ params, int n_tab) {
    const T bc0 = params[0];
    const T inv_bw = params[1];
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool hit;
    if (VEC == 2) {
        using V2 = typename Vec2<T>::type;
        const int64_t n2 = n / 2;
        for (; i < n2; i += stride) {
            const V2 t = reinterpret_cast<const V2 *>(dt)[i];
            V2 e = reinterpret_cast<V2 *>(dE)[i];
            T d = interp_kick_delta(t.x, vk, fac, bc0, inv_bw, n_tab, hit);
            if (hit) e.x = e.x + d;
            d = interp_kick_delta(t.y, vk, fac, bc0, inv_bw, n_tab, hit);
            if (hit) e.y = e.y + d;
            reinterpret_cast<V2 *>(dE)[i] = e;
        }
        if (blockIdx.x == 0 && threadIdx.x == 0 && (n & 1)) {
            const int64_t k = n - 1;
            const T d = interp_kick_delta(dt[k], vk, fac, bc0, inv_bw, n_tab, hit);
            if (hit) dE[k] = dE[k] + d;
        }
    } else {
        for (; i < n; i += stride) {
            const T d = interp_kick_delta(dt[i], vk, fac, bc0, inv_bw, n_tab, hit);
            if (hit) dE[i] = dE[i] + d;
        }
    }
}

// kick.cpp:37-51 : harmonic-major accumulation into rf_voltage
template <typename T>
__global__ void rf_volt_comp_kernel(const T *__restrict__ V, const T *__restrict__ w,
                                    const T *__restrict__ phi, const T *__restrict__ bc,
                                    int n_rf, int n_bins, T *__restrict__ out) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n_bins) return;
    T acc = out[b];
    const T x = bc[b];
    for (int j = 0; j < n_rf; ++j) acc = acc + V[j] * Trig<T>::sin(w[j] * x + phi[j]);
    out[b] = acc;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline bool aligned16(const void *p) { return (((uintptr_t)p) & 15u) == 0; }

int stream_grid(int64_t work_items, int block) {
    int64_t g = (work_items + block - 1) / block;
    const int64_t cap = (int64_t)device_cu_count() * 16;
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    return (int)g;
}

template <typename T>
int launch_kick(const T *dt, T *dE, int64_t n, int n_rf, T charge, const T *V, const T *w,
                const T *phi, T acc, hipStream_t s) {
    if (n <= 0) return BLMI_OK;
    if (aligned16(dt) && aligned16(dE)) {
        hipLaunchKernelGGL((kick_kernel<T, 2>), dim3(stream_grid((n + 1) / 2, 256)), dim3(256), 0, s,
                           dt, dE, n, n_rf, charge, V, w, phi, acc);
    } else {
        hipLaunchKernelGGL((kick_kernel<T, 1>), dim3(stream_grid(n, 256)), dim3(256), 0, s, dt, dE,
                           n, n_rf, charge, V, w, phi, acc);
    }
    return blmi_check_launch("kick");
}

template <typename T>
DriftCoeffs<T> make_drift_coeffs(int solver, T t_rev, T length_ratio, T alpha_order, T eta_0,
                                 T eta_1, T eta_2, T alpha_0, T alpha_1, T alpha_2, T beta,
                                 T energy) {
    // drift.cpp:30-61, evaluated on the host in the same order
    DriftCoeffs<T> c{};
    c.solver = solver;
    c.alpha_order = (alpha_order == T(0)) ? 0 : (alpha_order == T(1) ? 1 : 2);
    c.tl = t_rev * length_ratio;
    c.coeff = c.tl * eta_0 / (beta * beta * energy);
    const T k = T(1) / (beta * beta * energy);
    c.e0 = eta_0 * k;
    c.e1 = eta_1 * k * k;
    c.e2 = eta_2 * k * k * k;
    c.ib2 = T(1) / (beta * beta);
    c.ie2 = T(1) / (energy * energy);
    c.energy = energy;
    c.a0 = alpha_0;
    c.a1 = alpha_1;
    c.a2 = alpha_2;
    return c;
}

template <typename T>
int launch_drift(T *dt, const T *dE, int64_t n, const DriftCoeffs<T> &c, hipStream_t s) {
    if (n <= 0) return BLMI_OK;
    if (aligned16(dt) && aligned16(dE)) {
        hipLaunchKernelGGL((drift_kernel<T, 2>), dim3(stream_grid((n + 1) / 2, 256)), dim3(256), 0,
                           s, dt, dE, n, c);
    } else {
        hipLaunchKernelGGL((drift_kernel<T, 1>), dim3(stream_grid(n, 256)), dim3(256), 0, s, dt, dE,
                           n, c);
    }
    return blmi_check_launch("drift");
}

template <typename T>
int launch_interp_table(const T *volt, const T *bc, int n_slices, T charge, T acc, T *vk,
                        T *fac, T *params, hipStream_t s) {
    const int blocks = (n_slices + 255) / 256;
    hipLaunchKernelGGL((interp_table_kernel<T>), dim3(blocks), dim3(256), 0, s, volt, bc,
                       n_slices, charge, acc, vk, fac, params);
    return blmi_check_launch("interp_table");
}

template <typename T>
int launch_interp_kick(const T *dt, T *dE, int64_t n, const T *vk, const T *fac,
                       const T *params, int n_tab, hipStream_t s) {
    if (n <= 0) return BLMI_OK;
    if (aligned16(dt) && aligned16(dE)) {
        hipLaunchKernelGGL((interp_kick_kernel<T, 2>), dim3(stream_grid((n + 1) / 2, 256)),
                           dim3(256), 0, s, dt, dE, n, vk, fac, params, n_tab);
    } else {
        hipLaunchKernelGGL((interp_kick_kernel<T, 1>), dim3(stream_grid(n, 256)), dim3(256), 0, s,
                           dt, dE, n, vk, fac, params, n_tab);
    }
    return blmi_check_launch("linear_interp_kick");
}

template <typename T>
int launch_rf_volt_comp(const T *V, const T *w, const T *phi, const T *bc, int n_rf, int n_bins,
                        T *out, hipStream_t s) {
    if (n_bins <= 0) return BLMI_OK;
    hipLaunchKernelGGL((rf_volt_comp_kernel<T>), dim3((n_bins + 255) / 256), dim3(256), 0, s, V, w,
                       phi, bc, n_rf, n_bins, out);
    return blmi_check_launch("rf_volt_comp");
}

#define BLMI_INST(T)                                                                          \
    template int launch_kick<T>(const T *, T *, int64_t, int, T, const T *, const T *,        \
                                const T *, T, hipStream_t);                                   \
    template DriftCoeffs<T> make_drift_coeffs<T>(int, T, T, T, T, T, T, T, T, T, T, T);       \
    template int launch_drift<T>(T *, const T *, int64_t, const DriftCoeffs<T> &, hipStream_t); \
    template int launch_interp_table<T>(const T *, const T *, int, T, T, T *, T *, T *,       \
                                        hipStream_t);                                         \
    template int launch_interp_kick<T>(const T *, T *, int64_t, const T *, const T *,         \
                                       const T *, int, hipStream_t);                          \
    template int launch_rf_volt_comp<T>(const T *, const T *, const T *, const T *, int, int, \
                                        T *, hipStream_t);
BLMI_INST(double)
BLMI_INST(float)
#undef BLMI_INST


// ---------------------------------------------------------------------------
// Bunch moments (beam.py:430-458: mean and r.m.s. with ddof = 0 over the
// alive particles, id != 0), two passes like numpy's std: sums, then the
// squared deviations from the mean.  Per-thread sums in grid-stride order,
// a fixed tree inside each workgroup and a fixed order over workgroups:
// the result does not depend on scheduling.
// ---------------------------------------------------------------------------
constexpr int kMomBlock = 256;

__device__ __forceinline__ double block_sum(double v, double *red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double s = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w];
    return s;
}

template <typename T, bool DEV>
__global__ void __launch_bounds__(kMomBlock) moments_pass_kernel(const T *__restrict__ dt,
                                                                 const T *__restrict__ dE,
                                                                 const int64_t *__restrict__ id,
                                                                 int64_t n, const double *res,
                                                                 double *__restrict__ part) {
    __shared__ double red[kMomBlock / 64];
    double a = 0.0, b = 0.0, c = 0.0;
    double mt = 0.0, me = 0.0;
    if (DEV) {
        mt = res[1];
        me = res[3];
    }
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        if (id && id[i] == 0) continue;
        const double x = (double)dt[i], e = (double)dE[i];
        if (DEV) {
            a += (x - mt) * (x - mt);
            b += (e - me) * (e - me);
        } else {
            a += x;
            b += e;
            c += 1.0;
        }
    }
    a = block_sum(a, red);
    b = block_sum(b, red);
    c = block_sum(c, red);
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x] = a;
        part[3 * blockIdx.x + 1] = b;
        part[3 * blockIdx.x + 2] = c;
    }
}

// res = {n_alive, mean_dt, sigma_dt, mean_dE, sigma_dE}
__global__ void moments_final_kernel(const double *__restrict__ part, int grid, bool dev,
                                     double *__restrict__ res) {
    if (threadIdx.x != 0) return;
    double a = 0.0, b = 0.0, c = 0.0;
    for (int w = 0; w < grid; ++w) {
        a += part[3 * w];
        b += part[3 * w + 1];
        c += part[3 * w + 2];
    }
    if (!dev) {
        res[0] = c;
        res[1] = c > 0.0 ? a / c : 0.0;
        res[3] = c > 0.0 ? b / c : 0.0;
    } else {
        const double nn = res[0];
        res[2] = nn > 0.0 ? sqrt(a / nn) : 0.0;
        res[4] = nn > 0.0 ? sqrt(b / nn) : 0.0;
    }
}

template <typename T>
int launch_moments(const T *dt, const T *dE, const int64_t *id, int64_t n, double *res,
                   hipStream_t s) {
    int grid = stream_grid(n, kMomBlock);
    WorkspaceLease ws(sizeof(double) * 3 * (size_t)grid, 3, s);
    double *part = static_cast<double *>(ws.ptr());
    if (!part) return blmi_set_error(BLMI_ERR_ALLOC, "moments: workspace");
    hipLaunchKernelGGL((moments_pass_kernel<T, false>), dim3(grid), dim3(kMomBlock), 0, s, dt, dE,
                       id, n, res, part);
    hipLaunchKernelGGL(moments_final_kernel, dim3(1), dim3(64), 0, s, part, grid, false, res);
    hipLaunchKernelGGL((moments_pass_kernel<T, true>), dim3(grid), dim3(kMomBlock), 0, s, dt, dE,
                       id, n, res, part);
    hipLaunchKernelGGL(moments_final_kernel, dim3(1), dim3(64), 0, s, part, grid, true, res);
    return blmi_check_launch("moments");
}
template int launch_moments<double>(const double *, const double *, const int64_t *, int64_t,
                                    double *, hipStream_t);
template int launch_moments<float>(const float *, const float *, const int64_t *, int64_t, double *,
                                   hipStream_t);

}  // namespace blmi
