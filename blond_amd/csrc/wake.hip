// Slice-axis wake kernels: numpy interp / gradient / diff on the device, the
// multi-turn wake memory shift (impedance.py:491-550) and the engine's
// per-turn finishing step (memory, inductive source, kick table).
#include "wake.h"

namespace blmi {

constexpr int kWakeBlock = 256;

static int grid_for(int64_t n) { return (int)((n + kWakeBlock - 1) / kWakeBlock); }

// ---------------------------------------------------------------------------
// numpy helpers
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kWakeBlock)
interp_kernel(const double *__restrict__ x, int64_t n, const double *__restrict__ xp,
              const double *__restrict__ fp, int64_t m, double left, double right,
              double *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double xi = x[i];
    // guess from the end points (exact for a uniform xp up to rounding)
    int64_t g = -1;
    const double span = xp[m - 1] - xp[0];
    if (m > 1 && span > 0.0) {
        const double u = (xi - xp[0]) / span * (double)(m - 1);
        if (u >= 0.0 && u < (double)m) g = (int64_t)u;
    }
    out[i] = np_interp_at(xi, xp, fp, m, left, right, g);
}

__global__ void __launch_bounds__(kWakeBlock)
gradient_kernel(const double *__restrict__ f, int64_t n, double h, double *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // numpy.gradient(f, h): uniform spacing, edge_order 1
    if (i == 0) out[i] = (f[1] - f[0]) / h;
    else if (i == n - 1) out[i] = (f[n - 1] - f[n - 2]) / h;
    else out[i] = (f[i + 1] - f[i - 1]) / (2.0 * h);
}

__global__ void __launch_bounds__(kWakeBlock)
diff_kernel(const double *__restrict__ f, int64_t n, double *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n - 1) out[i] = f[i + 1] - f[i];
}

int launch_interp(const double *x, int64_t n, const double *xp, const double *fp, int64_t m,
                  double left, double right, double *out, hipStream_t s) {
    if (n <= 0) return BLMI_OK;
    hipLaunchKernelGGL(interp_kernel, dim3(grid_for(n)), dim3(kWakeBlock), 0, s, x, n, xp, fp, m,
                       left, right, out);
    return blmi_check_launch("interp");
}

int launch_gradient(const double *f, int64_t n, double h, double *out, hipStream_t s) {
    hipLaunchKernelGGL(gradient_kernel, dim3(grid_for(n)), dim3(kWakeBlock), 0, s, f, n, h, out);
    return blmi_check_launch("gradient");
}

int launch_diff(const double *f, int64_t n, double *out, hipStream_t s) {
    if (n < 2) return BLMI_OK;
    hipLaunchKernelGGL(diff_kernel, dim3(grid_for(n - 1)), dim3(kWakeBlock), 0, s, f, n, out);
    return blmi_check_launch("diff");
}

// ---------------------------------------------------------------------------
// multi-turn wake memory
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kWakeBlock)
mtw_shift_add_kernel(const double *__restrict__ mem_in, double *__restrict__ mem_out,
                     const double *__restrict__ time, int64_t n_mem, double t_rev,
                     const double *__restrict__ add, int64_t n_add) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_mem) return;
    double v = mtw_value(mem_in, time, n_mem, t_rev, i);
    if (add && i < n_add) v = v + add[i];
    mem_out[i] = v;
}

int launch_mtw_shift_add(const double *mem_in, double *mem_out, const double *time, int64_t n_mem,
                         double t_rev, const double *add, int64_t n_add, hipStream_t s) {
    if (n_mem <= 0) return BLMI_OK;
    hipLaunchKernelGGL(mtw_shift_add_kernel, dim3(grid_for(n_mem)), dim3(kWakeBlock), 0, s, mem_in,
                       mem_out, time, n_mem, t_rev, add, n_add);
    return blmi_check_launch("mtw_shift_add");
}

// ---------------------------------------------------------------------------
// engine finishing step
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kWakeBlock) wake_finish_kernel(WakeArgs a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int t = *a.turn_base + a.turn_off;
    const int ns = a.n_slices;
    const double *mem_in = (t & 1) ? a.mem1 : a.mem0;
    double *mem_out = (t & 1) ? a.mem0 : a.mem1;
    const unsigned *cin = (t & 1) ? a.counts1 : a.counts0;
    unsigned *cclr = (t & 1) ? a.counts0 : a.counts1;
    const bool from_counts = a.v1 == nullptr;
    const double trev = a.n_mem > 0 ? a.t_rev[t] : 0.0;
    const double coef = a.ind_coef ? a.ind_coef[t] : 0.0;

    auto prof = [&](int j) -> double {
        if (!from_counts) return a.profile[j];
        unsigned c = 0;
        for (int g = 0; g < a.groups; ++g) c += cin[(size_t)g * ns + j];
        return (double)c;
    };
    // the multi-turn memory after this turn at j (j < n_mem)
    auto mem_at = [&](int64_t j) -> double {
        const double v = mtw_value(mem_in, a.time, a.n_mem, trev, j);
        return (a.v1 && j < a.n_add) ? v + a.v1[j] : v;
    };
    // the turn's total induced voltage at bin j: the sources' sum starting
    // from 0 (impedance.py:117-127)
    auto total_at = [&](int j) -> double {
        double v = 0.0;
        if (a.n_mem > 0) v = v + mem_at(j);
        else if (a.v1) v = v + a.v1[j];
        if (a.ind_coef) {
            const double d = profile_derivative_at(prof, a.bc, ns, a.bc[1] - a.bc[0],
                                                   a.deriv_mode, j);
            v = v + (-(coef * d));
        }
        return v;
    };

    if (i < a.n_mem) mem_out[i] = mem_at(i);
    if (i < ns) {
        if (from_counts) {
            a.profile[i] = prof((int)i);
            for (int g = 0; g < a.groups; ++g) cclr[(size_t)g * ns + i] = 0u;
        }
        const double vi = total_at((int)i);
        a.volt[i] = vi;
        // linear_interp_kick.cpp:33-48 on the total (acceleration kick 0)
        const double inv_bw = (double)(ns - 1) / (a.bc[ns - 1] - a.bc[0]);
        if (i == 0) {
            a.tab_params[0] = a.bc[0];
            a.tab_params[1] = inv_bw;
        }
        if (i < ns - 1) {
            const double vn = total_at((int)i + 1);
            const double k = a.charge * (vn - vi) * inv_bw;
            a.tab[i] = make_double2(k, (a.charge * vi - a.bc[i] * k) + 0.0);
        }
    }
}

int launch_wake_finish(const WakeArgs &a, hipStream_t s, const KernelEvents &ev) {
    const int64_t n = a.n_mem > a.n_slices ? a.n_mem : (int64_t)a.n_slices;
    if (a.ind_coef && a.n_slices < 2)
        return blmi_set_error(BLMI_ERR_ARG, "inductive source: at least 2 slices");
    launch(wake_finish_kernel, dim3(grid_for(n)), dim3(kWakeBlock), 0, s, ev, a);
    return blmi_check_launch("wake_finish");
}

__global__ void mtw_copy_kernel(const int *turn_base, double *mem0, double *mem1, int64_t n,
                                const double *in, double *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double *cur = (*turn_base & 1) ? mem1 : mem0;
    if (in) cur[i] = in[i];
    else out[i] = cur[i];
}

int launch_mtw_publish(const int *turn_base, const double *mem0, const double *mem1, int64_t n_mem,
                       double *out, hipStream_t s) {
    if (n_mem <= 0) return BLMI_OK;
    hipLaunchKernelGGL(mtw_copy_kernel, dim3(grid_for(n_mem)), dim3(kWakeBlock), 0, s, turn_base,
                       const_cast<double *>(mem0), const_cast<double *>(mem1), n_mem,
                       (const double *)nullptr, out);
    return blmi_check_launch("mtw_publish");
}

int launch_mtw_load(const int *turn_base, double *mem0, double *mem1, int64_t n_mem,
                    const double *in, hipStream_t s) {
    if (n_mem <= 0) return BLMI_OK;
    hipLaunchKernelGGL(mtw_copy_kernel, dim3(grid_for(n_mem)), dim3(kWakeBlock), 0, s, turn_base,
                       mem0, mem1, n_mem, in, (double *)nullptr);
    return blmi_check_launch("mtw_load");
}

}  // namespace blmi
