// Bin-ordered working set of the fused engine (large slice counts).
//
// With 2^16 slices a workgroup's particles, taken in the caller's order,
// touch bins and interpolation-table entries all over the bunch: the table
// (1 MiB) is gathered from L2 one 128-byte line per particle, and the
// histogram needs the u16-pair partial rows.  Here the engine instead keeps
// the particles in its own double-buffered arrays grouped by coarse bin
// (64-bin digits), re-grouped every `sort_every` turns by a counting sort:
//   digits:  per workgroup (contiguous chunk) histogram of the digits,
//            stored digit-major;
//   scan:    one workgroup per digit scans its column over workgroups, one
//            workgroup scans the digit totals;
//   scatter: each particle to digit_start + column offset + rank within the
//            workgroup (wavefront-aggregated LDS counters), with its
//            original index (perm);
//   window:  each workgroup's first digit gives its bin window.
// A workgroup's particles then span a few thousand bins at most between
// sorts, so the particle kernel keeps a window of kSortWindow bins of both
// the histogram (u32, flushed with atomics) and the table (staged in LDS);
// particles outside the window use global memory.  The window follows the
// workgroup's particles: each turn it is re-centred on the range of the new
// bins.  track() copies the caller's arrays in (perm = identity) and
// scatters them back through perm at the end, so the caller sees its own
// order; every particle's arithmetic is that of the plain kernel, so results
// are bit-identical.
#include <string.h>

#include <vector>

#include "blmi_common.h"
#include "engine.h"
#include "particles.h"
#include "spectral.h"
#include "track_device.h"

namespace blmi {

struct WinArgs {
    int n_rf;
    double charge;
    const double *rf;
    const double *acc;
    const DriftCoeffs<double> *drift;
    const int *base;
    int off;
    int last_turn;
    const double2 *tab;
    const double *tab_params;
    int n_tab;
    unsigned *counts0, *counts1;
    int groups;
    int n_slices;
    double cut_left, inv_bw;
    int max_peel;
    double *profile;
    int *pending;
    int *win;  // [grid] window base per workgroup (read, then replaced)
};

__device__ __forceinline__ int window_base(int lo, int ns) {
    if (ns <= kSortWindow) return 0;
    lo = lo < 0 ? 0 : lo;
    return lo > ns - kSortWindow ? ns - kSortWindow : lo;
}

template <int SOLVER, bool INTERP, bool PUBLISH>
__global__ void __launch_bounds__(1024) turn_window_kernel(double *__restrict__ dt,
                                                           double *__restrict__ dE, int64_t n,
                                                           WinArgs a) {
    extern __shared__ unsigned wl_hist[];
    double2 *wl_tab = reinterpret_cast<double2 *>(wl_hist + kSortWindow);
    __shared__ int red_min[16], red_max[16];
    int t = *a.base + a.off;
    if (t < 0) t = 0;
    if (t > a.last_turn) t = a.last_turn;
    const double *V = a.rf + (size_t)t * 3 * a.n_rf;
    const double *w = V + a.n_rf;
    const double *phi = w + a.n_rf;
    const RfProg<double> rfp(V, w, phi, a.n_rf);
    const double acc = a.acc[t];
    const DriftCoeffs<double> dc = a.drift[t + 1];
    const int ns = a.n_slices;
    unsigned *counts = (((t + 1) & 1) ? a.counts1 : a.counts0) +
                       (size_t)(blockIdx.x % a.groups) * ns;
    if (PUBLISH) {
        if (a.pending[t & 1] == t) {
            unsigned *prev = (t & 1) ? a.counts1 : a.counts0;
            const int per = (ns + (int)gridDim.x - 1) / (int)gridDim.x;
            const int b0 = (int)blockIdx.x * per;
            const int b1 = min(b0 + per, ns);
            publish_range(prev, a.profile, ns, a.groups, b0, b1);
        }
        if (blockIdx.x == 0 && threadIdx.x == 0) a.pending[(t + 1) & 1] = t + 1;
    }
    const int lo = window_base(a.win[blockIdx.x], ns);
    double bc0 = 0.0, ibw = 0.0;
    if (INTERP) {
        bc0 = a.tab_params[0];
        ibw = a.tab_params[1];
    }
    for (int b = threadIdx.x; b < kSortWindow; b += blockDim.x) {
        wl_hist[b] = 0u;
        if (INTERP && lo + b < a.n_tab) wl_tab[b] = a.tab[lo + b];
    }
    __syncthreads();
    int bmin = 0x7fffffff, bmax = -1;
    auto step = [&](double &x, double &e) -> int {
        if (INTERP) {
            // linear_interp_kick.cpp:44-64 (acceleration kick 0)
            const double f = floor((x - bc0) * ibw);
            if (f >= 0.0 && f < (double)a.n_tab) {
                const int k = (int)f;
                const double2 tk = (k >= lo && k < lo + kSortWindow) ? wl_tab[k - lo] : a.tab[k];
                e = e + (x * tk.x + tk.y);
            }
        }
        e = e + rfp.kick(x, a.charge, acc);
        x = x + e_drift(dc, e, SOLVER);
        const double f = floor((x - a.cut_left) * a.inv_bw);
        return (f >= 0.0 && f < (double)ns) ? (int)f : -1;
    };
    auto count = [&](int bin) {
        const int rel = bin - lo;
        const bool in = bin >= 0 && rel >= 0 && rel < kSortWindow;
        if (bin >= 0) {
            bmin = min(bmin, bin);
            bmax = max(bmax, bin);
        }
        lds_count<false>(wl_hist, in ? rel : -1, a.max_peel);
        if (bin >= 0 && !in) atomic_add_agent(&counts[bin], 1u);
    };
    // contiguous chunk of pairs per workgroup (the ordered layout)
    const int64_t n2 = n / 2;
    const int64_t per = (n2 + gridDim.x - 1) / gridDim.x;
    const int64_t p0 = (int64_t)blockIdx.x * per;
    const int64_t p1 = min(p0 + per, n2);
    for (int64_t i = p0 + threadIdx.x; i < p1; i += blockDim.x) {
        double2 x = reinterpret_cast<const double2 *>(dt)[i];
        double2 e = reinterpret_cast<const double2 *>(dE)[i];
        const int b0 = step(x.x, e.x);
        const int b1 = step(x.y, e.y);
        reinterpret_cast<double2 *>(dt)[i] = x;
        reinterpret_cast<double2 *>(dE)[i] = e;
        count(b0);
        count(b1);
    }
    if ((n & 1) && blockIdx.x == gridDim.x - 1 && threadIdx.x < 64) {
        int b = -1;
        if (threadIdx.x == 0) {
            double x = dt[n - 1], e = dE[n - 1];
            b = step(x, e);
            dt[n - 1] = x;
            dE[n - 1] = e;
        }
        count(b);
    }
    // new window: centred on the range of this turn's bins
    for (int o = 32; o > 0; o >>= 1) {
        bmin = min(bmin, __shfl_xor(bmin, o));
        bmax = max(bmax, __shfl_xor(bmax, o));
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        red_min[wid] = bmin;
        red_max[wid] = bmax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int mn = 0x7fffffff, mx = -1;
        for (int q = 0; q < (int)(blockDim.x >> 6); ++q) {
            mn = min(mn, red_min[q]);
            mx = max(mx, red_max[q]);
        }
        if (mx >= 0) a.win[blockIdx.x] = (mn + mx) / 2 - kSortWindow / 2;
    }
    for (int b = threadIdx.x; b < kSortWindow; b += blockDim.x) {
        const unsigned c = wl_hist[b];
        if (c) atomic_add_agent(&counts[lo + b], c);
    }
}

// ---- counting sort by coarse bin ------------------------------------------
__device__ __forceinline__ int coarse_digit(double x, double cl, double ibw, int ns, int ndig) {
    const double f = floor((x - cl) * ibw);
    int b = (f >= 0.0 && f < (double)ns) ? (int)f : (f < 0.0 ? 0 : ns - 1);
    if (!(f == f)) b = 0;  // NaN
    const int d = b >> kSortDigitShift;
    return d < ndig ? d : ndig - 1;
}

// Workgroup w's chunk: particles [w * per2, min((w+1) * per2, n)), per2 even.
__global__ void __launch_bounds__(1024) sort_digits_kernel(const double *__restrict__ dt, int64_t n,
                                                           int64_t per2, double cl, double ibw,
                                                           int ns, int ndig,
                                                           unsigned *__restrict__ cnt) {
    extern __shared__ unsigned sh[];
    for (int d = threadIdx.x; d < ndig; d += blockDim.x) sh[d] = 0u;
    __syncthreads();
    const int64_t i0 = (int64_t)blockIdx.x * per2, i1 = min(i0 + per2, n);
    for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x)
        lds_count<false>(sh, coarse_digit(dt[i], cl, ibw, ns, ndig), 4);
    __syncthreads();
    for (int d = threadIdx.x; d < ndig; d += blockDim.x)
        cnt[(size_t)d * gridDim.x + blockIdx.x] = sh[d];
}

// One workgroup per digit: exclusive scan of its column over the workgroups
// (in place) and the digit's total.
__global__ void __launch_bounds__(1024) sort_column_kernel(unsigned *__restrict__ cnt, int grid,
                                                           unsigned *__restrict__ total) {
    __shared__ unsigned s[1024];
    unsigned *col = cnt + (size_t)blockIdx.x * grid;
    unsigned run = 0;
    for (int base = 0; base < grid; base += blockDim.x) {
        const int w = base + threadIdx.x;
        const unsigned v = w < grid ? col[w] : 0u;
        s[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < (int)blockDim.x; o <<= 1) {
            const unsigned add = threadIdx.x >= (unsigned)o ? s[threadIdx.x - o] : 0u;
            __syncthreads();
            s[threadIdx.x] += add;
            __syncthreads();
        }
        if (w < grid) col[w] = run + s[threadIdx.x] - v;
        run += s[blockDim.x - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) total[blockIdx.x] = run;
}

// One workgroup: exclusive scan of the digit totals -> digit starts.
__global__ void __launch_bounds__(1024) sort_total_kernel(unsigned *__restrict__ total, int ndig) {
    __shared__ unsigned s[1024];
    unsigned run = 0;
    for (int base = 0; base < ndig; base += blockDim.x) {
        const int d = base + threadIdx.x;
        const unsigned v = d < ndig ? total[d] : 0u;
        s[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < (int)blockDim.x; o <<= 1) {
            const unsigned add = threadIdx.x >= (unsigned)o ? s[threadIdx.x - o] : 0u;
            __syncthreads();
            s[threadIdx.x] += add;
            __syncthreads();
        }
        if (d < ndig) total[d] = run + s[threadIdx.x] - v;
        run += s[blockDim.x - 1];
        __syncthreads();
    }
}

__global__ void __launch_bounds__(1024) sort_scatter_kernel(
    const double *__restrict__ xi, const double *__restrict__ ei, const unsigned *__restrict__ pi,
    double *__restrict__ xo, double *__restrict__ eo, unsigned *__restrict__ po, int64_t n,
    int64_t per2, double cl, double ibw, int ns, int ndig, const unsigned *__restrict__ col_off,
    const unsigned *__restrict__ dstart) {
    extern __shared__ unsigned rank[];
    for (int d = threadIdx.x; d < ndig; d += blockDim.x) rank[d] = 0u;
    __syncthreads();
    const int lane = (int)__lane_id();
    const int64_t i0 = (int64_t)blockIdx.x * per2, i1 = min(i0 + per2, n);
    for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        const double x = xi[i];
        const int d = coarse_digit(x, cl, ibw, ns, ndig);
        // wavefront-aggregated rank: lanes of one digit take consecutive slots
        unsigned long long pending = __ballot(1);
        unsigned pos = 0;
        while (pending) {
            const int leader = __builtin_ctzll(pending);
            const int ld = __builtin_amdgcn_readlane(d, leader);
            const unsigned long long same = __ballot(d == ld) & pending;
            unsigned base = 0;
            if (lane == leader) base = atomicAdd(&rank[ld], (unsigned)__popcll(same));
            base = __builtin_amdgcn_readlane(base, leader);
            if (d == ld && ((pending >> lane) & 1ull))
                pos = dstart[ld] + col_off[(size_t)ld * gridDim.x + blockIdx.x] + base +
                      (unsigned)__popcll(same & ((1ull << lane) - 1ull));
            pending &= ~same;
        }
        xo[pos] = x;
        eo[pos] = ei[i];
        po[pos] = pi[i];
    }
}

// Window of workgroup w after a sort: the digit of its first particle.
__global__ void sort_window_kernel(const unsigned *__restrict__ dstart, int ndig, int grid,
                                   int64_t per_pairs, int *__restrict__ win) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= grid) return;
    const unsigned long long first = (unsigned long long)(2 * per_pairs) * (unsigned long long)w;
    int lo = 0, hi = ndig - 1;  // last digit with dstart <= first
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if ((unsigned long long)dstart[mid] <= first) lo = mid; else hi = mid - 1;
    }
    win[w] = (lo << kSortDigitShift) - kSortWindow / 4;
}

__global__ void order_in_kernel(const double *__restrict__ dt, const double *__restrict__ dE,
                                double *__restrict__ xo, double *__restrict__ eo,
                                unsigned *__restrict__ po, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        xo[i] = dt[i];
        eo[i] = dE[i];
        po[i] = (unsigned)i;
    }
}

__global__ void order_out_kernel(const double *__restrict__ xi, const double *__restrict__ ei,
                                 const unsigned *__restrict__ pi, double *__restrict__ dt,
                                 double *__restrict__ dE, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const unsigned j = pi[i];
        dt[j] = xi[i];
        dE[j] = ei[i];
    }
}

// ---------------------------------------------------------------------------
bool Engine::ordered() const {
    return sort_every > 0 && !f32 && !fb_on && !stats_on && n_slices > kSortWindow &&
           n < (int64_t)1 << 32 &&
           n >= 2;
}

int Engine::order_grid() const { return device_cu_count() * blocks_per_cu; }

int Engine::ensure_order_buffers() {
    const int grid = order_grid();
    const int ndig = (n_slices + (1 << kSortDigitShift) - 1) >> kSortDigitShift;
    if (d_ord && ord_n == n && ord_grid == grid && ord_ndig == ndig) return BLMI_OK;
    invalidate_graph();
    (void)hipFree(d_ord);
    (void)hipFree(d_perm);
    (void)hipFree(d_win);
    (void)hipFree(d_digcnt);
    d_ord = nullptr;
    d_perm = nullptr;
    d_win = nullptr;
    d_digcnt = nullptr;
    BLMI_HIP_TRY(hipMalloc(&d_ord, sizeof(double) * 4 * (size_t)n));
    BLMI_HIP_TRY(hipMalloc(&d_perm, sizeof(unsigned) * 2 * (size_t)n));
    BLMI_HIP_TRY(hipMalloc(&d_win, sizeof(int) * grid));
    BLMI_HIP_TRY(hipMalloc(&d_digcnt, sizeof(unsigned) * ((size_t)ndig * grid + ndig)));
    ord_n = n;
    ord_grid = grid;
    ord_ndig = ndig;
    return BLMI_OK;
}

double *Engine::ord_x(int which) const { return d_ord + (size_t)which * 2 * n; }
double *Engine::ord_e(int which) const { return d_ord + (size_t)which * 2 * n + n; }
unsigned *Engine::ord_p(int which) const { return d_perm + (size_t)which * n; }

int Engine::order_in(hipStream_t s) {
    int rc = ensure_order_buffers();
    if (rc) return rc;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(order_in_kernel, dim3(blocks), dim3(256), 0, s, dt, dE, ord_x(0), ord_e(0),
                       ord_p(0), n);
    ord_cur = 0;
    rc = blmi_check_launch("order_in");
    if (rc) return rc;
    return order_sort(s);
}

int Engine::order_sort(hipStream_t s) {
    const int grid = ord_grid, ndig = ord_ndig;
    const int64_t n2 = n / 2;
    const int64_t per = (n2 + grid - 1) / grid;  // pairs per workgroup, as the turn kernel
    const int64_t per2 = 2 * per;
    unsigned *cnt = d_digcnt, *tot = d_digcnt + (size_t)ndig * grid;
    const size_t lds = sizeof(unsigned) * ndig;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void *)sort_digits_kernel,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e == hipSuccess)
            e = hipFuncSetAttribute((const void *)sort_scatter_kernel,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess)
            return blmi_set_error(BLMI_ERR_HIP, "sort LDS attribute: %s", hipGetErrorString(e));
    }
    const int c = ord_cur;
    hipLaunchKernelGGL(sort_digits_kernel, dim3(grid), dim3(1024), lds, s, ord_x(c), n, per2,
                       cut_left, inv_bw, n_slices, ndig, cnt);
    hipLaunchKernelGGL(sort_column_kernel, dim3(ndig), dim3(256), 0, s, cnt, grid, tot);
    hipLaunchKernelGGL(sort_total_kernel, dim3(1), dim3(1024), 0, s, tot, ndig);
    hipLaunchKernelGGL(sort_scatter_kernel, dim3(grid), dim3(1024), lds, s, ord_x(c), ord_e(c),
                       ord_p(c), ord_x(c ^ 1), ord_e(c ^ 1), ord_p(c ^ 1), n, per2, cut_left,
                       inv_bw, n_slices, ndig, cnt, tot);
    hipLaunchKernelGGL(sort_window_kernel, dim3((grid + 255) / 256), dim3(256), 0, s, tot, ndig,
                       grid, per, d_win);
    ord_cur = c ^ 1;
    return blmi_check_launch("order_sort");
}

int Engine::order_out(hipStream_t s) {
    const unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(order_out_kernel, dim3(blocks), dim3(256), 0, s, ord_x(ord_cur),
                       ord_e(ord_cur), ord_p(ord_cur), dt, dE, n);
    return blmi_check_launch("order_out");
}

template <int S>
static int launch_window_s(const Engine &e, const WinArgs &a, bool interp, bool publish,
                           size_t lds, hipStream_t s, const KernelEvents &ev) {
    using Fn = void (*)(double *, double *, int64_t, WinArgs);
    Fn fn = interp ? (publish ? turn_window_kernel<S, true, true> : turn_window_kernel<S, true, false>)
                   : (publish ? turn_window_kernel<S, false, true> : turn_window_kernel<S, false, false>);
    hipError_t err = hipFuncSetAttribute((const void *)fn,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (err != hipSuccess)
        return blmi_set_error(BLMI_ERR_HIP, "window kernel LDS attribute: %s", hipGetErrorString(err));
    launch(fn, dim3(e.ord_grid), dim3(e.block_size), lds, s, ev, e.ord_x(e.ord_cur),
           e.ord_e(e.ord_cur), e.n, a);
    return blmi_check_launch("turn_window");
}

int Engine::launch_window_turn(hipStream_t s, int off, bool interp, const KernelEvents &ev) {
    WinArgs a{};
    a.n_rf = n_rf;
    a.charge = charge;
    a.rf = d_rf;
    a.acc = d_acc;
    a.drift = d_drift;
    a.base = d_base;
    a.off = off;
    a.last_turn = n_turns - 1;
    a.tab = reinterpret_cast<const double2 *>(d_table);
    a.tab_params = d_table ? d_table + 2 * n_slices : nullptr;
    a.n_tab = n_slices - 1;
    a.counts0 = d_counts;
    a.counts1 = d_counts + (size_t)groups * n_slices;
    a.groups = groups;
    a.n_slices = n_slices;
    a.cut_left = cut_left;
    a.inv_bw = inv_bw;
    a.max_peel = max_peel;
    a.profile = profile;
    a.pending = d_base + 1;
    a.win = d_win;
    const size_t lds = sizeof(unsigned) * kSortWindow + sizeof(double2) * kSortWindow;
    const bool publish = !interp;
    switch (solver) {
        case 0: return launch_window_s<0>(*this, a, interp, publish, lds, s, ev);
        case 1: return launch_window_s<1>(*this, a, interp, publish, lds, s, ev);
        default: return launch_window_s<2>(*this, a, interp, publish, lds, s, ev);
    }
}

}  // namespace blmi
