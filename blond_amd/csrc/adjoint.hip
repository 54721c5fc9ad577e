// Differentiable tracking on the MI355X: forward turns, checkpoints and the
// hand-written reverse (adjoint) pass.
//
// The map per turn t (the intensity map of the fused engine with a smooth
// profile), in one of two forms chosen at compile time (MAP):
//
// MAP = kMapReference, the reference's own functions:
//   p_t     = smooth_histogram(dt_t)   histogram.cpp:75-132 with its quirk:
//                            f = (dt - cut_left) inv_bw, k = (int) f,
//                            d = f - k; bin k gets 0.5 - d, bin
//                            k2 = (int)(f + 1) (d > 1/2) or (int)(f - 1)
//                            gets 0.5 + d; particles outside
//                            [cut_left + bw/2, cut_right - bw/2] dropped
//   e1      = linear_interp_kick(dt, dE, V_t, bin_centers, q, 0)
//                            (linear_interp_kick.cpp:21-71: slope and
//                            intercept of each bin, formed per workgroup in
//                            LDS from V_t in the reference's order)
//   The weights are piecewise linear in dt, with jumps where a particle
//   crosses a bin centre or edge; the reverse pass uses their derivatives on
//   the pieces: d p[k]/d dt = -inv_bw, d p[k2]/d dt = +inv_bw, d kick/d dt =
//   the bin's slope, d kick/d V[k], V[k+1] = q (1 - s), q s with
//   s = (dt - bc[k]) inv_bw_bc.
//
// MAP = kMapSpline, a variant with continuous first derivatives:
//   p_t     = D(dt_t)        quadratic B-spline deposit on bin centres: a
//                            particle at bin-centre coordinate
//                            u = (dt - cut_left) inv_bw - 1/2 = k + d,
//                            k = round(u), adds (1/2 - d)^2 / 2, 3/4 - d^2,
//                            (1/2 + d)^2 / 2 to bins k-1, k, k+1
//   e1      = dE + q sum_b V_t[b] w_b(dt)   the same spline weights
//
// Both:
//   V_t     = -scale * (g conv p_t)[:ns],  g = irfft(Z, n_fft) (circular,
//                            as the FFT route of impedance.py:457-489)
//   e2      = e1 + q sum_j V_j sin(w_j dt + phi_j) + acc   (kick.cpp:16-35)
//   dt_{t+1} = dt + coeff_{t+1} e2,  dE_{t+1} = e2          (drift.cpp "simple")
// and the loss depends on p_T.
//
// Reverse turn t, per particle (state at t recomputed from a checkpoint):
//   lx <- lx + sum_b lp_{t+1}[b] dp_b/d dt                    deposit of p_{t+1}
//   le2 = le + coeff lx
//   lx <- lx + le2 (q sum_j V_j w_j cos(.) + d kick / d dt)   RF + induced kick
//   lV_t[b] += le2 d kick / d V_t[b]                           induced kick
//   dL/dV_j += le2 q sin(.),  dL/dphi_j += le2 q V_j cos(.)
//   le <- le2
// then, once per turn over a few workgroups: lp_t = -scale * (g corr lV_t)
// and corr[m] += sum_i lV_t[i] p_t[(i - m) mod n_fft], from which the host
// forms dL/dZ-parameters as -scale * sum_m irfft(dZ/dtheta)[m] corr[m].
//
// Traffic per macroparticle-turn: 32 B forward, 32 B recompute (slot to
// slot), 48 B reverse (state read, two adjoints read and written).
#include <math.h>
#include <string.h>

#include <vector>

#include "adjoint.h"
#include "blmi_common.h"
#include "particles.h"

namespace blmi {

int smooth_scale_bits(int64_t n);

constexpr int kAdjMaxRf = 4;
constexpr int kAdjBlock = 1024;

struct AdjArgs {
    int n_rf;
    double charge;
    const double *rf;  // [turn][3][n_rf]
    const double *acc;
    const DriftCoeffs<double> *drift;
    int ns;
    double cut_left, inv_bw_cut, fix_scale;
    const double *vt;    // [turn][ns] induced voltage V_t
    unsigned long long *dep;  // [copies][ns]
    double *lv;               // [copies][ns]
    const double *lp;         // [ns]
    double *gpart;            // [turn][grid][2 n_rf]
    // the reference map: smooth_histogram's range and the interpolation's
    // bin centres (histogram.cpp:82-85, linear_interp_kick.cpp:33)
    double lo, hi;
    const double *bc;  // [ns]
    double bc0, inv_bw_bc;
};

constexpr int kMapSpline = 0;
constexpr int kMapReference = 1;

// Quadratic B-spline: u = (x - cut_left) inv_bw - 1/2 = k + d, k = round(u),
// d in [-1/2, 1/2); weights of bins k-1, k, k+1 and their u-derivatives.
// k = kNoBin when none of the three bins is inside [0, ns).
constexpr int kNoBin = -1000000;
__device__ __forceinline__ int spline(double x, const AdjArgs &a, double w[3], double dw[3]) {
    const double u = (x - a.cut_left) * a.inv_bw_cut - 0.5;
    const double kf = floor(u + 0.5);
    if (!(kf >= -1.0 && kf <= (double)a.ns)) return kNoBin;  // also NaN
    const double d = u - kf;
    const double l = 0.5 - d, r = 0.5 + d;
    w[0] = 0.5 * l * l;
    w[1] = 0.75 - d * d;
    w[2] = 0.5 * r * r;
    dw[0] = -l;
    dw[1] = -2.0 * d;
    dw[2] = r;
    return (int)kf;
}

__device__ __forceinline__ void deposit_lds(unsigned long long *h, double x, const AdjArgs &a) {
    double w[3], dw[3];
    const int k = spline(x, a, w, dw);
    if (k == kNoBin) return;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const int b = k - 1 + q;
        if (b >= 0 && b < a.ns)
            atomicAdd(&h[b], (unsigned long long)__double2ll_rn(w[q] * a.fix_scale));
    }
}

// smooth_histogram's two bins of a particle (histogram.cpp:100-113): bin k
// gets 0.5 - d, bin k2 gets 0.5 + d.  False outside [lo, hi] (NaN
// included); a bin outside [0, ns) (a rounding corner at hi) is skipped by
// the callers.
__device__ __forceinline__ bool smooth_bins(double x, const AdjArgs &a, int &k, int &k2,
                                            double &d) {
    if (!(x >= a.lo && x <= a.hi)) return false;
    const double f = (x - a.cut_left) * a.inv_bw_cut;
    k = (int)f;
    d = f - (double)k;
    k2 = d > 0.5 ? (int)(f + 1.0) : (int)(f - 1.0);
    return true;
}

__device__ __forceinline__ void deposit_smooth(unsigned long long *h, double x, const AdjArgs &a) {
    int k, k2;
    double d;
    if (!smooth_bins(x, a, k, k2, d)) return;
    if (k >= 0 && k < a.ns)
        atomicAdd(&h[k], (unsigned long long)__double2ll_rn((0.5 - d) * a.fix_scale));
    if (k2 >= 0 && k2 < a.ns)
        atomicAdd(&h[k2], (unsigned long long)__double2ll_rn((0.5 + d) * a.fix_scale));
}

template <int MAP>
__device__ __forceinline__ void deposit(unsigned long long *h, double x, const AdjArgs &a) {
    if (MAP == kMapReference) deposit_smooth(h, x, a);
    else deposit_lds(h, x, a);
}

// linear_interp_kick.cpp:44-48 on V_t (acceleration 0, as
// TotalInducedVoltage.track passes it): slope and intercept of bin b, into
// LDS, in the reference's order of operations.
__device__ __forceinline__ void interp_table_lds(double *vk, double *fac, const double *vt,
                                                 const AdjArgs &a) {
    for (int b = threadIdx.x; b < a.ns - 1; b += blockDim.x) {
        const double v0 = vt[b], v1 = vt[b + 1];
        const double k = a.charge * (v1 - v0) * a.inv_bw_bc;
        vk[b] = k;
        fac[b] = (a.charge * v0 - a.bc[b] * k) + 0.0;
    }
}

// linear_interp_kick.cpp:56-63: the bin a particle is kicked by, or -1 (a
// signed range test in place of the reference's unsigned cast: negative,
// NaN and last-bin positions get no kick).
__device__ __forceinline__ int interp_bin(double x, const AdjArgs &a) {
    const double f = floor((x - a.bc0) * a.inv_bw_bc);
    return (f >= 0.0 && f < (double)(a.ns - 1)) ? (int)f : -1;
}

// into deposit parity `par` (the profile p_t of turn t is parity t & 1)
__device__ __forceinline__ void deposit_flush(unsigned long long *h, const AdjArgs &a, int par) {
    __syncthreads();
    unsigned long long *out = a.dep + ((size_t)par * kAdjCopies + blockIdx.x % kAdjCopies) * a.ns;
    for (int b = threadIdx.x; b < a.ns; b += blockDim.x)
        if (h[b]) atomic_add_agent(&out[b], h[b]);
}

template <int MAP>
__global__ void __launch_bounds__(kAdjBlock) adj_deposit_kernel(const double *__restrict__ x,
                                                                int64_t n, AdjArgs a) {
    extern __shared__ unsigned long long hdep[];
    for (int b = threadIdx.x; b < a.ns; b += blockDim.x) hdep[b] = 0ull;
    __syncthreads();
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        deposit<MAP>(hdep, x[i], a);
    deposit_flush(hdep, a, 0);
}

// q sum_b V[b] w_b(x) and (with D) its u-derivative, V in LDS
__device__ __forceinline__ double induced_kick(const double *V, double x, const AdjArgs &a,
                                               double *dkick, double w[3], int &k) {
    double dw[3];
    k = spline(x, a, w, dw);
    double s = 0.0, ds = 0.0;
    if (k != kNoBin) {
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const int b = k - 1 + q;
            if (b >= 0 && b < a.ns) {
                s = s + V[b] * w[q];
                ds = ds + V[b] * dw[q];
            }
        }
    }
    if (dkick) *dkick = a.charge * ds * a.inv_bw_cut;
    return a.charge * s;
}

// LDS of the forward turn kernel: the deposit (u64) plus V_t (spline) or
// the interpolation table's slopes and intercepts (reference).
static size_t adj_turn_lds(int map, int ns) {
    return sizeof(unsigned long long) * ns + sizeof(double) * (map == kMapReference ? 2 : 1) * ns;
}

// Forward turn t: (xi, ei) -> (xo, eo) (may alias); DEPOSIT adds p_{t+1}.
template <bool DEPOSIT, int MAP>
__global__ void __launch_bounds__(kAdjBlock) adj_turn_kernel(const double *xi, const double *ei,
                                                             double *xo, double *eo, int64_t n,
                                                             int t, AdjArgs a) {
    extern __shared__ unsigned long long hdep[];
    double *vl = reinterpret_cast<double *>(hdep + a.ns);  // V_t, or the table's slopes
    double *fac = vl + a.ns;                               // the table's intercepts
    // software pipeline: a thread's next particle is loaded while the
    // current one is computed (the first before the LDS set-up)
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t i_first = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double xn = 0.0, en = 0.0;
    if (i_first < n) {
        xn = xi[i_first];
        en = ei[i_first];
    }
    if (DEPOSIT)
        for (int b = threadIdx.x; b < a.ns; b += blockDim.x) hdep[b] = 0ull;
    if (MAP == kMapReference)
        interp_table_lds(vl, fac, a.vt + (size_t)t * a.ns, a);
    else
        for (int b = threadIdx.x; b < a.ns; b += blockDim.x) vl[b] = a.vt[(size_t)t * a.ns + b];
    __syncthreads();
    const double *V = a.rf + (size_t)t * 3 * a.n_rf;
    const double *w = V + a.n_rf;
    const double *phi = w + a.n_rf;
    const double acc = a.acc[t];
    const double coeff = a.drift[t + 1].coeff;
    for (int64_t i = i_first; i < n; i += stride) {
        double x = xn, e = en;
        if (i + stride < n) {
            xn = xi[i + stride];
            en = ei[i + stride];
        }
        if (MAP == kMapReference) {
            const int k = interp_bin(x, a);
            if (k >= 0) e = e + (x * vl[k] + fac[k]);  // linear_interp_kick.cpp:60
        } else {
            double wk[3];
            int k;
            e = e + induced_kick(vl, x, a, nullptr, wk, k);
        }
        double s = 0.0;
        for (int j = 0; j < a.n_rf; ++j) s = s + V[j] * cephes_sin(w[j] * x + phi[j]);
        e = e + (a.charge * s + acc);
        x = x + coeff * e;
        xo[i] = x;
        eo[i] = e;
        if (DEPOSIT) deposit<MAP>(hdep, x, a);
    }
    if (DEPOSIT) deposit_flush(hdep, a, (t + 1) & 1);
}

// Lanes that share one output of a circular product (each a residue class
// of the k range); 16 with 1024-thread workgroups (c5: 2.51e10 against
// 2.42e10 macroparticle-turns/s with 4 lanes and 256 threads).
constexpr int kLanes = 16;

// Circular dot product over k = r, r + kLanes, ... < k_end of G[m(k)] X[k],
// m(k) = (s0 + SGN k) mod n_fft (s0 + SGN k within (-n_fft, 2 n_fft)), G
// read as zero from g_len on; four accumulators so the LDS reads of
// consecutive terms overlap.  Lane r of a group takes residue r: the group
// reads consecutive words.
template <int SGN>
__device__ __forceinline__ double circ_dot(const double *G, int n_fft, int g_len, const double *X,
                                           int r, int k_end, int s0) {
    auto gat = [&](int k) -> double {
        int m = s0 + SGN * k;
        m += m < 0 ? n_fft : 0;
        m -= m >= n_fft ? n_fft : 0;
        const double g = G[m < g_len ? m : 0];
        return m < g_len ? g : 0.0;
    };
    constexpr int L = kLanes;
    double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
    int k = r;
    for (; k + 3 * L < k_end; k += 4 * L) {
        v0 = fma(gat(k), X[k], v0);
        v1 = fma(gat(k + L), X[k + L], v1);
        v2 = fma(gat(k + 2 * L), X[k + 2 * L], v2);
        v3 = fma(gat(k + 3 * L), X[k + 3 * L], v3);
    }
    for (; k < k_end; k += L) v0 = fma(gat(k), X[k], v0);
    return (v0 + v1) + (v2 + v3);
}

// An output of a circular product is computed by a group of kLanes lanes;
// the groups stay together because every trip count below is a multiple of
// kLanes.  The group's sum, in every lane.
__device__ __forceinline__ double lanes_sum(double v) {
#pragma unroll
    for (int o = 1; o < kLanes; o <<= 1) v += __shfl_xor(v, o);
    return v;
}

// Outputs per workgroup of the slice-axis kernels (kAdjSliceBlock / kLanes).
constexpr int kAdjSliceBlock = 1024;
constexpr int kAdjSliceOut = kAdjSliceBlock / kLanes;

// p_t from the deposit copies of parity t & 1, stored; with `voltage`,
// V_t = -scale * (g conv p_t) stored too.  Workgroup b owns outputs
// [64 b, 64 b + 64): it stores them and clears the same bins of the other
// parity's copies (read by the previous launch, written by the next turn).
__global__ void __launch_bounds__(kAdjSliceBlock) adj_profile_kernel(AdjArgs a, int t, bool voltage,
                                                                     double *prof, const double *g,
                                                                     int n_fft, double scale,
                                                                     double *vt) {
    extern __shared__ double sh[];
    double *p = sh, *gs = sh + a.ns;
    const int ns = a.ns;
    const double inv = 1.0 / a.fix_scale;
    const unsigned long long *din = a.dep + (size_t)(t & 1) * kAdjCopies * ns;
    unsigned long long *dclr = a.dep + (size_t)((t + 1) & 1) * kAdjCopies * ns;
    const int o0 = blockIdx.x * kAdjSliceOut, o1 = min(ns, o0 + kAdjSliceOut);
    for (int j = threadIdx.x; j < ns; j += blockDim.x) {
        // every copy is loaded before the sum: one memory round trip
        unsigned long long v[kAdjCopies];
#pragma unroll
        for (int c = 0; c < kAdjCopies; ++c) v[c] = din[(size_t)c * ns + j];
        unsigned long long s = 0;
#pragma unroll
        for (int c = 0; c < kAdjCopies; ++c) s += v[c];
        p[j] = (double)(long long)s * inv;
        if (j >= o0 && j < o1) {
            prof[(size_t)t * ns + j] = p[j];
#pragma unroll
            for (int c = 0; c < kAdjCopies; ++c) dclr[(size_t)c * ns + j] = 0ull;
        }
    }
    if (!voltage) return;
    for (int m = threadIdx.x; m < n_fft; m += blockDim.x) gs[m] = g[m];
    __syncthreads();
    // V_t[i] = -scale sum_j g[(i - j) mod n_fft] p[j]
    for (int q = threadIdx.x; q < (o1 - o0) * kLanes; q += blockDim.x) {
        const int i = o0 + q / kLanes, r = q % kLanes;
        const double v = lanes_sum(circ_dot<-1>(gs, n_fft, n_fft, p, r, ns, i));
        if (r == 0) vt[(size_t)t * ns + i] = -(scale * v);
    }
}

// LDS of the reverse kernel: lp_{t+1} and lV_t, plus V_t (spline) or the
// interpolation table's slopes, intercepts and bin centres (reference).
static size_t adj_reverse_lds(int map, int ns) {
    return sizeof(double) * (map == kMapReference ? 5 : 3) * (size_t)ns;
}

// Reverse turn t (state (x, e) at turn t): see the file comment.
template <int MAP>
__global__ void __launch_bounds__(kAdjBlock) adj_reverse_kernel(const double *__restrict__ xs,
                                                                const double *__restrict__ es,
                                                                double *__restrict__ lam_x,
                                                                double *__restrict__ lam_e,
                                                                int64_t n, int t, AdjArgs a) {
    extern __shared__ double shr[];
    double *lp = shr, *lvs = shr + a.ns, *vl = shr + 2 * a.ns;
    double *fac = shr + 3 * a.ns, *bcs = shr + 4 * a.ns;  // reference map only
    __shared__ double red[kAdjBlock / 64][2 * kAdjMaxRf];
    // software pipeline as in adj_turn_kernel
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t i_first = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double xn = 0.0, en = 0.0, lxn = 0.0, len = 0.0;
    if (i_first < n) {
        xn = xs[i_first];
        en = es[i_first];
        lxn = lam_x[i_first];
        len = lam_e[i_first];
    }
    for (int b = threadIdx.x; b < a.ns; b += blockDim.x) {
        lp[b] = a.lp[b];
        lvs[b] = 0.0;
        if (MAP == kMapReference) bcs[b] = a.bc[b];
        else vl[b] = a.vt[(size_t)t * a.ns + b];
    }
    if (MAP == kMapReference) interp_table_lds(vl, fac, a.vt + (size_t)t * a.ns, a);
    __syncthreads();
    const double *V = a.rf + (size_t)t * 3 * a.n_rf;
    const double *w = V + a.n_rf;
    const double *phi = w + a.n_rf;
    const double acc = a.acc[t];
    const double coeff = a.drift[t + 1].coeff;
    double gsum[2 * kAdjMaxRf];
#pragma unroll
    for (int j = 0; j < 2 * kAdjMaxRf; ++j) gsum[j] = 0.0;
    for (int64_t i = i_first; i < n; i += stride) {
        const double x = xn, e = en, lam_x_i = lxn, lam_e_i = len;
        if (i + stride < n) {
            xn = xs[i + stride];
            en = es[i + stride];
            lxn = lam_x[i + stride];
            len = lam_e[i + stride];
        }
        // forward step t, as adj_turn_kernel
        double wk[3], dkick;
        int k;
        double e1;
        if (MAP == kMapReference) {
            k = interp_bin(x, a);
            e1 = k >= 0 ? e + (x * vl[k] + fac[k]) : e;
            dkick = k >= 0 ? vl[k] : 0.0;
        } else {
            e1 = e + induced_kick(vl, x, a, &dkick, wk, k);
        }
        // constant trip counts with a guard keep sn, cs and gsum in registers
        double sn[kAdjMaxRf], cs[kAdjMaxRf];
        double s = 0.0, ds = 0.0;
#pragma unroll
        for (int j = 0; j < kAdjMaxRf; ++j) {
            sn[j] = cs[j] = 0.0;
            if (j < a.n_rf) {
                cephes_sincos(w[j] * x + phi[j], sn[j], cs[j]);
                s = s + V[j] * sn[j];
                ds = ds + V[j] * w[j] * cs[j];
            }
        }
        const double e2 = e1 + (a.charge * s + acc);
        const double x1 = x + coeff * e2;
        // deposit adjoint of p_{t+1} at x1
        double lx = lam_x_i;
        if (MAP == kMapReference) {
            int k1, k2;
            double d;
            if (smooth_bins(x1, a, k1, k2, d)) {
                const double l1 = (k1 >= 0 && k1 < a.ns) ? lp[k1] : 0.0;
                const double l2 = (k2 >= 0 && k2 < a.ns) ? lp[k2] : 0.0;
                lx = lx + (l2 - l1) * a.inv_bw_cut;
            }
        } else {
            double sw[3], sdw[3];
            const int kk = spline(x1, a, sw, sdw);
            if (kk != kNoBin) {
                double s3 = 0.0;
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    const int b = kk - 1 + q;
                    if (b >= 0 && b < a.ns) s3 += lp[b] * sdw[q];
                }
                lx = lx + s3 * a.inv_bw_cut;
            }
        }
        const double le2 = lam_e_i + coeff * lx;
        lx = lx + le2 * (a.charge * ds);
#pragma unroll
        for (int j = 0; j < kAdjMaxRf; ++j) {
            if (j < a.n_rf) {
                gsum[2 * j] += le2 * a.charge * sn[j];
                gsum[2 * j + 1] += le2 * a.charge * V[j] * cs[j];
            }
        }
        lx = lx + le2 * dkick;
        if (MAP == kMapReference) {
            if (k >= 0) {
                const double sk = (x - bcs[k]) * a.inv_bw_bc;
                atomicAdd(&lvs[k], le2 * a.charge * (1.0 - sk));
                atomicAdd(&lvs[k + 1], le2 * a.charge * sk);
            }
        } else if (k != kNoBin) {
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const int b = k - 1 + q;
                if (b >= 0 && b < a.ns) atomicAdd(&lvs[b], le2 * a.charge * wk[q]);
            }
        }
        lam_x[i] = lx;
        lam_e[i] = le2;
    }
    // per-workgroup RF gradient sums (fixed order: wavefront tree, then waves)
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < 2 * kAdjMaxRf; ++j) {
        if (j < 2 * a.n_rf) {
            double v = gsum[j];
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
            if (lane == 0) red[wid][j] = v;
        }
    }
    __syncthreads();
    if ((int)threadIdx.x < 2 * a.n_rf) {
        double v = 0.0;
        for (int q = 0; q < kAdjBlock / 64; ++q) v += red[q][threadIdx.x];
        a.gpart[((size_t)t * gridDim.x + blockIdx.x) * 2 * a.n_rf + threadIdx.x] = v;
    }
    double *out = a.lv + ((size_t)(t & 1) * kAdjCopies + blockIdx.x % kAdjCopies) * a.ns;
    for (int b = threadIdx.x; b < a.ns; b += blockDim.x)
        if (lvs[b] != 0.0) atomic_add_agent(&out[b], lvs[b]);
}

// lV_t (copies of parity t & 1 summed) -> lp_t = -scale (g corr lV_t) and
// corr[m] += sum_i lV_t[i] p_t[(i - m) mod n_fft].  The outputs (ns values
// of lp_t, then the corr entries that have a term: m < ns and
// m > n_fft - ns) are split 64 per workgroup; each workgroup also clears
// its share of the other parity's copies (read by the previous launch,
// written by the next reverse turn).
__global__ void __launch_bounds__(kAdjSliceBlock) adj_iv_kernel(AdjArgs a, int t, const double *prof,
                                                                const double *g, int n_fft,
                                                                double scale, double *lp_out,
                                                                double *corr) {
    extern __shared__ double sh[];
    const int ns = a.ns;
    double *lvs = sh, *p = sh + ns, *gs = sh + 2 * ns;
    const double *lin = a.lv + (size_t)(t & 1) * kAdjCopies * ns;
    double *lclr = a.lv + (size_t)((t + 1) & 1) * kAdjCopies * ns;
    const int per = (ns + (int)gridDim.x - 1) / (int)gridDim.x;
    const int c0 = blockIdx.x * per, c1 = min(ns, c0 + per);
    for (int i = threadIdx.x; i < ns; i += blockDim.x) {
        // every copy is loaded before the sum: one memory round trip
        double v[kAdjCopies];
#pragma unroll
        for (int c = 0; c < kAdjCopies; ++c) v[c] = lin[(size_t)c * ns + i];
        double s = 0.0;
#pragma unroll
        for (int c = 0; c < kAdjCopies; ++c) s += v[c];
        lvs[i] = s;
        p[i] = prof[(size_t)t * ns + i];
        if (i >= c0 && i < c1)
#pragma unroll
            for (int c = 0; c < kAdjCopies; ++c) lclr[(size_t)c * ns + i] = 0.0;
    }
    for (int m = threadIdx.x; m < n_fft; m += blockDim.x) gs[m] = g[m];
    __syncthreads();
    const int hi0 = max(ns, n_fft - ns + 1);
    const int n_out = ns + ns + (n_fft - hi0);
    const int o0 = blockIdx.x * kAdjSliceOut, o1 = min(n_out, o0 + kAdjSliceOut);
    for (int q = threadIdx.x; q < (o1 - o0) * kLanes; q += blockDim.x) {
        const int o = o0 + q / kLanes, r = q % kLanes;
        int m = 0;
        double v;
        if (o < ns) {
            v = circ_dot<1>(gs, n_fft, n_fft, lvs, r, ns, -o);  // lp_t[o]
        } else {
            m = o - ns < ns ? o - ns : hi0 + (o - 2 * ns);
            v = circ_dot<1>(p, n_fft, ns, lvs, r, ns, -m);  // corr[m]
        }
        v = lanes_sum(v);
        if (r == 0) {
            if (o < ns) lp_out[o] = -(scale * v);
            else corr[m] += v;
        }
    }
}

// Workgroups of the slice-axis kernels: 64 outputs each.
static int adj_profile_grid(int ns) { return (ns + kAdjSliceOut - 1) / kAdjSliceOut; }
static int adj_iv_grid(int ns, int64_t n_fft) {
    const int64_t hi0 = ns > n_fft - ns + 1 ? ns : n_fft - ns + 1;
    const int64_t n_out = 2 * (int64_t)ns + (n_fft - hi0);
    return (int)((n_out + kAdjSliceOut - 1) / kAdjSliceOut);
}

// ---------------------------------------------------------------------------
Adjoint::~Adjoint() { release(); }

void Adjoint::invalidate_graphs() {
    if (fwd_exec) (void)hipGraphExecDestroy(fwd_exec);
    if (bwd_exec) (void)hipGraphExecDestroy(bwd_exec);
    fwd_exec = bwd_exec = nullptr;
    fwd_graph_turns = bwd_graph_turns = -1;
}

int Adjoint::ensure_stream() {
    if (own) return BLMI_OK;
    BLMI_HIP_TRY(hipStreamCreateWithFlags(&own, hipStreamNonBlocking));
    BLMI_HIP_TRY(hipEventCreateWithFlags(&ev_in, hipEventDisableTiming));
    BLMI_HIP_TRY(hipEventCreateWithFlags(&ev_out, hipEventDisableTiming));
    return BLMI_OK;
}

// dst[i] = src[i] (a kernel node: the passes are graph-captured)
__global__ void adj_copy_kernel(double *__restrict__ dst, const double *__restrict__ src,
                                int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        dst[i] = src[i];
}

static void adj_copy(double *dst, const double *src, int64_t n, int grid, hipStream_t s) {
    hipLaunchKernelGGL(adj_copy_kernel, dim3(grid), dim3(256), 0, s, dst, src, n);
}

void Adjoint::release() {
    invalidate_graphs();
    if (own) (void)hipStreamDestroy(own);
    if (ev_in) (void)hipEventDestroy(ev_in);
    if (ev_out) (void)hipEventDestroy(ev_out);
    own = nullptr;
    ev_in = ev_out = nullptr;
    void *bufs[] = {d_rf, d_acc, d_drift, d_x0, d_work, d_ckpt, d_slots, d_lam, d_dep,
                    d_prof, d_vt, d_g, d_lv, d_lp, d_corr, d_gpart, d_bc};
    for (void *p : bufs) (void)hipFree(p);
    d_rf = d_acc = d_x0 = d_work = d_ckpt = d_slots = d_lam = nullptr;
    d_prof = d_vt = d_g = d_lv = d_lp = d_corr = d_gpart = d_bc = nullptr;
    d_drift = nullptr;
    d_dep = nullptr;
}

int Adjoint::init(int64_t n_, int n_rf_, int n_turns_, int n_slices_, int every_) {
    n = n_;
    n_rf = n_rf_;
    n_turns = n_turns_;
    n_slices = n_slices_;
    every = every_;
    BLMI_HIP_TRY(hipGetDevice(&device));
    grid = device_cu_count();
    const size_t nt = (size_t)n_turns + 1;
    const size_t pn = 2 * (size_t)n;
    const int n_ckpt = (n_turns + every - 1) / every;
    BLMI_HIP_TRY(hipMalloc(&d_rf, sizeof(double) * 3 * n_rf * nt));
    BLMI_HIP_TRY(hipMalloc(&d_acc, sizeof(double) * nt));
    BLMI_HIP_TRY(hipMalloc(&d_drift, sizeof(DriftCoeffs<double>) * nt));
    BLMI_HIP_TRY(hipMalloc(&d_x0, sizeof(double) * pn));
    BLMI_HIP_TRY(hipMalloc(&d_work, sizeof(double) * pn));
    BLMI_HIP_TRY(hipMalloc(&d_ckpt, sizeof(double) * pn * n_ckpt));
    BLMI_HIP_TRY(hipMalloc(&d_slots, sizeof(double) * pn * every));
    BLMI_HIP_TRY(hipMalloc(&d_lam, sizeof(double) * pn));
    BLMI_HIP_TRY(hipMalloc(&d_dep, sizeof(unsigned long long) * 2 * kAdjCopies * n_slices));
    BLMI_HIP_TRY(hipMemset(d_dep, 0, sizeof(unsigned long long) * 2 * kAdjCopies * n_slices));
    BLMI_HIP_TRY(hipMalloc(&d_prof, sizeof(double) * nt * n_slices));
    BLMI_HIP_TRY(hipMalloc(&d_vt, sizeof(double) * nt * n_slices));
    BLMI_HIP_TRY(hipMemset(d_vt, 0, sizeof(double) * nt * n_slices));
    BLMI_HIP_TRY(hipMalloc(&d_lv, sizeof(double) * 2 * kAdjCopies * n_slices));
    BLMI_HIP_TRY(hipMemset(d_lv, 0, sizeof(double) * 2 * kAdjCopies * n_slices));
    BLMI_HIP_TRY(hipMalloc(&d_lp, sizeof(double) * n_slices));
    BLMI_HIP_TRY(hipMalloc(&d_gpart, sizeof(double) * (size_t)n_turns * grid * 2 * n_rf));
    fixed_bits = smooth_scale_bits(n);
    return BLMI_OK;
}

int Adjoint::set_particles(const double *dt, const double *dE) {
    BLMI_HIP_TRY(hipMemcpy(d_x0, dt, sizeof(double) * n, hipMemcpyDeviceToDevice));
    BLMI_HIP_TRY(hipMemcpy(d_x0 + n, dE, sizeof(double) * n, hipMemcpyDeviceToDevice));
    forward_done = false;
    return BLMI_OK;
}

int Adjoint::set_programs(const double *rf_host, const double *acc_host, const double *t_rev,
                          double length_ratio, double alpha_order, const double *eta0,
                          const double *eta1, const double *eta2, const double *alpha0,
                          const double *alpha1, const double *alpha2, const double *beta,
                          const double *energy, double charge_) {
    const size_t nt = (size_t)n_turns + 1;
    // the captured passes hold scalars by value: a changed one re-captures
    if (charge_ != charge) invalidate_graphs();
    charge = charge_;
    BLMI_HIP_TRY(hipMemcpy(d_rf, rf_host, sizeof(double) * 3 * n_rf * nt, hipMemcpyHostToDevice));
    BLMI_HIP_TRY(hipMemcpy(d_acc, acc_host, sizeof(double) * nt, hipMemcpyHostToDevice));
    std::vector<DriftCoeffs<double>> dc(nt);
    for (size_t i = 0; i < nt; ++i)
        dc[i] = make_drift_coeffs<double>(0, t_rev[i], length_ratio, alpha_order, eta0[i], eta1[i],
                                          eta2[i], alpha0[i], alpha1[i], alpha2[i], beta[i],
                                          energy[i]);
    BLMI_HIP_TRY(hipMemcpy(d_drift, dc.data(), sizeof(DriftCoeffs<double>) * nt,
                           hipMemcpyHostToDevice));
    forward_done = false;
    return BLMI_OK;
}

int Adjoint::set_slices(double cl, double cr) {
    const double ibw = (double)n_slices / (cr - cl);  // histogram.cpp:82
    if (cl != cut_left || ibw != inv_bw_cut) invalidate_graphs();
    cut_left = cl;
    inv_bw_cut = ibw;
    const double bw = (cr - cl) / (double)n_slices;  // histogram.cpp:83-85
    lo = cl + bw * 0.5;
    hi = cr - bw * 0.5;
    forward_done = false;
    return BLMI_OK;
}

int Adjoint::set_map(int map_, const double *bc_host) {
    if (map_ != kMapSpline && map_ != kMapReference)
        return blmi_set_error(BLMI_ERR_ARG, "adjoint: unknown map %d", map_);
    if (map_ == kMapReference) {
        if (!bc_host || n_slices < 2)
            return blmi_set_error(BLMI_ERR_ARG, "adjoint: the reference map needs >= 2 bin centres");
        if (!d_bc) BLMI_HIP_TRY(hipMalloc(&d_bc, sizeof(double) * n_slices));
        BLMI_HIP_TRY(hipMemcpy(d_bc, bc_host, sizeof(double) * n_slices, hipMemcpyHostToDevice));
        bc0 = bc_host[0];
        // linear_interp_kick.cpp:33-35
        inv_bw_bc = (double)(n_slices - 1) / (bc_host[n_slices - 1] - bc_host[0]);
    }
    map = map_;
    invalidate_graphs();
    forward_done = false;
    return BLMI_OK;
}

int Adjoint::set_wake(int64_t n_fft_, const double *g_host, double scale_) {
    if (n_fft_ < n_slices || n_fft_ > 8192)
        return blmi_set_error(BLMI_ERR_UNSUPPORTED, "adjoint: n_fft %lld outside [n_slices, 8192]",
                              (long long)n_fft_);
    if (n_fft_ != n_fft || scale_ != scale) invalidate_graphs();
    if (n_fft_ != n_fft) {
        (void)hipFree(d_g);
        (void)hipFree(d_corr);
        d_g = d_corr = nullptr;
        BLMI_HIP_TRY(hipMalloc(&d_g, sizeof(double) * n_fft_));
        BLMI_HIP_TRY(hipMalloc(&d_corr, sizeof(double) * n_fft_));
    }
    n_fft = n_fft_;
    scale = scale_;
    BLMI_HIP_TRY(hipMemcpy(d_g, g_host, sizeof(double) * n_fft, hipMemcpyHostToDevice));
    // slice-axis kernels: profile + wake (+ adjoint) in LDS, up to 80 KiB
    const int lds_p = (int)(sizeof(double) * ((size_t)n_slices + n_fft));
    const int lds_i = (int)(sizeof(double) * (2 * (size_t)n_slices + n_fft));
    BLMI_HIP_TRY(hipFuncSetAttribute((const void *)adj_profile_kernel,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, lds_p));
    BLMI_HIP_TRY(hipFuncSetAttribute((const void *)adj_iv_kernel,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, lds_i));
    forward_done = false;
    return BLMI_OK;
}

static AdjArgs adj_args(const Adjoint &j) {
    AdjArgs a{};
    a.n_rf = j.n_rf;
    a.charge = j.charge;
    a.rf = j.d_rf;
    a.acc = j.d_acc;
    a.drift = j.d_drift;
    a.ns = j.n_slices;
    a.cut_left = j.cut_left;
    a.inv_bw_cut = j.inv_bw_cut;
    a.fix_scale = ldexp(1.0, j.fixed_bits);
    a.vt = j.d_vt;
    a.dep = j.d_dep;
    a.lv = j.d_lv;
    a.lp = j.d_lp;
    a.gpart = j.d_gpart;
    a.lo = j.lo;
    a.hi = j.hi;
    a.bc = j.d_bc;
    a.bc0 = j.bc0;
    a.inv_bw_bc = j.inv_bw_bc;
    return a;
}

template <int MAP>
static int enqueue_forward_map(Adjoint &j, int turns, hipStream_t s) {
    const int64_t n = j.n;
    const int n_slices = j.n_slices, every = j.every, grid = j.grid;
    const int64_t n_fft = j.n_fft;
    const double scale = j.scale;
    double *d_dep_ = reinterpret_cast<double *>(j.d_dep);
    double *d_ckpt = j.d_ckpt, *d_work = j.d_work, *d_x0 = j.d_x0, *d_prof = j.d_prof;
    double *d_g = j.d_g, *d_vt = j.d_vt;
    const AdjArgs a = adj_args(j);
    const size_t pn = 2 * (size_t)n;
    const size_t lds_dep = adj_turn_lds(MAP, n_slices);
    const size_t lds_one = sizeof(double) * ((size_t)n_slices + n_fft);
    const int pgrid = adj_profile_grid(n_slices);
    double *x = d_work, *e = d_work + n;
    // both deposit parities start cleared (a previous pass leaves its last
    // profile's copies behind)
    int rc = launch_fill<double>(d_dep_, 2 * (int64_t)kAdjCopies * n_slices, 0.0, s);
    if (rc) return rc;
    if (every == 1) {
        // every state is kept (n_turns x 16 B per macroparticle of HBM): turn t
        // reads checkpoint t and writes checkpoint t + 1 (the last state goes
        // to the work array), and the reverse pass recomputes nothing
        adj_copy(d_ckpt, d_x0, (int64_t)pn, grid, s);
        hipLaunchKernelGGL(adj_deposit_kernel<MAP>, dim3(grid), dim3(kAdjBlock), lds_dep, s, d_ckpt,
                           n, a);
        hipLaunchKernelGGL(adj_profile_kernel, dim3(pgrid), dim3(kAdjSliceBlock), lds_one, s, a, 0,
                           true, d_prof, d_g, (int)n_fft, scale, d_vt);
        for (int t = 0; t < turns; ++t) {
            const double *si = d_ckpt + (size_t)t * pn;
            double *so = t + 1 < turns ? d_ckpt + (size_t)(t + 1) * pn : d_work;
            hipLaunchKernelGGL((adj_turn_kernel<true, MAP>), dim3(grid), dim3(kAdjBlock), lds_dep, s,
                               si, si + n, so, so + n, n, t, a);
            hipLaunchKernelGGL(adj_profile_kernel, dim3(pgrid), dim3(kAdjSliceBlock), lds_one, s, a,
                               t + 1, t + 1 < turns, d_prof, d_g, (int)n_fft, scale, d_vt);
        }
        return blmi_check_launch("adjoint_forward");
    }
    adj_copy(d_work, d_x0, (int64_t)pn, grid, s);
    hipLaunchKernelGGL(adj_deposit_kernel<MAP>, dim3(grid), dim3(kAdjBlock), lds_dep, s, x, n, a);
    hipLaunchKernelGGL(adj_profile_kernel, dim3(pgrid), dim3(kAdjSliceBlock), lds_one, s, a, 0,
                       true, d_prof, d_g, (int)n_fft, scale, d_vt);
    for (int t = 0; t < turns; ++t) {
        if (t % every == 0) adj_copy(d_ckpt + (size_t)(t / every) * pn, d_work, (int64_t)pn, grid, s);
        hipLaunchKernelGGL((adj_turn_kernel<true, MAP>), dim3(grid), dim3(kAdjBlock), lds_dep, s, x,
                           e, x, e, n, t, a);
        hipLaunchKernelGGL(adj_profile_kernel, dim3(pgrid), dim3(kAdjSliceBlock), lds_one, s, a,
                           t + 1, t + 1 < turns, d_prof, d_g, (int)n_fft, scale, d_vt);
    }
    return blmi_check_launch("adjoint_forward");
}

int Adjoint::enqueue_forward(int turns, hipStream_t s) {
    return map == kMapReference ? enqueue_forward_map<kMapReference>(*this, turns, s)
                                : enqueue_forward_map<kMapSpline>(*this, turns, s);
}

template <int MAP>
static int enqueue_backward_map(Adjoint &j, int T, hipStream_t s) {
    const int64_t n = j.n;
    const int n_slices = j.n_slices, every = j.every, grid = j.grid;
    const int64_t n_fft = j.n_fft;
    const double scale = j.scale;
    double *d_lam = j.d_lam, *d_corr = j.d_corr, *d_lv = j.d_lv, *d_ckpt = j.d_ckpt;
    double *d_slots = j.d_slots, *d_prof = j.d_prof, *d_g = j.d_g, *d_lp = j.d_lp;
    const size_t pn = 2 * (size_t)n;
    const size_t lds_dep = adj_turn_lds(MAP, n_slices);
    const size_t lds_rev = adj_reverse_lds(MAP, n_slices);
    const size_t lds_one = sizeof(double) * (2 * (size_t)n_slices + n_fft);
    AdjArgs a = adj_args(j);
    const int igrid = adj_iv_grid(n_slices, n_fft);
    int rc = launch_fill<double>(d_lam, (int64_t)pn, 0.0, s);
    if (!rc) rc = launch_fill<double>(d_corr, n_fft, 0.0, s);
    if (!rc) rc = launch_fill<double>(d_lv, 2 * (int64_t)kAdjCopies * n_slices, 0.0, s);
    if (rc) return rc;
    if (every == 1) {
        for (int t = T - 1; t >= 0; --t) {
            const double *st = d_ckpt + (size_t)t * pn;
            hipLaunchKernelGGL(adj_reverse_kernel<MAP>, dim3(grid), dim3(kAdjBlock), lds_rev, s, st,
                               st + n, d_lam, d_lam + n, n, t, a);
            hipLaunchKernelGGL(adj_iv_kernel, dim3(igrid), dim3(kAdjSliceBlock), lds_one, s, a, t,
                               d_prof, d_g, (int)n_fft, scale, d_lp, d_corr);
        }
        return blmi_check_launch("adjoint_backward");
    }
    for (int seg = (T - 1) / every; seg >= 0; --seg) {
        const int t0 = seg * every, t1 = (t0 + every < T) ? t0 + every : T;
        // states t0 .. t1-1 into slots 0 .. t1-t0-1
        adj_copy(d_slots, d_ckpt + (size_t)seg * pn, (int64_t)pn, grid, s);
        for (int t = t0; t < t1 - 1; ++t) {
            double *si = d_slots + (size_t)(t - t0) * pn, *so = si + pn;
            hipLaunchKernelGGL((adj_turn_kernel<false, MAP>), dim3(grid), dim3(kAdjBlock), lds_dep, s,
                               si, si + n, so, so + n, n, t, a);
        }
        for (int t = t1 - 1; t >= t0; --t) {
            const double *st = d_slots + (size_t)(t - t0) * pn;
            hipLaunchKernelGGL(adj_reverse_kernel<MAP>, dim3(grid), dim3(kAdjBlock), lds_rev, s, st,
                               st + n, d_lam, d_lam + n, n, t, a);
            hipLaunchKernelGGL(adj_iv_kernel, dim3(igrid), dim3(kAdjSliceBlock), lds_one, s, a, t,
                               d_prof, d_g, (int)n_fft, scale, d_lp, d_corr);
        }
    }
    return blmi_check_launch("adjoint_backward");
}

int Adjoint::enqueue_backward(int T, hipStream_t s) {
    return map == kMapReference ? enqueue_backward_map<kMapReference>(*this, T, s)
                                : enqueue_backward_map<kMapSpline>(*this, T, s);
}

// Capture `enqueue` on the own stream into *exec (once per turn count).
template <typename F>
static int capture(hipStream_t own, hipGraphExec_t *exec, F enqueue) {
    BLMI_HIP_TRY(hipStreamBeginCapture(own, hipStreamCaptureModeThreadLocal));
    int rc = enqueue();
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(own, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess)
        return blmi_set_error(BLMI_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    hipError_t ie = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess)
        return blmi_set_error(BLMI_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(ie));
    return BLMI_OK;
}

int Adjoint::forward(int turns, double *profile_out, hipStream_t user) {
    if (!d_g) return blmi_set_error(BLMI_ERR_ARG, "adjoint: set the wake first");
    if (turns < 1 || turns > n_turns)
        return blmi_set_error(BLMI_ERR_ARG, "adjoint: turns %d outside [1, %d]", turns, n_turns);
    int rc = ensure_stream();
    if (rc) return rc;
    if (!fwd_exec || fwd_graph_turns != turns) {
        if (fwd_exec) (void)hipGraphExecDestroy(fwd_exec);
        fwd_exec = nullptr;
        rc = capture(own, &fwd_exec, [&] { return enqueue_forward(turns, own); });
        if (rc) return rc;
        fwd_graph_turns = turns;
    }
    BLMI_HIP_TRY(hipEventRecord(ev_in, user));
    BLMI_HIP_TRY(hipStreamWaitEvent(own, ev_in, 0));
    BLMI_HIP_TRY(hipGraphLaunch(fwd_exec, own));
    BLMI_HIP_TRY(hipMemcpyAsync(profile_out, d_prof + (size_t)turns * n_slices,
                                sizeof(double) * n_slices, hipMemcpyDeviceToDevice, own));
    BLMI_HIP_TRY(hipEventRecord(ev_out, own));
    BLMI_HIP_TRY(hipStreamWaitEvent(user, ev_out, 0));
    forward_done = true;
    forward_turns = turns;
    return BLMI_OK;
}

int Adjoint::backward(const double *dprofile, double *grad_rf_host, double *corr_host,
                      hipStream_t user) {
    if (!forward_done) return blmi_set_error(BLMI_ERR_ARG, "adjoint: run forward first");
    const int T = forward_turns;
    int rc = ensure_stream();
    if (rc) return rc;
    if (!bwd_exec || bwd_graph_turns != T) {
        if (bwd_exec) (void)hipGraphExecDestroy(bwd_exec);
        bwd_exec = nullptr;
        rc = capture(own, &bwd_exec, [&] { return enqueue_backward(T, own); });
        if (rc) return rc;
        bwd_graph_turns = T;
    }
    BLMI_HIP_TRY(hipEventRecord(ev_in, user));
    BLMI_HIP_TRY(hipStreamWaitEvent(own, ev_in, 0));
    BLMI_HIP_TRY(hipMemcpyAsync(d_lp, dprofile, sizeof(double) * n_slices, hipMemcpyDeviceToDevice,
                                own));
    BLMI_HIP_TRY(hipGraphLaunch(bwd_exec, own));
    std::vector<double> gp((size_t)T * grid * 2 * n_rf);
    BLMI_HIP_TRY(hipMemcpyAsync(gp.data(), d_gpart, sizeof(double) * gp.size(),
                                hipMemcpyDeviceToHost, own));
    BLMI_HIP_TRY(hipMemcpyAsync(corr_host, d_corr, sizeof(double) * n_fft, hipMemcpyDeviceToHost,
                                own));
    BLMI_HIP_TRY(hipStreamSynchronize(own));
    for (int j = 0; j < 2 * n_rf; ++j) {
        double v = 0.0;
        for (int t = 0; t < T; ++t)
            for (int b = 0; b < grid; ++b) v += gp[((size_t)t * grid + b) * 2 * n_rf + j];
        grad_rf_host[j] = v;
    }
    return BLMI_OK;
}

}  // namespace blmi
