// Slice-axis wake steps beyond the one-turn impedance convolution: the
// multi-turn wake memory (blond/impedances/impedance.py:491-550), the
// inductive Z/n source (impedance.py:943-1000) and the numpy helpers they
// are written with (interp, gradient, diff).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "blmi_common.h"

namespace blmi {

// numpy.interp(x, xp, fp, left, right) -- the function the reference binds as
// bm.interp (blond/utils/bmath_backends.py:373) -- evaluated at one x.  For
// xp[0] <= x <= xp[m-1] it takes the largest j with xp[j] <= x; j == m-1 or
// xp[j] == x gives fp[j], otherwise slope*(x - xp[j]) + fp[j] with
// slope = (fp[j+1] - fp[j]) / (xp[j+1] - xp[j]) (a NaN result retries from
// the right end point).  x below xp[0] gives left, above xp[m-1] right, a NaN
// x gives NaN (m == 1: fp[0] unless x is strictly outside).  `guess` is a
// hint for j (any value; the result does not depend on it).
__device__ __forceinline__ double np_interp_at(double x, const double *__restrict__ xp,
                                               const double *__restrict__ fp, int64_t m,
                                               double left, double right, int64_t guess) {
    if (m == 1) return x < xp[0] ? left : (x > xp[0] ? right : fp[0]);
    if (isnan(x)) return x;
    if (x > xp[m - 1]) return right;
    if (x < xp[0]) return left;
    // invariant: xp[lo] <= x and (hi == m or xp[hi] > x)
    int64_t lo = 0, hi = m;
    if (guess >= 0 && guess < m) {
        if (xp[guess] <= x) {
            lo = guess;
            if (guess + 1 < m) {
                if (xp[guess + 1] > x) {
                    hi = guess + 1;
                } else {
                    lo = guess + 1;
                    if (guess + 2 < m && xp[guess + 2] > x) hi = guess + 2;
                }
            }
        } else {
            hi = guess;
            if (guess >= 1 && xp[guess - 1] <= x) lo = guess - 1;
        }
    }
    while (hi - lo > 1) {
        const int64_t mid = lo + ((hi - lo) >> 1);
        if (xp[mid] <= x) lo = mid;
        else hi = mid;
    }
    const int64_t j = lo;
    if (j == m - 1 || xp[j] == x) return fp[j];
    const double slope = (fp[j + 1] - fp[j]) / (xp[j + 1] - xp[j]);
    double r = slope * (x - xp[j]) + fp[j];
    if (isnan(r)) {
        r = slope * (x - xp[j + 1]) + fp[j + 1];
        if (isnan(r) && fp[j] == fp[j + 1]) r = fp[j];
    }
    return r;
}

// Profile derivative modes of Profile.beam_profile_derivative
// (blond/beam/profile.py:905-937) that run on the device.
constexpr int kDerivGradient = 0;  // numpy.gradient(p, h)
constexpr int kDerivDiff = 1;      // interp(bc, bc[:-1] + h/2, diff(p) / h)

// d(profile)/dt at bin i with spacing h = bin_centers[1] - bin_centers[0]
// (profile.py:914-932); p(j) returns profile value j.
template <typename P>
__device__ __forceinline__ double profile_derivative_at(P p, const double *__restrict__ bc, int n,
                                                        double h, int mode, int i) {
    if (mode == kDerivGradient) {
        // numpy.gradient, uniform spacing, edge_order 1 (n >= 2)
        if (i == 0) return (p(1) - p(0)) / h;
        if (i == n - 1) return (p(n - 1) - p(n - 2)) / h;
        return (p(i + 1) - p(i - 1)) / (2.0 * h);
    }
    // diff mode: numpy.interp(bc[i], dc, d) with dc[k] = bc[k] + h/2 and
    // d[k] = (p[k+1] - p[k]) / h, k < n-1, default left = d[0], right = d[n-2]
    const int m = n - 1;
    const double x = bc[i];
    auto dc = [&](int k) { return bc[k] + h / 2; };
    auto d = [&](int k) { return (p(k + 1) - p(k)) / h; };
    if (m == 1) return d(0);  // left, right and fp[0] are all d[0]
    if (isnan(x)) return x;
    if (x > dc(m - 1)) return d(m - 1);
    if (x < dc(0)) return d(0);
    int lo = 0, hi = m;
    // bc is uniform: the centre below bc[i] is dc[i-1]
    const int g = i - 1;
    if (g >= 0 && g < m && dc(g) <= x && (g + 1 == m || dc(g + 1) > x)) {
        lo = g;
        hi = g + 1;
    }
    while (hi - lo > 1) {
        const int mid = lo + ((hi - lo) >> 1);
        if (dc(mid) <= x) lo = mid;
        else hi = mid;
    }
    const int j = lo;
    if (j == m - 1 || dc(j) == x) return d(j);
    const double slope = (d(j + 1) - d(j)) / (dc(j + 1) - dc(j));
    double r = slope * (x - dc(j)) + d(j);
    if (isnan(r)) {
        r = slope * (x - dc(j + 1)) + d(j + 1);
        if (isnan(r) && d(j) == d(j + 1)) r = d(j);
    }
    return r;
}

// shift_trev_time (impedance.py:536-550): the memory seen one revolution
// later, interp(time + t_rev, time, memory, left=0, right=0), then the
// current turn's voltage added to its first n_add points (impedance.py:510-512).
__device__ __forceinline__ double mtw_value(const double *__restrict__ mem,
                                            const double *__restrict__ time, int64_t n_mem,
                                            double t_rev, int64_t i) {
    const double x = time[i] + t_rev;
    int64_t g = -1;
    if (n_mem > 1) {
        const double step = time[1] - time[0];
        if (step > 0.0) {
            const double sh = floor(t_rev / step);
            if (sh >= 0.0 && sh < (double)n_mem) g = i + (int64_t)sh;
        }
    }
    return np_interp_at(x, time, mem, n_mem, 0.0, 0.0, g);
}

int launch_interp(const double *x, int64_t n, const double *xp, const double *fp, int64_t m,
                  double left, double right, double *out, hipStream_t s);
int launch_gradient(const double *f, int64_t n, double h, double *out, hipStream_t s);
int launch_diff(const double *f, int64_t n, double *out, hipStream_t s);
// mem_out[i] = interp(time[i] + t_rev, time, mem_in, 0, 0) + (i < n_add ? add[i] : 0)
int launch_mtw_shift_add(const double *mem_in, double *mem_out, const double *time, int64_t n_mem,
                         double t_rev, const double *add, int64_t n_add, hipStream_t s);

// Engine step after the impedance convolution when the turn has a
// multi-turn wake or an inductive source: the total induced voltage of the
// turn t = *turn_base + turn_off, written to volt[:n_slices] with its
// linear_interp_kick table.
struct WakeArgs {
    const int *turn_base;
    int turn_off;
    const double *v1;       // this turn's convolution V (n_slices, or n_mem with a memory); may be null
    // multi-turn wake: memory[t & 1] is read, memory[(t + 1) & 1] written
    int64_t n_mem;          // 0: no memory
    int64_t n_add;          // v1 enters the memory at i < n_add (front-wake cut, impedance.py:506-508)
    double *mem0, *mem1;
    const double *time;     // [n_mem] memory time axis
    const double *t_rev;    // [n_turns + 1] revolution period program
    // inductive source: V += -(ind_coef[t] * d(profile)/dt)
    const double *ind_coef; // [n_turns + 1] or null
    int deriv_mode;
    // this turn's profile: read from `profile` when a convolution kernel ran
    // before (it wrote the profile); with v1 == null it is the sum of the
    // `groups` count copies at counts[t & 1], written to `profile` here, and
    // counts[(t + 1) & 1] are cleared for the next particle pass
    double *profile;        // [n_slices]
    unsigned *counts0, *counts1;
    int groups;
    int n_slices;
    const double *bc;
    double charge;
    double *volt;
    double2 *tab;           // (slope, intercept) per bin, tab_params = (bc[0], inv_bw)
    double *tab_params;
};
int launch_wake_finish(const WakeArgs &a, hipStream_t s, const KernelEvents &ev = KernelEvents());
// memory[(*turn_base) & 1] -> out[n_mem] (publish the current memory)
int launch_mtw_publish(const int *turn_base, const double *mem0, const double *mem1, int64_t n_mem,
                       double *out, hipStream_t s);
// in[n_mem] -> memory[(*turn_base) & 1]
int launch_mtw_load(const int *turn_base, double *mem0, double *mem1, int64_t n_mem,
                    const double *in, hipStream_t s);

}  // namespace blmi
