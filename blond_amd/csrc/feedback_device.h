// Device code of the beam phase loop shared by its own kernel and by the
// induced-voltage launch that runs the loop as one extra workgroup.
#pragma once

#include "feedback.h"

namespace blmi {

constexpr int kFbBlock = 1024;

// Sum of a and b over the workgroup in a fixed order (butterfly within each
// wavefront, then the wavefronts in index order by thread 0): the result is
// valid in thread 0.
__device__ inline void block_sum2(double &a, double &b) {
    __shared__ double wa[kFbBlock / kWave], wb[kFbBlock / kWave];
    for (int off = kWave / 2; off > 0; off >>= 1) {
        a += __shfl_xor(a, off);
        b += __shfl_xor(b, off);
    }
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    if (lane == 0) {
        wa[w] = a;
        wb[w] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sa = 0.0, sb = 0.0;
        for (int k = 0; k < (int)blockDim.x / kWave; ++k) {
            sa += wa[k];
            sb += wb[k];
        }
        a = sa;
        b = sb;
    }
}

// trapz_const_delta (blondmath.cpp:496-512) of f_s and f_c over n points:
// bin_size * ((f[0] + f[n-1]) / 2 + sum of the interior), for both sums;
// returns scoeff / ccoeff in thread 0.  f(i, &s, &c) evaluates point i.
template <typename F>
__device__ inline double trapz_ratio(F f, int n, double bin_size) {
    double ps = 0.0, pc = 0.0;
    for (int i = 1 + (int)threadIdx.x; i < n - 1; i += blockDim.x) {
        double s, c;
        f(i, s, c);
        ps += s;
        pc += c;
    }
    block_sum2(ps, pc);
    double out = 0.0;
    if (threadIdx.x == 0) {
        double s0, c0, s1, c1;
        f(0, s0, c0);
        f(n - 1, s1, c1);
        const double psum_s = (s0 + s1) / 2. + ps;
        const double psum_c = (c0 + c1) / 2. + pc;
        out = (bin_size * psum_s) / (bin_size * psum_c);
    }
    return out;
}

// beam_phase.cpp:16-24 / 49-53 at bin b: base * sin(a), base * cos(a) with
// a = omega_rf * bc + phi_rf and base = exp(alpha * bc) * p (or p).
__device__ __forceinline__ void phase_terms(double bc, double p, double alpha, double omega,
                                            double phi, bool fast, double &s, double &c) {
    const double base = fast ? p : cephes_exp(alpha * bc) * p;
    const double a = omega * bc + phi;
    double sn, cs;
    cephes_sincos(a, sn, cs);
    s = base * sn;
    c = base * cs;
}

// ---------------------------------------------------------------------------
// one turn of the loop (any workgroup size of whole wavefronts, at least n_rf
// threads), as workgroup `part` of a.n_parts (one when n_parts <= 1)
// ---------------------------------------------------------------------------
__device__ inline void feedback_turn_body(const FeedbackArgs &a, int part = 0) {
    __shared__ double domega_sh;
    __shared__ int last_sh;
    const int t = *a.turn_base + a.turn_off;
    const int nr = a.n_rf;
    const bool active = t >= a.delay;
    const int n_parts = a.n_parts > 1 ? a.n_parts : 1;
    if (!active && part != 0) return;  // the RF update below is part 0's
    double *dphi_rf = a.state + kFbState;
    // inputs of the serial tail, read before the sums so that their latency
    // overlaps them: nothing in this launch writes them before their use
    const int j = threadIdx.x;
    double nx_w = 0.0, nx_p = 0.0, hj = 0.0, h0 = 0.0, wdj = 0.0, dpj = 0.0;
    if (j < nr) {
        const double *nxt = a.rf + (size_t)(t + 1) * 3 * nr;
        const double *h = a.harmonic + (size_t)(t + 1) * nr;
        nx_w = nxt[nr + j];
        nx_p = nxt[2 * nr + j];
        hj = h[j];
        h0 = h[0];
        wdj = a.omega_d[(size_t)(t + 1) * nr + j];
        dpj = dphi_rf[j];
    }
    double on_time_pc = 0.0;  // PSB: the loop turn the counter points at
    if (active && a.machine == kFbPSB && threadIdx.x == 0) {
        const int pc = (int)a.state[kFbPLCounter];
        if (pc < a.n_on_time) on_time_pc = a.on_time[pc];
    }
    if (active) {
        const unsigned *cin = nullptr;
        if (a.from_counts || (a.pending && a.pending[t & 1] == t))
            cin = (t & 1) ? a.counts1 : a.counts0;
        const int ns = a.n_slices;
        auto prof = [&](int b) -> double {
            if (!cin) return a.profile[b];
            unsigned c = 0;
            for (int g = 0; g < a.groups; ++g) c += cin[(size_t)g * ns + b];
            return (double)c;
        };
        const double *rft = a.rf + (size_t)t * 3 * nr;
        const double omega = rft[nr], phi = rft[2 * nr];
        const int b0 = a.first_bin;
        auto f = [&](int i, double &s, double &c) {
            phase_terms(a.bc[b0 + i], prof(b0 + i), a.alpha, omega, phi, false, s, c);
        };
        double coeff;
        if (n_parts == 1) {
            coeff = trapz_ratio(f, ns - b0, a.bin_size);
        } else {
            // trapz_const_delta split by interior index over the parts
            const int n = ns - b0;
            double ps = 0.0, pc = 0.0;
            for (int i = 1 + part * (int)blockDim.x + (int)threadIdx.x; i < n - 1;
                 i += n_parts * (int)blockDim.x) {
                double s, c;
                f(i, s, c);
                ps += s;
                pc += c;
            }
            block_sum2(ps, pc);
            if (threadIdx.x == 0) {
                __hip_atomic_store(&a.parts[2 * part], ps, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&a.parts[2 * part + 1], pc, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
                // release of this part's sums, acquire of the earlier parts'
                const unsigned prev = __hip_atomic_fetch_add(a.arrive, 1u, __ATOMIC_ACQ_REL,
                                                             __HIP_MEMORY_SCOPE_AGENT);
                last_sh = prev == (unsigned)(n_parts - 1);
            }
            __syncthreads();
            if (!last_sh) return;  // uniform over the workgroup
            coeff = 0.0;
            if (threadIdx.x == 0) {
                double ss = 0.0, sc = 0.0;
                for (int k = 0; k < n_parts; ++k) {
                    ss += __hip_atomic_load(&a.parts[2 * k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    sc += __hip_atomic_load(&a.parts[2 * k + 1], __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_AGENT);
                }
                __hip_atomic_store(a.arrive, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                double s0, c0, s1, c1;
                f(0, s0, c0);
                f(n - 1, s1, c1);
                const double psum_s = (s0 + s1) / 2. + ss;
                const double psum_c = (c0 + c1) / 2. + sc;
                coeff = (a.bin_size * psum_s) / (a.bin_size * psum_c);
            }
        }
        if (threadIdx.x == 0) {
            const double *pr = a.prog + (size_t)t * kFbCols;
            double *st = a.state;
            // beam_feedback.py:282-286, 317-322
            const double phi_beam = atan(coeff) + M_PI;
            const double dphi = phi_beam - pr[kFbPhiS];
            const double wd = a.omega_d[(size_t)t * nr];
            double domega = 0.0;
            if (a.machine == kFbLHC) {
                // beam_feedback.py:545-561
                const double x = dphi_rf[0] + a.reference;
                domega = -pr[kFbGain] * dphi - pr[kFbGain2P] * (st[kFbLhcY] + pr[kFbLhcA] * x);
                st[kFbLhcY] = (1 - pr[kFbLhcT]) * st[kFbLhcY] + (1 - pr[kFbLhcA]) * pr[kFbLhcT] * x;
            } else if (a.machine == kFbLHC_F) {
                // beam_feedback.py:414-441
                domega = -pr[kFbGain] * dphi - pr[kFbGain2P] * (omega - wd + a.reference);
            } else {
                // beam_feedback.py:563-622
                st[kFbDphiSum] += dphi;
                const int pc = (int)st[kFbPLCounter];
                if (pc < a.n_on_time && (double)t == on_time_pc && t >= a.delay) {
                    const int prev = pc - 1 < 0 ? a.n_on_time - 1 : pc - 1;
                    st[kFbDphiAv] = st[kFbDphiSum] / (a.on_time[pc] - a.on_time[prev]);
                    st[kFbDomegaPL] = 0.99803799 * st[kFbDomegaPL] +
                                      pr[kFbGain] * (0.99901903 * st[kFbDphiAv] -
                                                     0.99901003 * st[kFbDphiAvPrev]);
                    st[kFbDphiAvPrev] = st[kFbDphiAv];
                    st[kFbDphiSum] = 0.0;
                    const double g = pr[kFbGamma];
                    st[kFbDROverR] = (omega - wd) / (wd * (1.0 / (pr[kFbAlpha0] * (g * g)) - 1.0));
                    st[kFbDomegaRL] = st[kFbDomegaRL] +
                                      pr[kFbGain2P] * (st[kFbDROverR] - st[kFbDROverRPrev]) +
                                      pr[kFbGain2I] * st[kFbDROverR];
                    st[kFbDROverRPrev] = st[kFbDROverR];
                    st[kFbPLCounter] = (double)(pc + 1);
                }
                domega = -st[kFbDomegaPL] - st[kFbDomegaRL];
            }
            st[kFbDomegaRf] = domega;
            st[kFbPhiBeam] = phi_beam;
            st[kFbDphi] = dphi;
            domega_sh = domega;
        }
    }
    __syncthreads();
    // beam_feedback.py:246-251: the frequency of the next turn; then
    // tracker.py:551-565: the accumulated phase offset and the next phase
    if (j < nr) {
        double *nxt = a.rf + (size_t)(t + 1) * 3 * nr;
        if (active) nx_w += domega_sh * hj / h0;
        nxt[nr + j] = nx_w;
        dpj += 2.0 * M_PI * hj * (nx_w - wdj) / wdj;
        dphi_rf[j] = dpj;
        nxt[2 * nr + j] = nx_p + dpj;
    }
}

}  // namespace blmi
