// Slice-axis kernels: LDS Stockham FFT (radix 2/3/4/5/8), the real-FFT
// split/merge, the one-turn induced voltage and the resonator impedance.
//
// A real transform of length n = 2M runs as a complex transform of length M
// on z[m] = x[2m] + i x[2m+1] held in one workgroup's LDS (M <= 10240 complex
// fits the 160 KiB of a CDNA4 workgroup).  Each Stockham stage reads all of
// its butterfly inputs into registers, synchronises, and writes the outputs
// back to the same LDS buffer, so one buffer suffices.  Twiddles come from an
// exact table computed on the host in extended precision.  Power-of-two
// lengths up to 1024 (the four-step passes' columns and rows, the batched
// bunches' transforms) run as one wavefront a transform instead: radix-4
// stages with compile-time lengths and no workgroup barrier (wave_fft).
#include <math.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "blmi_common.h"
#include "particles.h"
#include "spectral.h"
#include "feedback_device.h"

namespace blmi {

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) {
    return make_double2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ double2 csub(double2 a, double2 b) {
    return make_double2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 cscale(double2 a, double s) {
    return make_double2(a.x * s, a.y * s);
}
// multiply by -i (forward) or +i (inverse)
template <bool INV> __device__ __forceinline__ double2 mul_mi(double2 a) {
    return INV ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
}

template <int R, bool INV> struct Dft;

template <bool INV> struct Dft<2, INV> {
    __device__ __forceinline__ static void run(double2 *v) {
        const double2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <bool INV> struct Dft<4, INV> {
    __device__ __forceinline__ static void run(double2 *v) {
        const double2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
        const double2 t2 = cadd(v[1], v[3]), t3 = mul_mi<INV>(csub(v[1], v[3]));
        v[0] = cadd(t0, t2);
        v[2] = csub(t0, t2);
        v[1] = cadd(t1, t3);
        v[3] = csub(t1, t3);
    }
};

template <bool INV> struct Dft<8, INV> {
    __device__ __forceinline__ static void run(double2 *v) {
        double2 e[4] = {v[0], v[2], v[4], v[6]};
        double2 o[4] = {v[1], v[3], v[5], v[7]};
        Dft<4, INV>::run(e);
        Dft<4, INV>::run(o);
        const double h = 0.70710678118654752440;  // sqrt(1/2)
        // w8^k o_k, w8 = exp(-+ 2 pi i / 8)
        const double2 o1 = INV ? make_double2(h * (o[1].x - o[1].y), h * (o[1].x + o[1].y))
                               : make_double2(h * (o[1].x + o[1].y), h * (o[1].y - o[1].x));
        const double2 o2 = mul_mi<INV>(o[2]);
        const double2 o3 = INV ? make_double2(-h * (o[3].x + o[3].y), h * (o[3].x - o[3].y))
                               : make_double2(h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y));
        v[0] = cadd(e[0], o[0]);
        v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], o1);
        v[5] = csub(e[1], o1);
        v[2] = cadd(e[2], o2);
        v[6] = csub(e[2], o2);
        v[3] = cadd(e[3], o3);
        v[7] = csub(e[3], o3);
    }
};

template <bool INV> struct Dft<3, INV> {
    __device__ __forceinline__ static void run(double2 *v) {
        const double s3 = 0.86602540378443864676;  // sqrt(3)/2
        const double2 s = cadd(v[1], v[2]), d = csub(v[1], v[2]);
        const double2 m = csub(v[0], cscale(s, 0.5));
        const double2 r = mul_mi<INV>(cscale(d, s3));
        v[0] = cadd(v[0], s);
        v[1] = cadd(m, r);
        v[2] = csub(m, r);
    }
};

template <bool INV> struct Dft<5, INV> {
    __device__ __forceinline__ static void run(double2 *v) {
        const double c1 = 0.30901699437494742410;   // cos(2pi/5)
        const double c2 = -0.80901699437494742410;  // cos(4pi/5)
        const double s1 = 0.95105651629515357212;   // sin(2pi/5)
        const double s2 = 0.58778525229247312917;   // sin(4pi/5)
        const double2 a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
        const double2 a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
        const double2 x0 = v[0];
        const double2 p1 = cadd(x0, cadd(cscale(a1, c1), cscale(a2, c2)));
        const double2 p2 = cadd(x0, cadd(cscale(a1, c2), cscale(a2, c1)));
        const double2 q1 = mul_mi<INV>(cadd(cscale(b1, s1), cscale(b2, s2)));
        const double2 q2 = mul_mi<INV>(csub(cscale(b1, s2), cscale(b2, s1)));
        v[0] = cadd(x0, cadd(a1, a2));
        v[1] = cadd(p1, q1);
        v[4] = csub(p1, q1);
        v[2] = cadd(p2, q2);
        v[3] = csub(p2, q2);
    }
};

// One Stockham stage over `batch` transforms of length L, transform f at
// buf + f * ld.  Ns = product of the radices already applied.  Twiddles come
// from a table of exp(-2 pi i j / tw_len), L dividing tw_len.  The host
// guarantees batch * L <= 15 * blockDim.x (at most 16/R butterflies a thread).
template <int R, bool INV>
__device__ __forceinline__ void stockham_stage(double2 *buf, int L, int batch, int ld, int Ns,
                                               const double2 *__restrict__ tw, int tw_len, int T) {
    constexpr int BMAX = (16 / R) > 0 ? (16 / R) : 1;
    const int nb = L / R;
    const int total = nb * batch;
    const int tstride = tw_len / (Ns * R);
    double2 v[BMAX][R];
#pragma unroll
    for (int q = 0; q < BMAX; ++q) {
        const int j = threadIdx.x + q * T;
        if (j < total) {
            const int f = batch == 1 ? 0 : j / nb;
            const int jj = j - f * nb;
            const double2 *b = buf + f * ld;
            const int k = jj % Ns;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                double2 x = b[jj + r * nb];
                if (r > 0) {
                    double2 w = tw[r * k * tstride];
                    if (INV) w.y = -w.y;
                    x = cmul(x, w);
                }
                v[q][r] = x;
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < BMAX; ++q) {
        const int j = threadIdx.x + q * T;
        if (j < total) {
            Dft<R, INV>::run(v[q]);
            const int f = batch == 1 ? 0 : j / nb;
            const int jj = j - f * nb;
            double2 *b = buf + f * ld;
            const int k = jj % Ns;
            const int base = (jj - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) b[base + r * Ns] = v[q][r];
        }
    }
    __syncthreads();
}

struct StageList {
    int n;
    int radix[kMaxStages];
};

template <bool INV>
// nthreads: the workgroup size.  The FFT passes give it as a constant: read
// from blockDim.x inside the stage loop it is reloaded from the dispatch
// packet with a vector load and a wait at every stage.
__device__ void fft_lds_batch(double2 *buf, int L, int batch, int ld, const StageList &st,
                              const double2 *__restrict__ tw, int tw_len, int nthreads) {
    int Ns = 1;
    for (int s = 0; s < st.n; ++s) {
        switch (st.radix[s]) {
            case 8: stockham_stage<8, INV>(buf, L, batch, ld, Ns, tw, tw_len, nthreads); break;
            case 4: stockham_stage<4, INV>(buf, L, batch, ld, Ns, tw, tw_len, nthreads); break;
            case 2: stockham_stage<2, INV>(buf, L, batch, ld, Ns, tw, tw_len, nthreads); break;
            case 5: stockham_stage<5, INV>(buf, L, batch, ld, Ns, tw, tw_len, nthreads); break;
            default: stockham_stage<3, INV>(buf, L, batch, ld, Ns, tw, tw_len, nthreads); break;
        }
        Ns *= st.radix[s];
    }
}

// One transform of length L (a power of two, 16..1024) by one wavefront:
// the Stockham stages of stockham_stage (a radix-2 stage first when log2 L
// is odd, then radix 4) with compile-time lengths, each lane taking
// butterflies lane, lane + 64, ...  A wavefront's LDS operations are
// performed in the order it issues them, so a stage's writes land before the
// next stage's reads without a workgroup barrier; the wavefront fences only
// keep the compiler from moving LDS accesses across a stage boundary.  (The
// workgroup-wide stages spend two barriers a stage and keep only L / R
// threads busy.)
__device__ __forceinline__ void wave_lds_fence() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int R, int L, int NS, bool INV>
__device__ __forceinline__ void wave_stage(double2 *b, const double2 *__restrict__ tw, int lane) {
    constexpr int NB = L / R;            // butterflies
    constexpr int Q = (NB + 63) / 64;    // a lane's share
    constexpr int TS = L / (NS * R);     // twiddle stride in the length-L table
    double2 v[Q][R];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int j = lane + 64 * q;
        if (NB % 64 == 0 || j < NB) {
            const int k = j & (NS - 1);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                double2 x = b[j + r * NB];
                if (r > 0 && NS > 1) {  // first stage: every twiddle is 1
                    double2 w = tw[r * k * TS];
                    if (INV) w.y = -w.y;
                    x = cmul(x, w);
                }
                v[q][r] = x;
            }
        }
    }
    wave_lds_fence();
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int j = lane + 64 * q;
        if (NB % 64 == 0 || j < NB) {
            const int k = j & (NS - 1);
            Dft<R, INV>::run(v[q]);
            const int base = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) b[base + r * NS] = v[q][r];
        }
    }
    wave_lds_fence();
}

template <int L, int NS, bool INV>
__device__ __forceinline__ void wave_radix4_stages(double2 *b, const double2 *__restrict__ tw,
                                                   int lane) {
    if constexpr (NS < L) {
        wave_stage<4, L, NS, INV>(b, tw, lane);
        wave_radix4_stages<L, NS * 4, INV>(b, tw, lane);
    }
}

template <int L, bool INV>
__device__ __forceinline__ void wave_fft(double2 *b, const double2 *__restrict__ tw, int lane) {
    static_assert(L >= 16 && L <= 1024 && (L & (L - 1)) == 0, "power of two in [16, 1024]");
    if constexpr ((__builtin_ctz(L) & 1) != 0) {
        wave_stage<2, L, 1, INV>(b, tw, lane);
        wave_radix4_stages<L, 2, INV>(b, tw, lane);
    } else {
        wave_radix4_stages<L, 1, INV>(b, tw, lane);
    }
}

// `batch` transforms of length L, transform f at buf + f * ld, wavefront w
// taking f = w, w + nthreads / 64, ...; ends with a workgroup barrier.
template <int L, bool INV>
__device__ __forceinline__ void wave_fft_batch(double2 *buf, int batch, int ld,
                                               const double2 *__restrict__ tw, int nthreads) {
    const int lane = threadIdx.x & 63;
    for (int f = threadIdx.x >> 6; f < batch; f += nthreads >> 6)
        wave_fft<L, INV>(buf + f * ld, tw, lane);
    __syncthreads();
}

template <bool INV>
__device__ void fft_lds(double2 *buf, int M, const StageList &st, const double2 *__restrict__ tw) {
    fft_lds_batch<INV>(buf, M, 1, M, st, tw, M, (int)blockDim.x);
}

// X[k] and X[M-k] of the real transform from the half-length complex
// transform Z (k in [0, M/2]); for k == 0 returns X[0] and X[M].
__device__ __forceinline__ void real_split(double2 Zk, double2 Zmk, double2 Wk, double2 Wmk,
                                           double2 &Xk, double2 &Xmk) {
    // X[k] = 1/2 [(Z[k] + conj Z[M-k]) - i W^k (Z[k] - conj Z[M-k])]
    // (mul_mi<false>(a) = -i a)
    const double2 s1 = cadd(Zk, cconj(Zmk)), d1 = csub(Zk, cconj(Zmk));
    Xk = cscale(cadd(s1, mul_mi<false>(cmul(Wk, d1))), 0.5);
    const double2 s2 = cadd(Zmk, cconj(Zk)), d2 = csub(Zmk, cconj(Zk));
    Xmk = cscale(cadd(s2, mul_mi<false>(cmul(Wmk, d2))), 0.5);
}

// U[k] = 1/2 [(Y[k] + conj Y[M-k]) + i conj(W^k) (Y[k] - conj Y[M-k])]
__device__ __forceinline__ double2 real_merge(double2 Yk, double2 Ymk, double2 Wk) {
    const double2 s = cadd(Yk, cconj(Ymk)), d = csub(Yk, cconj(Ymk));
    return cscale(cadd(s, mul_mi<true>(cmul(cconj(Wk), d))), 0.5);
}

// ---------------------------------------------------------------------------
// standalone rfft / irfft (numpy semantics for even n)
// ---------------------------------------------------------------------------
__global__ void rfft_lds_kernel(const double *__restrict__ in, int64_t n_in, double2 *__restrict__ out,
                                int M, StageList st, const double2 *__restrict__ tw,
                                const double2 *__restrict__ tw2) {
    extern __shared__ double2 buf[];
    const int n = 2 * M;
    for (int m = threadIdx.x; m < M; m += blockDim.x) {
        const int64_t j0 = 2 * (int64_t)m, j1 = j0 + 1;
        buf[m] = make_double2(j0 < n_in ? in[j0] : 0.0, j1 < n_in ? in[j1] : 0.0);
    }
    (void)n;
    __syncthreads();
    fft_lds<false>(buf, M, st, tw);
    for (int k = threadIdx.x; k <= M / 2; k += blockDim.x) {
        const double2 Zk = buf[k % M], Zmk = buf[(M - k) % M];
        double2 Xk, Xmk;
        real_split(Zk, Zmk, tw2[k], tw2[M - k], Xk, Xmk);
        out[k] = Xk;
        out[M - k] = Xmk;
    }
}

__global__ void irfft_lds_kernel(const double2 *__restrict__ in, int64_t m_in, double *__restrict__ out,
                                 int64_t n_out, int M, StageList st, const double2 *__restrict__ tw,
                                 const double2 *__restrict__ tw2) {
    extern __shared__ double2 buf[];
    auto Y = [&](int k) -> double2 {
        double2 y = k < m_in ? in[k] : make_double2(0.0, 0.0);
        if (k == 0 || k == M) y.y = 0.0;  // c2r ignores imag of DC and Nyquist
        return y;
    };
    for (int k = threadIdx.x; k <= M / 2; k += blockDim.x) {
        const double2 Yk = Y(k), Ymk = Y(M - k);
        buf[k % M] = real_merge(Yk, Ymk, tw2[k]);
        if (k != 0 && 2 * k != M) buf[M - k] = real_merge(Ymk, Yk, tw2[M - k]);
    }
    __syncthreads();
    fft_lds<true>(buf, M, st, tw);
    // U carries the 1/2 of the split, so 1/M here gives numpy's 1/n overall
    const double fct = 1.0 / (double)M;
    for (int m = threadIdx.x; m < M; m += blockDim.x) {
        const double2 u = buf[m];
        const int64_t j0 = 2 * (int64_t)m;
        if (j0 < n_out) out[j0] = u.x * fct;
        if (j0 + 1 < n_out) out[j0 + 1] = u.y * fct;
    }
}

// ---------------------------------------------------------------------------
// one-turn induced voltage on one workgroup
// ---------------------------------------------------------------------------
// The profile is read from a.src, or as the sum of a.groups copies of u32
// counts at cin (then cleared at cclr, which may equal cin: each element is
// read before it is cleared, by the same thread).  FL: M as a constant when
// it is a power of two in [16, 1024] (one wavefront a transform, twiddles
// staged in LDS after the data), else 0 (workgroup-wide stages).
template <int FL>
__device__ void induced_lds_body(const InducedArgs &a, const unsigned *cin, unsigned *cclr, int M,
                                 const StageList &st, const double2 *__restrict__ tw,
                                 const double2 *__restrict__ tw2) {
    extern __shared__ double2 buf[];
    double2 *twl = buf + M;
    if constexpr (FL > 0)
        for (int j = threadIdx.x; j < M; j += blockDim.x) twl[j] = tw[j];
    const int ns = a.n_slices;
    // load the profile as z[m] = p[2m] + i p[2m+1], zero padded to n = 2M
    auto prof = [&](int j) -> double {
        if (!cin) return a.src[j];
        unsigned s = 0;
        for (int g = 0; g < a.groups; ++g) s += cin[(size_t)g * ns + j];
        return (double)s;
    };
    for (int m = threadIdx.x; m < M; m += blockDim.x) {
        double p0 = 0.0, p1 = 0.0;
        const int j0 = 2 * m, j1 = j0 + 1;
        if (j0 < ns) p0 = prof(j0);
        if (j1 < ns) p1 = prof(j1);
        if (a.profile_out) {
            if (j0 < ns) a.profile_out[j0] = p0;
            if (j1 < ns) a.profile_out[j1] = p1;
        }
        if (cclr) {
            for (int g = 0; g < a.groups; ++g) {
                if (j0 < ns) cclr[(size_t)g * ns + j0] = 0u;
                if (j1 < ns) cclr[(size_t)g * ns + j1] = 0u;
            }
        }
        buf[m] = make_double2(p0, p1);
    }
    __syncthreads();
    if constexpr (FL > 0)
        wave_fft_batch<FL, false>(buf, 1, M, twl, (int)blockDim.x);
    else
        fft_lds<false>(buf, M, st, tw);
    // spectrum -> times impedance -> half-length inverse input, pairwise.
    // All pairs are read into registers before any is overwritten.
    constexpr int PMAX = 8;  // host guarantees M/2+1 <= PMAX * blockDim.x
    double2 ua[PMAX], ub[PMAX];
#pragma unroll
    for (int q = 0; q < PMAX; ++q) {
        const int k = threadIdx.x + q * blockDim.x;
        if (k <= M / 2) {
            const double2 Zk = buf[k % M], Zmk = buf[(M - k) % M];
            double2 Xk, Xmk;
            real_split(Zk, Zmk, tw2[k], tw2[M - k], Xk, Xmk);
            // impedance.py:479-484: total_impedance * beam_spectrum
            double2 Yk = cmul(a.Z[k], Xk), Ymk = cmul(a.Z[M - k], Xmk);
            if (k == 0) {  // c2r ignores the imaginary parts of DC and Nyquist
                Yk.y = 0.0;
                Ymk.y = 0.0;
            }
            ua[q] = real_merge(Yk, Ymk, tw2[k]);
            ub[q] = real_merge(Ymk, Yk, tw2[M - k]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PMAX; ++q) {
        const int k = threadIdx.x + q * blockDim.x;
        if (k <= M / 2) {
            buf[k % M] = ua[q];
            if (k != 0 && 2 * k != M) buf[M - k] = ub[q];
        }
    }
    __syncthreads();
    if constexpr (FL > 0)
        wave_fft_batch<FL, true>(buf, 1, M, twl, (int)blockDim.x);
    else
        fft_lds<true>(buf, M, st, tw);
    // V = -(charge*e*ratio * irfft(...))[:n_slices]   (impedance.py:475-489)
    // Each thread rewrites only its own buf[m] as the two reals vl[2m], vl[2m+1].
    // U carries the 1/2 of the split, so 1/M here gives numpy's 1/n overall
    const double fct = 1.0 / (double)M;
    double *vl = reinterpret_cast<double *>(buf);
    for (int m = threadIdx.x; m < M; m += blockDim.x) {
        const double2 u = buf[m];
        const double v0 = -(a.scale * (u.x * fct)), v1 = -(a.scale * (u.y * fct));
        vl[2 * m] = v0;
        vl[2 * m + 1] = v1;
        if (2 * m < a.n_out) a.volt[2 * m] = v0;
        if (2 * m + 1 < a.n_out) a.volt[2 * m + 1] = v1;
    }
    if (a.tab) {
        __syncthreads();
        // linear_interp_kick.cpp:33-48 on V (acc = 0 for the induced kick)
        const double inv_bw = interp_inv_bw(a.bc, ns, a.f32);
        if (threadIdx.x == 0) {
            a.tab_params[0] = a.bc[0];
            a.tab_params[1] = inv_bw;
        }
        for (int i = threadIdx.x; i < ns - 1; i += blockDim.x)
            store_interp_entry(a.tab, i,
                               interp_entry(vl[i], vl[i + 1], a.bc[i], a.charge, inv_bw, a.f32),
                               a.f32);
    }
}

// FL > 0 launches take 512 threads (lds_threads of M <= 1024): the bound
// leaves their registers room (no scratch spills)
constexpr int kLdsWaveThreads = 512;

template <int FL>
__global__ void __launch_bounds__(FL > 0 ? kLdsWaveThreads : 1024)
    induced_lds_kernel(InducedArgs a, int M, StageList st,
                                   const double2 *__restrict__ tw, const double2 *__restrict__ tw2) {
    const unsigned *cin = nullptr;
    unsigned *cclr = nullptr;
    if (a.counts0) {
        const int t = *a.turn_base + a.turn_off;
        cin = (t & 1) ? a.counts1 : a.counts0;
        cclr = (t & 1) ? a.counts0 : a.counts1;
    }
    induced_lds_body<FL>(a, cin, cclr, M, st, tw, tw2);
}

// One workgroup per bunch: bunch b's arguments are args[b]; its counts (one
// copy) are read and cleared in place.
template <int FL>
__global__ void __launch_bounds__(FL > 0 ? kLdsWaveThreads : 1024)
    induced_lds_batch_kernel(const InducedArgs *__restrict__ args, int M, StageList st,
                                         const double2 *__restrict__ tw,
                                         const double2 *__restrict__ tw2) {
    const InducedArgs a = args[blockIdx.x];
    induced_lds_body<FL>(a, a.counts0, a.counts0, M, st, tw, tw2);
}

// ---------------------------------------------------------------------------
// one-turn induced voltage by direct convolution over all CUs
//   V[i] = sum_m q[m] gl[i + m],  q[m] = p[ns-1-m],  gl[k] = g[(k-(ns-1)) mod n_fft]
// for i < n_out (n_slices for one turn, the multi-turn memory's length for a
// wake).  Workgroup b produces the C outputs i0 = b C .. i0 + C - 1 plus output
// i0 + C (needed by the last table entry).  The reversed profile q and the
// workgroup's window w[k] = gl[i0 + k] sit in LDS with one pad word per 8
// (entry k at k + k/8).  Thread l takes CHUNK consecutive samples
// m = s + CHUNK l + j of every span s of CHUNK*TH samples: it loads
// w[m .. m + CHUNK + C) into registers once and reuses each value for up to
// C + 1 outputs, so a sample costs about two LDS reads instead of C + 2, and
// with the padding lanes CHUNK samples apart fall on distinct bank pairs.
// A wavefront's C + 1 sums are reduced by a butterfly that halves the values
// each lane carries at every exchange, then across wavefronts in LDS, in a
// fixed order (deterministic).
// LDS: pad(nq) + pad(nq + C - 1) + 1 doubles, nq = ns rounded up to a span.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int conv_pad(int k) { return k + (k >> 3); }

// The wake step of the turn on this workgroup's outputs, as the engine's
// separate wake step computes it (wake.hip): the memory after the turn
// mem(i) = interp(old memory one revolution later) + V(i) for i < n_add
// (impedance.py:491-512), written for i < n_mem; the total induced voltage
// 0 + mem(i) (or 0 + V(i) without a memory) - coef * d(profile)/dt (i)
// (impedance.py:117-127, 943-1000) and its linear_interp_kick table
// (linear_interp_kick.cpp:33-48).  q holds the reversed profile in LDS.
template <int C>
__device__ __forceinline__ void direct_wake_epilogue(const InducedArgs &a, const double *q,
                                                     const double *vout, double *memv,
                                                     double mprev, int t, int i0,
                                                     unsigned *cclr) {
    const WakeArgs &w = a.wk;
    const int ns = a.n_slices;
    if ((int)threadIdx.x <= C) {
        const int64_t i = (int64_t)i0 + threadIdx.x;
        double v = vout[threadIdx.x];
        if (w.n_mem > 0) {
            v = i < w.n_add ? mprev + vout[threadIdx.x] : mprev;
            if ((int)threadIdx.x < C && i < w.n_mem) ((t & 1) ? w.mem0 : w.mem1)[i] = v;
        }
        memv[threadIdx.x] = v;
    }
    __syncthreads();
    if ((int)threadIdx.x >= C) return;
    const int i = i0 + threadIdx.x;
    if (i >= ns) return;
    const double coef = w.ind_coef ? w.ind_coef[t] : 0.0;
    auto prof = [&](int k) -> double { return q[conv_pad(ns - 1 - k)]; };
    auto total = [&](int c, int j) -> double {
        double v = 0.0;
        v = v + memv[c];
        if (w.ind_coef) {
            const double d = profile_derivative_at(prof, w.bc, ns, w.bc[1] - w.bc[0],
                                                   w.deriv_mode, j);
            v = v + (-(coef * d));
        }
        return v;
    };
    const double vi = total(threadIdx.x, i);
    a.volt[i] = vi;
    if (a.profile_out) a.profile_out[i] = prof(i);
    if (cclr)
        for (int g = 0; g < a.groups; ++g) cclr[(size_t)g * ns + i] = 0u;
    if (a.tab) {
        const double inv_bw = (double)(ns - 1) / (a.bc[ns - 1] - a.bc[0]);
        if (i == 0) {
            a.tab_params[0] = a.bc[0];
            a.tab_params[1] = inv_bw;
        }
        if (i < ns - 1) {
            const double vn = total(threadIdx.x + 1, i + 1);
            const double k = a.charge * (vn - vi) * inv_bw;
            a.tab[i] = make_double2(k, (a.charge * vi - a.bc[i] * k) + 0.0);
        }
    }
}

template <int C, int TH, int CHUNK>
__global__ void __launch_bounds__(TH) induced_direct_kernel(InducedArgs a) {
    static_assert(C >= 2 && C <= 32 && (C & (C - 1)) == 0, "C is a power of two up to 32");
    static_assert(CHUNK % 8 == 0, "a thread's samples start on a pad period");
    const int nfb = a.with_fb ? a.fb.n_parts : 0;
    if ((int)blockIdx.x < nfb) {
        // the beam phase loop of this turn, independent of the convolution
        feedback_turn_body(a.fb, blockIdx.x);
        return;
    }
    extern __shared__ double conv_lds[];
    __shared__ double red[TH / 64][C + 1];
    __shared__ double vout[C + 1];
    __shared__ double memv[C + 1];
    constexpr int kSpan = CHUNK * TH;
    const int ns = a.n_slices;
    // with the loop, workgroups 0 .. nfb - 1 run it (dispatched first: they
    // are the longest) and workgroup b >= nfb takes outputs of b - nfb
    const int i0 = (blockIdx.x - nfb) * C;
    // wake step: the previous memory's value one revolution later at this
    // workgroup's outputs (impedance.py:536-550), read first so that its
    // dependent loads overlap the staging and the products below
    double mprev = 0.0;
    int wt = 0;
    if (a.with_wake) {
        wt = *a.wk.turn_base + a.wk.turn_off;
        const int64_t i = (int64_t)i0 + threadIdx.x;
        if ((int)threadIdx.x <= C && i < a.wk.n_mem)
            mprev = mtw_value((wt & 1) ? a.wk.mem1 : a.wk.mem0, a.wk.time, a.wk.n_mem,
                              a.wk.t_rev[wt], i);
    }
    const int lim = a.n_out + ns - 1;  // gl entries used (host checks gl_len)
    const int nq = (ns + kSpan - 1) / kSpan * kSpan;
    double *q = conv_lds;                  // q[pad(m)], zero for m >= ns
    double *wv = conv_lds + conv_pad(nq);  // wv[pad(k)] = gl[i0 + k], k < nq + C
    const unsigned *cin = nullptr;
    unsigned *cclr = nullptr;
    if (a.counts0) {
        const int t = *a.turn_base + a.turn_off;
        cin = (t & 1) ? a.counts1 : a.counts0;
        cclr = (t & 1) ? a.counts0 : a.counts1;
    }
    // Staging in batches of 8 entries per thread: every load of a batch
    // (window, profile, and the window's last C entries with the last batch)
    // is issued before the first LDS store, with clamped in-bounds indexes
    // and the out-of-range values replaced afterwards, so the batch costs
    // one memory round trip (plus one per extra count copy).
    constexpr int kStage = 8;
    for (int b = 0; b < nq; b += kStage * TH) {
        double wx[kStage], qx[kStage];
#pragma unroll
        for (int u = 0; u < kStage; ++u) {
            const int k = b + u * TH + (int)threadIdx.x;
            const bool ok = k < ns + C && i0 + k < lim;
            const double x = a.gl[ok ? i0 + k : 0];
            wx[u] = ok ? x : 0.0;
        }
        const bool tail = b + kStage * TH >= nq && (int)threadIdx.x < C;
        const int kt = nq + (int)threadIdx.x;
        const bool okt = tail && kt < ns + C && i0 + kt < lim;
        const double xt = a.gl[okt ? i0 + kt : 0];
        if (cin) {
            unsigned cs[kStage];
#pragma unroll
            for (int u = 0; u < kStage; ++u) cs[u] = 0u;
            for (int g = 0; g < a.groups; ++g) {
                const unsigned *row = cin + (size_t)g * ns;
#pragma unroll
                for (int u = 0; u < kStage; ++u) {
                    const int m = b + u * TH + (int)threadIdx.x;
                    const unsigned x = row[m < ns ? ns - 1 - m : 0];
                    cs[u] += m < ns ? x : 0u;
                }
            }
#pragma unroll
            for (int u = 0; u < kStage; ++u) qx[u] = (double)cs[u];
        } else {
#pragma unroll
            for (int u = 0; u < kStage; ++u) {
                const int m = b + u * TH + (int)threadIdx.x;
                const double x = a.src[m < ns ? ns - 1 - m : 0];
                qx[u] = m < ns ? x : 0.0;
            }
        }
#pragma unroll
        for (int u = 0; u < kStage; ++u) {
            const int k = b + u * TH + (int)threadIdx.x;  // k < nq: nq is a multiple of 8 TH
            wv[conv_pad(k)] = wx[u];
            q[conv_pad(k)] = qx[u];
        }
        if (tail) wv[conv_pad(kt)] = okt ? xt : 0.0;
    }
    __syncthreads();
    double acc[C + 1];
#pragma unroll
    for (int c = 0; c <= C; ++c) acc[c] = 0.0;
    for (int s0 = 0; s0 < nq; s0 += kSpan) {
        const int m0 = s0 + CHUNK * (int)threadIdx.x;
        if (m0 < ns) {
            const int pb = conv_pad(m0);  // m0 % 8 == 0: pad(m0 + j) = pb + j + j / 8
            double w[CHUNK + C];
#pragma unroll
            for (int j = 0; j < CHUNK + C; ++j) w[j] = wv[pb + j + (j >> 3)];
#pragma unroll
            for (int j = 0; j < CHUNK; ++j) {
                const double qj = q[pb + j + (j >> 3)];
#pragma unroll
                for (int c = 0; c <= C; ++c) acc[c] = fma(qj, w[j + c], acc[c]);
            }
        }
    }
    // butterfly over lane bits 5, 4, ...: at each exchange a lane keeps half
    // of its sums and adds its partner's copy of them; after log2(C)
    // exchanges a lane holds one sum (index idx) over 64/C lanes' worth
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int idx = 0, bit = 32;
#pragma unroll
    for (int half = C / 2; half >= 1; half >>= 1) {
        const bool hi = lane & bit;
#pragma unroll
        for (int k = 0; k < half; ++k) {
            const double send = hi ? acc[k] : acc[k + half];
            const double keep = hi ? acc[k + half] : acc[k];
            acc[k] = keep + __shfl_xor(send, bit);
        }
        if (hi) idx += half;
        bit >>= 1;
    }
    for (int b = bit; b > 0; b >>= 1) acc[0] += __shfl_xor(acc[0], b);
    double last = acc[C];
    for (int off = 32; off > 0; off >>= 1) last += __shfl_xor(last, off);
    if ((lane & (64 / C - 1)) == 0) red[wid][idx] = acc[0];
    if (lane == 0) red[wid][C] = last;
    __syncthreads();
    if (threadIdx.x <= C) {
        double v = 0.0;
        for (int w = 0; w < TH / 64; ++w) v += red[w][threadIdx.x];
        vout[threadIdx.x] = v;
    }
    __syncthreads();
    if (a.with_wake) {
        direct_wake_epilogue<C>(a, q, vout, memv, mprev, wt, i0, cclr);
        return;
    }
    if (threadIdx.x < C) {
        const int i = i0 + threadIdx.x;
        if (i < a.n_out) a.volt[i] = vout[threadIdx.x];
        if (i < ns) {
            if (a.profile_out) a.profile_out[i] = q[conv_pad(ns - 1 - i)];
            if (cclr)
                for (int g = 0; g < a.groups; ++g) cclr[(size_t)g * ns + i] = 0u;
        }
    }
    if (a.tab) {
        // linear_interp_kick.cpp:33-48 on V (acc = 0 for the induced kick)
        const double inv_bw = interp_inv_bw(a.bc, ns, a.f32);
        if (i0 == 0 && threadIdx.x == 0) {
            a.tab_params[0] = a.bc[0];
            a.tab_params[1] = inv_bw;
        }
        if (threadIdx.x < C) {
            const int i = i0 + threadIdx.x;
            if (i < ns - 1)
                store_interp_entry(a.tab, i,
                                   interp_entry(vout[threadIdx.x], vout[threadIdx.x + 1], a.bc[i],
                                                a.charge, inv_bw, a.f32),
                                   a.f32);
        }
    }
}

static int set_lds_limit(const void *fn, size_t bytes);

template <int C, int TH, int CHUNK>
static int launch_direct_shape(const InducedArgs &a0, hipStream_t s, const KernelEvents &ev) {
    InducedArgs a = a0;
    int nfb = 0;
    if (a.with_fb) {
        // the loop's sums over ceil(bins / TH) workgroups (one bin a thread),
        // when the caller provides the partial-sum buffer and counter
        nfb = 1;
        if (a.fb.parts && a.fb.arrive) {
            const int nb = a.fb.n_slices - a.fb.first_bin;
            nfb = std::min(kFbMaxParts, std::max(1, (nb + TH - 1) / TH));
        }
        a.fb.n_parts = nfb;
    }
    const int blocks = (a.n_out + C - 1) / C + nfb;
    if (a.with_fb && a.fb.n_rf > TH)
        return blmi_set_error(BLMI_ERR_ARG, "direct induced voltage: too many RF systems for the loop");
    constexpr int kSpan = CHUNK * TH;
    const int nq = (a.n_slices + kSpan - 1) / kSpan * kSpan;
    auto pad = [](int k) { return (size_t)(k + (k >> 3)); };
    const size_t lds = sizeof(double) * (pad(nq) + pad(nq + C - 1) + 1);
    // the kernel's static LDS: its reduction rows and, inlined, the beam
    // phase loop's (the compiler's figure; the rows alone if it is unknown)
    size_t static_lds = sizeof(double) * (TH / 64 + 1) * (C + 1);
    hipFuncAttributes fa{};
    if (hipFuncGetAttributes(&fa, (const void *)induced_direct_kernel<C, TH, CHUNK>) == hipSuccess)
        static_lds = fa.sharedSizeBytes;
    else
        (void)hipGetLastError();
    if (lds + static_lds > (size_t)device_max_lds())
        return blmi_set_error(BLMI_ERR_UNSUPPORTED, "direct induced voltage: %d slices exceed LDS",
                              a.n_slices);
    int rc = set_lds_limit((const void *)induced_direct_kernel<C, TH, CHUNK>, lds);
    if (rc) return rc;
    launch(induced_direct_kernel<C, TH, CHUNK>, dim3(blocks), dim3(TH), lds, s, ev, a);
    return blmi_check_launch("induced_direct");
}

// Shape of the direct convolution (outputs per workgroup, workgroup size,
// samples per thread), measured on one MI355X with a 4096-slice profile and
// its neighbours: 1024 slices 5.0 us (16, 256, 8); 2048 slices 6.1 us and
// 4096 slices 6.8 us (16, 512, 8); 8192 slices 11.6 us (32, 512, 16).  A
// multi-turn memory (n_out = 4 x 4096 values of 4096 slices) takes (32, 512,
// 8): 21.8 against 22.5 us for (16, 512, 8) with its wake step.
int launch_induced_direct(const InducedArgs &a, hipStream_t s, const KernelEvents &ev) {
    if (!a.gl || a.n_slices < 1 || a.n_out < a.n_slices ||
        (int64_t)a.n_out + a.n_slices - 1 > a.gl_len ||
        (a.tab && a.n_out != a.n_slices && !a.with_wake) ||
        (a.with_wake && (a.wk.n_mem > a.n_out || (a.wk.ind_coef && a.n_slices < 2))))
        return blmi_set_error(BLMI_ERR_ARG, "direct induced voltage: bad arguments");
    if (a.n_out > a.n_slices && a.n_slices > 1024) return launch_direct_shape<32, 512, 8>(a, s, ev);
    if (a.n_slices <= 1024) return launch_direct_shape<16, 256, 8>(a, s, ev);
    if (a.n_slices <= 4096) return launch_direct_shape<16, 512, 8>(a, s, ev);
    return launch_direct_shape<32, 512, 16>(a, s, ev);
}

// ---------------------------------------------------------------------------
// resonator impedance (fast_resonator.cpp:45-71): one thread per frequency,
// resonators accumulated in order
// ---------------------------------------------------------------------------
__global__ void fast_resonator_kernel(double *__restrict__ re, double *__restrict__ im,
                                      const double *__restrict__ freq, const double *__restrict__ R,
                                      const double *__restrict__ Q, const double *__restrict__ fr,
                                      int n_res, int n_f) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < 1 || k >= n_f) return;
    double zr = re[k], zi = im[k];
    const double f = freq[k];
    for (int r = 0; r < n_res; ++r) {
        const double q2 = Q[r] * Q[r];
        const double u = f / fr[r] - fr[r] / f;
        zr = zr + R[r] / (1.0 + q2 * u * u);
        zi = zi - R[r] * (Q[r] * u) / (1.0 + q2 * u * u);
    }
    re[k] = zr;
    im[k] = zi;
}

// ---------------------------------------------------------------------------
// plans
// ---------------------------------------------------------------------------
static bool factor_radices(int M, StageList &st) {
    st.n = 0;
    int m = M;
    const int order[3] = {8, 4, 2};
    for (int r : order)
        while (m % r == 0) {
            if (st.n >= kMaxStages) return false;
            st.radix[st.n++] = r;
            m /= r;
        }
    const int odd[2] = {5, 3};
    for (int r : odd)
        while (m % r == 0) {
            if (st.n >= kMaxStages) return false;
            st.radix[st.n++] = r;
            m /= r;
        }
    return m == 1;
}

// Threads of the one-workgroup LDS transforms: enough for 15 elements each,
// at least 512 (the c4 ensemble's 512-point transforms, one workgroup per
// bunch: 29.7 us a turn with 64 threads, 17.3 with 256, 14.9 with 512).
static int lds_threads(int M) {
    int t = (M + 14) / 15;
    t = ((t + 63) / 64) * 64;
    if (t < 512) t = 512;
    if (t > 1024) t = 1024;
    return t;
}

int64_t fft_lds_max_real() {
    const int64_t lds = device_max_lds();
    return 2 * (lds / (int64_t)sizeof(double2));
}

struct PlanCache {
    std::mutex mu;
    std::map<std::pair<int, int>, FftPlan> plans;  // (device, M)
};
static PlanCache &plan_cache() {
    static PlanCache c;
    return c;
}

int fft_plan_get(int64_t n_real, FftPlan *plan) {
    if (n_real < 2 || (n_real & 1))
        return blmi_set_error(BLMI_ERR_UNSUPPORTED, "FFT length %lld must be even and >= 2",
                              (long long)n_real);
    if (n_real > fft_lds_max_real())
        return blmi_set_error(BLMI_ERR_UNSUPPORTED,
                              "FFT length %lld exceeds the single-workgroup LDS transform (%lld)",
                              (long long)n_real, (long long)fft_lds_max_real());
    const int M = (int)(n_real / 2);
    StageList st;
    if (!factor_radices(M, st))
        return blmi_set_error(BLMI_ERR_UNSUPPORTED,
                              "FFT length %lld is not of the form 2*2^a*3^b*5^c", (long long)n_real);
    int dev = 0;
    BLMI_HIP_TRY(hipGetDevice(&dev));
    PlanCache &pc = plan_cache();
    std::lock_guard<std::mutex> lk(pc.mu);
    auto key = std::make_pair(dev, M);
    auto it = pc.plans.find(key);
    if (it != pc.plans.end()) {
        *plan = it->second;
        return BLMI_OK;
    }
    FftPlan p;
    p.M = M;
    p.n_real = 2 * M;
    p.n_stages = st.n;
    for (int s = 0; s < st.n; ++s) p.radix[s] = st.radix[s];
    p.threads = lds_threads(M);
    // twiddles from x87 extended-precision cosl/sinl, then rounded once to
    // double, so each entry is (nearly always) the correctly rounded value
    auto cis = [](long double num, long double den) {
        // exp(-2 pi i num/den)
        const long double two_pi = 6.283185307179586476925286766559L;
        long double a = two_pi * num / den;
        return make_double2((double)cosl(a), (double)(-sinl(a)));
    };
    std::vector<double2> tw(M), tw2(M + 1);
    for (int m = 0; m < M; ++m) tw[m] = cis(m, M);
    for (int k = 0; k <= M; ++k) tw2[k] = cis(k, 2.0L * M);
    BLMI_HIP_TRY(hipMalloc(&p.tw, sizeof(double2) * M));
    BLMI_HIP_TRY(hipMalloc(&p.tw2, sizeof(double2) * (M + 1)));
    BLMI_HIP_TRY(hipMemcpy(p.tw, tw.data(), sizeof(double2) * M, hipMemcpyHostToDevice));
    BLMI_HIP_TRY(hipMemcpy(p.tw2, tw2.data(), sizeof(double2) * (M + 1), hipMemcpyHostToDevice));
    pc.plans[key] = p;
    *plan = p;
    return BLMI_OK;
}

static StageList stages_of(const FftPlan &p) {
    StageList st;
    st.n = p.n_stages;
    for (int s = 0; s < p.n_stages; ++s) st.radix[s] = p.radix[s];
    return st;
}

static int set_lds_limit(const void *fn, size_t bytes) {
    if (bytes > 65536) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
        if (e != hipSuccess)
            return blmi_set_error(BLMI_ERR_HIP, "hipFuncSetAttribute(%zu B LDS): %s", bytes,
                                  hipGetErrorString(e));
    }
    return BLMI_OK;
}

int launch_rfft(const double *in, int64_t n_in, double2 *out, const FftPlan &p, hipStream_t s) {
    const size_t lds = sizeof(double2) * p.M;
    int rc = set_lds_limit((const void *)rfft_lds_kernel, lds);
    if (rc) return rc;
    hipLaunchKernelGGL(rfft_lds_kernel, dim3(1), dim3(p.threads), lds, s, in, n_in, out, p.M,
                       stages_of(p), p.tw, p.tw2);
    return blmi_check_launch("rfft");
}

int launch_irfft(const double2 *in, int64_t m_in, double *out, int64_t n_out, const FftPlan &p,
                 hipStream_t s) {
    const size_t lds = sizeof(double2) * p.M;
    int rc = set_lds_limit((const void *)irfft_lds_kernel, lds);
    if (rc) return rc;
    hipLaunchKernelGGL(irfft_lds_kernel, dim3(1), dim3(p.threads), lds, s, in, m_in, out, n_out,
                       p.M, stages_of(p), p.tw, p.tw2);
    return blmi_check_launch("irfft");
}

// The one-workgroup induced-voltage kernels for a plan's length: one
// wavefront a transform (twiddles in LDS after the data, twice the LDS)
// where M is a power of two in [16, 1024].
static int wave_len(int M) { return M >= 16 && M <= 1024 && (M & (M - 1)) == 0 ? M : 0; }

template <int FL>
static void pick_lds_kernels(int M, decltype(&induced_lds_kernel<0>) &one,
                             decltype(&induced_lds_batch_kernel<0>) &batch) {
    if constexpr (FL <= 1024) {
        if (M == FL) {
            one = induced_lds_kernel<FL>;
            batch = induced_lds_batch_kernel<FL>;
            return;
        }
        pick_lds_kernels<FL * 2>(M, one, batch);
    }
}

struct LdsKernels {
    decltype(&induced_lds_kernel<0>) one = induced_lds_kernel<0>;
    decltype(&induced_lds_batch_kernel<0>) batch = induced_lds_batch_kernel<0>;
    size_t lds = 0;
};

static LdsKernels lds_kernels(int M, int threads) {
    LdsKernels k;
    const bool wave = wave_len(M) && threads == kLdsWaveThreads;
    if (wave) pick_lds_kernels<16>(M, k.one, k.batch);
    k.lds = sizeof(double2) * (size_t)M * (wave ? 2 : 1);
    return k;
}

int launch_induced_voltage_lds(const InducedArgs &a, const FftPlan &p, hipStream_t s,
                               const KernelEvents &ev) {
    if (p.M / 2 + 1 > 8 * p.threads || p.M > 16 * p.threads)
        return blmi_set_error(BLMI_ERR_UNSUPPORTED, "FFT plan M=%d too large for %d threads", p.M,
                              p.threads);
    if (a.n_slices > p.n_real || a.n_out < 0 || a.n_out > p.n_real)
        return blmi_set_error(BLMI_ERR_ARG, "n_slices %d / n_out %d exceed n_fft %d", a.n_slices,
                              a.n_out, p.n_real);
    const LdsKernels k = lds_kernels(p.M, p.threads);
    int rc = set_lds_limit((const void *)k.one, k.lds);
    if (rc) return rc;
    launch(k.one, dim3(1), dim3(p.threads), k.lds, s, ev, a, p.M, stages_of(p), p.tw, p.tw2);
    return blmi_check_launch("induced_voltage");
}

int launch_induced_lds_batch(const InducedArgs *d_args, int n_bunches, int n_slices,
                             const FftPlan &p, hipStream_t s, const KernelEvents &ev) {
    if (p.M / 2 + 1 > 8 * p.threads || p.M > 16 * p.threads)
        return blmi_set_error(BLMI_ERR_UNSUPPORTED, "FFT plan M=%d too large for %d threads", p.M,
                              p.threads);
    if (n_slices > p.n_real)
        return blmi_set_error(BLMI_ERR_ARG, "n_slices %d exceeds n_fft %d", n_slices, p.n_real);
    if (n_bunches <= 0) return BLMI_OK;
    const LdsKernels k = lds_kernels(p.M, p.threads);
    int rc = set_lds_limit((const void *)k.batch, k.lds);
    if (rc) return rc;
    launch(k.batch, dim3(n_bunches), dim3(p.threads), k.lds, s, ev, d_args, p.M, stages_of(p), p.tw,
           p.tw2);
    return blmi_check_launch("induced_voltage_batch");
}

int launch_fast_resonator(double *re, double *im, const double *freq, const double *R,
                          const double *Q, const double *fr, int n_res, int n_f, hipStream_t s) {
    if (n_f <= 1) return BLMI_OK;
    hipLaunchKernelGGL(fast_resonator_kernel, dim3((n_f + 255) / 256), dim3(256), 0, s, re, im,
                       freq, R, Q, fr, n_res, n_f);
    return blmi_check_launch("fast_resonator");
}


// ---------------------------------------------------------------------------
// four-step real-FFT convolution for n_fft beyond one workgroup's LDS
//   complex length M = M1 * M2, z[m] = p[2m] + i p[2m+1], m = M2 m1 + m2
//   A: per column m2, T[k1][m2] = W_M^(m2 k1) sum_m1 z[M2 m1 + m2] W_M1^(m1 k1)
//   B: per row k1, Zc[k1 + M1 k2] = sum_m2 T[k1][m2] W_M2^(m2 k2); rows k1 and
//      M1 - k1 hold each other's mirror bins M - k, so a workgroup owns such a
//      pair, does the real split, the impedance product and the real merge,
//      then the inverse M2-point transform and the conjugate twiddle -> S
//   C: per column, inverse M1-point transform of S -> u[M2 m1 + m2], and
//      V[2m], V[2m+1] = -(scale * (u / M))   (the merge carries the 1/2)
// Columns are processed `cols` at a time so that each row access is a
// contiguous run (coalesced); LDS rows are padded by one element.
// ---------------------------------------------------------------------------
constexpr int kMaxCols = 8;

// Workgroups are dispatched to the 8 XCDs round-robin; the column passes
// give each XCD a contiguous range of columns, so that the columns sharing a
// 128-byte line of T or S (adjacent m2) meet in one L2.
__device__ __forceinline__ int xcd_block(int b, int n) {
    return (n % 8) ? b : (b % 8) * (n / 8) + b / 8;
}

__device__ __forceinline__ void count_buffers(const InducedArgs &a, const unsigned *&cin,
                                              unsigned *&cclr) {
    cin = nullptr;
    cclr = nullptr;
    if (a.counts0) {
        const int t = *a.turn_base + a.turn_off;
        cin = (t & 1) ? a.counts1 : a.counts0;
        cclr = (t & 1) ? a.counts0 : a.counts1;
    }
}

// FL, FL2: the column / row length as a constant when it is a power of two
// in [16, 1024] (one wavefront a transform), else 0 (workgroup-wide stages).
template <int FL>
__global__ void __launch_bounds__(kColThreads) fft2_a_kernel(InducedArgs a, int M1, int M2,
                                                             int cols, StageList st1,
                                                             const double2 *__restrict__ twM,
                                                             const double2 *__restrict__ twA,
                                                             double2 *__restrict__ T) {
    extern __shared__ double2 buf[];
    const int ld = M1 + 1;
    double2 *twl = buf + cols * ld;  // exp(-2 pi i j / M1), j < M1
    const int ns = a.n_slices;
    const int c0 = xcd_block(blockIdx.x, gridDim.x) * cols;
    const unsigned *cin;
    unsigned *cclr;
    count_buffers(a, cin, cclr);
    for (int j = threadIdx.x; j < M1; j += kColThreads) twl[j] = twM[(size_t)j * M2];
    // transposed partial rows: thread (g, q) sums words 4q'..4q'+3 (q' the
    // quad q within the batch's columns) of rows g, g + G, ... with 16-byte
    // loads (consecutive lanes, consecutive words); the G row groups are
    // then summed per element below
    unsigned *red = reinterpret_cast<unsigned *>(twl + M1);
    const int L = a.part ? 1 << a.part_lshift : 0;
    const int QC = cols * (L / 4);  // quads of the batch (host: QC divides blockDim)
    if (a.part) {
        const int G = kColThreads / QC;
        const int q = threadIdx.x % QC, g = threadIdx.x / QC;
        const int c = q / (L / 4), w = 4 * (q % (L / 4));
        unsigned lo[4] = {0u, 0u, 0u, 0u}, hi[4] = {0u, 0u, 0u, 0u};
        const unsigned *src = a.part + (size_t)(c0 + c) * a.part_rows * L + w;
        auto add = [&](const uint4 v) {
            lo[0] += v.x & 0xffffu;
            hi[0] += v.x >> 16;
            lo[1] += v.y & 0xffffu;
            hi[1] += v.y >> 16;
            lo[2] += v.z & 0xffffu;
            hi[2] += v.z >> 16;
            lo[3] += v.w & 0xffffu;
            hi[3] += v.w >> 16;
        };
        // kRowBatch rows' loads are issued before any is summed: a loop of
        // one load and one wait per row leaves each thread with a single
        // request in flight and the pass bound by memory latency
        constexpr int kRowBatch = 16;
        int r = g;
        for (; r + (kRowBatch - 1) * G < a.part_rows; r += kRowBatch * G) {
            uint4 v[kRowBatch];
#pragma unroll
            for (int u = 0; u < kRowBatch; ++u)
                v[u] = *reinterpret_cast<const uint4 *>(src + (size_t)(r + u * G) * L);
#pragma unroll
            for (int u = 0; u < kRowBatch; ++u) add(v[u]);
        }
        for (; r < a.part_rows; r += G) add(*reinterpret_cast<const uint4 *>(src + (size_t)r * L));
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            red[(g * QC + q) * 8 + 2 * i] = lo[i];
            red[(g * QC + q) * 8 + 2 * i + 1] = hi[i];
        }
        __syncthreads();
    }
    for (int j = threadIdx.x; j < cols * M1; j += kColThreads) {
        const int c = j % cols, m1 = j / cols;
        const int m = M2 * m1 + c0 + c;
        const int j0 = 2 * m, j1 = j0 + 1;
        double p[2] = {0.0, 0.0};
        const int jj[2] = {j0, j1};
        if (a.part) {
            // bins 2m (low halves) and 2m + 1 (high halves) of column c's word
            // m1, held by the rows at word m1 - part_m1_lo; with windowed rows
            // the bins outside are in cnt_out (read and cleared here)
            const int mw = m1 - a.part_m1_lo;
            unsigned s0 = 0u, s1 = 0u;
            if (mw >= 0 && mw < L) {
                const int qm = c * (L / 4) + mw / 4, i = mw % 4;
                for (int g = 0; g < (int)kColThreads / QC; ++g) {
                    s0 += red[(g * QC + qm) * 8 + 2 * i];
                    s1 += red[(g * QC + qm) * 8 + 2 * i + 1];
                }
            }
            if (a.cnt_out) {
                if (j0 < ns) {
                    s0 += a.cnt_out[j0];
                    a.cnt_out[j0] = 0u;
                }
                if (j1 < ns) {
                    s1 += a.cnt_out[j1];
                    a.cnt_out[j1] = 0u;
                }
            }
            p[0] = (double)s0;
            p[1] = (double)s1;
            if (a.profile_out) {
                if (j0 < ns) a.profile_out[j0] = p[0];
                if (j1 < ns) a.profile_out[j1] = p[1];
            }
            buf[c * ld + m1] = make_double2(p[0], p[1]);
            continue;
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            if (jj[h] >= ns) continue;
            if (cin) {
                unsigned sum = 0;
                for (int g = 0; g < a.groups; ++g) sum += cin[(size_t)g * ns + jj[h]];
                p[h] = (double)sum;
                for (int g = 0; g < a.groups; ++g) cclr[(size_t)g * ns + jj[h]] = 0u;
            } else {
                p[h] = a.src[jj[h]];
            }
            if (a.profile_out) a.profile_out[jj[h]] = p[h];
        }
        buf[c * ld + m1] = make_double2(p[0], p[1]);
    }
    // with one output element a thread (the usual split), its twiddle is
    // loaded before the transforms, so that load's latency overlaps them
    const bool one = cols * M1 <= kColThreads;
    double tw_x = 0.0, tw_y = 0.0;  // two doubles: a double2 here went to scratch
    if (one && (int)threadIdx.x < cols * M1) {
        const int c = threadIdx.x % cols, k1 = threadIdx.x / cols;
        const double2 t = twA[(size_t)(c0 + c) * M1 + k1];
        tw_x = t.x;
        tw_y = t.y;
    }
    __syncthreads();
    if constexpr (FL > 0)
        wave_fft_batch<FL, false>(buf, cols, ld, twl, kColThreads);
    else
        fft_lds_batch<false>(buf, M1, cols, ld, st1, twl, M1, kColThreads);
    for (int j = threadIdx.x; j < cols * M1; j += kColThreads) {
        const int c = j % cols, k1 = j / cols;
        const int m2 = c0 + c;
        const double2 w = one ? make_double2(tw_x, tw_y) : twA[(size_t)m2 * M1 + k1];
        T[(size_t)k1 * M2 + m2] = cmul(buf[c * ld + k1], w);
    }
}

constexpr int kPassBThreads = 512;

template <int FL2>
__global__ void __launch_bounds__(kPassBThreads) fft2_b_kernel(InducedArgs a, int M1, int M2,
                                                               StageList st2,
                              const double2 *__restrict__ twM, const double2 *__restrict__ twB,
                              const double2 *__restrict__ T, double2 *__restrict__ S) {
    extern __shared__ double2 buf[];
    double2 *twl = buf + 2 * M2;  // exp(-2 pi i j / M2), j < M2
    const int M = M1 * M2;
    const int q = blockIdx.x;  // rows q and M1 - q
    const int r0 = q, r1 = (M1 - q) % M1;
    const int nrow = r1 == r0 ? 1 : 2;
    const int ne = nrow * M2;
    constexpr int EMAX = 4;  // host guarantees ne <= EMAX * kPassBThreads
    // everything the split/merge needs is requested up front, so its latency
    // overlaps the row loads and the forward transforms
    double2 zk[EMAX], zmk[EMAX], wk[EMAX], wmk[EMAX], wo[EMAX];
#pragma unroll
    for (int e = 0; e < EMAX; ++e) {
        const int j = threadIdx.x + e * kPassBThreads;
        if (j < ne) {
            const int row = j >= M2, k2 = j - row * M2;
            const int r = row ? r1 : r0;
            const size_t rk = (size_t)r * M2 + k2;  // the tables in this pass's order
            if (a.Zp) {
                zk[e] = a.Zp[rk];
                zmk[e] = a.Zp[M + rk];
            } else {
                const int k = r + M1 * k2;
                zk[e] = a.Z[k];
                zmk[e] = a.Z[M - k];
            }
            wk[e] = twB[rk];
            wmk[e] = twB[M + rk];
            wo[e] = twB[2 * (size_t)M + rk];  // output twiddle of element (row, m2 = k2)
        }
    }
    for (int j = threadIdx.x; j < M2; j += kPassBThreads) twl[j] = twM[(size_t)j * M1];
    for (int j = threadIdx.x; j < ne; j += kPassBThreads) {
        const int row = j >= M2, k2 = j - row * M2;
        buf[j] = T[(size_t)(row ? r1 : r0) * M2 + k2];
    }
    __syncthreads();
    if constexpr (FL2 > 0)
        wave_fft_batch<FL2, false>(buf, nrow, M2, twl, kPassBThreads);
    else
        fft_lds_batch<false>(buf, M2, nrow, M2, st2, twl, M2, kPassBThreads);
    // buf[row][k2] = Zc[r + M1 k2]; the mirror bin M - k sits at
    // row 0, (M2 - k2) % M2 for r = 0 and at row of M1 - r, M2 - 1 - k2 otherwise
    double2 u[EMAX];
#pragma unroll
    for (int e = 0; e < EMAX; ++e) {
        const int j = threadIdx.x + e * kPassBThreads;
        if (j < ne) {
            const int row = j >= M2, k2 = j - row * M2;
            const int r = row ? r1 : r0;
            const int k = r + M1 * k2;
            int pj;
            if (r == 0) {
                pj = (M2 - k2) % M2;
            } else {
                const int pr = M1 - r;
                pj = (pr == r0 ? 0 : M2) + (M2 - 1 - k2);
            }
            double2 Xk, Xmk;
            real_split(buf[j], buf[pj], wk[e], wmk[e], Xk, Xmk);
            // impedance.py:479-484: total_impedance * beam_spectrum
            double2 Yk = cmul(zk[e], Xk), Ymk = cmul(zmk[e], Xmk);
            if (k == 0) {  // c2r ignores the imaginary parts of DC and Nyquist
                Yk.y = 0.0;
                Ymk.y = 0.0;
            }
            u[e] = real_merge(Yk, Ymk, wk[e]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < EMAX; ++e) {
        const int j = threadIdx.x + e * kPassBThreads;
        if (j < ne) buf[j] = u[e];
    }
    __syncthreads();
    if constexpr (FL2 > 0)
        wave_fft_batch<FL2, true>(buf, nrow, M2, twl, kPassBThreads);
    else
        fft_lds_batch<true>(buf, M2, nrow, M2, st2, twl, M2, kPassBThreads);
#pragma unroll
    for (int e = 0; e < EMAX; ++e) {
        const int j = threadIdx.x + e * kPassBThreads;
        if (j < ne) {
            const int row = j >= M2, m2 = j - row * M2;
            const int r = row ? r1 : r0;
            S[(size_t)r * M2 + m2] = cmul(buf[j], cconj(wo[e]));
        }
    }
}

// Pass C also transforms the column after its batch (column 0, one row down,
// for the last batch), so every V[2m+2] a table entry needs is at hand.
template <int FL>
__global__ void __launch_bounds__(kColThreads) fft2_c_kernel(InducedArgs a, int M1, int M2,
                                                             int cols, StageList st1,
                                                             const double2 *__restrict__ twM,
                                                             const double2 *__restrict__ S) {
    extern __shared__ double2 buf[];
    const int NC = cols + 1;
    const int ld = M1 + 1;
    double2 *twl = buf + NC * ld;
    const int M = M1 * M2;
    const int c0 = xcd_block(blockIdx.x, gridDim.x) * cols;
    const bool last = c0 + cols == M2;
    for (int j = threadIdx.x; j < M1; j += kColThreads) twl[j] = twM[(size_t)j * M2];
    for (int j = threadIdx.x; j < NC * M1; j += kColThreads) {
        const int c = j % NC, k1 = j / NC;
        const int col = (c0 + c) % M2;
        buf[c * ld + k1] = S[(size_t)k1 * M2 + col];
    }
    // with one output element a thread (the usual split), the bin centres
    // its table entries need are loaded before the transforms
    const int ns = a.n_slices;
    const bool one = cols * M1 <= kColThreads;
    double bc_pre[2] = {0.0, 0.0};
    if (one && a.tab && (int)threadIdx.x < cols * M1) {
        const int c = threadIdx.x % cols, m1 = threadIdx.x / cols;
        const int m = M2 * m1 + c0 + c;
        if (2 * m < ns - 1) bc_pre[0] = a.bc[2 * m];
        if (2 * m + 1 < ns - 1) bc_pre[1] = a.bc[2 * m + 1];
    }
    // linear_interp_kick.cpp:33-48 on V (acc = 0 for the induced kick); the
    // scale is formed before the transforms (its loads overlap them)
    double inv_bw = 0.0;
    if (a.tab) inv_bw = interp_inv_bw(a.bc, ns, a.f32);
    __syncthreads();
    if constexpr (FL > 0)
        wave_fft_batch<FL, true>(buf, NC, ld, twl, kColThreads);
    else
        fft_lds_batch<true>(buf, M1, NC, ld, st1, twl, M1, kColThreads);
    // V = -(charge*e*ratio * irfft(...))[:n_out]   (impedance.py:475-489)
    const double fct = 1.0 / (double)M;
    if (a.tab && blockIdx.x == 0 && threadIdx.x == 0) {
        a.tab_params[0] = a.bc[0];
        a.tab_params[1] = inv_bw;
    }
    for (int j = threadIdx.x; j < cols * M1; j += kColThreads) {
        const int c = j % cols, m1 = j / cols;
        const int m = M2 * m1 + c0 + c;
        const double2 u = buf[c * ld + m1];
        const double v0 = -(a.scale * (u.x * fct)), v1 = -(a.scale * (u.y * fct));
        if (2 * m < a.n_out) a.volt[2 * m] = v0;
        if (2 * m + 1 < a.n_out) a.volt[2 * m + 1] = v1;
        if (a.tab && 2 * m < ns - 1) {
            const double b0 = one ? bc_pre[0] : a.bc[2 * m];
            store_interp_entry(a.tab, 2 * m, interp_entry(v0, v1, b0, a.charge, inv_bw, a.f32),
                               a.f32);
            if (2 * m + 1 < ns - 1) {
                // V[2m+2]: next column, same row (column 0, next row after the last)
                const double2 un = c + 1 < cols ? buf[(c + 1) * ld + m1]
                                   : buf[cols * ld + (last ? m1 + 1 : m1)];
                const double v2 = -(a.scale * (un.x * fct));
                const double b1 = one ? bc_pre[1] : a.bc[2 * m + 1];
                store_interp_entry(a.tab, 2 * m + 1,
                                   interp_entry(v1, v2, b1, a.charge, inv_bw, a.f32), a.f32);
            }
        }
    }
}

static bool factor_list(int L, int *radix, int &n) {
    StageList st;
    if (!factor_radices(L, st)) return false;
    n = st.n;
    for (int s = 0; s < st.n; ++s) radix[s] = st.radix[s];
    return true;
}

static StageList stage_list(const int *radix, int n) {
    StageList st;
    st.n = n;
    for (int s = 0; s < n; ++s) st.radix[s] = radix[s];
    return st;
}

int fft2_plan_create(int64_t n_real, Fft2Plan *plan) {
    *plan = Fft2Plan();
    if (n_real < 4 || (n_real & 1))
        return blmi_set_error(BLMI_ERR_UNSUPPORTED, "FFT length %lld must be even and >= 4",
                              (long long)n_real);
    if (n_real > (int64_t)1 << 30)
        return blmi_set_error(BLMI_ERR_UNSUPPORTED, "FFT length %lld too large", (long long)n_real);
    const int M = (int)(n_real / 2);
    // Split M = M1 * M2 so that every pass has work for many workgroups
    // and little per workgroup (the passes are latency-bound): M2 a divisor
    // in [16, 1024] as close to sqrt(M) as possible, pass A/C batches of
    // `cols` columns with about 512 elements per workgroup.
    int M1 = 0, M2 = 0, cols = 1;
    double best = 1e300;
    for (int d = 16; d <= 1024 && d <= M / 2; ++d) {
        if (M % d) continue;
        const int m1 = M / d;
        int c = 1;
        while (c < kMaxCols && 2 * c * m1 <= 512 && d % (2 * c) == 0) c *= 2;
        if ((c + 1) * m1 > 15 * kColThreads) continue;  // pass C's batch
        int r[kMaxStages], n1, n2;
        if (!factor_list(m1, r, n1) || !factor_list(d, r, n2)) continue;
        const double score = fabs(log((double)d) - 0.5 * log((double)M));
        if (score < best) {
            best = score;
            M1 = m1;
            M2 = d;
            cols = c;
        }
    }
    if (!M1)
        return blmi_set_error(BLMI_ERR_UNSUPPORTED,
                              "FFT length %lld has no four-step split (needs 2*2^a*3^b*5^c)",
                              (long long)n_real);
    Fft2Plan p;
    p.M = M;
    p.M1 = M1;
    p.M2 = M2;
    p.cols = cols;
    p.n_real = 2 * M;
    factor_list(M1, p.r1, p.n1);
    factor_list(M2, p.r2, p.n2);
    // pass B: at most 4 elements a thread (registers hold their operands)
    int tb = (2 * M2 + 3) / 4;
    tb = ((tb + 63) / 64) * 64;
    // at least 512 threads (c3's 2^17 transform: passes at 512 threads take
    // the three launches from 33.1 to 31.6 us, with kColThreads = 512)
    p.threads_b = tb < 512 ? 512 : tb;  // <= kPassBThreads for M2 <= 1024
    if (p.threads_b != kPassBThreads)
        return blmi_set_error(BLMI_ERR_UNSUPPORTED, "four-step: pass B needs %d threads", p.threads_b);
    auto cis = [](long double num, long double den) {
        const long double two_pi = 6.283185307179586476925286766559L;
        long double x = two_pi * num / den;
        return make_double2((double)cosl(x), (double)(-sinl(x)));
    };
    std::vector<double2> twM(M), twN(M + 1);
    for (int j = 0; j < M; ++j) twM[j] = cis(j, M);
    for (int k = 0; k <= M; ++k) twN[k] = cis(k, 2.0L * M);
    std::vector<double2> twA((size_t)M), twB(3 * (size_t)M);
    for (int m2 = 0; m2 < M2; ++m2)
        for (int k1 = 0; k1 < M1; ++k1) twA[(size_t)m2 * M1 + k1] = twM[(size_t)m2 * k1];
    for (int r = 0; r < M1; ++r)
        for (int k2 = 0; k2 < M2; ++k2) {
            const size_t rk = (size_t)r * M2 + k2;
            const int k = r + M1 * k2;
            twB[rk] = twN[k];
            twB[M + rk] = twN[M - k];
            twB[2 * (size_t)M + rk] = twM[(size_t)r * k2];
        }
    hipError_t e = hipMalloc(&p.twM, sizeof(double2) * M);
    if (e == hipSuccess) e = hipMalloc(&p.twN, sizeof(double2) * (M + 1));
    if (e == hipSuccess) e = hipMalloc(&p.twA, sizeof(double2) * M);
    if (e == hipSuccess) e = hipMalloc(&p.twB, sizeof(double2) * 3 * (size_t)M);
    if (e == hipSuccess)
        e = hipMemcpy(p.twA, twA.data(), sizeof(double2) * M, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(p.twB, twB.data(), sizeof(double2) * 3 * (size_t)M, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&p.T, sizeof(double2) * M);
    if (e == hipSuccess) e = hipMalloc(&p.S, sizeof(double2) * M);
    if (e == hipSuccess)
        e = hipMemcpy(p.twM, twM.data(), sizeof(double2) * M, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(p.twN, twN.data(), sizeof(double2) * (M + 1), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        fft2_plan_destroy(&p);
        return blmi_set_error(BLMI_ERR_ALLOC, "four-step plan: %s", hipGetErrorString(e));
    }
    *plan = p;
    return BLMI_OK;
}

void fft2_permute_z(const Fft2Plan &p, const double2 *Z, double2 *Zp) {
    const size_t M = (size_t)p.M;
    for (int r = 0; r < p.M1; ++r)
        for (int k2 = 0; k2 < p.M2; ++k2) {
            const size_t rk = (size_t)r * p.M2 + k2;
            const size_t k = (size_t)r + (size_t)p.M1 * k2;
            Zp[rk] = Z[k];
            Zp[M + rk] = Z[M - k];
        }
}

void fft2_plan_destroy(Fft2Plan *p) {
    (void)hipFree(p->twM);
    (void)hipFree(p->twN);
    (void)hipFree(p->twA);
    (void)hipFree(p->twB);
    (void)hipFree(p->T);
    (void)hipFree(p->S);
    *p = Fft2Plan();
}

// The pass kernels for the plan's lengths: one wavefront a transform where
// a length is a power of two in [16, 1024], workgroup-wide stages otherwise.
struct Fft2Kernels {
    decltype(&fft2_a_kernel<0>) a;
    decltype(&fft2_b_kernel<0>) b;
    decltype(&fft2_c_kernel<0>) c;
};

static Fft2Kernels fft2_kernels(int M1, int M2) {
    Fft2Kernels k{fft2_a_kernel<0>, fft2_b_kernel<0>, fft2_c_kernel<0>};
    switch (M1) {
        case 16: k.a = fft2_a_kernel<16>; k.c = fft2_c_kernel<16>; break;
        case 32: k.a = fft2_a_kernel<32>; k.c = fft2_c_kernel<32>; break;
        case 64: k.a = fft2_a_kernel<64>; k.c = fft2_c_kernel<64>; break;
        case 128: k.a = fft2_a_kernel<128>; k.c = fft2_c_kernel<128>; break;
        case 256: k.a = fft2_a_kernel<256>; k.c = fft2_c_kernel<256>; break;
        case 512: k.a = fft2_a_kernel<512>; k.c = fft2_c_kernel<512>; break;
        case 1024: k.a = fft2_a_kernel<1024>; k.c = fft2_c_kernel<1024>; break;
        default: break;
    }
    switch (M2) {
        case 16: k.b = fft2_b_kernel<16>; break;
        case 32: k.b = fft2_b_kernel<32>; break;
        case 64: k.b = fft2_b_kernel<64>; break;
        case 128: k.b = fft2_b_kernel<128>; break;
        case 256: k.b = fft2_b_kernel<256>; break;
        case 512: k.b = fft2_b_kernel<512>; break;
        case 1024: k.b = fft2_b_kernel<1024>; break;
        default: break;
    }
    return k;
}

int launch_induced_fourstep(const InducedArgs &a, const Fft2Plan &p, hipStream_t s,
                            const KernelEvents &ev) {
    if (!p.M) return blmi_set_error(BLMI_ERR_ARG, "four-step plan not created");
    if (a.n_slices > p.n_real || a.n_out < 0 || a.n_out > p.n_real)
        return blmi_set_error(BLMI_ERR_ARG, "n_slices %d / n_out %d exceed n_fft %d", a.n_slices,
                              a.n_out, p.n_real);
    if (a.counts0 && a.groups > kMaxCountGroups)
        return blmi_set_error(BLMI_ERR_ARG, "too many count groups");
    const int cols_a = p.cols;
    if (a.part) {
        const int L = 1 << a.part_lshift;
        if (L < 4 || kColThreads % (cols_a * (L / 4)) || a.part_m1_lo < 0 ||
            a.part_m1_lo + L > p.M1 ||
            (!a.cnt_out && (a.part_m1_lo != 0 || (int64_t)L * p.M2 * 2 < a.n_slices)))
            return blmi_set_error(BLMI_ERR_ARG, "four-step: partial-row layout does not fit");
    }
    // the transposed partial rows add the row groups' sums (8 words a thread)
    const size_t lds_a = sizeof(double2) * (cols_a * (size_t)(p.M1 + 1) + p.M1) +
                         (a.part ? sizeof(unsigned) * 8 * kColThreads : 0);
    const size_t lds_c = sizeof(double2) * ((p.cols + 1) * (size_t)(p.M1 + 1) + p.M1);
    const size_t lds_b = sizeof(double2) * 3 * (size_t)p.M2;
    const Fft2Kernels k = fft2_kernels(p.M1, p.M2);
    int rc = set_lds_limit((const void *)k.a, lds_a);
    if (!rc) rc = set_lds_limit((const void *)k.c, lds_c);
    if (!rc) rc = set_lds_limit((const void *)k.b, lds_b);
    if (rc) return rc;
    const StageList st1 = stage_list(p.r1, p.n1), st2 = stage_list(p.r2, p.n2);
    const int cols = p.M2 / p.cols;
    const KernelEvents first{ev.start, nullptr}, last{nullptr, ev.stop};
    launch(k.a, dim3(p.M2 / cols_a), dim3(kColThreads), lds_a, s, first, a, p.M1, p.M2,
           cols_a, st1, p.twM, p.twA, p.T);
    hipLaunchKernelGGL(k.b, dim3(p.M1 / 2 + 1), dim3(p.threads_b), lds_b, s, a, p.M1,
                       p.M2, st2, p.twM, p.twB, p.T, p.S);
    launch(k.c, dim3(cols), dim3(kColThreads), lds_c, s, last, a, p.M1, p.M2, p.cols,
           st1, p.twM, p.S);
    return blmi_check_launch("induced_fourstep");
}

}  // namespace blmi
