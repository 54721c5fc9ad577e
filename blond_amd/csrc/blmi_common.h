// Shared device helpers of the MI355X tracking library (gfx950 only).
//
// Arithmetic contract: every kernel that must agree bit for bit with the
// reference's strict C++ build evaluates its expressions in the same order
// and with FMA contraction disabled (the translation units are compiled with
// -ffp-contract=off).  FP64 division and square root are the IEEE correctly
// rounded expansions the compiler emits without fast-math flags.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#define BLMI_OK 0
#define BLMI_ERR_ARG 1
#define BLMI_ERR_HIP 2
#define BLMI_ERR_UNSUPPORTED 3
#define BLMI_ERR_ALLOC 4

// Host-side error plumbing (defined in capi.hip).
int blmi_set_error(int code, const char *fmt, ...);
int blmi_check_launch(const char *what);
extern "C" const char *blmi_last_error(void);

#define BLMI_HIP_TRY(expr)                                                   \
    do {                                                                     \
        hipError_t _e = (expr);                                              \
        if (_e != hipSuccess)                                                \
            return blmi_set_error(BLMI_ERR_HIP, "%s: %s", #expr,             \
                                  hipGetErrorString(_e));                    \
    } while (0)

namespace blmi {

// Wavefront width of CDNA4.
constexpr int kWave = 64;

// Optional start/stop events recorded by the kernel's own dispatch
// (hipExtLaunchKernelGGL), i.e. the kernel's execution interval without the
// queue gaps a separate hipEventRecord would add.
struct KernelEvents {
    hipEvent_t start = nullptr, stop = nullptr;
};

template <typename F, typename... Args>
inline void launch(F kernel, dim3 grid, dim3 block, size_t lds, hipStream_t s,
                   const KernelEvents &ev, Args... args) {
    // either event may be null (a span over several launches)
    if (ev.start || ev.stop)
        hipExtLaunchKernelGGL(kernel, grid, block, (std::uint32_t)lds, s, ev.start, ev.stop, 0u,
                              args...);
    else
        hipLaunchKernelGGL(kernel, grid, block, lds, s, args...);
}

// ---------------------------------------------------------------------------
// Sine and cosine by the Cephes method (S. L. Moshier, Cephes Math Library,
// sin.c), arranged as in the VDT library the reference vendors
// (blond/cpp_routines/sincos.h:72-160): even octant index j = (int(4|x|/pi)+1)&~1,
// three-part Cody-Waite reduction r = ((|x| - j*DP1) - j*DP2) - j*DP3, the
// Cephes degree-6 sine and cosine polynomials on r, then the octant picks the
// polynomial and the sign.  Same constants and the same operation order as
// the reference, so with contraction off the result is bit-identical to the
// reference's vdt::fast_sin on x86-64.
// ---------------------------------------------------------------------------
struct CephesD {
    static constexpr double kFourOverPi = 4.0 / 3.14159265358979323846;
    static constexpr double kDP1 = 7.853981554508209228515625E-1;
    static constexpr double kDP2 = 7.94662735614792836714E-9;
    static constexpr double kDP3 = 3.06161699786838294307E-17;
    static constexpr double kS0 = 1.58962301576546568060E-10;
    static constexpr double kS1 = -2.50507477628578072866E-8;
    static constexpr double kS2 = 2.75573136213857245213E-6;
    static constexpr double kS3 = -1.98412698295895385996E-4;
    static constexpr double kS4 = 8.33333333332211858878E-3;
    static constexpr double kS5 = -1.66666666666666307295E-1;
    static constexpr double kC0 = -1.13585365213876817300E-11;
    static constexpr double kC1 = 2.08757008419747316778E-9;
    static constexpr double kC2 = -2.75573141792967388112E-7;
    static constexpr double kC3 = 2.48015872888517045348E-5;
    static constexpr double kC4 = -1.38888888888730564116E-3;
    static constexpr double kC5 = 4.16666666666665929218E-2;
};

// The reference's branch test ((j - 2) & 2) == 0 (sincos.h: the cosine
// polynomial serves the octants j = 2, 6 mod 8).  j is even there (rounded
// up to even), and for even j, modulo 2^32 too, bit 1 of j - 2 is bit 1 of
// j inverted, so the test is (j & 2) != 0: the same branch, one subtraction
// fewer.
__device__ __forceinline__ bool cephes_cos_branch(int j) { return (j & 2) != 0; }

// Reduced-argument polynomials: returns sin-branch and cos-branch values of r
// and the octant index of |x|.
__device__ __forceinline__ void cephes_reduce(double x, double &sp, double &cp,
                                              int &oct) {
    const double ax = fabs(x);
    int j = (int)(CephesD::kFourOverPi * ax);
    j = (j + 1) & ~1;
    const double y = (double)j;
    const double r = ((ax - y * CephesD::kDP1) - y * CephesD::kDP2) - y * CephesD::kDP3;
    const double zz = r * r;
    double ps = CephesD::kS0;
    ps = ps * zz; ps = ps + CephesD::kS1;
    ps = ps * zz; ps = ps + CephesD::kS2;
    ps = ps * zz; ps = ps + CephesD::kS3;
    ps = ps * zz; ps = ps + CephesD::kS4;
    ps = ps * zz; ps = ps + CephesD::kS5;
    double pc = CephesD::kC0;
    pc = pc * zz; pc = pc + CephesD::kC1;
    pc = pc * zz; pc = pc + CephesD::kC2;
    pc = pc * zz; pc = pc + CephesD::kC3;
    pc = pc * zz; pc = pc + CephesD::kC4;
    pc = pc * zz; pc = pc + CephesD::kC5;
    sp = r + r * zz * ps;
    cp = 1.0 - zz * 0.5 + zz * zz * pc;
    oct = j;
}

__device__ __forceinline__ double cephes_sin(double x) {
    double sp, cp;
    int j;
    cephes_reduce(x, sp, cp, j);
    const double out = cephes_cos_branch(j) ? cp : sp;
    // the reference's two negations (octant j & 4, then x < 0) as one flip
    // of the sign bit; x < 0 (not the sign bit of x) keeps sin(-0.0) = +0.0
    const unsigned flip = ((unsigned)j << 29) ^ (x < 0.0 ? 0x80000000u : 0u);
    return __hiloint2double(__double2hiint(out) ^ (int)(flip & 0x80000000u),
                            __double2loint(out));
}

// cephes_sin with one polynomial a lane, for kernels that keep the
// coefficient table in LDS (kCephesTabDoubles doubles, 16-byte aligned:
// kS0..kS5 of the sine branch, then kC0..kC5 of the cosine branch;
// cephes_tab_fill writes it).  The lane's octant picks its branch's six
// coefficients, evaluated in cephes_reduce's Horner order; the branch's last
// steps, r + (r zz) P and (1 - zz/2) + (zz zz) P, become one A + B P with
// A and B selected.  Every operation is one cephes_sin performs on the same
// operands, so the result is bit-identical, and no lane evaluates the
// polynomial its octant discards.
constexpr int kCephesTabDoubles = 12;
using LdsConstDoublePtr = __attribute__((address_space(3))) const double *;
using LdsDoublePtr = __attribute__((address_space(3))) double *;

__device__ __forceinline__ double cephes_coef(int i) {
    // a select chain, not an indexed local array (which would live in
    // scratch memory)
    switch (i) {
        case 0: return CephesD::kS0;
        case 1: return CephesD::kS1;
        case 2: return CephesD::kS2;
        case 3: return CephesD::kS3;
        case 4: return CephesD::kS4;
        case 5: return CephesD::kS5;
        case 6: return CephesD::kC0;
        case 7: return CephesD::kC1;
        case 8: return CephesD::kC2;
        case 9: return CephesD::kC3;
        case 10: return CephesD::kC4;
        default: return CephesD::kC5;
    }
}

__device__ __forceinline__ void cephes_tab_fill(LdsDoublePtr tab) {
    for (int i = (int)threadIdx.x; i < kCephesTabDoubles; i += (int)blockDim.x)
        tab[i] = cephes_coef(i);
}

__device__ __forceinline__ double cephes_sin_tab(double x, LdsConstDoublePtr tab) {
    const double ax = fabs(x);
    int j = (int)(CephesD::kFourOverPi * ax);
    j = (j + 1) & ~1;
    const double y = (double)j;
    const double r = ((ax - y * CephesD::kDP1) - y * CephesD::kDP2) - y * CephesD::kDP3;
    const double zz = r * r;
    const bool cosb = cephes_cos_branch(j);
    using D2 = double __attribute__((ext_vector_type(2)));
    using LdsD2 = __attribute__((address_space(3))) const D2 *;
    const LdsD2 c = reinterpret_cast<LdsD2>(tab + (cosb ? 6 : 0));
    const D2 c01 = c[0], c23 = c[1], c45 = c[2];
    double p = c01.x;
    p = p * zz; p = p + c01.y;
    p = p * zz; p = p + c23.x;
    p = p * zz; p = p + c23.y;
    p = p * zz; p = p + c45.x;
    p = p * zz; p = p + c45.y;
    const double A = cosb ? 1.0 - zz * 0.5 : r;
    const double B = (cosb ? zz : r) * zz;
    const double out = A + B * p;
    const unsigned flip = ((unsigned)j << 29) ^ (x < 0.0 ? 0x80000000u : 0u);
    return __hiloint2double(__double2hiint(out) ^ (int)(flip & 0x80000000u),
                            __double2loint(out));
}

// Both functions from one reduction (the adjoint needs the cosine).
__device__ __forceinline__ void cephes_sincos(double x, double &s, double &c) {
    double sp, cp;
    int j;
    cephes_reduce(x, sp, cp, j);
    const bool swap = cephes_cos_branch(j);
    s = swap ? cp : sp;
    c = swap ? sp : cp;
    if (j & 4) s = -s;
    if (((j - 2) & 4) == 0) c = -c;
    if (x < 0.0) s = -s;
}

// Exponential by the Cephes method (exp.c), arranged as in the VDT library
// the reference vendors (blond/cpp_routines/exp.h:70-113): n = floor'(x
// log2(e) + 1/2) with floor'(y) = trunc(y) - signbit((float)y)
// (vdtcore_common.h:218-224), the two-part reduction x - n C1 - n C2, the
// Pade form 1 + 2 x P(x^2) / (Q(x^2) - x P(x^2)) with Cephes's P and Q, times
// 2^n; +inf above 708 and 0 below -708.
__device__ __forceinline__ double cephes_exp(double x0) {
    double x = x0;
    const double t = 1.4426950408889634073599 * x + 0.5;
    double px = (double)((int)t - (signbit((float)t) ? 1 : 0));
    const int n = (int)px;
    x -= px * 6.93145751953125E-1;
    x -= px * 1.42860682030941723212E-6;
    const double xx = x * x;
    px = 1.26177193074810590878E-4;
    px *= xx;
    px += 3.02994407707441961300E-2;
    px *= xx;
    px += 9.99999999999999999910E-1;
    px *= x;
    double qx = 3.00198505138664455042E-6;
    qx *= xx;
    qx += 2.52448340349684104192E-3;
    qx *= xx;
    qx += 2.27265548208155028766E-1;
    qx *= xx;
    qx += 2.00000000000000000009E0;
    x = px / (qx - px);
    x = 1.0 + 2.0 * x;
    x *= __longlong_as_double((long long)((((uint64_t)(int64_t)n) + 1023) << 52));
    if (x0 > 708.0) x = __longlong_as_double(0x7ff0000000000000LL);
    if (x0 < -708.0) x = 0.0;
    return x;
}

// float version (sincos.h:166-233, Cephes sinf)
__device__ __forceinline__ float cephes_sinf(float x) {
    const float kFourOverPiF = (float)(4.0 / 3.14159265358979323846);
    const float ax = fabsf(x);
    int j = (int)(kFourOverPiF * ax);
    j = (j + 1) & ~1;
    const float y = (float)j;
    const float r = ((ax - y * 0.78515625f) - y * 2.4187564849853515625e-4f) -
                    y * 3.77489497744594108e-8f;
    const float z = r * r;
    const float s = (((-1.9515295891E-4f * z + 8.3321608736E-3f) * z -
                      1.6666654611E-1f) * z * r) + r;
    const float c = ((2.443315711809948E-005f * z - 1.388731625493765E-003f) * z +
                     4.166664568298827E-002f) * z * z - 0.5f * z + 1.0f;
    const float out = cephes_cos_branch(j) ? c : s;
    // the two sign changes (octant, then x < 0) as one: they cancel when
    // both apply
    const bool neg = ((j & 4) != 0) != (x < 0.0f);
    return neg ? -out : out;
}

template <typename T> struct Trig;
template <> struct Trig<double> {
    __device__ __forceinline__ static double sin(double x) { return cephes_sin(x); }
};
template <> struct Trig<float> {
    __device__ __forceinline__ static float sin(float x) { return cephes_sinf(x); }
};
// the sine with an LDS coefficient table (FP64; the FP32 sine ignores it)
template <typename T> struct TrigTab;
template <> struct TrigTab<double> {
    __device__ __forceinline__ static double sin(double x, LdsConstDoublePtr tab) {
        return cephes_sin_tab(x, tab);
    }
};
template <> struct TrigTab<float> {
    __device__ __forceinline__ static float sin(float x, LdsConstDoublePtr) { return cephes_sinf(x); }
};

// linear_interp_kick.cpp:33-48: the interpolation table of a voltage on
// the bin centres (acceleration kick 0), its scale and one bin's (slope,
// intercept).  With f32 the reference's single-precision build (real_t =
// float): the voltage and the bin centres are float arrays there, so V is
// rounded to float and the scale and each entry are formed in float, in the
// same order, and stored exactly as doubles.
__device__ __forceinline__ double interp_inv_bw(const double *bc, int ns, bool f32) {
    if (f32) return (double)((float)(ns - 1) / ((float)bc[ns - 1] - (float)bc[0]));
    return (double)(ns - 1) / (bc[ns - 1] - bc[0]);
}

__device__ __forceinline__ double2 interp_entry(double v0, double v1, double bcb, double charge,
                                                double inv_bw, bool f32) {
    if (f32) {
        const float a = (float)v0, b = (float)v1, q = (float)charge, ib = (float)inv_bw;
        const float k = q * (b - a) * ib;
        return make_double2((double)k, (double)((q * a - (float)bcb * k) + 0.0f));
    }
    const double k = charge * (v1 - v0) * inv_bw;
    return make_double2(k, (charge * v0 - bcb * k) + 0.0);
}

// A table entry as stored: a (slope, intercept) pair of doubles, or with f32
// of floats, which hold the single build's FP32 values exactly and halve the
// bytes the kick gathers (the table then holds float2 entries; its
// (bin_centers[0], inv_bw) parameters stay doubles after 2 * n_slices doubles)
__device__ __forceinline__ void store_interp_entry(double2 *tab, int i, double2 e, bool f32) {
    if (f32)
        reinterpret_cast<float2 *>(tab)[i] = make_float2((float)e.x, (float)e.y);
    else
        tab[i] = e;
}

// Agent-scope atomics (all XCDs of the device see one coherent value).
__device__ __forceinline__ void atomic_add_agent(unsigned *p, unsigned v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void atomic_add_agent(unsigned long long *p,
                                                 unsigned long long v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void atomic_add_agent(double *p, double v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void atomic_add_agent(float *p, float v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace blmi
