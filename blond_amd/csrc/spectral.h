// FFT plans and slice-axis (spectral) launchers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "blmi_common.h"
#include "feedback.h"
#include "wake.h"

namespace blmi {

constexpr int kMaxStages = 24;
// zero padding after the direct method's unrolled wake kernel
constexpr int kConvPad = 64;
// copies of the global histogram counts (workgroup b adds into copy b % groups)
constexpr int kMaxCountGroups = 16;

// Complex FFT of length M = prod(radix[s]), radices in {2,3,4,5,8}, done as a
// Stockham autosort sequence inside one workgroup's LDS.
struct FftPlan {
    int M = 0;              // complex length (n_fft / 2 for the real transforms)
    int n_real = 0;         // real length n = 2M
    int n_stages = 0;
    int radix[kMaxStages];
    int threads = 0;        // workgroup size used by the LDS kernels
    double2 *tw = nullptr;  // exp(-2 pi i m / M), m in [0, M)
    double2 *tw2 = nullptr; // exp(-2 pi i k / 2M), k in [0, M]
};

// Returns BLMI_OK and fills *plan (device twiddles cached per M).
int fft_plan_get(int64_t n_real, FftPlan *plan);
// Largest real length the single-workgroup LDS transform accepts.
int64_t fft_lds_max_real();

int launch_rfft(const double *in, int64_t n_in, double2 *out, const FftPlan &p, hipStream_t s);
int launch_irfft(const double2 *in, int64_t m_in, double *out, int64_t n_out, const FftPlan &p,
                 hipStream_t s);

// One-turn induced voltage (impedance.py:457-489 + 113-131):
// profile -> rfft(n_fft) -> * Z -> irfft -> -(scale * x)[:n_out] -> V.
// `src` is the profile read; when profile_out != nullptr it is also copied
// there (the engine's staged profile becomes the caller's).  When
// vk != nullptr the linear_interp_kick table of V (charge, bin_centers) is
// built as well (linear_interp_kick.cpp:33-48, acceleration kick 0).
struct InducedArgs {
    const double *src;            // profile (nullptr: read the parity counts below)
    // engine path: t = *turn_base + turn_off; the profile is counts[t & 1]
    // and counts[(t + 1) & 1] are cleared for the next particle pass
    const int *turn_base;
    int turn_off;
    unsigned *counts0, *counts1;  // each [groups][n_slices]; the profile is the sum
    int groups;
    double *profile_out;
    int n_slices;
    const double2 *Z;
    double scale;     // charge * e * ratio (sign applied inside)
    double *volt;     // out, n_out values
    int n_out;        // values of V written (n_slices <= n_out <= n_fft)
    const double *bc; // bin centres (table only)
    double charge;
    // table outputs (optional): tab[i] = (slope, intercept) of bin i, and
    // tab_params = (bin_centers[0], inv_bw)
    double2 *tab;
    double *tab_params;
    int f32;          // table in the reference's single-precision arithmetic
    // four-step method: the profile as the sum of part_rows u16-pair rows
    // stored transposed (column m2's word m1 of row r at
    // part[(m2 part_rows + r) 2^part_lshift + m1]); nullptr: the counts
    const unsigned *part;
    int part_rows, part_lshift;
    // windowed rows: they hold column words part_m1_lo .. part_m1_lo +
    // 2^part_lshift - 1; the other bins' counts are in cnt_out ([n_slices],
    // added and cleared here; nullptr: the rows cover every bin)
    int part_m1_lo;
    unsigned *cnt_out;
    // four-step method: Z in pass B's order, Zp[r M2 + k2] = Z[k] and
    // Zp[M + r M2 + k2] = Z[M - k], k = r + M1 k2 (nullptr: read Z)
    const double2 *Zp;
    const double *gl; // direct method: g[(k - (ns-1)) mod n_fft], g = -scale * irfft(Z)
    int64_t gl_len;   // entries of gl (at least n_out + n_slices - 1)
    // direct method: one extra workgroup runs this turn's beam phase loop
    // (feedback.h) beside the convolution
    int with_fb;
    FeedbackArgs fb;
    // direct method: the wake step of the turn in the same launch (the
    // multi-turn memory shift, the inductive term and the table of the
    // total, from this launch's V); volt and tab above are then the total's
    int with_wake;
    WakeArgs wk;
};
// Single workgroup: LDS Stockham rfft, impedance product, irfft.
int launch_induced_voltage_lds(const InducedArgs &a, const FftPlan &p, hipStream_t s,
                               const KernelEvents &ev = KernelEvents());
// One workgroup per bunch (d_args: device array of n_bunches argument sets,
// each with one copy of counts in counts0, read and cleared in place; all
// with this plan's n_fft and n_slices).
int launch_induced_lds_batch(const InducedArgs *d_args, int n_bunches, int n_slices,
                             const FftPlan &p, hipStream_t s,
                             const KernelEvents &ev = KernelEvents());
// All CUs: V[i] = sum_j p[j] g[(i - j) mod n_fft], the same linear map as the
// FFT route (Z[0] = 0 and the c2r Nyquist convention included in g).  Only
// n_out == n_slices.
int launch_induced_direct(const InducedArgs &a, hipStream_t s,
                          const KernelEvents &ev = KernelEvents());

int launch_fast_resonator(double *re, double *im, const double *freq, const double *R,
                          const double *Q, const double *fr, int n_res, int n_f, hipStream_t s);

// Threads of the four-step transform's column passes.
constexpr int kColThreads = 512;

// Multi-workgroup real-FFT convolution for n_fft beyond one workgroup's LDS:
// complex length M = n_fft/2 = M1 * M2 as a four-step transform in three
// launches -- (A) M1-point column transforms with twiddles, (B) per pair of
// rows (k1, M1-k1): M2-point transforms, real split, impedance product, real
// merge, inverse M2-point transforms and twiddles, (C) inverse M1-point
// column transforms writing V.  The plan owns its twiddles and the two
// M-complex scratch arrays (one plan per user: launches on one stream).
struct Fft2Plan {
    int M = 0, M1 = 0, M2 = 0, n_real = 0;
    int cols = 1;                 // columns per workgroup in passes A and C
    int threads_b = 0;            // pass B workgroup size (passes A, C: 256)
    int n1 = 0, n2 = 0;           // stage counts of the M1- and M2-point transforms
    int r1[kMaxStages], r2[kMaxStages];
    double2 *twM = nullptr;       // exp(-2 pi i j / M), j < M (all sub-transforms)
    double2 *twN = nullptr;       // exp(-2 pi i k / 2M), k <= M (real split / merge)
    // the same twiddles in the order the passes read them, so that every
    // lane's load is contiguous: twA[m2 M1 + k1] = twM[m2 k1] (pass A's
    // output twiddle), twB[3][r M2 + k2] = twN[k], twN[M - k], twM[r k2]
    // with k = r + M1 k2 (pass B's split/merge and output twiddles)
    double2 *twA = nullptr, *twB = nullptr;
    double2 *T = nullptr, *S = nullptr;  // M-complex intermediates
};
int fft2_plan_create(int64_t n_real, Fft2Plan *plan);
void fft2_plan_destroy(Fft2Plan *plan);
// Z[n_fft/2 + 1] (host) in pass B's order (InducedArgs::Zp), 2 M values.
void fft2_permute_z(const Fft2Plan &p, const double2 *Z_host, double2 *Zp_host);
// Same contract as launch_induced_voltage_lds (InducedArgs, including the
// interpolation table).  `ev`, when given, spans the three
// launches (start at pass A, stop at pass C).
int launch_induced_fourstep(const InducedArgs &a, const Fft2Plan &p, hipStream_t s,
                            const KernelEvents &ev = KernelEvents());

}  // namespace blmi
