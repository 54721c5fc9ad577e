// FFT plans and slice-axis (spectral) launchers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace blmi {

constexpr int kMaxStages = 24;

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

// One-turn induced voltage (impedance.py:457-489 + 113-131) on one workgroup:
// profile -> rfft(n_fft) -> * Z -> irfft -> -(scale * x)[:n_slices] -> V.
// When counts != nullptr the profile is read from u32 counts, written to
// `profile` as real values, and the counts are cleared.  When vk != nullptr
// the linear_interp_kick table of V (charge, bin_centers) is built too.
struct InducedArgs {
    const int *turn;            // optional device turn counter selecting
    unsigned *counts0, *counts1;  // counts[turn & 1] (nullptr: read `profile`)
    double *profile;
    int n_slices;
    const double2 *Z;
    double scale;     // charge * e * ratio (sign applied inside)
    double *volt;     // out, n_out values
    int n_out;        // values of V written (n_slices <= n_out <= n_fft)
    const double *bc; // bin centres (table only)
    double charge;
    double *vk, *fac, *tab_params;  // table outputs (optional)
};
int launch_induced_voltage_lds(const InducedArgs &a, const FftPlan &p, hipStream_t s);

int launch_fast_resonator(double *re, double *im, const double *freq, const double *R,
                          const double *Q, const double *fr, int n_res, int n_f, hipStream_t s);

}  // namespace blmi
