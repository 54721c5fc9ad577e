// Beam phase loop kernels (see feedback.h).
#include "feedback_device.h"

namespace blmi {

__global__ void __launch_bounds__(kFbBlock)
beam_phase_kernel(const double *__restrict__ bc, const double *__restrict__ profile, int n,
                  double alpha, double omega, double phi, double bin_size, bool fast,
                  double *__restrict__ out) {
    auto f = [&](int i, double &s, double &c) {
        phase_terms(bc[i], profile[i], alpha, omega, phi, fast, s, c);
    };
    const double r = trapz_ratio(f, n, bin_size);
    if (threadIdx.x == 0) out[0] = r;
}

int launch_beam_phase(const double *bc, const double *profile, int n_bins, double alpha,
                      double omega_rf, double phi_rf, double bin_size, bool fast, double *out,
                      hipStream_t s) {
    if (n_bins < 1) return blmi_set_error(BLMI_ERR_ARG, "beam_phase: no bins");
    hipLaunchKernelGGL(beam_phase_kernel, dim3(1), dim3(kFbBlock), 0, s, bc, profile, n_bins,
                       alpha, omega_rf, phi_rf, bin_size, fast, out);
    return blmi_check_launch("beam_phase");
}

__global__ void __launch_bounds__(kFbBlock)
strided_mean_kernel(const double *__restrict__ x, int64_t n, int64_t stride,
                    double *__restrict__ out) {
    const int64_t count = (n + stride - 1) / stride;
    double a = 0.0, b = 0.0;
    for (int64_t k = threadIdx.x; k < count; k += blockDim.x) a += x[k * stride];
    block_sum2(a, b);
    if (threadIdx.x == 0) out[0] = a / (double)count;
}

int launch_strided_mean(const double *x, int64_t n, int64_t stride, double *out, hipStream_t s) {
    if (n < 1 || stride < 1) return blmi_set_error(BLMI_ERR_ARG, "strided_mean: bad arguments");
    hipLaunchKernelGGL(strided_mean_kernel, dim3(1), dim3(kFbBlock), 0, s, x, n, stride, out);
    return blmi_check_launch("strided_mean");
}

// ---------------------------------------------------------------------------
// one turn of the loop
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kFbBlock) feedback_kernel(FeedbackArgs a) {
    feedback_turn_body(a);
}

int launch_feedback(const FeedbackArgs &a, hipStream_t s) {
    if (a.n_rf > kFbBlock || a.first_bin < 0 || a.first_bin >= a.n_slices)
        return blmi_set_error(BLMI_ERR_ARG, "feedback: bad bins or harmonics");
    hipLaunchKernelGGL(feedback_kernel, dim3(1), dim3(kFbBlock), 0, s, a);
    return blmi_check_launch("feedback");
}

}  // namespace blmi
