// Differentiable tracking: forward turns with a smooth profile (the
// reference's smooth_histogram and linear_interp_kick, or a quadratic
// B-spline variant), turn checkpoints, and the hand-written reverse pass.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "blmi_common.h"
#include "particles.h"

namespace blmi {

// Largest slice count of the differentiable tracker (each workgroup of its
// induced-voltage kernels holds the profile, the wake and the adjoints).
constexpr int kAdjMaxSlices = 1024;
constexpr int kAdjCopies = 16;  // copies of the global deposit / adjoint accumulators

struct Adjoint {
    int device = 0;
    int64_t n = 0;
    int n_rf = 0, n_turns = 0, n_slices = 0, every = 50;
    int64_t n_fft = 0;
    double charge = 1.0, scale = 0.0;
    double cut_left = 0, inv_bw_cut = 0;  // deposit: bin = (dt - cut_left) * inv_bw_cut - 1/2
    double lo = 0, hi = 0;                // smooth_histogram's particle range
    int map = 0;                          // 0: spline variant, 1: the reference's functions
    double *d_bc = nullptr;               // [n_slices] bin centres (reference map)
    double bc0 = 0, inv_bw_bc = 0;        // linear_interp_kick's table origin and scale
    int fixed_bits = 40;                  // deposit fixed point 2^-S (exact, order-free sums)
    // programs
    double *d_rf = nullptr, *d_acc = nullptr;
    DriftCoeffs<double> *d_drift = nullptr;
    // coordinates: initial (owned copy), working, checkpoints, recompute slots
    double *d_x0 = nullptr;    // [2][n]: dt, dE at turn 0
    double *d_work = nullptr;  // [2][n]
    double *d_ckpt = nullptr;  // [n_ckpt][2][n]
    double *d_slots = nullptr; // [every][2][n]
    double *d_lam = nullptr;   // [2][n]: adjoints of dt, dE
    // slice axis
    unsigned long long *d_dep = nullptr;  // [2 parity][copies][n_slices] fixed-point deposit
    double *d_prof = nullptr;  // [n_turns + 1][n_slices] profiles p_t
    double *d_vt = nullptr;    // [n_turns + 1][n_slices] induced voltage V_t
    double *d_g = nullptr;     // wake g = irfft(Z) [n_fft]
    double *d_lv = nullptr;    // [2 parity][copies][n_slices] adjoint of V_t (accumulated)
    double *d_lp = nullptr;    // [n_slices] adjoint of the profile
    double *d_corr = nullptr;  // [n_fft] sum_t sum_i lambda_V,t[i] p_t[(i - m) mod n_fft]
    double *d_gpart = nullptr; // [n_turns][grid][2 n_rf] per-workgroup RF gradient sums
    int grid = 0, block = 256;
    bool forward_done = false;
    int forward_turns = 0;
    // the forward and reverse passes are captured once per turn count and
    // replayed as HIP graphs on the adjoint's own stream
    hipStream_t own = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    hipGraphExec_t fwd_exec = nullptr, bwd_exec = nullptr;
    int fwd_graph_turns = -1, bwd_graph_turns = -1;

    ~Adjoint();
    void release();
    int init(int64_t n, int n_rf, int n_turns, int n_slices, int every);
    int set_particles(const double *dt, const double *dE);
    int set_programs(const double *rf_host, const double *acc_host, const double *t_rev,
                     double length_ratio, double alpha_order, const double *eta0,
                     const double *eta1, const double *eta2, const double *alpha0,
                     const double *alpha1, const double *alpha2, const double *beta,
                     const double *energy, double charge);
    int set_slices(double cut_left, double cut_right);
    int set_map(int map, const double *bin_centers_host);
    int set_wake(int64_t n_fft, const double *g_host, double scale);
    void invalidate_graphs();
    int ensure_stream();
    int enqueue_forward(int turns, hipStream_t s);
    int enqueue_backward(int T, hipStream_t s);
    int forward(int turns, double *profile_out, hipStream_t s);
    int backward(const double *dprofile, double *grad_rf_host, double *corr_host, hipStream_t s);
};

}  // namespace blmi
