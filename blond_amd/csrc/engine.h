// Fused tracking-turn engine (device-resident per-turn tables, graph replay).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "particles.h"
#include "spectral.h"

namespace blmi {

struct Engine {
    int device = 0;
    // particle arrays (borrowed device pointers, 16-byte aligned)
    double *dt = nullptr, *dE = nullptr;
    int64_t n = 0;
    // programs: [turn][3][n_rf] (V, omega, phi), acc[turn], drift coefficients[turn]
    int n_rf = 0, n_turns = 0, solver = 0;
    double charge = 1.0;
    double *d_rf = nullptr, *d_acc = nullptr;
    DriftCoeffs<double> *d_drift = nullptr;
    int *d_turn = nullptr;
    int host_turn = 0;             // mirror of *d_turn, kept by set_turn/track
    unsigned *d_ticket = nullptr;
    // slicing
    int n_slices = 0;
    double cut_left = 0, cut_right = 0, inv_bw = 0;
    unsigned *d_counts = nullptr;  // 2 x n_slices staging counts
    double *profile = nullptr;     // borrowed: f64 profile seen by the caller
    // induced voltage
    bool intensity = false;
    FftPlan plan{};
    double2 *d_Z = nullptr;
    double scale = 0;
    double *volt = nullptr;        // borrowed: induced voltage seen by the caller
    const double *bc = nullptr;    // borrowed: bin centres
    double *d_table = nullptr;     // vk[n_slices], fac[n_slices], (bc0, inv_bw)
    // launch shape (tunable)
    int block_size = 256;
    int blocks_per_cu = 8;
    // graph replay
    int graph_chunk = 50;
    int graph_turns = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    hipStream_t own_stream = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;

    ~Engine();
    void release();
    int init(int64_t n_particles, int n_rf, int n_turns, int solver);
    int set_programs(const double *rf_host, const double *acc_host, const double *t_rev,
                     double length_ratio, double alpha_order, const double *eta0,
                     const double *eta1, const double *eta2, const double *alpha0,
                     const double *alpha1, const double *alpha2, const double *beta,
                     const double *energy);
    int set_slices(int n_slices, double cut_left, double cut_right, double *profile);
    int set_induced(int64_t n_fft, const double *Z_host, double scale, double *volt,
                    const double *bc, double charge);
    int set_turn(int turn);
    int get_turn(int *turn);
    int prime(hipStream_t s);
    int track(int turns, hipStream_t user, bool use_graph);
    int launch_one_turn(hipStream_t s);
    int launch_turn_kernel(hipStream_t s, bool interp);
    InducedArgs induced_args() const;
    int time_kernels(int turns, hipStream_t user, float *ms_induced, float *ms_turn);
    void invalidate_graph();
};

}  // namespace blmi
