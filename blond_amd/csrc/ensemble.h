// Batched ensemble: many independent bunches of uneven particle counts, each
// with its own RF settings, profile, impedance and induced voltage, tracked
// together -- two launches a turn for all bunches (one workgroup per bunch
// for the induced voltage, then one streaming pass over every particle),
// captured as a HIP graph per chunk of turns.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "blmi_common.h"
#include "particles.h"
#include "spectral.h"

namespace blmi {

// A workgroup's share of one bunch: particle pairs [p0, p1) of the shared
// coordinate arrays; particles at index >= end (the odd bunch's pad) are
// moved but not counted.
struct EnsSeg {
    int bunch;
    int pad;
    int64_t p0, p1, end;
};

struct Ensemble {
    int device = 0;
    int n_bunches = 0, n_rf = 0, n_turns = 0, solver = 0;
    double charge = 1.0;
    // coordinates: all bunches in one dt and one dE array (borrowed);
    // bunch b holds [off[b], off[b] + cnt[b]), off[b] even
    double *dt = nullptr, *dE = nullptr;
    std::vector<int64_t> off, cnt;
    // programs: rf [bunch][turn][3][n_rf]; acc[turn] and drift[turn] shared (one ring)
    double *d_rf = nullptr, *d_acc = nullptr;
    DriftCoeffs<double> *d_drift = nullptr;
    int *d_base = nullptr;
    int host_turn = 0;
    // slicing: per-bunch cuts, one u32 count array per bunch
    int n_slices = 0;
    double *d_cut_left = nullptr, *d_inv_bw = nullptr;
    double *profiles = nullptr;  // borrowed [bunch][n_slices]
    unsigned *d_counts = nullptr;
    // induced voltage
    bool intensity = false;
    int64_t n_fft = 0;
    FftPlan plan{};
    double2 *d_Z = nullptr;      // [bunch][n_fft/2+1]
    double *volt = nullptr;      // borrowed [bunch][n_slices]
    const double *bc = nullptr;  // borrowed [bunch][n_slices]
    double *d_table = nullptr;   // [bunch][2 n_slices + 2]: (slope, intercept)[n_slices], bc0, inv_bw
    InducedArgs *d_iargs = nullptr;
    // launch shape (autotuned)
    int block = 1024, wg_per_cu = 2, graph_chunk = 50;
    int max_peel = 0;
    int grid = 0;
    EnsSeg *d_segs = nullptr;
    int *d_seg_ptr = nullptr;
    // graph replay
    int graph_turns = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    hipStream_t own_stream = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;

    ~Ensemble();
    void release();
    int init(int n_bunches, int n_rf, int n_turns, int solver);
    int set_particles(double *dt, double *dE, const int64_t *off, const int64_t *cnt);
    int set_programs(const double *rf_host, const double *acc_host, const double *t_rev,
                     double length_ratio, double alpha_order, const double *eta0,
                     const double *eta1, const double *eta2, const double *alpha0,
                     const double *alpha1, const double *alpha2, const double *beta,
                     const double *energy, double charge);
    int set_slices(int n_slices, const double *cut_left, const double *cut_right,
                   double *profiles);
    int set_induced(int64_t n_fft, const double *Z_host, const double *scale, double *volt,
                    const double *bc);
    int set_launch(int block, int wg_per_cu, int graph_chunk, int max_peel);
    int set_turn(int turn);
    int prime(hipStream_t s);
    int track(int turns, hipStream_t user, bool use_graph);
    int time_turns(int turns, hipStream_t user, float *ms_induced, float *ms_turn);
    int autotune(int turns, hipStream_t user, int *best_block, int *best_wg_per_cu,
                 float *best_ms);
    int64_t total_particles() const;

    int build_segments();
    int launch_turns(hipStream_t s, int count, bool advance);
    int launch_turn(hipStream_t s, int off, const KernelEvents &ev);
    int launch_induced(hipStream_t s, const KernelEvents &ev);
    int launch_publish(hipStream_t s);
    void invalidate_graph();
};

}  // namespace blmi
