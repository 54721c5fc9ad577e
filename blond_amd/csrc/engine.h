// Fused tracking-turn engine (device-resident per-turn tables, graph replay).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "blmi_common.h"
#include "particles.h"
#include "spectral.h"
#include "wake.h"
#include "feedback.h"

namespace blmi {

// LDS budget of the turn kernel's private histogram: u32 per bin when it
// fits, else two u16 counters per u32 word (a workgroup then takes at most
// 65535 particles, so no counter can carry into its neighbour).
constexpr int kLdsHistBytes = 131072;    // histogram words in LDS at most
constexpr int kLdsHistU32Bytes = 65536;  // u32 counts up to this, u16 pairs above
constexpr int kMaxParticlesPerGroup16 = 65535;
// turns a plain-tracking launch takes at most, by default (odd)
constexpr int kTurnBlockDefault = 25;
// the particle kernels' sink: 2 pairs (dt, dE) of doubles for each lane of a wavefront
constexpr size_t kSinkBytes = 64 * 2 * 2 * sizeof(double);
// LDS the particle kernel may use for its histogram plus a staged copy of the
// interpolation table (random 16-byte gathers served by LDS instead of L2)
constexpr int kLdsTableBudget = 144 * 1024;
// LDS kept free of dynamic use beside the particle kernel's static arrays
constexpr size_t kStaticLdsMargin = 1024;
// MI355X's last-level (Infinity) cache: particle arrays above it stream from HBM
constexpr double kLastLevelCacheBytes = 256.0 * 1024 * 1024;
// Bin-ordered working set (slice counts above kSortWindow): bins of a
// workgroup's LDS window (u32 histogram + interpolation table) and the
// coarse-digit width of the counting sort (2^shift bins).
constexpr int kSortWindow = 4096;
constexpr int kSortDigitShift = 6;
// Largest slice count for which the induced voltage uses the direct
// (Toeplitz) convolution: ns^2 FP64 FMAs spread over all CUs.
constexpr int kDirectMaxSlices = 8192;
constexpr double kDirectMaxWork = 67108864.0;  // outputs x slices of the direct method

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
    int *d_base = nullptr;         // [0] turn base, [1..2] pending flags of the count buffers
    void *d_sink = nullptr;        // kSinkBytes: the particle kernels' store target for idle lanes
    int host_turn = 0;             // mirror of the turn base, kept by set_turn/track
    // slicing: counts double buffered by turn parity
    int n_slices = 0;
    double cut_left = 0, cut_right = 0, inv_bw = 0;
    int max_peel = 0;
    int groups = 1;                // copies of the counts (same-address atomics / groups)
    bool groups_auto = true;       // groups from the slice count (4 up to 1024 slices, else 1)
    unsigned *d_counts = nullptr;  // [2][groups][n_slices]
    unsigned *d_partials = nullptr;  // u16-pair mode: [grid][words] per-workgroup histograms
    size_t partials_bytes = 0;
    // u16-pair rows for the four-step transform above 2^15 slices: the LDS
    // histograms (and the rows) cover a window of half the columns' words,
    // rows win_m1 .. win_m1 + L/2 - 1 of the transform's columns (-1: no
    // window, set at prime() around the bunch); bins outside it are counted
    // in d_cnt_out with device atomics ([n_slices], cleared by pass A)
    int win_m1 = -1;
    unsigned *d_cnt_out = nullptr;
    double *profile = nullptr;     // borrowed: f64 profile seen by the caller
    // induced voltage
    bool intensity = false;
    int64_t n_fft = 0;
    // 0: LDS FFT (one workgroup), 1: direct convolution, 2: four-step FFT
    // over many workgroups (the particle kernel then interpolates V itself)
    int method = 0;
    bool have_fft = false, have_direct = false, have_fourstep = false;
    FftPlan plan{};
    Fft2Plan plan2{};
    double2 *d_Z = nullptr;
    double2 *d_Zp = nullptr;       // four-step method: Z in its second pass's order [2][M]
    double *d_gl = nullptr;        // direct method: -scale*irfft(Z), unrolled
    int64_t gl_len = 0;            // entries of d_gl
    double scale = 0;
    double *volt = nullptr;        // borrowed: induced voltage seen by the caller
    const double *bc = nullptr;    // borrowed: bin centres
    double *d_table = nullptr;     // (slope, intercept)[n_slices] as double2, then (bc0, inv_bw)
    double *d_trev = nullptr;      // [n_turns + 1] revolution period program
    // wake step after the convolution (multi-turn memory and/or an inductive
    // Z/n source): the convolution then writes d_v1 and the step writes the
    // caller's voltage and the table
    bool have_conv = true;         // false: no impedance convolution (inductive sources only)
    int64_t n_mem = 0;             // multi-turn memory length (0: none)
    int64_t n_add = 0;             // the turn's voltage enters the memory at i < n_add
    double *d_mem = nullptr;       // [2][n_mem], by turn parity
    double *d_time = nullptr;      // [n_mem] memory time axis
    double *mem_user = nullptr;    // borrowed: the caller's memory, loaded at track start and
                                   // published after every chunk
    double *d_v1 = nullptr;        // [max(n_mem, n_slices)] convolution output
    double *d_indc = nullptr;      // [n_turns + 1] inductive coefficients (null: none)
    int deriv_mode = kDerivGradient;
    // bunch moments after every turn, accumulated by the particle kernel
    // (per-workgroup rows) and reduced once per chunk
    bool stats_on = false;
    double stat_shift_t = 0, stat_shift_e = 0;
    int stats_grid = 0;
    double *d_stats = nullptr;      // [n_turns + 1][stats_grid][4]
    double *d_moments = nullptr;    // [n_turns + 1][5]
    // beam phase loop (feedback.h): one workgroup per turn after the induced
    // voltage, before the particle kernel; it updates the next turn's
    // omega_rf and phi_rf in d_rf
    bool fb_on = false;
    int fb_machine = 0, fb_delay = 0, fb_first_bin = 0, fb_n_on = 0;
    double fb_alpha = 0, fb_reference = 0;
    const double *fb_bc = nullptr;  // borrowed: bin centres
    double *d_fb = nullptr;         // harmonic, omega_rf_d [nt][n_rf]; prog [nt][kFbCols];
                                    // on_time [n_on]; state [kFbState + n_rf]
    double *fb_h = nullptr, *fb_wd = nullptr, *fb_prog = nullptr, *fb_on_time = nullptr,
           *fb_state = nullptr, *fb_parts = nullptr;
    unsigned *fb_arrive = nullptr;  // the split loop's arrival counter (zero between launches)
    // FP32 mode: float particles and programs; the slice axis (profile,
    // induced voltage, table) stays FP64 in engine-owned buffers and is
    // converted to the caller's float arrays at the end of every chunk
    bool f32 = false;
    float *dt32 = nullptr, *dE32 = nullptr;
    float *d_rf32 = nullptr, *d_acc32 = nullptr;
    DriftCoeffs<float> *d_drift32 = nullptr;
    float cut_left32 = 0.f, inv_bw32 = 0.f;
    float *profile32 = nullptr, *volt32 = nullptr;
    double *d_prof64 = nullptr, *d_volt64 = nullptr, *d_bc64 = nullptr;
    // bin-ordered working set: engine-owned double-buffered copies of the
    // particles grouped by coarse bin, re-grouped every sort_every turns
    int sort_every = 0;            // 0: off (the default: slower at 1e7 particles and 65536 slices)
    double *d_ord = nullptr;       // [2][2][n]: dt, dE per buffer
    unsigned *d_perm = nullptr;    // [2][n]: caller's index of each slot
    int *d_win = nullptr;          // [grid]: window base per workgroup
    unsigned *d_digcnt = nullptr;  // [ndig][grid] column offsets, then [ndig] digit starts
    int64_t ord_n = 0;
    int ord_grid = 0, ord_ndig = 0, ord_cur = 0, ord_since = 0;
    // interpolation=True (tracker.py:585-601): every turn one launch forms
    // the table from the RF voltage on the bin centres (plus the induced
    // voltage) with the acceleration kick in its intercepts; the particle
    // kernel then kicks by the table only (no RF systems, no acceleration)
    bool interp_rf = false;
    const double *interp_bc = nullptr;  // borrowed: bin centres
    double *d_zero_acc = nullptr;       // [n_turns + 1] zeros
    // launch shape (tunable)
    int block_size = 1024;
    int blocks_per_cu = 1;
    // plain tracking: at most this many turns per particle-kernel launch
    // (the particles stay in registers between them; 1 = a launch a turn)
    int turn_block = kTurnBlockDefault;
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
    int set_particles_f32(float *dt, float *dE);
    int set_slices_f32(int n_slices, double cut_left, double cut_right, float *profile);
    int set_induced_f32(int64_t n_fft, const double *Z_host, double scale, float *volt,
                        const float *bc, double charge);
    int set_method(int method);
    int set_interpolation(const double *bc);
    int launch_interp_table(hipStream_t s, int off);
    // one turn's launches in map order (all but the wake/feedback-free case
    // of the caller's events)
    int launch_turn_seq(hipStream_t s, int k, const KernelEvents &ei, const KernelEvents &et);
    int preferred_fft() const;
    int set_wake(int64_t n_mem, int64_t n_add, const double *time, double *memory,
                 const double *ind_coef_host, int deriv_mode);
    bool wake_on() const { return n_mem > 0 || d_indc != nullptr; }
    WakeArgs wake_args(int off) const;
    int set_feedback(int machine, int delay, int first_bin, double alpha, double reference,
                     const double *bc, const double *harmonic, const double *omega_d,
                     const double *prog, const double *on_time, int n_on_time,
                     const double *state);
    int get_feedback_state(double *host);
    int get_programs(int turn0, int turns, double *host);
    int set_statistics(bool on, double shift_t, double shift_e);
    int ensure_stats();
    int get_statistics(int turn0, int turns, double *host);
    int launch_feedback_turn(hipStream_t s, int off);
    FeedbackArgs feedback_args(int off) const;
    // the loop runs as one extra workgroup of the direct convolution's launch
    bool fb_in_induced() const { return fb_on && intensity && have_conv && method == 1 && !wake_on(); }
    int set_hist(int groups, int max_peel);
    int alloc_counts();
    int set_turn(int turn);
    int get_turn(int *turn);
    int prime(hipStream_t s);
    int track(int turns, hipStream_t user, bool use_graph);
    int launch_turns(hipStream_t s, int count, bool advance);
    int finish_chunk(hipStream_t s, int count);
    int launch_turn_kernel(hipStream_t s, int off, bool interp, const KernelEvents &ev);
    // whether plain turns may run several to a launch (nothing observes the
    // coordinates or the histogram between turns), and the turns of the
    // next launch of a chunk with `left` turns to go (odd, so the last turn's
    // counts use the buffer parity a launch a turn would use)
    bool multi_turn_ok() const;
    int multi_turn_piece(int left) const;
    int launch_turns_block(hipStream_t s, int off, int turns, const KernelEvents &ev);
    // histogram mode of the particle kernel (0 global atomics, 1 LDS u32,
    // 2 LDS u16 pairs) and its grid and LDS bytes for `block` threads
    int hist_layout(int block, int *grid, size_t *lds) const;
    bool ordered() const;
    int order_grid() const;
    int ensure_order_buffers();
    double *ord_x(int which) const;
    double *ord_e(int which) const;
    unsigned *ord_p(int which) const;
    int order_in(hipStream_t s);
    int order_sort(hipStream_t s);
    int order_out(hipStream_t s);
    int launch_window_turn(hipStream_t s, int off, bool interp, const KernelEvents &ev);
    int ensure_partials();
    // whether the u16-pair rows are stored transposed for the four-step
    // transform's first pass, which then sums them itself (no reduction
    // launch); shift = log2 M2 and lshift = log2 of the words per column
    bool part_perm(int *shift, int *lshift, int *m1_lo = nullptr) const;
    // whether the windowed rows apply (part_perm holds and the full rows
    // exceed 64 KiB of LDS), and the window's size in column words
    bool part_window_possible(int *lw) const;
    int choose_window(hipStream_t s);
    int launch_reduce(hipStream_t s, int off);
    int launch_induced(hipStream_t s, int off, const KernelEvents &ev);
    InducedArgs induced_args(int off) const;
    int time_kernels(int turns, hipStream_t user, float *ms_induced, float *ms_turn);
    void invalidate_graph();
};

}  // namespace blmi
