/*
 * blond_mi355x.h -- C ABI of libblond_mi355x.so, the MI355X (gfx950) engine
 * for BLonD's per-turn macroparticle tracking path.
 *
 * Conventions
 *   - Every array argument is a DEVICE pointer (HBM), except where a name ends
 *     in _host.  Particle arrays are structure-of-arrays dt[n], dE[n].
 *   - Particle counts are int64 (the reference passes int32, butils_wrap_cpp.py:79-80).
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream).
 *     Calls are asynchronous with respect to the host unless noted.
 *   - Every function returns 0 (BLMI_OK) or a BLMI_ERR_* code; the message of
 *     the last failure on the calling thread is blmi_last_error().
 *   - The reference entry point each function replaces is cited as
 *     path:line of blond-admin/BLonD.  Argument order and meaning follow it;
 *     differences are listed per function.
 *   - real_t: functions without suffix take double (the reference's
 *     libblond_double.so); the _f32 variants take float (libblond_single.so,
 *     built with -DUSEFLOAT, blond/compile.py:167).
 */
#ifndef BLOND_MI355X_H
#define BLOND_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BLMI_OK 0
#define BLMI_ERR_ARG 1
#define BLMI_ERR_HIP 2
#define BLMI_ERR_UNSUPPORTED 3
#define BLMI_ERR_ALLOC 4

/* Message of the last failed call on this thread. */
const char *blmi_last_error(void);
int blmi_version(void);
/* Select the device for subsequent calls from this thread (hipSetDevice). */
int blmi_init(int device);
int blmi_device_info(int *cus, int *max_lds_per_workgroup);

/* RF kick: dE[i] += charge * sum_j voltage[j] sin(omega_rf[j] dt[i] + phi_rf[j]) + acc_kick.
 * Replaces kick(), blond/cpp_routines/kick.cpp:16-35 (bound by
 * blond/utils/butils_wrap_cpp.py:863-898).  Bit-identical to the reference's
 * strict build (same Cephes-method sine, same operation order). */
int blmi_kick(const double *dt, double *dE, int n_rf, double charge, const double *voltage,
              const double *omega_rf, const double *phi_rf, int64_t n, double acc_kick,
              void *stream);
int blmi_kick_f32(const float *dt, float *dE, int n_rf, float charge, const float *voltage,
                  const float *omega_rf, const float *phi_rf, int64_t n, float acc_kick,
                  void *stream);

/* Drift: dt[i] += f(dE[i]) with solver 0 simple, 1 legacy, 2 exact.
 * Replaces drift(), blond/cpp_routines/drift.cpp:19-78 (butils_wrap_cpp.py:901-943). */
int blmi_drift(double *dt, const double *dE, int solver, double t_rev, double length_ratio,
               double alpha_order, double eta_0, double eta_1, double eta_2, double alpha_0,
               double alpha_1, double alpha_2, double beta, double energy, int64_t n,
               void *stream);
int blmi_drift_f32(float *dt, const float *dE, int solver, float t_rev, float length_ratio,
                   float alpha_order, float eta_0, float eta_1, float eta_2, float alpha_0,
                   float alpha_1, float alpha_2, float beta, float energy, int64_t n,
                   void *stream);

/* Nearest-bin histogram of dt over [cut_left, cut_right) into profile[n_slices]
 * (overwritten; counts stored as real).  Replaces histogram(),
 * blond/cpp_routines/histogram.cpp:22-73 (slice_beam, butils_wrap_cpp.py:971-984).
 * Counts are identical to the reference for identical dt. */
int blmi_histogram(const double *dt, double *profile, double cut_left, double cut_right,
                   int n_slices, int64_t n, void *stream);
int blmi_histogram_f32(const float *dt, float *profile, float cut_left, float cut_right,
                       int n_slices, int64_t n, void *stream);

/* Per-bucket profiles of a filling pattern: profile[n_filled][n_slices_bucket]
 * (overwritten) from dt, with the filled buckets' cuts cut_left/right[n_filled]
 * and bunch_indexes[n_pattern] (the row of each RF bucket, -1 when empty; the
 * reference passes it as real_t).  Replaces sparse_histogram(),
 * blond/beam/sparse_histogram.cpp:22-87 (butils_wrap_cpp.py:1003-1023); the
 * extra n_pattern bounds the bucket lookup where the reference reads past its
 * array (a particle at the last right cut). */
int blmi_sparse_histogram(const double *dt, double *profile, const double *cut_left,
                          const double *cut_right, const double *bunch_indexes,
                          int n_slices_bucket, int n_filled_buckets, int n_pattern, int64_t n,
                          void *stream);
int blmi_sparse_histogram_f32(const float *dt, float *profile, const float *cut_left,
                              const float *cut_right, const float *bunch_indexes,
                              int n_slices_bucket, int n_filled_buckets, int n_pattern, int64_t n,
                              void *stream);

/* Two-bin "smooth" histogram with the reference's weight rule.  Replaces
 * smooth_histogram(), histogram.cpp:75-132 (slice_smooth, butils_wrap_cpp.py:987-1000).
 * Weights are summed in 64-bit fixed point: deterministic, within ~1e-13 of
 * the exact sum per particle. */
int blmi_smooth_histogram(const double *dt, double *profile, double cut_left, double cut_right,
                          int n_slices, int64_t n, void *stream);
int blmi_smooth_histogram_f32(const float *dt, float *profile, float cut_left, float cut_right,
                              int n_slices, int64_t n, void *stream);

/* Kick by linear interpolation of voltage[n_slices] on bin_centers.
 * Replaces linear_interp_kick(), linear_interp_kick.cpp:21-71
 * (butils_wrap_cpp.py:946-968).  dt is const here (the reference's pointer is
 * non-const but never written). */
int blmi_linear_interp_kick(const double *dt, double *dE, const double *voltage,
                            const double *bin_centers, double charge, int n_slices, int64_t n,
                            double acc_kick, void *stream);
int blmi_linear_interp_kick_f32(const float *dt, float *dE, const float *voltage,
                                const float *bin_centers, float charge, int n_slices, int64_t n,
                                float acc_kick, void *stream);
/* blmi_linear_interp_kick without any device allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * The slope/intercept table lives in a fixed per-device array of at most
 * 65536 slices (1 MiB per device, held until the process ends); larger
 * n_slices is refused with an error.  Every call on a device shares that
 * array, so calls that overlap on one device race.  Same arguments and
 * results as blmi_linear_interp_kick. */
int blmi_linear_interp_kick_noalloc(const double *dt, double *dE, const double *voltage,
                                    const double *bin_centers, double charge, int n_slices,
                                    int64_t n, double acc_kick, void *stream);

/* RF voltage on bin centres, accumulated into rf_voltage[n_bins].
 * Replaces rf_volt_comp(), kick.cpp:37-51 (butils_wrap_cpp.py:835-860). */
int blmi_rf_volt_comp(const double *voltage, const double *omega_rf, const double *phi_rf,
                      const double *bin_centers, int n_rf, int n_bins, double *rf_voltage,
                      void *stream);
int blmi_rf_volt_comp_f32(const float *voltage, const float *omega_rf, const float *phi_rf,
                          const float *bin_centers, int n_rf, int n_bins, float *rf_voltage,
                          void *stream);

/* Resonator impedance accumulated into (re, im)[n_freq] for freq[1:].
 * Replaces fast_resonator_real_imag(), fast_resonator.cpp:19-73
 * (fast_resonator, butils_wrap_cpp.py:1133-1171). */
int blmi_fast_resonator(double *re, double *im, const double *freq, const double *R_S,
                        const double *Q, const double *f_r, int n_res, int n_freq, void *stream);

/* numpy.fft.rfft(in[:n_in], n) -> out (n/2+1 complex, interleaved re/im) and
 * numpy.fft.irfft(in[:m_in], n) -> out[n].  These replace the numpy FFTs the
 * reference binds as bm.rfft/bm.irfft (blond/utils/bmath_backends.py:447-449).
 * n must be even and 5-smooth (2*2^a*3^b*5^c), at most blmi_fft_max_lds_length. */
int blmi_fft_max_lds_length(int64_t *n);
int blmi_rfft(const double *in, int64_t n_in, double *out, int64_t n, void *stream);
int blmi_irfft(const double *in, int64_t m_in, double *out, int64_t n, void *stream);

/* Bunch moments of the alive particles (id != 0; id may be NULL for all),
 * replacing Beam.statistics, blond/beam/beam.py:430-458 (numpy mean and std
 * with ddof = 0, two passes): out is a device double[5] = {n_alive,
 * mean_dt, sigma_dt, mean_dE, sigma_dE}.  Deterministic reduction order. */
int blmi_beam_statistics(const double *dt, const double *dE, const int64_t *id, int64_t n,
                         double *out, void *stream);
int blmi_beam_statistics_f32(const float *dt, const float *dE, const int64_t *id, int64_t n,
                             double *out, void *stream);

/* One-turn induced voltage: volt = -(scale * irfft(Z * rfft(profile, n_fft)))[:n_out]
 * with scale = charge * e * ratio and n_out = n_induced_voltage (or n_slices
 * for the cut of impedance.py:113-131).  Replaces
 * _InducedVoltage.induced_voltage_1turn, blond/impedances/impedance.py:457-489.
 * Z: n_fft/2+1 complex (interleaved), already divided by bin_size (impedance.py:890).
 * n_fft up to blmi_fft_max_lds_length runs on one workgroup; longer ones run
 * as a four-step transform whose scratch is kept per (device, n_fft, stream):
 * calls with the same n_fft on one stream are ordered, so they never race. */
int blmi_induced_voltage(const double *profile, int n_slices, const double *Z, int64_t n_fft,
                         double scale, double *volt, int n_out, void *stream);

/* numpy.interp(x, xp, fp, left, right) -> out[n] (xp increasing, m >= 1;
 * pass fp[0] / fp[m-1] for numpy's default left / right).  Replaces bm.interp
 * (blond/utils/bmath_backends.py:373 np.interp, 484 cp.interp) as called by
 * _InducedVoltage.shift_trev_time (blond/impedances/impedance.py:533-550) and
 * the "diff" profile derivative (blond/beam/profile.py:929-932). */
int blmi_interp(const double *x, int64_t n, const double *xp, const double *fp, int64_t m,
                double left, double right, double *out, void *stream);

/* numpy.gradient(f, h), uniform spacing, edge order 1 -> out[n] (n >= 2).
 * Replaces bm.gradient (bmath_backends.py:429) in
 * Profile.beam_profile_derivative("gradient") (profile.py:927-928). */
int blmi_gradient(const double *f, int64_t n, double h, double *out, void *stream);

/* numpy.diff(f) -> out[n-1].  Replaces bm.diff (bmath_backends.py:441) in
 * Profile.beam_profile_derivative("diff") (profile.py:929-932). */
int blmi_diff(const double *f, int64_t n, double *out, void *stream);

/* Multi-turn wake memory, time mode, in one pass:
 *   mem_out[i] = interp(time[i] + t_rev, time, mem_in, left=0, right=0)
 *                + (i < n_add ? add[i] : 0),  i < n_mem,
 * i.e. _InducedVoltage.shift_trev_time followed by the memory update of
 * induced_voltage_mtw (impedance.py:491-512, 533-550).  mem_in != mem_out. */
int blmi_mtw_shift_add(const double *mem_in, double *mem_out, const double *time, int64_t n_mem,
                       double t_rev, const double *add, int64_t n_add, void *stream);

/* Beam phase of a profile: *coeff = scoeff / ccoeff, the trapezoid sums of
 * exp(alpha bc) p sin(w bc + phi) and ... cos(...) (beam_phase) or of
 * p sin / p cos (beam_phase_fast); coeff is a device double.  Replaces
 * beam_phase() / beam_phase_fast(), blond/cpp_routines/beam_phase.cpp:6-65
 * (bm.beam_phase, butils_wrap_cpp.py:785-832).  Sines, cosines and the
 * exponential are the VDT (Cephes) functions the reference uses; the
 * interior sum runs in a fixed tree order. */
int blmi_beam_phase(const double *bin_centers, const double *profile, int n_bins, double alpha,
                    double omega_rf, double phi_rf, double bin_size, double *coeff, void *stream);
int blmi_beam_phase_fast(const double *bin_centers, const double *profile, int n_bins,
                         double omega_rf, double phi_rf, double bin_size, double *coeff,
                         void *stream);

/* *out = mean(x[::stride]) (device double), fixed summation order: the
 * radial loop's bm.mean(beam.dE[::sample_dE]), beam_feedback.py:375-378. */
int blmi_strided_mean(const double *x, int64_t n, int64_t stride, double *out, void *stream);

/* Fused turn engine.  Runs the reference's map
 *   with intensity effects:  [Profile, TotalInducedVoltage, RingAndRFTracker]
 *                            (EX_18_robinson_instability.py:226; tracker.py:521-615,
 *                             profile.py:700-719, impedance.py:113-146)
 *   without:                 [RingAndRFTracker, Profile]  (EX_01_Acceleration.py:152-153)
 * as at most two kernels per turn, optionally replayed from a captured HIP graph.
 * Per-turn programs are uploaded once; the turn counter lives on the device.
 * No reference counterpart as a single call: it replaces the per-object
 * .track() loop of the user's main script.
 *   rf_host: [n_turns+1][3][n_rf] doubles (voltage, omega_rf, phi_rf per turn);
 *   acc_host: [n_turns+1] acceleration kicks (-delta_E, tracker.py:331);
 *   t_rev ... energy: [n_turns+1] ring programs (drift index turn+1, tracker.py:606). */
void *blmi_engine_create(int64_t n, int n_rf, int n_turns, int solver);
int blmi_engine_destroy(void *engine);
int blmi_engine_set_particles(void *engine, double *dt, double *dE, int64_t n);
int blmi_engine_set_programs(void *engine, const double *rf_host, const double *acc_host,
                             const double *t_rev_host, double length_ratio, double alpha_order,
                             const double *eta0_host, const double *eta1_host,
                             const double *eta2_host, const double *alpha0_host,
                             const double *alpha1_host, const double *alpha2_host,
                             const double *beta_host, const double *energy_host, double charge);
int blmi_engine_set_slices(void *engine, int n_slices, double cut_left, double cut_right,
                           double *profile);
/* FP32 mode (the reference's libblond_single, utils/__init__.py:18-55): float
 * dt/dE (8-byte aligned), float programs (the double host programs cast to
 * float, as the single library receives them), float cuts; the profile,
 * induced voltage and interpolation table are formed in FP64 inside the
 * engine and written to the caller's float arrays after each track(). */
int blmi_engine_set_particles_f32(void *engine, float *dt, float *dE, int64_t n);
int blmi_engine_set_slices_f32(void *engine, int n_slices, double cut_left, double cut_right,
                               float *profile);
int blmi_engine_set_induced_f32(void *engine, int64_t n_fft, const double *Z_host, double scale,
                                float *volt, const float *bin_centers);
int blmi_engine_set_induced(void *engine, int64_t n_fft, const double *Z_host, double scale,
                            double *volt, const double *bin_centers);
/* Wake step of the fused turn, after the impedance convolution.
 *   Multi-turn wake (mtw_mode "time"; induced_voltage_mtw, impedance.py:491-512,
 *   shift_trev_time 533-550): n_mem > 0 points with the memory time axis
 *   `time` and the caller's memory buffer `memory` (device arrays of n_mem
 *   doubles, n_slices <= n_mem <= n_fft); the turn's voltage enters the
 *   memory at i < n_add (n_mem minus the front-wake buffer, impedance.py:506-508).
 *   The memory is read from `memory`
 *   at the start of every blmi_engine_track / time_kernels call and written
 *   back after every chunk; the shift uses the t_rev program of set_programs.
 *   Inductive source (InductiveImpedance.induced_voltage_1turn,
 *   impedance.py:986-1006): ind_coef_host[n_turns+1] holds, per turn,
 *   charge*e/(2 pi)*ratio*Z_over_n[t]*t_rev[t]/bin_size, and deriv_mode picks
 *   the profile derivative (0 "gradient", 1 "diff"; profile.py:905-937);
 *   NULL for none.
 * Call after blmi_engine_set_induced, which drops it; set_induced with
 * Z_host = NULL and n_fft = 0 means no convolution (inductive sources only).
 * FP64 particles only. */
int blmi_engine_set_wake(void *engine, int64_t n_mem, int64_t n_add, const double *time,
                         double *memory, const double *ind_coef_host, int deriv_mode);
/* Beam phase loop of the fused turn (BeamFeedback with RingAndRFTracker,
 * blond/llrf/beam_feedback.py:236-251, 275-367, 414-441, 514-622;
 * tracker.py:551-565): one workgroup per turn, after the induced voltage and
 * before the particle kernel, computes the beam phase of the turn's profile
 * over bins [first_bin, n_slices) (window coefficient alpha), the machine's
 * frequency correction (machine 0 LHC, 1 LHC_F, 2 PSB; -1 switches the loop
 * off), the next turn's omega_rf and the accumulated phase offset into the
 * next turn's phi_rf, in place in the engine's RF program.  The programs of
 * set_programs must then hold phi_rf without the accumulated offsets.
 *   bin_centers: device [n_slices];
 *   harmonic_host, omega_rf_d_host: [n_turns+1][n_rf];
 *   prog_host: [n_turns+1][8] = PL gain, gain2 (SL / FL / PSB RL P), PSB RL I,
 *              LHC a, LHC tau, phi_s, alpha_0, gamma per turn;
 *   on_time_host: PSB loop turns (beam_feedback.py:253-273);
 *   state_host: [16 + n_rf] = lhc_y, dphi_sum, dphi_av, dphi_av_prev,
 *              domega_PL, domega_RL, dR_over_R, dR_over_R_prev, PL_counter,
 *              domega_rf, phi_beam, dphi, 4 spare, then dphi_rf[n_rf].
 * FP64 particles only; turns the bin-ordered mode off.
 * blmi_engine_get_feedback copies the state back (synchronous);
 * blmi_engine_get_programs copies `turns` rows [3][n_rf] of the RF program
 * from turn0 (synchronous). */
int blmi_engine_set_feedback(void *engine, int machine, int delay, int first_bin, double alpha,
                             double reference, const double *bin_centers,
                             const double *harmonic_host, const double *omega_rf_d_host,
                             const double *prog_host, const double *on_time_host, int n_on_time,
                             const double *state_host);
int blmi_engine_get_feedback(void *engine, double *state_host);
int blmi_engine_get_programs(void *engine, int turn0, int turns, double *rf_host);
/* Bunch moments after every turn (Beam.statistics, beam.py:430-458, for the
 * whole bunch: the engine has no lost particles), accumulated inside the
 * particle kernel as sums of (x - shift) and (x - shift)^2 per workgroup in a
 * fixed order and reduced once per chunk: no extra pass over the particles.
 * blmi_engine_get_statistics copies rows [turn0, turn0 + turns) of
 * {n, mean_dt, sigma_dt, mean_dE, sigma_dE} (synchronous); the row of turn
 * t + 1 is written by turn t.  Shifts near the means keep the one-pass
 * variance accurate.  Turns the bin-ordered mode off. */
int blmi_engine_set_statistics(void *engine, int enable, double shift_dt, double shift_dE);
int blmi_engine_get_statistics(void *engine, int turn0, int turns, double *moments_host);
/* Launch shape of the particle kernel (block in {64..1024}, workgroups per CU)
 * and the number of turns per captured graph. */
int blmi_engine_set_launch(void *engine, int block, int blocks_per_cu, int graph_chunk);
/* Plain tracking (no intensity effects, statistics, phase loop,
 * interpolation table or sort, at most 65536/4 slices' u32 histogram): up to
 * max_turns turns (rounded down to odd) in one particle-kernel launch, the
 * particles kept in registers between them and only the last turn's
 * histogram formed -- the coordinates are those of one launch a turn, bit
 * for bit, and nothing reads them in between within a track() call.  1 = a
 * launch a turn; at most 63 turns are taken a launch.  Default 25. */
int blmi_engine_set_turn_block(void *engine, int max_turns);
/* Induced-voltage method: 0 = one-workgroup LDS FFT (rfft, x Z, irfft),
 * 1 = direct convolution over all CUs (n_slices <= 8192; with a multi-turn
 * memory it forms all n_mem values and runs the wake step in the same
 * launch), 2 = four-step FFT over many workgroups (any 5-smooth n_fft with a
 * factor 16).  set_induced picks 1, else 0 below 8192 points, else 2; a
 * multi-turn memory with n_mem * n_slices > 2^26 moves 1 to the FFT. */
int blmi_engine_set_method(void *engine, int method);
int blmi_engine_get_method(void *engine, int *method);
/* The tracker's interpolation=True (tracker.py:585-601: rf_voltage_calculation,
 * then linear_interp_kick of the RF voltage on the bin centres plus the
 * induced voltage, with the acceleration kick, in place of kick()).  Every
 * turn one launch forms the table from bin_centers (device, n_slices f64,
 * borrowed) and the turn's programs; the particle kernel kicks by the table
 * alone.  Needs set_slices first; FP64 particles only. */
int blmi_engine_set_interpolation(void *engine, const double *bin_centers);
/* Histogram: `groups` copies of the global counts (1..16, or 0 = automatic:
 * 4 up to 1024 slices, else 1; workgroup b adds into copy b % groups, the
 * consumer sums them) and `max_peel` wavefront-aggregation
 * steps (0..64) before plain LDS atomics.  Changing `groups` drops the staged
 * counts of the current turn (prime again). */
int blmi_engine_set_hist(void *engine, int groups, int max_peel);
/* Bin-ordered working set for slice counts above 4096 (FP64): the engine
 * tracks copies of the particles grouped by coarse bin (a counting sort
 * every `sort_every` turns; 0 = off, the default) so that each workgroup's
 * histogram and interpolation table fit a 4096-bin LDS window; the caller's
 * arrays are copied in and written back in their own order by every
 * track().  Results are bit-identical to the unordered engine.  Off by
 * default: at 1e7 particles and 65536 slices it measured slower (the sort
 * and the window kernel cost more than the gathers they save). */
int blmi_engine_set_order(void *engine, int sort_every);
int blmi_engine_set_turn(void *engine, int turn);   /* synchronous */
int blmi_engine_get_turn(void *engine, int *turn);  /* synchronous */
int blmi_engine_prime(void *engine, void *stream);
int blmi_engine_track(void *engine, int turns, int use_graph, void *stream);
/* Track `turns` eager turns with HIP events around each kernel on the engine's
 * stream; returns the average device time (ms) of the induced-voltage kernel
 * and of the particle kernel.  Synchronous. */
int blmi_engine_time_kernels(void *engine, int turns, float *ms_induced, float *ms_turn,
                             void *stream);

/* ---- batched ensemble ------------------------------------------------------
 * Many independent bunches (own particle count, RF program, cuts, profile,
 * impedance and induced voltage; one ring) tracked together: per turn one
 * launch computes every bunch's induced voltage (one workgroup each, LDS
 * FFT) and one launch streams all particles, split evenly over the grid;
 * chunks of turns replay as a captured graph.  This replaces running the
 * reference's per-bunch map (EX_18_robinson_instability.py:226) once per
 * bunch; each bunch's arithmetic is that of blmi_engine_* with method 0.
 * Order: create, set_particles, set_programs, set_slices, set_induced
 * (optional: without it the map is tracker then profile), then
 * set_turn/prime/track.
 *   offsets[b], counts[b]: bunch b occupies dt/dE[offsets[b] .. +counts[b]);
 *     offsets even, bunches not overlapping (an odd bunch owns one pad slot).
 *   rf_host: [bunch][n_turns+1][3][n_rf] (V, omega_rf, phi_rf); acc and the
 *     ring programs are per turn as for the engine.
 *   cut_left/cut_right: [bunch] (host); profiles, volt, bin_centers:
 *     device [bunch][n_slices]; Z_host [bunch][n_fft/2+1] complex; scale
 *     [bunch] = charge * e * ratio_b. */
void *blmi_ensemble_create(int n_bunches, int n_rf, int n_turns, int solver);
int blmi_ensemble_destroy(void *ensemble);
int blmi_ensemble_set_particles(void *ensemble, double *dt, double *dE, const int64_t *offsets,
                                const int64_t *counts);
int blmi_ensemble_set_programs(void *ensemble, const double *rf_host, const double *acc_host,
                               const double *t_rev_host, double length_ratio, double alpha_order,
                               const double *eta0_host, const double *eta1_host,
                               const double *eta2_host, const double *alpha0_host,
                               const double *alpha1_host, const double *alpha2_host,
                               const double *beta_host, const double *energy_host, double charge);
int blmi_ensemble_set_slices(void *ensemble, int n_slices, const double *cut_left,
                             const double *cut_right, double *profiles);
int blmi_ensemble_set_induced(void *ensemble, int64_t n_fft, const double *Z_host,
                              const double *scale, double *volt, const double *bin_centers);
/* Launch shape: workgroup size (multiple of 64 up to 1024), workgroups per
 * CU, turns per captured graph, wavefront-aggregation steps of the histogram. */
int blmi_ensemble_set_launch(void *ensemble, int block, int wg_per_cu, int graph_chunk,
                             int max_peel);
int blmi_ensemble_get_launch(void *ensemble, int *block, int *wg_per_cu);
int blmi_ensemble_set_turn(void *ensemble, int turn);
int blmi_ensemble_prime(void *ensemble, void *stream);
int blmi_ensemble_track(void *ensemble, int turns, int use_graph, void *stream);
int blmi_ensemble_time_turns(void *ensemble, int turns, float *ms_induced, float *ms_turn,
                             void *stream);
/* Times `turns` eager turns for each launch shape (block 256/512/1024 x
 * workgroups per CU up to 2048 threads), keeps the fastest and restores the
 * coordinates, counts and turn it started from. */
int blmi_ensemble_autotune(void *ensemble, int turns, int *block, int *wg_per_cu,
                           float *ms_per_turn, void *stream);

/* ---- differentiable tracking (adjoint) -------------------------------------
 * The intensity map of the fused engine with a smooth profile -- profile,
 * induced voltage, induced kick, RF kick (kick.cpp:16-35), "simple" drift
 * (drift.cpp:34-37) -- with V = -scale * (g circular-convolved with the
 * profile), g = irfft(Z, n_fft) (impedance.py:457-489).  The profile and the
 * induced kick are, by set_map: map 1, the reference's smooth_histogram
 * (histogram.cpp:75-132, its quirk kept) and linear_interp_kick
 * (linear_interp_kick.cpp:21-71) on bin_centers (host array of n_slices),
 * differentiated on their piecewise-linear pieces; map 0 (the default), one
 * quadratic B-spline on bin centres for both (continuous first derivative).
 * forward() tracks `turns` turns from the particles given to set_particles
 * (copied), keeping (dt, dE) every `checkpoint_every` turns (1: every state,
 * and backward() recomputes nothing; n_turns x 16 B per macroparticle of
 * HBM); backward() takes dL/d(final profile) and returns dL/dV_j, dL/dphi_j
 * summed over turns (grad_rf[2j], grad_rf[2j+1], host) and corr[m] =
 * sum_t sum_i dL/dV_t[i] p_t[(i - m) mod n_fft] (host), from which
 * dL/dtheta of any impedance parameter is
 * -scale * sum_m irfft(dZ/dtheta)[m] corr[m].  n_rf 1..4, 2..1024 slices,
 * n_fft <= 8192.  No reference counterpart: the reference has no adjoint. */
void *blmi_adjoint_create(int64_t n, int n_rf, int n_turns, int n_slices, int checkpoint_every);
int blmi_adjoint_destroy(void *adjoint);
int blmi_adjoint_set_particles(void *adjoint, const double *dt, const double *dE);
int blmi_adjoint_set_programs(void *adjoint, const double *rf_host, const double *acc_host,
                              const double *t_rev_host, double length_ratio, double alpha_order,
                              const double *eta0_host, const double *eta1_host,
                              const double *eta2_host, const double *alpha0_host,
                              const double *alpha1_host, const double *alpha2_host,
                              const double *beta_host, const double *energy_host, double charge);
int blmi_adjoint_set_slices(void *adjoint, double cut_left, double cut_right);
int blmi_adjoint_set_map(void *adjoint, int map, const double *bin_centers_host);
int blmi_adjoint_set_wake(void *adjoint, int64_t n_fft, const double *g_host, double scale);
int blmi_adjoint_forward(void *adjoint, int turns, double *profile_out, void *stream);
int blmi_adjoint_backward(void *adjoint, const double *dprofile, double *grad_rf_host,
                          double *corr_host, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* BLOND_MI355X_H */
