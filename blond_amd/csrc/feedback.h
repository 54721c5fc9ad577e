// Beam phase loop on the device: the beam phase of a profile
// (beam_phase.cpp:6-65) and the per-turn loop of BeamFeedback
// (blond/llrf/beam_feedback.py:236-251, 317-355, 414-465, 514-622) with the
// RF phase accumulation of RingAndRFTracker.track (tracker.py:551-565).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "blmi_common.h"

namespace blmi {

// machines of the device loop
constexpr int kFbLHC = 0;
constexpr int kFbLHC_F = 1;
constexpr int kFbPSB = 2;

// per-turn program columns: [n_turns + 1][kFbCols]
constexpr int kFbGain = 0;     // PL_gain (PSB: its per-turn array)
constexpr int kFbGain2P = 1;   // SL / FL gain (PSB: RL_gain[0] per turn)
constexpr int kFbGain2I = 2;   // PSB: RL_gain[1] per turn
constexpr int kFbLhcA = 3;     // LHC synchro loop a(turn)
constexpr int kFbLhcT = 4;     // LHC synchro loop tau(turn)
constexpr int kFbPhiS = 5;     // design synchronous phase phi_s(turn)
constexpr int kFbAlpha0 = 6;   // ring alpha_0[0, turn] (PSB radial loop)
constexpr int kFbGamma = 7;    // gamma(turn) (PSB radial loop)
constexpr int kFbCols = 8;

// loop state (doubles), then dphi_rf[n_rf] (the tracker's accumulated phase)
constexpr int kFbLhcY = 0;
constexpr int kFbDphiSum = 1;
constexpr int kFbDphiAv = 2;
constexpr int kFbDphiAvPrev = 3;
constexpr int kFbDomegaPL = 4;
constexpr int kFbDomegaRL = 5;
constexpr int kFbDROverR = 6;
constexpr int kFbDROverRPrev = 7;
constexpr int kFbPLCounter = 8;
constexpr int kFbDomegaRf = 9;
constexpr int kFbPhiBeam = 10;
constexpr int kFbDphi = 11;
constexpr int kFbState = 16;

struct FeedbackArgs {
    const int *turn_base;
    int turn_off;
    int machine, delay;
    int first_bin, n_slices, n_rf;
    double alpha, bin_size, reference;
    const double *bc;             // [n_slices]
    // the turn's profile: counts[t & 1] while pending[t & 1] == t (the
    // plain map before the particle kernel publishes them), else `profile`
    const double *profile;
    const unsigned *counts0, *counts1;
    const int *pending;           // null: always `profile`
    int from_counts;              // 1: always counts[t & 1] (the loop runs beside the
                                  // launch that writes `profile`)
    int groups;
    double *rf;                   // [n_turns + 1][3][n_rf]: V, omega_rf, phi_rf (updated in place)
    const double *harmonic;       // [n_turns + 1][n_rf]
    const double *omega_d;        // [n_turns + 1][n_rf]
    const double *prog;           // [n_turns + 1][kFbCols]
    const double *on_time;        // [n_on_time] PSB loop turns
    int n_on_time;
    double *state;                // [kFbState + n_rf]
    // n_parts > 1: the trapezoid sums are split over that many workgroups;
    // each leaves its interior sums in parts[2 part], parts[2 part + 1] and
    // the last to arrive (arrive counts them, zero between launches) adds
    // them in part order and runs the control law
    int n_parts;
    double *parts;                // [2 kFbMaxParts]
    unsigned *arrive;
};
constexpr int kFbMaxParts = 16;
int launch_feedback(const FeedbackArgs &a, hipStream_t s);

// beam_phase.cpp:6-39 (alpha_mode 1) and 42-65 (beam_phase_fast, no window
// exponential): out[0] = scoeff / ccoeff over bins [0, n_bins).
int launch_beam_phase(const double *bc, const double *profile, int n_bins, double alpha,
                      double omega_rf, double phi_rf, double bin_size, bool fast, double *out,
                      hipStream_t s);
// out[0] = sum_i x[i * stride] / count  (count = ceil(n / stride)), the mean
// of numpy's x[::stride] (beam_feedback.py:375-378), fixed summation order.
int launch_strided_mean(const double *x, int64_t n, int64_t stride, double *out, hipStream_t s);

}  // namespace blmi
