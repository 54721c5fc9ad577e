// Fused tracking turn.
//
// One turn of the reference's map is, with intensity effects
// (EX_18_robinson_instability.py:226: profile, total induced voltage, tracker):
//   profile(dt_t) -> V_t -> linear_interp_kick(V_t) -> RF kick(t) -> drift(t+1)
// and without them (EX_01_Acceleration.py:152-153: tracker, profile):
//   RF kick(t) -> drift(t+1) -> profile(dt_{t+1}).
//
// Here a turn is the induced-voltage launch(es) and one particle launch:
//   induced(t): counts of dt_t -> caller's profile, V_t (caller's induced
//               voltage) and the interpolation table (slope, intercept);
//               direct Toeplitz convolution over all CUs for moderate slice
//               counts, the one-workgroup LDS FFT, or the three-launch
//               four-step FFT for long transforms.
//   turn(t):    per particle: interp kick, RF kick, drift, bin of dt_{t+1}
//               counted into a workgroup-private LDS histogram (wavefront-
//               aggregated increments) and flushed to global counts with
//               relaxed device-scope atomics -- or, for slice counts whose
//               u32 histogram exceeds LDS, counted as u16 pairs and stored
//               as the workgroup's row of partial histograms, summed by one
//               more launch (plain loads and stores instead of millions of
//               scattered atomics).
// dt and dE are read once and written once per turn (32 B/particle in FP64).
//
// No kernel synchronises workgroups with each other (a device-scope
// release/acquire pair writes back and invalidates the per-XCD L2 caches);
// ordering comes from kernel boundaries only:
//   * the turn index is t = *base + k: k is baked into each launch (and into
//     the captured graph of a chunk of turns), *base is advanced by a
//     one-thread kernel after each chunk;
//   * counts are double buffered by turn parity: turn(t) accumulates into
//     counts[(t+1)&1]; induced(t+1) reads them and clears counts[t&1] for
//     turn(t+1); without intensity effects workgroup 0 of turn(t+1) publishes
//     counts[(t+1)&1] as the caller's profile (and clears them), and every
//     track() ends with a publish launch for its last turn.
// Every per-turn scalar lives in device tables indexed by t, so a captured
// graph replays unchanged for any turn range, and no memset node is recorded.
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "blmi_common.h"
#include "engine.h"
#include "particles.h"
#include "spectral.h"
#include "track_device.h"

namespace blmi {

// T: the particle precision (double, or float for the FP32 mode, whose
// slice-axis work -- profile, induced voltage, table -- stays in FP64).
template <typename T>
struct TurnArgs {
    int n_rf;
    T charge;
    const T *rf;                       // [turn][3][n_rf]: V, omega, phi
    const T *acc;                      // [turn]
    const DriftCoeffs<T> *drift;       // [turn]
    const int *base;                   // device turn base
    int off;                           // this launch's turn offset
    int last_turn;                     // n_turns - 1
    // interpolation kick (intensity effects): tab[k] = (slope, intercept)
    // of bin k, tab_params = (bin_centers[0], inv_bw)
    const double2 *tab;
    const double *tab_params;
    int n_tab;
    // slicing: counts[parity][group][bin]; workgroup b adds into group b % groups
    unsigned *counts0, *counts1;
    int groups;
    int n_slices;
    T cut_left, inv_bw;
    int max_peel;
    // plain mode: the previous turn's counts are published by all workgroups
    double *profile;
    int *pending;                      // [2]: turn whose histogram counts[p] hold, or -1
    // u16-pair histograms: workgroup b stores its LDS words at partials + b * part_words,
    // or (part_shift < 31) padded in LDS and stored transposed for the
    // four-step transform's first pass (store_partial_perm)
    unsigned *partials;
    int part_words;
    int part_shift, part_lshift;
    // windowed rows: LDS counts bins [win_lo, win_lo + win_bins), the others
    // go to cnt_out with device atomics (win_bins 0: no window)
    int win_lo, win_bins;
    unsigned *cnt_out;
    // interpolation table from L2: entries [tw_lo, tw_lo + tw_n) staged in
    // LDS after the histogram (tw_n 0: none)
    int tw_lo, tw_n;
    // bunch moments of the coordinates after the drift: workgroup b writes
    // sum(dt - shift_t), sum((dt - shift_t)^2), sum(dE - shift_e),
    // sum((dE - shift_e)^2) of its particles to stats[(t + 1) * grid + b]
    double *stats;
    double shift_t, shift_e;
    // store target of the lanes that hold no pair in a wavefront's last
    // iterations (every lane of a wavefront runs the same iteration count);
    // at least 64 pairs
    void *sink;
};

// Workgroup sums of the four moment accumulators in a fixed order (wavefront
// butterfly, then wavefronts in index order), written by thread 0.
__device__ __forceinline__ void stats_flush(double s[4], double *__restrict__ out) {
    __shared__ double wred[1024 / kWave][4];
    for (int k = 0; k < 4; ++k)
        for (int off = kWave / 2; off > 0; off >>= 1) s[k] += __shfl_xor(s[k], off);
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    if (lane == 0)
        for (int k = 0; k < 4; ++k) wred[w][k] = s[k];
    __syncthreads();
    if (threadIdx.x < 4) {
        double v = 0.0;
        for (int q = 0; q < (int)blockDim.x / kWave; ++q) v += wred[q][threadIdx.x];
        out[threadIdx.x] = v;
    }
}

// Moments of turns *base + 1 .. *base + gridDim.x from the workgroup rows:
// moments[t] = (n, mean_dt, sigma_dt, mean_dE, sigma_dE), sigma with ddof 0.
__global__ void moments_reduce_kernel(const double *__restrict__ part, int grid, const int *base,
                                      double n, double shift_t, double shift_e,
                                      double *__restrict__ moments) {
    __shared__ double s[4];
    const int t = *base + 1 + blockIdx.x;
    if (threadIdx.x < 4) {
        double v = 0.0;
        for (int g = 0; g < grid; ++g) v += part[((size_t)t * grid + g) * 4 + threadIdx.x];
        s[threadIdx.x] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const double mt = s[0] / n, me = s[2] / n;
        const double vt = s[1] / n - mt * mt, ve = s[3] / n - me * me;
        double *m = moments + (size_t)t * 5;
        m[0] = n;
        m[1] = shift_t + mt;
        m[2] = sqrt(vt > 0.0 ? vt : 0.0);
        m[3] = shift_e + me;
        m[4] = sqrt(ve > 0.0 ? ve : 0.0);
    }
}

// Copy the workgroup's LDS words to its row of partial histograms (16-byte
// stores; part_words is a multiple of 4).
__device__ __forceinline__ void store_partial(const unsigned *__restrict__ h,
                                              unsigned *__restrict__ row, int words) {
    for (int w = 4 * (int)threadIdx.x; w < words; w += 4 * (int)blockDim.x)
        *reinterpret_cast<uint4 *>(row + w) = make_uint4(h[w], h[w + 1], h[w + 2], h[w + 3]);
}

// counts[(t) & 1][bin] for t = *base + off: sum over `rows` partial
// histograms of u16 pairs (plain stores: every bin is written).  Thread
// (q, g) adds words 4q..4q+3 of rows g, g + 32, ... with 16-byte loads; the
// 32 row groups are then summed in LDS (512 threads: at c3's 256 rows x
// 32768 words the turn takes 108.7 against 110.2 us with 256).
constexpr int kRedWords = 64;
constexpr int kRedGroups = 32;
__global__ void __launch_bounds__(512) partials_reduce_kernel(const unsigned *__restrict__ P,
                                                              int rows, int W, const int *base,
                                                              int off, unsigned *counts0,
                                                              unsigned *counts1, int n_slices) {
    __shared__ unsigned red[kRedGroups][2 * kRedWords];
    const int t = *base + off;
    unsigned *out = (t & 1) ? counts1 : counts0;
    const int q = threadIdx.x % (kRedWords / 4), g = threadIdx.x / (kRedWords / 4);
    const int w = blockIdx.x * kRedWords + 4 * q;
    unsigned lo[4] = {0u, 0u, 0u, 0u}, hi[4] = {0u, 0u, 0u, 0u};
    if (w < W) {
#pragma unroll 8
        for (int r = g; r < rows; r += kRedGroups) {
            const uint4 v = *reinterpret_cast<const uint4 *>(P + (size_t)r * W + w);
            lo[0] += v.x & 0xffffu;
            hi[0] += v.x >> 16;
            lo[1] += v.y & 0xffffu;
            hi[1] += v.y >> 16;
            lo[2] += v.z & 0xffffu;
            hi[2] += v.z >> 16;
            lo[3] += v.w & 0xffffu;
            hi[3] += v.w >> 16;
        }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        red[g][2 * (4 * q + i)] = lo[i];
        red[g][2 * (4 * q + i) + 1] = hi[i];
    }
    __syncthreads();
    if (threadIdx.x < 2 * kRedWords) {
        unsigned s = 0;
        for (int k = 0; k < kRedGroups; ++k) s += red[k][threadIdx.x];
        const int bin = 2 * blockIdx.x * kRedWords + (int)threadIdx.x;
        if (bin < n_slices) out[bin] = s;
    }
}

// a pointer into global memory (address space 1): stores through it are
// global_store, which unlike flat stores leave the LDS wait counter alone
template <typename V>
using GlobalPtr = __attribute__((address_space(1))) V *;

// an interpolation-table entry as the particle kernel reads it: (slope,
// intercept) doubles, or floats for FP32 particles (store_interp_entry)
template <typename T> struct TabEntry { using type = double2; };
template <> struct TabEntry<float> { using type = float2; };

// NT: the coordinates are read and written with non-temporal loads and
// stores (for bunches larger than the last-level cache, where they stream
// from HBM either way and, kept out of the caches' recently-used lines,
// leave L2 to the interpolation table's gathers)
template <typename T, int SOLVER, int INTERP, bool HIST, int LDSH, bool PUBLISH, bool NT = false>
__global__ void __launch_bounds__(1024) turn_kernel(T *__restrict__ dt, T *__restrict__ dE,
                                                    int64_t n, TurnArgs<T> a) {
    using V2 = typename Vec2<T>::type;
    extern __shared__ unsigned lds_hist[];
    __shared__ __align__(16) double sin_tab[kCephesTabDoubles];
    const LdsConstDoublePtr stab = (LdsConstDoublePtr)sin_tab;
    cephes_tab_fill((LdsDoublePtr)sin_tab);
    int t = *a.base + a.off;
    if (t < 0) t = 0;
    if (t > a.last_turn) t = a.last_turn;  // host checks the range; never read past the tables
    const T *V = a.rf + (size_t)t * 3 * a.n_rf;
    const T *w = V + a.n_rf;
    const T *phi = w + a.n_rf;
    const T acc = a.acc[t];
    const DriftCoeffs<T> dc = a.drift[t + 1];
    unsigned *counts = (((t + 1) & 1) ? a.counts1 : a.counts0) +
                       (size_t)(blockIdx.x % a.groups) * a.n_slices;
    if (PUBLISH) {
        // counts[t&1] hold the histogram of dt_t when the previous turn ran in
        // this chunk: every workgroup publishes its share of the bins
        if (a.pending[t & 1] == t) {
            unsigned *prev = (t & 1) ? a.counts1 : a.counts0;
            const int per = (a.n_slices + (int)gridDim.x - 1) / (int)gridDim.x;
            const int b0 = (int)blockIdx.x * per;
            const int b1 = min(b0 + per, a.n_slices);
            publish_range(prev, a.profile, a.n_slices, a.groups, b0, b1);
        }
        if (blockIdx.x == 0 && threadIdx.x == 0) a.pending[(t + 1) & 1] = t + 1;
    }
    RfProg<T> rfp(V, w, phi, a.n_rf);
    T bc0 = T(0), ibw = T(0);
    if (INTERP) {
        bc0 = (T)a.tab_params[0];
        ibw = (T)a.tab_params[1];
    }
    // Software pipeline over a thread's pairs i_first, i_first + stride, ...
    // with two register slots used in turn (no register rotation, which
    // would force a wait on data still in flight).  A slot holds a pair and
    // its table entries; while one slot's pair is computed, the other slot's
    // pair is already loaded and its table loads are in flight.  Every wait
    // targets a load issued one pair's computation earlier: vector-memory
    // waits complete in issue order, so a load waited for too soon would
    // also wait for everything issued before it (the previous stores among
    // them).  For the compiler to place those waits exactly, the loop's
    // control is wavefront-uniform: every lane runs the iteration count of
    // the wavefront's first lane, loads past a lane's last pair reload pair
    // n2 - 1, and that lane's results go to the sink.  The first pair is
    // loaded before the LDS set-up below, so its latency overlaps the
    // clearing and the table staging.
    const int64_t i_first = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n2 = n / 2;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t i_wave = __builtin_amdgcn_readfirstlane(i_first);
    const int64_t iters = i_wave < n2 ? (n2 - 1 - i_wave) / stride + 1 : 0;
    auto clamp = [&](int64_t j) { return j < n2 ? j : n2 - 1; };
    V2 xa{}, ea{}, xb{}, eb{};
    if (iters > 0) {
        xa = reinterpret_cast<const V2 *>(dt)[clamp(i_first)];
        ea = reinterpret_cast<const V2 *>(dE)[clamp(i_first)];
    }
    // the histogram words (with windowed rows, the window's); the padded
    // layout adds one word per 2^part_shift (none when 31)
    const int hwords = (HIST && LDSH) ? ((LDSH == 2 && a.win_bins) ? a.win_bins / 2
                                                                  : hist_words(LDSH, a.n_slices))
                                      : 0;
    const int lwords = LDSH == 2 ? lds_word(hwords, a.part_shift) : hwords;
    // after them in LDS: INTERP == 2 the whole table; INTERP == 1 with u16
    // rows a window of it (entries tw_lo .. tw_lo + tw_n - 1)
    using TabE = typename TabEntry<T>::type;
    const TabE *gtab = reinterpret_cast<const TabE *>(a.tab);
    TabE *ltab = reinterpret_cast<TabE *>(lds_hist + ((lwords + 3) & ~3));
    constexpr bool kTabWindow = INTERP == 1 && HIST && LDSH == 2;
    const bool tw_on = kTabWindow && a.tw_n > 0;
    if ((HIST && LDSH) || INTERP == 2) {
        for (int b = threadIdx.x; b < lwords; b += blockDim.x) lds_hist[b] = 0u;
        if (INTERP == 2)
            for (int k = threadIdx.x; k < a.n_tab; k += blockDim.x) ltab[k] = gtab[k];
        if (tw_on)
            for (int k = threadIdx.x; k < a.tw_n; k += blockDim.x) ltab[k] = gtab[a.tw_lo + k];
    }
    __syncthreads();  // the LDS set-up, and the sine's coefficient table
    // the table window (lds_hist itself, never read, when there is none)
    const TabE *ltw = tw_on ? ltab : reinterpret_cast<const TabE *>(lds_hist);
    // linear_interp_kick.cpp:44-64 (acceleration kick 0): the table entry of
    // x, loaded unconditionally (bin 0 stands in outside the table; `in`
    // says whether the kick applies)
    struct Tab {
        TabE k;
        bool in;
    };
    // the per-particle constants, moved into vector registers like the RF
    // programs (scalar ones are spilled in this loop and read back with a
    // v_readlane at every use)
    T q_charge = a.charge, q_acc = acc, q_cl = a.cut_left, q_ibw = a.inv_bw;
    T q_ntab = (T)a.n_tab, q_ns = (T)a.n_slices, q_bc0 = bc0, q_tibw = ibw;
    int q_twlo = a.tw_lo, q_twn = a.tw_n, q_wlo = a.win_lo, q_wn = a.win_bins;
    uint64_t q_tabp = (uint64_t)a.tab, q_cntp = (uint64_t)a.cnt_out;
    auto pin_constants = [&]() {
        asm volatile("" : "+v"(q_charge), "+v"(q_acc), "+v"(q_cl), "+v"(q_ibw));
        if (INTERP) asm volatile("" : "+v"(q_ntab), "+v"(q_bc0), "+v"(q_tibw));
        if (HIST) asm volatile("" : "+v"(q_ns));
        if (kTabWindow) asm volatile("" : "+v"(q_twlo), "+v"(q_twn), "+v"(q_tabp));
        if (HIST && LDSH == 2) asm volatile("" : "+v"(q_wlo), "+v"(q_wn), "+v"(q_cntp));
    };
    auto gather = [&](T x) -> Tab {
        Tab g{TabE{}, false};
        if (INTERP) {
            const T f = floor((x - q_bc0) * q_tibw);
            g.in = f >= T(0) && f < q_ntab;
            const int fi = g.in ? (int)f : 0;
            if (kTabWindow) {
                // both loads unconditional: lanes served by the LDS window
                // all ask L2 for entry 0, one request for all of them
                const int kw = fi - q_twlo;
                const bool intw = (unsigned)kw < (unsigned)q_twn;
                using NE = T __attribute__((ext_vector_type(2)));
                const NE kv = reinterpret_cast<GlobalPtr<const NE>>(q_tabp)[intw ? 0 : fi];
                TabE kg;
                kg.x = kv.x;
                kg.y = kv.y;
                const TabE kl = ltw[intw ? kw : 0];
                g.k = intw ? kl : kg;
            } else {
                g.k = INTERP == 2 ? ltab[fi] : gtab[fi];
            }
        }
        return g;
    };
    // the rest of the turn of one particle given its table entry; the sums
    // keep the order e + interp + rf
    auto finish = [&](T &x, T &e, const Tab &g, auto nrf) -> int {
        constexpr int NRF = decltype(nrf)::value;
        const T rf = rfp.template kick_tab<NRF>(x, q_charge, q_acc, stab);
        if (INTERP) {
            // a select, not a branch: nothing waits on the gather before the sines
            const T ei = e + (x * (T)g.k.x + (T)g.k.y);
            e = g.in ? ei : e;
        }
        e = e + rf;
        x = x + e_drift(dc, e, SOLVER);
        if (HIST) {
            const T f = floor((x - q_cl) * q_ibw);
            return (f >= T(0) && f < q_ns) ? (int)f : -1;
        }
        return -1;
    };
    double st[4] = {0.0, 0.0, 0.0, 0.0};
    auto moment = [&](T x, T e) {
        const double u = (double)x - a.shift_t, v = (double)e - a.shift_e;
        st[0] += u;
        st[1] += u * u;
        st[2] += v;
        st[3] += v * v;
    };
    auto count = [&](int bin) {
        if (!HIST) return;
        if (LDSH == 2 && a.win_bins) {
            // windowed rows: bins outside the window with device atomics.
            // No count (-1) is never inside: with 0 <= win_lo < n_slices,
            // -1 - win_lo read unsigned is at least 2^32 - n_slices, above
            // any window's width
            const int wb = bin - q_wlo;
            const bool in = (unsigned)wb < (unsigned)q_wn;
            if (a.max_peel > 0)
                lds_count<true>(lds_hist, in ? wb : -1, a.max_peel, a.part_shift);
            else if (in)
                lds_add1<true>(lds_hist, wb, a.part_shift);
            if (bin >= 0 && !in)
                __hip_atomic_fetch_add(&reinterpret_cast<GlobalPtr<unsigned>>(q_cntp)[bin], 1u,
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else if (LDSH) {
            lds_count<LDSH == 2>(lds_hist, bin, a.max_peel, a.part_shift);
        } else if (bin >= 0) {
            atomic_add_agent(&counts[bin], 1u);
        }
    };
    // one copy of the particle loop per RF-system count (see with_rf_count)
    with_rf_count(a.n_rf, [&](auto nrf) {
    rfp.template pin<decltype(nrf)::value>();
    pin_constants();
    // one pair: RF kick, interpolated kick, drift, store, count; then the
    // other slot's table loads and this slot's reload two pairs ahead
    V2 *const sink = reinterpret_cast<V2 *>(a.sink) + 2 * (threadIdx.x % kWave);
    auto pair = [&](V2 &x, V2 &e, Tab &g0, Tab &g1, const V2 &xo, Tab &go0, Tab &go1,
                    int64_t i) {
        const bool own = i < n2;
        int b0 = finish(x.x, e.x, g0, nrf);
        int b1 = finish(x.y, e.y, g1, nrf);
        // the target is chosen as an integer and hidden from the optimiser,
        // which would otherwise split the store into two guarded stores:
        // with a divergent branch around them the waits above lose count
        uint64_t at = own ? (uint64_t)(reinterpret_cast<V2 *>(dt) + i) : (uint64_t)sink;
        uint64_t ae = own ? (uint64_t)(reinterpret_cast<V2 *>(dE) + i) : (uint64_t)(sink + 1);
        asm volatile("" : "+v"(at), "+v"(ae));
        using NV2 = T __attribute__((ext_vector_type(2)));
        if constexpr (NT) {
            __builtin_nontemporal_store(NV2{x.x, x.y}, reinterpret_cast<GlobalPtr<NV2>>(at));
            __builtin_nontemporal_store(NV2{e.x, e.y}, reinterpret_cast<GlobalPtr<NV2>>(ae));
        } else {
            *reinterpret_cast<GlobalPtr<NV2>>(at) = NV2{x.x, x.y};
            *reinterpret_cast<GlobalPtr<NV2>>(ae) = NV2{e.x, e.y};
        }
        if (own) {
            if (a.stats) {
                moment(x.x, e.x);
                moment(x.y, e.y);
            }
        } else {
            b0 = b1 = -1;
        }
        go0 = gather(xo.x);
        go1 = gather(xo.y);
        const int64_t j = clamp(i + 2 * stride);
        if constexpr (NT) {
            const NV2 xv =
                __builtin_nontemporal_load(reinterpret_cast<GlobalPtr<const NV2>>((uint64_t)dt) + j);
            const NV2 ev =
                __builtin_nontemporal_load(reinterpret_cast<GlobalPtr<const NV2>>((uint64_t)dE) + j);
            x = V2{xv.x, xv.y};
            e = V2{ev.x, ev.y};
        } else {
            x = reinterpret_cast<const V2 *>(dt)[j];
            e = reinterpret_cast<const V2 *>(dE)[j];
        }
        count(b0);
        count(b1);
    };
    if (iters > 0) {
        Tab ga0 = gather(xa.x), ga1 = gather(xa.y), gb0{}, gb1{};
        xb = reinterpret_cast<const V2 *>(dt)[clamp(i_first + stride)];
        eb = reinterpret_cast<const V2 *>(dE)[clamp(i_first + stride)];
        // both slots every iteration, no exit between them (a branch there
        // would again leave the waits conservative): with an odd count the
        // last iteration's second pair is past every lane's end and sunk
        for (int64_t k = 0; k < iters; k += 2) {
            pair(xa, ea, ga0, ga1, xb, gb0, gb1, i_first + k * stride);
            pair(xb, eb, gb0, gb1, xa, ga0, ga1, i_first + (k + 1) * stride);
        }
    }
    if ((n & 1) && blockIdx.x == gridDim.x - 1 && threadIdx.x < 64) {
        // odd tail: the first wavefront of the last block takes part in the
        // aggregated count, lane 0 owns the particle
        int b = -1;
        if (threadIdx.x == 0) {
            T x = dt[n - 1], e = dE[n - 1];
            const Tab g = gather(x);
            b = finish(x, e, g, nrf);
            dt[n - 1] = x;
            dE[n - 1] = e;
            if (a.stats) moment(x, e);
        }
        count(b);
    }
    });
    if (a.stats) stats_flush(st, a.stats + ((size_t)(t + 1) * gridDim.x + blockIdx.x) * 4);
    if (HIST && LDSH) {
        __syncthreads();
        if (LDSH == 2 && a.part_shift < 31)
            store_partial_perm(lds_hist, a.partials, blockIdx.x, gridDim.x, a.part_shift,
                               a.part_lshift);
        else if (LDSH == 2)
            store_partial(lds_hist, a.partials + (size_t)blockIdx.x * a.part_words, a.part_words);
        else
            hist_flush<LDSH>(lds_hist, counts, a.n_slices);
    }
}

// LDS bytes of one turn's programs in turns_kernel (at most two RF systems
// staged, plus the acceleration kick and the drift coefficients)
template <typename T>
constexpr size_t turn_prog_bytes() {
    return ((sizeof(T) * 7 + sizeof(DriftCoeffs<T>)) + 15) / 16 * 16;
}
constexpr int kTurnBlockMax = 63;  // turns a multi-turn launch stages in LDS at most

// Plain tracking, several turns a launch.  With no intensity effects,
// statistics, loop, interpolation table or sort, nothing observes a turn's
// coordinates or histogram before the last turn of a chunk, so each thread
// keeps its pair in registers across the launch's K turns and only the last
// turn is histogrammed: the particle data cross the memory system once per
// K turns instead of once per turn.  Every turn reads its own RF programs,
// acceleration kick and drift coefficients from the tables, and each
// particle's arithmetic is turn_kernel's (kick.cpp:27-33 then drift.cpp),
// in the same order, so the coordinates are bit-identical.  The counts of
// turn t0 + K go to buffer (t0 + K) & 1 = (t0 + 1) & 1 (K odd) and are
// published as turn_kernel's would be.
template <typename T, int SOLVER, int LDSH>
__global__ void __launch_bounds__(1024) turns_kernel(T *__restrict__ dt, T *__restrict__ dE,
                                                     int64_t n, TurnArgs<T> a, int K) {
    using V2 = typename Vec2<T>::type;
    extern __shared__ unsigned lds_hist[];
    int t0 = *a.base + a.off;
    if (t0 < 0) t0 = 0;
    if (t0 + K - 1 > a.last_turn) t0 = a.last_turn - K + 1;  // host checks the range
    const int tl = t0 + K;
    unsigned *counts = ((tl & 1) ? a.counts1 : a.counts0) +
                       (size_t)(blockIdx.x % a.groups) * a.n_slices;
    if (a.pending[t0 & 1] == t0) {
        unsigned *prev = (t0 & 1) ? a.counts1 : a.counts0;
        const int per = (a.n_slices + (int)gridDim.x - 1) / (int)gridDim.x;
        const int b0 = (int)blockIdx.x * per;
        const int b1 = min(b0 + per, a.n_slices);
        publish_range(prev, a.profile, a.n_slices, a.groups, b0, b1);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) a.pending[tl & 1] = tl;
    __shared__ __align__(16) double sin_tab[kCephesTabDoubles];
    const LdsConstDoublePtr stab = (LdsConstDoublePtr)sin_tab;
    cephes_tab_fill((LdsDoublePtr)sin_tab);
    if (LDSH)
        for (int b = threadIdx.x; b < a.n_slices; b += blockDim.x) lds_hist[b] = 0u;
    __syncthreads();
    const T q_charge = a.charge, q_cl = a.cut_left, q_ibw = a.inv_bw, q_ns = (T)a.n_slices;
    auto count = [&](int bin) {
        if (LDSH)
            lds_count<false>(lds_hist, bin, a.max_peel);
        else if (bin >= 0)
            atomic_add_agent(&counts[bin], 1u);
    };
    auto bin_of = [&](T x) -> int {
        const T f = floor((x - q_cl) * q_ibw);
        return (f >= T(0) && f < q_ns) ? (int)f : -1;
    };
    const int64_t n2 = n / 2;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    with_rf_count(a.n_rf, [&](auto nrf) {
        constexpr int NRF = decltype(nrf)::value;
        // a turn's programs: RF systems, acceleration kick, drift coefficients
        struct TurnProg {
            T V[NRF > 0 ? NRF : 1], w[NRF > 0 ? NRF : 1], phi[NRF > 0 ? NRF : 1];
            T acc;
            DriftCoeffs<T> dc;
        };
        static_assert(NRF < 0 || sizeof(TurnProg) <= turn_prog_bytes<T>(), "LDS program slot");
        // with a known RF-system count the launch's K turns of programs are
        // staged in LDS after the histogram words (one turn a thread): the
        // turn loop then reads them with broadcast LDS loads instead of a
        // global round trip at every turn
        TurnProg *progs = reinterpret_cast<TurnProg *>(
            reinterpret_cast<char *>(lds_hist) +
            ((LDSH ? sizeof(unsigned) * (size_t)a.n_slices : 0) + 15) / 16 * 16);
        if constexpr (NRF >= 0) {
            for (int k = threadIdx.x; k < K; k += blockDim.x) {
                const int t = t0 + k;
                const T *V = a.rf + (size_t)t * 3 * a.n_rf;
                TurnProg p;
#pragma unroll
                for (int j = 0; j < NRF; ++j) {
                    p.V[j] = V[j];
                    p.w[j] = V[a.n_rf + j];
                    p.phi[j] = V[2 * a.n_rf + j];
                }
                p.acc = a.acc[t];
                p.dc = a.drift[t + 1];
                progs[k] = p;
            }
            __syncthreads();
        }
        // turns t0 .. tl-1 of M particles in lock step: M independent chains
        // of dependent FP64 operations keep the vector pipeline busy; the sum
        // is RfProg::kick_n's, in the same order
        auto run = [&](T *x, T *e, auto m) {
            constexpr int M = decltype(m)::value;
            if constexpr (NRF >= 0) {
                for (int k = 0; k < K; ++k) {
                    const TurnProg p = progs[k];
#pragma unroll
                    for (int j = 0; j < M; ++j) {
                        T s = T(0);
#pragma unroll
                        for (int r = 0; r < NRF; ++r)
                            s = s + p.V[r] * TrigTab<T>::sin(p.w[r] * x[j] + p.phi[r], stab);
                        e[j] = e[j] + (q_charge * s + p.acc);
                        x[j] = x[j] + e_drift(p.dc, e[j], SOLVER);
                    }
                }
            } else {
                for (int t = t0; t < tl; ++t) {
                    const T *V = a.rf + (size_t)t * 3 * a.n_rf;
                    const RfProg<T> rfp(V, V + a.n_rf, V + 2 * a.n_rf, a.n_rf);
                    const T acc = a.acc[t];
                    const DriftCoeffs<T> dc = a.drift[t + 1];
#pragma unroll
                    for (int j = 0; j < M; ++j) {
                        e[j] = e[j] + rfp.kick(x[j], q_charge, acc);
                        x[j] = x[j] + e_drift(dc, e[j], SOLVER);
                    }
                }
            }
        };
        // two pairs a thread at a time, i and i + stride (past the end the
        // second slot repeats pair i, computed again and neither stored nor
        // counted)
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += 2 * stride) {
            const bool two = i + stride < n2;
            const int64_t j = two ? i + stride : i;
            const V2 xa = reinterpret_cast<const V2 *>(dt)[i], ea = reinterpret_cast<const V2 *>(dE)[i];
            const V2 xb = reinterpret_cast<const V2 *>(dt)[j], eb = reinterpret_cast<const V2 *>(dE)[j];
            T x[4] = {xa.x, xa.y, xb.x, xb.y}, e[4] = {ea.x, ea.y, eb.x, eb.y};
            run(x, e, std::integral_constant<int, 4>{});
            reinterpret_cast<V2 *>(dt)[i] = V2{x[0], x[1]};
            reinterpret_cast<V2 *>(dE)[i] = V2{e[0], e[1]};
            if (two) {
                reinterpret_cast<V2 *>(dt)[j] = V2{x[2], x[3]};
                reinterpret_cast<V2 *>(dE)[j] = V2{e[2], e[3]};
            }
            count(bin_of(x[0]));
            count(bin_of(x[1]));
            count(two ? bin_of(x[2]) : -1);
            count(two ? bin_of(x[3]) : -1);
        }
        if ((n & 1) && blockIdx.x == gridDim.x - 1 && threadIdx.x < 64) {
            // odd tail: lane 0 of the last block's first wavefront owns it
            int b = -1;
            if (threadIdx.x == 0) {
                T x[1] = {dt[n - 1]}, e[1] = {dE[n - 1]};
                run(x, e, std::integral_constant<int, 1>{});
                dt[n - 1] = x[0];
                dE[n - 1] = e[0];
                b = bin_of(x[0]);
            }
            count(b);
        }
    });
    if (LDSH) {
        __syncthreads();
        hist_flush<1>(lds_hist, counts, a.n_slices);
    }
}

// Plain mode, end of a track() call: publish the last turn's counts (one
// workgroup, so the pending flag is cleared after every bin is done).
__global__ void publish_kernel(const int *base, int off, unsigned *counts0, unsigned *counts1,
                               int groups, int *pending, double *profile, int n_slices) {
    const int t = *base + off;  // histogram of dt_t
    if (pending[t & 1] != t) return;
    publish_range((t & 1) ? counts1 : counts0, profile, n_slices, groups, 0, n_slices);
    __syncthreads();
    if (threadIdx.x == 0) pending[t & 1] = -1;
}

__global__ void to_float_kernel(const double *__restrict__ src, float *__restrict__ dst, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (float)src[i];
}

__global__ void to_double_kernel(const float *__restrict__ src, double *__restrict__ dst, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (double)src[i];
}

// *base += k, after every launch of the chunk has read it (stream order).
__global__ void advance_kernel(int *base, int k) { *base += k; }

// Histogram of the current coordinates (engine priming), in the particle
// kernel's layout: counts[t & 1] (cleared first by a fill kernel) for modes
// 0 and 1, the workgroups' partial rows for mode 2 (reduced afterwards).
template <typename T, int LDSH>
__global__ void __launch_bounds__(1024) prime_hist_kernel(const T *__restrict__ dt, int64_t n,
                                                          const int *base, unsigned *counts0,
                                                          unsigned *counts1, int n_slices,
                                                          T cut_left, T inv_bw,
                                                          unsigned *partials, int part_words,
                                                          int part_shift, int part_lshift,
                                                          int win_lo, int win_bins,
                                                          unsigned *cnt_out) {
    using V2 = typename Vec2<T>::type;
    extern __shared__ unsigned lds_hist[];
    const int t = *base;
    unsigned *counts = (t & 1) ? counts1 : counts0;
    if (LDSH) {
        const int hw = (LDSH == 2 && win_bins) ? win_bins / 2 : hist_words(LDSH, n_slices);
        const int words = LDSH == 2 ? lds_word(hw, part_shift) : hw;
        for (int b = threadIdx.x; b < words; b += blockDim.x) lds_hist[b] = 0u;
        __syncthreads();
    }
    auto bin_of = [&](T x) -> int {
        const T f = floor((x - cut_left) * inv_bw);
        return (f >= T(0) && f < (T)n_slices) ? (int)f : -1;
    };
    auto count = [&](int bin) {
        if (LDSH == 2 && win_bins) {
            // windowed rows: bins outside the window with device atomics
            const int wb = bin - win_lo;
            const bool in = bin >= 0 && (unsigned)wb < (unsigned)win_bins;
            lds_count<true>(lds_hist, in ? wb : -1, 0, part_shift);
            if (bin >= 0 && !in) atomic_add_agent(&cnt_out[bin], 1u);
        } else if (LDSH) {
            lds_count<LDSH == 2>(lds_hist, bin, 0, part_shift);
        } else if (bin >= 0) {
            atomic_add_agent(&counts[bin], 1u);
        }
    };
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t n2 = n / 2;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += stride) {
        const V2 x = reinterpret_cast<const V2 *>(dt)[i];
        count(bin_of(x.x));
        count(bin_of(x.y));
    }
    if ((n & 1) && blockIdx.x == gridDim.x - 1 && threadIdx.x < 64)
        count(threadIdx.x == 0 ? bin_of(dt[n - 1]) : -1);
    if (LDSH) {
        __syncthreads();
        if (LDSH == 2 && part_shift < 31)
            store_partial_perm(lds_hist, partials, blockIdx.x, gridDim.x, part_shift, part_lshift);
        else if (LDSH == 2)
            store_partial(lds_hist, partials + (size_t)blockIdx.x * part_words, part_words);
        else
            hist_flush<LDSH>(lds_hist, counts, n_slices);
    }
}

// ---------------------------------------------------------------------------
// Engine object
// ---------------------------------------------------------------------------
Engine::~Engine() { release(); }

void Engine::release() {
    invalidate_graph();
    (void)hipFree(d_rf);
    (void)hipFree(d_acc);
    (void)hipFree(d_drift);
    (void)hipFree(d_base);
    (void)hipFree(d_sink);
    d_sink = nullptr;
    (void)hipFree(d_counts);
    (void)hipFree(d_table);
    (void)hipFree(d_Z);
    (void)hipFree(d_Zp);
    d_Zp = nullptr;
    (void)hipFree(d_gl);
    (void)hipFree(d_partials);
    d_partials = nullptr;
    partials_bytes = 0;
    (void)hipFree(d_cnt_out);
    d_cnt_out = nullptr;
    win_m1 = -1;
    (void)hipFree(d_ord);
    (void)hipFree(d_perm);
    (void)hipFree(d_win);
    (void)hipFree(d_digcnt);
    d_ord = nullptr;
    d_perm = nullptr;
    d_win = nullptr;
    d_digcnt = nullptr;
    ord_n = 0;
    (void)hipFree(d_rf32);
    (void)hipFree(d_acc32);
    (void)hipFree(d_drift32);
    (void)hipFree(d_prof64);
    (void)hipFree(d_volt64);
    (void)hipFree(d_bc64);
    (void)hipFree(d_trev);
    (void)hipFree(d_mem);
    (void)hipFree(d_time);
    (void)hipFree(d_v1);
    (void)hipFree(d_indc);
    (void)hipFree(d_fb);
    d_fb = nullptr;
    fb_on = false;
    (void)hipFree(d_zero_acc);
    d_zero_acc = nullptr;
    interp_rf = false;
    (void)hipFree(d_stats);
    (void)hipFree(d_moments);
    d_stats = d_moments = nullptr;
    stats_on = false;
    stats_grid = 0;
    d_trev = d_mem = d_time = d_v1 = d_indc = nullptr;
    n_mem = 0;
    mem_user = nullptr;
    d_rf32 = d_acc32 = nullptr;
    d_drift32 = nullptr;
    d_prof64 = d_volt64 = d_bc64 = nullptr;
    fft2_plan_destroy(&plan2);
    d_rf = d_acc = d_table = d_gl = nullptr;
    d_drift = nullptr;
    d_base = nullptr;
    d_counts = nullptr;
    d_Z = nullptr;
    if (own_stream) (void)hipStreamDestroy(own_stream);
    own_stream = nullptr;
    if (ev_in) (void)hipEventDestroy(ev_in);
    if (ev_out) (void)hipEventDestroy(ev_out);
    ev_in = ev_out = nullptr;
}

int Engine::init(int64_t n_particles, int n_rf_, int n_turns_, int solver_) {
    n = n_particles;
    n_rf = n_rf_;
    n_turns = n_turns_;
    solver = solver_;
    BLMI_HIP_TRY(hipGetDevice(&device));
    BLMI_HIP_TRY(hipMalloc(&d_rf, sizeof(double) * 3 * n_rf * (size_t)(n_turns + 1)));
    BLMI_HIP_TRY(hipMalloc(&d_acc, sizeof(double) * (size_t)(n_turns + 1)));
    BLMI_HIP_TRY(hipMalloc(&d_drift, sizeof(DriftCoeffs<double>) * (size_t)(n_turns + 1)));
    BLMI_HIP_TRY(hipMalloc(&d_trev, sizeof(double) * (size_t)(n_turns + 1)));
    // d_base[0] = turn base, d_base[1..2] = turn held by each count buffer (-1: none)
    BLMI_HIP_TRY(hipMalloc(&d_base, sizeof(int) * 4));
    BLMI_HIP_TRY(hipMalloc(&d_sink, kSinkBytes));
    BLMI_HIP_TRY(hipMemset(d_base, 0, sizeof(int)));
    BLMI_HIP_TRY(hipMemset(d_base + 1, 0xff, sizeof(int) * 2));
    BLMI_HIP_TRY(hipStreamCreateWithFlags(&own_stream, hipStreamNonBlocking));
    BLMI_HIP_TRY(hipEventCreateWithFlags(&ev_in, hipEventDisableTiming));
    BLMI_HIP_TRY(hipEventCreateWithFlags(&ev_out, hipEventDisableTiming));
    return BLMI_OK;
}

int Engine::set_programs(const double *rf_host, const double *acc_host, const double *t_rev,
                         double length_ratio, double alpha_order, const double *eta0,
                         const double *eta1, const double *eta2, const double *alpha0,
                         const double *alpha1, const double *alpha2, const double *beta,
                         const double *energy) {
    const size_t nt = (size_t)n_turns + 1;
    BLMI_HIP_TRY(hipMemcpy(d_rf, rf_host, sizeof(double) * 3 * n_rf * nt, hipMemcpyHostToDevice));
    BLMI_HIP_TRY(hipMemcpy(d_acc, acc_host, sizeof(double) * nt, hipMemcpyHostToDevice));
    BLMI_HIP_TRY(hipMemcpy(d_trev, t_rev, sizeof(double) * nt, hipMemcpyHostToDevice));
    std::vector<DriftCoeffs<double>> dc(nt);
    for (size_t i = 0; i < nt; ++i)
        dc[i] = make_drift_coeffs<double>(solver, t_rev[i], length_ratio, alpha_order, eta0[i],
                                          eta1[i], eta2[i], alpha0[i], alpha1[i], alpha2[i],
                                          beta[i], energy[i]);
    BLMI_HIP_TRY(hipMemcpy(d_drift, dc.data(), sizeof(DriftCoeffs<double>) * nt,
                           hipMemcpyHostToDevice));
    // FP32 programs: the scalars as the single-precision library receives
    // them (cast to float), coefficients formed in float
    std::vector<float> rf32((size_t)3 * n_rf * nt), acc32(nt);
    for (size_t i = 0; i < rf32.size(); ++i) rf32[i] = (float)rf_host[i];
    for (size_t i = 0; i < nt; ++i) acc32[i] = (float)acc_host[i];
    std::vector<DriftCoeffs<float>> dc32(nt);
    for (size_t i = 0; i < nt; ++i)
        dc32[i] = make_drift_coeffs<float>(solver, (float)t_rev[i], (float)length_ratio,
                                           (float)alpha_order, (float)eta0[i], (float)eta1[i],
                                           (float)eta2[i], (float)alpha0[i], (float)alpha1[i],
                                           (float)alpha2[i], (float)beta[i], (float)energy[i]);
    if (!d_rf32) {
        BLMI_HIP_TRY(hipMalloc(&d_rf32, sizeof(float) * rf32.size()));
        BLMI_HIP_TRY(hipMalloc(&d_acc32, sizeof(float) * nt));
        BLMI_HIP_TRY(hipMalloc(&d_drift32, sizeof(DriftCoeffs<float>) * nt));
    }
    BLMI_HIP_TRY(hipMemcpy(d_rf32, rf32.data(), sizeof(float) * rf32.size(), hipMemcpyHostToDevice));
    BLMI_HIP_TRY(hipMemcpy(d_acc32, acc32.data(), sizeof(float) * nt, hipMemcpyHostToDevice));
    BLMI_HIP_TRY(hipMemcpy(d_drift32, dc32.data(), sizeof(DriftCoeffs<float>) * nt,
                           hipMemcpyHostToDevice));
    invalidate_graph();
    return BLMI_OK;
}

int Engine::set_particles_f32(float *x, float *e) {
    f32 = true;
    dt32 = x;
    dE32 = e;
    dt = dE = nullptr;
    invalidate_graph();
    return BLMI_OK;
}

int Engine::set_slices_f32(int ns, double cl, double cr, float *profile_user) {
    (void)hipFree(d_prof64);
    d_prof64 = nullptr;
    if (ns > 0) BLMI_HIP_TRY(hipMalloc(&d_prof64, sizeof(double) * ns));
    profile32 = profile_user;
    // histogram.cpp:29 in float: the cuts arrive as float
    cut_left32 = (float)cl;
    inv_bw32 = (float)ns / ((float)cr - (float)cl);
    return set_slices(ns, cl, cr, d_prof64);
}

int Engine::set_induced_f32(int64_t n_fft_, const double *Z_host, double scale_, float *volt_user,
                            const float *bc_user, double charge_) {
    (void)hipFree(d_volt64);
    (void)hipFree(d_bc64);
    d_volt64 = d_bc64 = nullptr;
    BLMI_HIP_TRY(hipMalloc(&d_volt64, sizeof(double) * n_slices));
    BLMI_HIP_TRY(hipMemset(d_volt64, 0, sizeof(double) * n_slices));
    BLMI_HIP_TRY(hipMalloc(&d_bc64, sizeof(double) * n_slices));
    hipLaunchKernelGGL(to_double_kernel, dim3((n_slices + 255) / 256), dim3(256), 0, nullptr,
                       bc_user, d_bc64, n_slices);
    int rc = blmi_check_launch("bin centres");
    if (rc) return rc;
    BLMI_HIP_TRY(hipDeviceSynchronize());
    volt32 = volt_user;
    return set_induced(n_fft_, Z_host, scale_, d_volt64, d_bc64, charge_);
}

int Engine::set_slices(int n_slices_, double cut_left_, double cut_right_, double *profile_) {
    n_slices = n_slices_;
    cut_left = cut_left_;
    cut_right = cut_right_;
    inv_bw = (double)n_slices / (cut_right - cut_left);  // histogram.cpp:29
    profile = profile_;
    // copies of the counts: few bins take many same-address atomics
    if (groups_auto) groups = n_slices <= 1024 ? 4 : 1;
    invalidate_graph();
    return alloc_counts();
}

int Engine::alloc_counts() {
    (void)hipFree(d_counts);
    d_counts = nullptr;
    if (n_slices > 0) {
        const size_t bytes = sizeof(unsigned) * 2 * (size_t)groups * (size_t)n_slices;
        BLMI_HIP_TRY(hipMalloc(&d_counts, bytes));
        BLMI_HIP_TRY(hipMemset(d_counts, 0, bytes));
    }
    BLMI_HIP_TRY(hipMemset(d_base + 1, 0xff, sizeof(int) * 2));
    return BLMI_OK;
}

int Engine::set_induced(int64_t n_fft_, const double *Z_host, double scale_, double *volt_,
                        const double *bc_, double charge_) {
    n_fft = n_fft_;
    scale = scale_;
    volt = volt_;
    bc = bc_;
    charge = charge_;
    const int64_t nz = n_fft / 2 + 1;
    (void)hipFree(d_Z);
    (void)hipFree(d_Zp);
    (void)hipFree(d_table);
    (void)hipFree(d_gl);
    d_Z = d_Zp = nullptr;
    d_table = d_gl = nullptr;
    gl_len = 0;
    have_fft = have_direct = have_fourstep = false;
    // a new source set drops the wake step (set_wake follows when needed)
    (void)hipFree(d_mem);
    (void)hipFree(d_time);
    (void)hipFree(d_v1);
    (void)hipFree(d_indc);
    d_mem = d_time = d_v1 = d_indc = nullptr;
    n_mem = 0;
    mem_user = nullptr;
    if (n_fft == 0 && !Z_host) {
        // no impedance convolution: the induced voltage comes from the wake
        // step alone (inductive sources, set_wake)
        BLMI_HIP_TRY(hipMalloc(&d_table, sizeof(double) * (2 * (size_t)n_slices + 2)));
        BLMI_HIP_TRY(hipMemset(d_table, 0, sizeof(double) * (2 * (size_t)n_slices + 2)));
        have_conv = false;
        method = -1;
        intensity = true;
        invalidate_graph();
        return BLMI_OK;
    }
    have_conv = true;
    BLMI_HIP_TRY(hipMalloc(&d_Z, sizeof(double2) * (size_t)nz));
    BLMI_HIP_TRY(hipMemcpy(d_Z, Z_host, sizeof(double2) * (size_t)nz, hipMemcpyHostToDevice));
    BLMI_HIP_TRY(hipMalloc(&d_table, sizeof(double) * (2 * (size_t)n_slices + 2)));
    BLMI_HIP_TRY(hipMemset(d_table, 0, sizeof(double) * (2 * (size_t)n_slices + 2)));
    have_fft = fft_plan_get(n_fft, &plan) == BLMI_OK;
    fft2_plan_destroy(&plan2);
    have_fourstep = n_slices <= n_fft && fft2_plan_create(n_fft, &plan2) == BLMI_OK;
    if (have_fourstep) {
        std::vector<double2> zp(2 * (size_t)plan2.M);
        fft2_permute_z(plan2, reinterpret_cast<const double2 *>(Z_host), zp.data());
        BLMI_HIP_TRY(hipMalloc(&d_Zp, sizeof(double2) * zp.size()));
        BLMI_HIP_TRY(hipMemcpy(d_Zp, zp.data(), sizeof(double2) * zp.size(),
                               hipMemcpyHostToDevice));
    }
    // the direct method's wake kernel g = -scale * irfft(Z) needs one
    // transform of length n_fft at set-up
    have_direct = false;
    if (have_fft && n_slices <= kDirectMaxSlices) {
        double *h = nullptr;
        BLMI_HIP_TRY(hipMalloc(&h, sizeof(double) * (size_t)n_fft));
        int rc = launch_irfft(d_Z, nz, h, n_fft, plan, nullptr);
        std::vector<double> hh((size_t)n_fft);
        if (rc == BLMI_OK) {
            hipError_t e = hipMemcpy(hh.data(), h, sizeof(double) * (size_t)n_fft,
                                     hipMemcpyDeviceToHost);
            if (e != hipSuccess) rc = blmi_set_error(BLMI_ERR_HIP, "wake copy: %s", hipGetErrorString(e));
        }
        (void)hipFree(h);
        if (rc) return rc;
        // gl[k] = g[(k - (ns-1)) mod n_fft], k in [0, n_fft + ns - 1), zero
        // padded: outputs up to n_fft (a multi-turn memory) can be formed
        const size_t len = (size_t)n_fft + n_slices + kConvPad;
        std::vector<double> gl(len, 0.0);
        for (int64_t k = 0; k < n_fft + (int64_t)n_slices - 1; ++k) {
            int64_t idx = (k - (n_slices - 1)) % n_fft;
            if (idx < 0) idx += n_fft;
            gl[(size_t)k] = -(scale * hh[(size_t)idx]);
        }
        BLMI_HIP_TRY(hipMalloc(&d_gl, sizeof(double) * len));
        BLMI_HIP_TRY(hipMemcpy(d_gl, gl.data(), sizeof(double) * len, hipMemcpyHostToDevice));
        gl_len = (int64_t)len;
        have_direct = true;
    }
    if (!have_fft && !have_direct && !have_fourstep) {
        const std::string why = blmi_last_error();  // the plan's own message
        return blmi_set_error(BLMI_ERR_UNSUPPORTED, "induced voltage: n_fft=%lld is not supported (%s)",
                              (long long)n_fft, why.c_str());
    }
    method = have_direct ? 1 : preferred_fft();
    intensity = true;
    invalidate_graph();
    return BLMI_OK;
}

// The one-workgroup LDS transform wins below 8192 points; from 8192 the
// four-step transform over many workgroups is faster (8192: 21 against
// 33 us; 16384: 24 against 60 us per turn, measured on the c2 bunch).
int Engine::preferred_fft() const {
    if (have_fourstep && (n_fft >= 8192 || !have_fft)) return 2;
    return have_fft ? 0 : 2;
}

int Engine::set_method(int method_) {
    if (method_ == 1 && !have_direct)
        return blmi_set_error(BLMI_ERR_UNSUPPORTED, "direct convolution not available (n_slices %d)",
                              n_slices);
    if (method_ == 0 && !have_fft)
        return blmi_set_error(BLMI_ERR_UNSUPPORTED, "LDS FFT not available (n_fft %lld)",
                              (long long)n_fft);
    if (method_ == 2 && !have_fourstep)
        return blmi_set_error(BLMI_ERR_UNSUPPORTED, "four-step FFT not available (n_fft %lld)",
                              (long long)n_fft);
    if (method_ < 0 || method_ > 2) return blmi_set_error(BLMI_ERR_ARG, "method must be 0, 1 or 2");
    if (!have_conv) return blmi_set_error(BLMI_ERR_ARG, "no impedance convolution is set");
    method = method_;
    invalidate_graph();
    return BLMI_OK;
}

int Engine::set_hist(int groups_, int max_peel_) {
    if (groups_ < 0 || groups_ > kMaxCountGroups)
        return blmi_set_error(BLMI_ERR_ARG, "groups in [0 (automatic), %d]", kMaxCountGroups);
    if (max_peel_ < 0 || max_peel_ > 64) return blmi_set_error(BLMI_ERR_ARG, "max_peel in [0, 64]");
    max_peel = max_peel_;
    invalidate_graph();
    groups_auto = groups_ == 0;
    if (groups_auto) groups_ = n_slices <= 1024 ? 4 : 1;
    if (groups_ != groups) {
        // the staged counts of the current turn are dropped: the caller
        // primes again (FusedTracker does so on its next track())
        groups = groups_;
        return alloc_counts();
    }
    return BLMI_OK;
}

void Engine::invalidate_graph() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph = nullptr;
    graph_turns = 0;
}

int Engine::set_turn(int turn) {
    if (turn < 0 || turn >= n_turns)
        return blmi_set_error(BLMI_ERR_ARG, "turn %d outside [0, %d)", turn, n_turns);
    BLMI_HIP_TRY(hipMemcpy(d_base, &turn, sizeof(int), hipMemcpyHostToDevice));
    host_turn = turn;
    // staged counts refer to the old turn: drop them (prime rebuilds them)
    if (d_counts)
        BLMI_HIP_TRY(hipMemset(d_counts, 0,
                               sizeof(unsigned) * 2 * (size_t)groups * (size_t)n_slices));
    BLMI_HIP_TRY(hipMemset(d_base + 1, 0xff, sizeof(int) * 2));
    return BLMI_OK;
}

int Engine::get_turn(int *turn) {
    BLMI_HIP_TRY(hipMemcpy(turn, d_base, sizeof(int), hipMemcpyDeviceToHost));
    return BLMI_OK;
}

int Engine::hist_layout(int block, int *grid_out, size_t *lds_out) const {
    int grid = device_cu_count() * blocks_per_cu;
    const int64_t n2 = n / 2;
    const int64_t need = (n2 + block - 1) / block;
    if (grid > need) grid = (int)(need > 0 ? need : 1);
    int mode = 0;
    size_t lds = 0;
    // u32 counts up to 64 KiB of LDS; above, u16 pairs and partial rows are
    // faster although u32 would still fit (1e7 macroparticles, whole turn:
    // 24576 slices 115.3 against 117.5 us, 32768 slices 117.5 against
    // 122.2 us; at 16384 the u32 counts win, 111.3 against 115.1 us)
    if ((size_t)n_slices * sizeof(unsigned) <= (size_t)kLdsHistU32Bytes) {
        mode = 1;
        lds = (size_t)n_slices * sizeof(unsigned);
    } else if ((size_t)hist_words(2, n_slices) * sizeof(unsigned) <= (size_t)kLdsHistBytes) {
        // a thread takes ceil(n2 / (grid*block)) pairs: keep a workgroup's
        // particles (+1 for an odd tail) within the u16 counters
        const int64_t max_pairs = (kMaxParticlesPerGroup16 - 1) / (2 * (int64_t)block);
        if (max_pairs >= 1) {
            int64_t g = (n2 + block * max_pairs - 1) / (block * max_pairs);
            if (g > grid) {
                // whole waves of workgroups: every CU runs the same number
                const int64_t wave = (int64_t)device_cu_count() * blocks_per_cu;
                g = (g + wave - 1) / wave * wave;
                grid = (int)g;
            }
            mode = 2;
            lds = (size_t)hist_words(2, n_slices) * sizeof(unsigned);
        }
    }
    *grid_out = grid;
    *lds_out = lds;
    return mode;
}

static int lds_attribute(const void *fn, size_t lds) {
    if (lds <= 65536) return BLMI_OK;
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess)
        return blmi_set_error(BLMI_ERR_HIP, "kernel LDS attribute (%zu B): %s", lds,
                              hipGetErrorString(e));
    return BLMI_OK;
}

// Partial-histogram rows for the u16-pair mode, sized for the current
// launch shape (allocated outside any graph capture).
int Engine::ensure_partials() {
    if (n_slices <= 0 || !d_counts || ordered()) return BLMI_OK;
    int grid = 0;
    size_t lds = 0;
    if (hist_layout(block_size, &grid, &lds) != 2) return BLMI_OK;
    const size_t need = sizeof(unsigned) * (size_t)grid * hist_words(2, n_slices);
    if (need > partials_bytes) {
        invalidate_graph();  // a captured graph holds the old rows
        (void)hipFree(d_partials);
        d_partials = nullptr;
        partials_bytes = 0;
        BLMI_HIP_TRY(hipMalloc(&d_partials, need));
        partials_bytes = need;
    }
    return BLMI_OK;
}

bool Engine::part_perm(int *shift, int *lshift, int *m1_lo) const {
    *shift = 31;
    *lshift = 0;
    if (m1_lo) *m1_lo = 0;
    if (!intensity || !have_conv || method != 2 || !have_fourstep || fb_on || ordered())
        return false;
    int grid = 0;
    size_t lds = 0;
    if (hist_layout(block_size, &grid, &lds) != 2) return false;
    const int W = hist_words(2, n_slices);
    const int M2 = plan2.M2;
    if (M2 < 2 || (M2 & (M2 - 1)) || W % M2) return false;
    const int L = W / M2;
    if (L < 4 || (L & (L - 1)) || L > plan2.M1 || kColThreads % (plan2.cols * (L / 4)))
        return false;
    // the padded LDS rows must still fit beside the kernel's static arrays
    if ((size_t)(W + W / M2) * sizeof(unsigned) + kStaticLdsMargin > (size_t)device_max_lds())
        return false;
    *shift = __builtin_ctz(M2);
    *lshift = __builtin_ctz(L);
    int lw = 0;
    if (win_m1 >= 0 && d_cnt_out && part_window_possible(&lw) && win_m1 + lw <= L) {
        *lshift = __builtin_ctz(lw);
        if (m1_lo) *m1_lo = win_m1;
    }
    return true;
}

bool Engine::part_window_possible(int *lw) const {
    // full rows above 64 KiB of LDS: half of each column's words in LDS,
    // which leaves room for a window of the interpolation table
    const int W = hist_words(2, n_slices);
    const int M2 = plan2.M2;
    if (!have_fourstep || M2 < 2 || W % M2 || (size_t)W * sizeof(unsigned) <= 65536) return false;
    const int L = W / M2;
    if (L < 8 || (L & (L - 1)) || kColThreads % (plan2.cols * (L / 8))) return false;
    *lw = L / 2;
    return true;
}

int Engine::launch_reduce(hipStream_t s, int off) {
    if (ordered()) return BLMI_OK;
    int ps = 31, pl = 0;
    if (part_perm(&ps, &pl)) return BLMI_OK;  // the four-step's first pass sums the rows
    int grid = 0;
    size_t lds = 0;
    if (hist_layout(block_size, &grid, &lds) != 2) return BLMI_OK;
    const int W = hist_words(2, n_slices);
    hipLaunchKernelGGL(partials_reduce_kernel, dim3((W + kRedWords - 1) / kRedWords), dim3(512), 0,
                       s, d_partials, grid, W, d_base, off, d_counts,
                       d_counts + (size_t)groups * n_slices, n_slices);
    return blmi_check_launch("partials_reduce");
}

template <typename T>
static int launch_prime(Engine *e, const T *x, T cl, T ibw, int mode, int grid, size_t lds,
                        unsigned *c1, int W, hipStream_t s) {
    int ps = 31, pl = 0, m1lo = 0, win_lo = 0, win_bins = 0;
    unsigned *cnt_out = nullptr;
    if (mode == 2 && e->part_perm(&ps, &pl, &m1lo)) {
        const int We = (1 << pl) << ps;  // the rows' words (the window's, if any)
        lds = sizeof(unsigned) * (size_t)lds_word(We, ps);
        if (We < W) {
            win_lo = 2 * (m1lo << ps);
            win_bins = 2 * We;
            cnt_out = e->d_cnt_out;
        }
    }
    const void *fn = mode == 2   ? (const void *)prime_hist_kernel<T, 2>
                     : mode == 1 ? (const void *)prime_hist_kernel<T, 1>
                                 : (const void *)prime_hist_kernel<T, 0>;
    int rc = lds_attribute(fn, lds);
    if (rc) return rc;
    const int blk = e->block_size;
    if (mode == 2)
        hipLaunchKernelGGL((prime_hist_kernel<T, 2>), dim3(grid), dim3(blk), lds, s, x, e->n,
                           e->d_base, e->d_counts, c1, e->n_slices, cl, ibw, e->d_partials, W,
                           ps, pl, win_lo, win_bins, cnt_out);
    else if (mode == 1)
        hipLaunchKernelGGL((prime_hist_kernel<T, 1>), dim3(grid), dim3(blk), lds, s, x, e->n,
                           e->d_base, e->d_counts, c1, e->n_slices, cl, ibw, e->d_partials, W,
                           ps, pl, win_lo, win_bins, cnt_out);
    else
        hipLaunchKernelGGL((prime_hist_kernel<T, 0>), dim3(grid), dim3(blk), 0, s, x, e->n,
                           e->d_base, e->d_counts, c1, e->n_slices, cl, ibw, e->d_partials, W,
                           ps, pl, win_lo, win_bins, cnt_out);
    return BLMI_OK;
}

// Counts of the current coordinates per transform column row (bin >>
// cshift), for choosing the windowed rows' position (nc <= 1024).
template <typename T>
__global__ void __launch_bounds__(1024) coarse_hist_kernel(const T *__restrict__ dt, int64_t n,
                                                           T cut_left, T inv_bw, int n_slices,
                                                           int cshift, int nc,
                                                           unsigned *__restrict__ out) {
    __shared__ unsigned h[1024];
    for (int b = threadIdx.x; b < nc; b += blockDim.x) h[b] = 0u;
    __syncthreads();
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const T f = floor((dt[i] - cut_left) * inv_bw);
        if (f >= T(0) && f < (T)n_slices) atomicAdd(&h[(int)f >> cshift], 1u);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nc; b += blockDim.x)
        if (h[b]) atomic_add_agent(&out[b], h[b]);
}

// The windowed rows' position: the half of the transform's column rows
// holding most of the bunch (one synchronous read of L counts, at prime).
int Engine::choose_window(hipStream_t s) {
    int lw = 0;
    int ps = 31, pl = 0;
    win_m1 = -1;
    if (!part_perm(&ps, &pl) || !part_window_possible(&lw)) return BLMI_OK;
    const int L = 1 << pl;  // full rows (no window set yet)
    if (L > 1024) return BLMI_OK;
    if (!d_cnt_out) BLMI_HIP_TRY(hipMalloc(&d_cnt_out, sizeof(unsigned) * (size_t)n_slices));
    int rc = launch_fill<unsigned>(d_cnt_out, n_slices, 0u, s);
    if (rc) return rc;
    const int cshift = 1 + ps;  // bins per column row: 2 M2
    const int grid = device_cu_count();
    if (f32)
        hipLaunchKernelGGL(coarse_hist_kernel<float>, dim3(grid), dim3(1024), 0, s, dt32, n,
                           cut_left32, inv_bw32, n_slices, cshift, L, d_cnt_out);
    else
        hipLaunchKernelGGL(coarse_hist_kernel<double>, dim3(grid), dim3(1024), 0, s, dt, n,
                           cut_left, inv_bw, n_slices, cshift, L, d_cnt_out);
    rc = blmi_check_launch("coarse_hist");
    if (rc) return rc;
    std::vector<unsigned> h(L);
    BLMI_HIP_TRY(hipMemcpyAsync(h.data(), d_cnt_out, sizeof(unsigned) * L, hipMemcpyDeviceToHost, s));
    BLMI_HIP_TRY(hipStreamSynchronize(s));
    rc = launch_fill<unsigned>(d_cnt_out, n_slices, 0u, s);
    if (rc) return rc;
    uint64_t cur = 0, best = 0;
    for (int k = 0; k < lw; ++k) cur += h[k];
    int at = 0;
    best = cur;
    for (int k = 1; k + lw <= L; ++k) {
        cur += h[k + lw - 1];
        cur -= h[k - 1];
        if (cur > best) {
            best = cur;
            at = k;
        }
    }
    win_m1 = at;
    invalidate_graph();
    return BLMI_OK;
}

int Engine::prime(hipStream_t s) {
    // counts of the current coordinates (the first pass of the map's
    // Profile.track()) into counts[t & 1]; the other buffer cleared
    if (!intensity || n_slices <= 0) return BLMI_OK;
    int rc = ensure_partials();
    if (rc) return rc;
    rc = choose_window(s);
    if (rc) return rc;
    rc = launch_fill<unsigned>(d_counts, 2 * (int64_t)groups * n_slices, 0u, s);
    if (rc) return rc;
    int grid = 0;
    size_t lds = 0;
    int mode = hist_layout(block_size, &grid, &lds);
    if (mode == 2 && ordered()) {
        // no partial rows in the ordered mode: count with device atomics
        mode = 0;
        lds = 0;
    }
    unsigned *c1 = d_counts + (size_t)groups * n_slices;
    const int W = hist_words(2, n_slices);
    if (f32)
        rc = launch_prime<float>(this, dt32, cut_left32, inv_bw32, mode, grid, lds, c1, W, s);
    else
        rc = launch_prime<double>(this, dt, cut_left, inv_bw, mode, grid, lds, c1, W, s);
    if (rc) return rc;
    rc = blmi_check_launch("engine_prime");
    if (rc) return rc;
    return mode == 2 ? launch_reduce(s, 0) : BLMI_OK;  // mode 2 only when not ordered
}

template <typename T>
using TurnFn = void (*)(T *, T *, int64_t, TurnArgs<T>);

template <typename T, int S>
static TurnFn<T> pick_turn(int interp, bool hist, int ldsh, bool publish, bool nt) {
    // publish: plain turns (no intensity effects), whose histogram is the
    // caller's profile; with intensity effects the induced-voltage launch
    // consumes the counts instead
    if (!hist)
        return interp == 2   ? turn_kernel<T, S, 2, false, 0, false>
               : interp == 1 ? turn_kernel<T, S, 1, false, 0, false>
                             : turn_kernel<T, S, 0, false, 0, false>;
    if (interp == 2) {
        if (publish)
            return ldsh == 1 ? turn_kernel<T, S, 2, true, 1, true> : turn_kernel<T, S, 2, true, 0, true>;
        return ldsh == 1 ? turn_kernel<T, S, 2, true, 1, false> : turn_kernel<T, S, 2, true, 0, false>;
    }
    if (interp == 1 && publish) {
        switch (ldsh) {
            case 0: return turn_kernel<T, S, 1, true, 0, true>;
            case 1: return turn_kernel<T, S, 1, true, 1, true>;
            default: return turn_kernel<T, S, 1, true, 2, true>;
        }
    }
    if (interp == 1 && nt) {
        switch (ldsh) {
            case 0: return turn_kernel<T, S, 1, true, 0, false, true>;
            case 1: return turn_kernel<T, S, 1, true, 1, false, true>;
            default: return turn_kernel<T, S, 1, true, 2, false, true>;
        }
    }
    switch (interp * 3 + ldsh) {
        case 0: return turn_kernel<T, S, 0, true, 0, true>;
        case 1: return turn_kernel<T, S, 0, true, 1, true>;
        case 2: return turn_kernel<T, S, 0, true, 2, true>;
        case 3: return turn_kernel<T, S, 1, true, 0, false>;
        case 4: return turn_kernel<T, S, 1, true, 1, false>;
        default: return turn_kernel<T, S, 1, true, 2, false>;
    }
}

template <typename T>
static int launch_turn_t(const Engine &e, T *x, T *de, const T *rf, const T *acc,
                         const DriftCoeffs<T> *drift, T cl, T ibw, hipStream_t s, int off,
                         bool interp, const KernelEvents &ev) {
    const bool hist = e.n_slices > 0 && e.d_counts != nullptr;
    TurnArgs<T> a{};
    a.n_rf = e.n_rf;
    a.charge = (T)e.charge;
    a.rf = rf;
    a.acc = acc;
    if (e.interp_rf) {
        // the table carries the RF voltage and the acceleration kick
        a.n_rf = 0;
        a.acc = reinterpret_cast<const T *>(e.d_zero_acc);
    }
    a.drift = drift;
    a.base = e.d_base;
    a.off = off;
    a.last_turn = e.n_turns - 1;
    a.tab = reinterpret_cast<const double2 *>(e.d_table);
    a.tab_params = e.d_table ? e.d_table + 2 * e.n_slices : nullptr;
    a.n_tab = e.n_slices - 1;
    a.counts0 = e.d_counts;
    a.counts1 = e.d_counts ? e.d_counts + (size_t)e.groups * e.n_slices : nullptr;
    a.groups = e.groups;
    a.n_slices = e.n_slices;
    a.cut_left = cl;
    a.inv_bw = ibw;
    a.max_peel = e.max_peel;
    a.profile = e.profile;
    a.pending = e.d_base + 1;
    a.partials = e.d_partials;
    a.part_words = hist_words(2, e.n_slices);
    int m1lo = 0;
    const bool perm = e.part_perm(&a.part_shift, &a.part_lshift, &m1lo);
    a.win_lo = a.win_bins = 0;
    a.cnt_out = nullptr;
    a.tw_lo = a.tw_n = 0;
    a.stats = e.stats_on ? e.d_stats : nullptr;
    a.shift_t = e.stat_shift_t;
    a.shift_e = e.stat_shift_e;
    a.sink = e.d_sink;
    int grid = 0;
    size_t lds = 0;
    int ldsh = e.hist_layout(e.block_size, &grid, &lds);
    if (!hist) {
        ldsh = 0;
        lds = 0;
    }
    if (ldsh == 2 && perm) {
        const int We = (1 << a.part_lshift) << a.part_shift;  // the window's words, if any
        lds = sizeof(unsigned) * (size_t)lds_word(We, a.part_shift);
        if (We < a.part_words) {
            a.win_lo = 2 * (m1lo << a.part_shift);
            a.win_bins = 2 * We;
            a.cnt_out = e.d_cnt_out;
        }
    }
    int imode = interp ? 1 : 0;
    // stage the interpolation table in LDS when it fits beside a u32 histogram
    const size_t entry = sizeof(typename TabEntry<T>::type);
    const size_t tab_bytes = entry * (size_t)(e.n_slices > 1 ? e.n_slices - 1 : 0);
    const size_t hist_bytes = ((lds + 15) / 16) * 16;
    if (interp && ldsh != 2 && hist_bytes + tab_bytes <= (size_t)kLdsTableBudget) {
        imode = 2;
        lds = hist_bytes + tab_bytes;
    }
    if (imode == 1 && hist && ldsh == 2 && a.n_tab > 0) {
        // the LDS the u16 rows leave holds a window of the table, centred on
        // the rows' window (else on the slice range)
        const size_t base = (lds + 15) / 16 * 16;
        const size_t cap = (size_t)device_max_lds();
        const size_t avail = cap > base + kStaticLdsMargin ? cap - base - kStaticLdsMargin : 0;
        int tw = (int)std::min<size_t>(avail / entry, (size_t)a.n_tab) & ~63;
        if (tw >= 256) {
            const int c = a.win_bins ? a.win_lo + a.win_bins / 2 : a.n_tab / 2;
            int lo = c - tw / 2;
            lo = std::max(0, std::min(lo, a.n_tab - tw));
            a.tw_lo = lo;
            a.tw_n = tw;
            lds = base + entry * (size_t)tw;
        }
    }
    const bool publish = !e.intensity;
    // non-temporal coordinates when they exceed the last-level cache (c3 at
    // 3e7 macroparticles, 480 MB: particle kernel 319 -> 285 us; at 1e7,
    // cache-resident, they would cost 68 -> 99 us)
    const bool nt = (double)sizeof(T) * 2.0 * (double)e.n > kLastLevelCacheBytes;
    TurnFn<T> fn = e.solver == 0   ? pick_turn<T, 0>(imode, hist, ldsh, publish, nt)
                   : e.solver == 1 ? pick_turn<T, 1>(imode, hist, ldsh, publish, nt)
                                   : pick_turn<T, 2>(imode, hist, ldsh, publish, nt);
    int rc = lds_attribute((const void *)fn, lds);
    if (rc) return rc;
    launch(fn, dim3(grid), dim3(e.block_size), lds, s, ev, x, de, e.n, a);
    return blmi_check_launch("turn");
}

template <typename T>
using TurnsFn = void (*)(T *, T *, int64_t, TurnArgs<T>, int);

template <typename T, int S>
static TurnsFn<T> pick_turns(int ldsh) {
    return ldsh == 1 ? turns_kernel<T, S, 1> : turns_kernel<T, S, 0>;
}

// `turns` plain turns from base + off in one launch (multi_turn_ok holds)
template <typename T>
static int launch_turns_t(const Engine &e, T *x, T *de, const T *rf, const T *acc,
                          const DriftCoeffs<T> *drift, T cl, T ibw, hipStream_t s, int off,
                          int turns, const KernelEvents &ev) {
    TurnArgs<T> a{};
    a.n_rf = e.n_rf;
    a.charge = (T)e.charge;
    a.rf = rf;
    a.acc = acc;
    a.drift = drift;
    a.base = e.d_base;
    a.off = off;
    a.last_turn = e.n_turns - 1;
    a.counts0 = e.d_counts;
    a.counts1 = e.d_counts + (size_t)e.groups * e.n_slices;
    a.groups = e.groups;
    a.n_slices = e.n_slices;
    a.cut_left = cl;
    a.inv_bw = ibw;
    a.max_peel = e.max_peel;
    a.profile = e.profile;
    a.pending = e.d_base + 1;
    int grid = 0;
    size_t lds = 0;
    const int ldsh = e.hist_layout(e.block_size, &grid, &lds);
    if (ldsh > 1) return blmi_set_error(BLMI_ERR_ARG, "multi-turn launch with u16 histogram rows");
    if (turns > kTurnBlockMax) return blmi_set_error(BLMI_ERR_ARG, "multi-turn launch of %d turns", turns);
    // the histogram words (mode 1), then the turns' programs
    lds = (ldsh == 1 ? (lds + 15) / 16 * 16 : 0) + turn_prog_bytes<T>() * (size_t)turns;
    TurnsFn<T> fn = e.solver == 0   ? pick_turns<T, 0>(ldsh)
                    : e.solver == 1 ? pick_turns<T, 1>(ldsh)
                                    : pick_turns<T, 2>(ldsh);
    int rc = lds_attribute((const void *)fn, lds);
    if (rc) return rc;
    launch(fn, dim3(grid), dim3(e.block_size), lds, s, ev, x, de, e.n, a, turns);
    return blmi_check_launch("turns");
}

bool Engine::multi_turn_ok() const {
    if (turn_block < 2 || intensity || interp_rf || fb_on || stats_on || ordered() || wake_on())
        return false;
    if (n_slices <= 0 || !d_counts) return false;
    int grid = 0;
    size_t lds = 0;
    return hist_layout(block_size, &grid, &lds) <= 1;
}

int Engine::multi_turn_piece(int left) const {
    const int cap = turn_block < kTurnBlockMax ? turn_block : kTurnBlockMax;
    int k = left < cap ? left : cap;
    if (k > 1 && !(k & 1)) --k;
    return k;
}

int Engine::launch_turns_block(hipStream_t s, int off, int turns, const KernelEvents &ev) {
    if (f32)
        return launch_turns_t<float>(*this, dt32, dE32, d_rf32, d_acc32, d_drift32, cut_left32,
                                     inv_bw32, s, off, turns, ev);
    return launch_turns_t<double>(*this, dt, dE, d_rf, d_acc, d_drift, cut_left, inv_bw, s, off,
                                  turns, ev);
}

InducedArgs Engine::induced_args(int off) const {
    InducedArgs ia{};
    ia.src = nullptr;
    ia.turn_base = d_base;
    ia.turn_off = off;
    ia.counts0 = d_counts;
    ia.counts1 = d_counts + (size_t)groups * n_slices;
    ia.groups = groups;
    ia.profile_out = profile;
    ia.n_slices = n_slices;
    ia.Z = reinterpret_cast<const double2 *>(d_Z);
    ia.Zp = d_Zp;
    ia.scale = scale;
    ia.volt = volt;
    ia.n_out = n_slices;
    ia.bc = bc;
    ia.charge = charge;
    ia.tab = reinterpret_cast<double2 *>(d_table);
    ia.tab_params = d_table + 2 * n_slices;
    ia.f32 = f32 ? 1 : 0;
    int ps = 31, pl = 0, m1lo = 0;
    ia.part = nullptr;
    ia.cnt_out = nullptr;
    ia.part_m1_lo = 0;
    if (part_perm(&ps, &pl, &m1lo)) {
        int grid = 0;
        size_t lds = 0;
        (void)hist_layout(block_size, &grid, &lds);
        ia.part = d_partials;
        ia.part_rows = grid;
        ia.part_lshift = pl;
        if (((1 << pl) << ps) < hist_words(2, n_slices)) {
            ia.part_m1_lo = m1lo;
            ia.cnt_out = d_cnt_out;
        }
    }
    ia.gl = d_gl;
    ia.gl_len = gl_len;
    return ia;
}

int Engine::launch_induced(hipStream_t s, int off, const KernelEvents &ev) {
    const bool w = wake_on();
    if (!have_conv && !w) return blmi_set_error(BLMI_ERR_ARG, "induced voltage: no source is set");
    if (have_conv) {
        InducedArgs ia = induced_args(off);
        KernelEvents e1 = ev;
        if (w && method == 1) {
            // the direct convolution runs the wake step in its own epilogue
            ia.n_out = (int)(n_mem > n_slices ? n_mem : n_slices);
            ia.with_wake = 1;
            ia.wk = wake_args(off);
            if (fb_in_induced()) {
                ia.with_fb = 1;
                ia.fb = feedback_args(off);
                ia.fb.from_counts = 1;
            }
            return launch_induced_direct(ia, s, e1);
        }
        if (w) {
            // the convolution of this turn goes to d_v1; the wake step
            // forms the caller's voltage and the table
            ia.volt = d_v1;
            ia.n_out = (int)(n_mem > n_slices ? n_mem : n_slices);
            ia.tab = nullptr;
            ia.tab_params = nullptr;
            e1.stop = nullptr;
        }
        if (fb_in_induced()) {
            ia.with_fb = 1;
            ia.fb = feedback_args(off);
            ia.fb.from_counts = 1;  // the profile is being written beside it
        }
        int rc = method == 1   ? launch_induced_direct(ia, s, e1)
                 : method == 2 ? launch_induced_fourstep(ia, plan2, s, e1)
                               : launch_induced_voltage_lds(ia, plan, s, e1);
        if (rc || !w) return rc;
    }
    KernelEvents e2;
    e2.start = have_conv ? nullptr : ev.start;
    e2.stop = ev.stop;
    return launch_wake_finish(wake_args(off), s, e2);
}

int Engine::set_feedback(int machine, int delay, int first_bin, double alpha, double reference,
                         const double *bc_dev, const double *harmonic, const double *omega_d,
                         const double *prog, const double *on_time, int n_on_time,
                         const double *state) {
    (void)hipFree(d_fb);
    d_fb = nullptr;
    fb_on = false;
    invalidate_graph();
    if (machine < 0) return BLMI_OK;
    if (machine > kFbPSB) return blmi_set_error(BLMI_ERR_ARG, "feedback: machine 0 (LHC), 1 (LHC_F) or 2 (PSB)");
    if (f32) return blmi_set_error(BLMI_ERR_UNSUPPORTED, "feedback: FP64 particles only");
    if (n_slices < 1 || !profile || !bc_dev)
        return blmi_set_error(BLMI_ERR_ARG, "feedback: needs the profile and its bin centres");
    if (first_bin < 0 || first_bin >= n_slices)
        return blmi_set_error(BLMI_ERR_ARG, "feedback: first bin %d outside [0, %d)", first_bin, n_slices);
    if (n_rf > 1024) return blmi_set_error(BLMI_ERR_ARG, "feedback: at most 1024 RF systems");
    if (n_on_time < 0 || (n_on_time > 0 && !on_time) || !harmonic || !omega_d || !prog || !state)
        return blmi_set_error(BLMI_ERR_ARG, "feedback: null program");
    const size_t nt = (size_t)n_turns + 1;
    const size_t nh = nt * (size_t)n_rf, np = nt * kFbCols, ns = (size_t)kFbState + n_rf;
    // after the state: the split loop's partial sums and its arrival counter
    const size_t total = 2 * nh + np + (size_t)n_on_time + ns + 2 * kFbMaxParts + 1;
    BLMI_HIP_TRY(hipMalloc(&d_fb, sizeof(double) * total));
    fb_h = d_fb;
    fb_wd = fb_h + nh;
    fb_prog = fb_wd + nh;
    fb_on_time = fb_prog + np;
    fb_state = fb_on_time + n_on_time;
    fb_parts = fb_state + ns;
    fb_arrive = reinterpret_cast<unsigned *>(fb_parts + 2 * kFbMaxParts);
    BLMI_HIP_TRY(hipMemset(fb_arrive, 0, sizeof(double)));
    BLMI_HIP_TRY(hipMemcpy(fb_h, harmonic, sizeof(double) * nh, hipMemcpyHostToDevice));
    BLMI_HIP_TRY(hipMemcpy(fb_wd, omega_d, sizeof(double) * nh, hipMemcpyHostToDevice));
    BLMI_HIP_TRY(hipMemcpy(fb_prog, prog, sizeof(double) * np, hipMemcpyHostToDevice));
    if (n_on_time > 0)
        BLMI_HIP_TRY(hipMemcpy(fb_on_time, on_time, sizeof(double) * n_on_time, hipMemcpyHostToDevice));
    BLMI_HIP_TRY(hipMemcpy(fb_state, state, sizeof(double) * ns, hipMemcpyHostToDevice));
    fb_machine = machine;
    fb_delay = delay;
    fb_first_bin = first_bin;
    fb_alpha = alpha;
    fb_reference = reference;
    fb_n_on = n_on_time;
    fb_bc = bc_dev;
    fb_on = true;
    return BLMI_OK;
}

int Engine::get_feedback_state(double *host) {
    if (!fb_on) return blmi_set_error(BLMI_ERR_ARG, "feedback: not set");
    BLMI_HIP_TRY(hipStreamSynchronize(own_stream));
    BLMI_HIP_TRY(hipMemcpy(host, fb_state, sizeof(double) * ((size_t)kFbState + n_rf),
                           hipMemcpyDeviceToHost));
    return BLMI_OK;
}

int Engine::get_programs(int turn0, int turns, double *host) {
    if (turn0 < 0 || turns < 0 || turn0 + turns > n_turns + 1)
        return blmi_set_error(BLMI_ERR_ARG, "get_programs: turns [%d, %d) outside [0, %d]", turn0,
                              turn0 + turns, n_turns + 1);
    BLMI_HIP_TRY(hipStreamSynchronize(own_stream));
    BLMI_HIP_TRY(hipMemcpy(host, d_rf + (size_t)turn0 * 3 * n_rf,
                           sizeof(double) * 3 * n_rf * (size_t)turns, hipMemcpyDeviceToHost));
    return BLMI_OK;
}

int Engine::set_statistics(bool on, double shift_t, double shift_e) {
    invalidate_graph();
    stats_on = on;
    stat_shift_t = shift_t;
    stat_shift_e = shift_e;
    if (!on) return BLMI_OK;
    if (!d_moments) {
        const size_t bytes = sizeof(double) * 5 * ((size_t)n_turns + 1);
        BLMI_HIP_TRY(hipMalloc(&d_moments, bytes));
        BLMI_HIP_TRY(hipMemset(d_moments, 0, bytes));
    }
    return ensure_stats();
}

int Engine::ensure_stats() {
    if (!stats_on) return BLMI_OK;
    int grid = 0;
    size_t lds = 0;
    hist_layout(block_size, &grid, &lds);
    if (grid == stats_grid && d_stats) return BLMI_OK;
    (void)hipFree(d_stats);
    d_stats = nullptr;
    BLMI_HIP_TRY(hipMalloc(&d_stats, sizeof(double) * 4 * (size_t)grid * ((size_t)n_turns + 1)));
    stats_grid = grid;
    return BLMI_OK;
}

int Engine::get_statistics(int turn0, int turns, double *host) {
    if (!d_moments) return blmi_set_error(BLMI_ERR_ARG, "statistics: not enabled");
    if (turn0 < 0 || turns < 0 || turn0 + turns > n_turns + 1)
        return blmi_set_error(BLMI_ERR_ARG, "statistics: turns [%d, %d) outside [0, %d]", turn0,
                              turn0 + turns, n_turns + 1);
    BLMI_HIP_TRY(hipStreamSynchronize(own_stream));
    BLMI_HIP_TRY(hipMemcpy(host, d_moments + (size_t)turn0 * 5, sizeof(double) * 5 * (size_t)turns,
                           hipMemcpyDeviceToHost));
    return BLMI_OK;
}

int Engine::launch_feedback_turn(hipStream_t s, int off) {
    return launch_feedback(feedback_args(off), s);
}

FeedbackArgs Engine::feedback_args(int off) const {
    FeedbackArgs a{};
    a.turn_base = d_base;
    a.turn_off = off;
    a.machine = fb_machine;
    a.delay = fb_delay;
    a.first_bin = fb_first_bin;
    a.n_slices = n_slices;
    a.n_rf = n_rf;
    a.alpha = fb_alpha;
    a.bin_size = (cut_right - cut_left) / (double)n_slices;
    a.reference = fb_reference;
    a.bc = fb_bc;
    a.profile = profile;
    // the plain map leaves the turn's histogram in the counts until the
    // particle kernel publishes it; with intensity effects the induced
    // voltage step has written the profile already
    a.counts0 = d_counts;
    a.counts1 = d_counts + (size_t)groups * n_slices;
    a.pending = intensity ? nullptr : d_base + 1;
    a.groups = groups;
    a.rf = d_rf;
    a.harmonic = fb_h;
    a.omega_d = fb_wd;
    a.prog = fb_prog;
    a.on_time = fb_on_time;
    a.n_on_time = fb_n_on;
    a.state = fb_state;
    a.from_counts = 0;
    a.n_parts = 1;  // the direct convolution's launch splits the sums
    a.parts = fb_parts;
    a.arrive = fb_arrive;
    return a;
}

WakeArgs Engine::wake_args(int off) const {
    WakeArgs a{};
    a.turn_base = d_base;
    a.turn_off = off;
    a.v1 = have_conv ? d_v1 : nullptr;
    a.n_mem = n_mem;
    a.n_add = n_add;
    a.mem0 = d_mem;
    a.mem1 = d_mem ? d_mem + n_mem : nullptr;
    a.time = d_time;
    a.t_rev = d_trev;
    a.ind_coef = d_indc;
    a.deriv_mode = deriv_mode;
    a.profile = profile;
    a.counts0 = d_counts;
    a.counts1 = d_counts + (size_t)groups * n_slices;
    a.groups = groups;
    a.n_slices = n_slices;
    a.bc = bc;
    a.charge = charge;
    a.volt = volt;
    a.tab = reinterpret_cast<double2 *>(d_table);
    a.tab_params = d_table + 2 * n_slices;
    return a;
}

int Engine::set_wake(int64_t n_mem_, int64_t n_add_, const double *time, double *memory,
                     const double *ind_host, int deriv_mode_) {
    if (!intensity) return blmi_set_error(BLMI_ERR_ARG, "set_wake: set the induced voltage first");
    if (f32)
        return blmi_set_error(BLMI_ERR_UNSUPPORTED,
                              "multi-turn wakes and inductive sources run with FP64 particles");
    if (n_mem_ < 0 || (n_mem_ > 0 && (!time || !memory)))
        return blmi_set_error(BLMI_ERR_ARG, "set_wake: memory needs its time axis and buffer");
    if (n_mem_ > 0 && !have_conv)
        return blmi_set_error(BLMI_ERR_ARG, "set_wake: a multi-turn memory needs a convolution");
    if (n_mem_ > 0 && (n_mem_ < n_slices || n_mem_ > n_fft))
        return blmi_set_error(BLMI_ERR_ARG, "set_wake: memory length %lld outside [n_slices %d, n_fft %lld]",
                              (long long)n_mem_, n_slices, (long long)n_fft);
    if (n_add_ < 0 || n_add_ > n_mem_)
        return blmi_set_error(BLMI_ERR_ARG, "set_wake: n_add %lld outside [0, n_mem]", (long long)n_add_);
    if (deriv_mode_ != kDerivGradient && deriv_mode_ != kDerivDiff)
        return blmi_set_error(BLMI_ERR_ARG, "set_wake: derivative mode 0 (gradient) or 1 (diff)");
    if (ind_host && n_slices < 2)
        return blmi_set_error(BLMI_ERR_ARG, "set_wake: the profile derivative needs 2 slices");
    if (!have_conv && !ind_host)
        return blmi_set_error(BLMI_ERR_ARG, "set_wake: no induced-voltage source");
    (void)hipFree(d_mem);
    (void)hipFree(d_time);
    (void)hipFree(d_v1);
    (void)hipFree(d_indc);
    d_mem = d_time = d_v1 = d_indc = nullptr;
    n_mem = n_mem_;
    n_add = n_add_;
    mem_user = memory;
    deriv_mode = deriv_mode_;
    if (n_mem > 0) {
        BLMI_HIP_TRY(hipMalloc(&d_mem, sizeof(double) * 2 * (size_t)n_mem));
        BLMI_HIP_TRY(hipMemset(d_mem, 0, sizeof(double) * 2 * (size_t)n_mem));
        BLMI_HIP_TRY(hipMalloc(&d_time, sizeof(double) * (size_t)n_mem));
        BLMI_HIP_TRY(hipMemcpy(d_time, time, sizeof(double) * (size_t)n_mem, hipMemcpyDefault));
        // the direct convolution also forms the memory's n_mem values: it
        // wins up to n_mem * n_slices = 2^26 (16384 x 4096: 22 against 33 us
        // for the four-step transform of 16384 points, with the wake step)
        if (method == 1 && (double)n_mem * n_slices > kDirectMaxWork) method = preferred_fft();
        if (method == 2 && !have_fourstep)
            return blmi_set_error(BLMI_ERR_UNSUPPORTED, "multi-turn memory: no FFT for n_fft %lld",
                                  (long long)n_fft);
    }
    if (have_conv) {
        const size_t nv = (size_t)(n_mem > n_slices ? n_mem : n_slices);
        BLMI_HIP_TRY(hipMalloc(&d_v1, sizeof(double) * nv));
    }
    if (ind_host) {
        const size_t nt = (size_t)n_turns + 1;
        BLMI_HIP_TRY(hipMalloc(&d_indc, sizeof(double) * nt));
        BLMI_HIP_TRY(hipMemcpy(d_indc, ind_host, sizeof(double) * nt, hipMemcpyHostToDevice));
    }
    invalidate_graph();
    return BLMI_OK;
}

// tracker.py:585-601, interpolation=True: the table of one turn from the
// RF voltage on the bin centres (rf_volt_comp, kick.cpp:37-51, summed from
// zero over the systems), plus the induced voltage when there is one, with
// the acceleration kick in the intercepts (linear_interp_kick.cpp:30-48).
__global__ void interp_rf_table_kernel(const double *__restrict__ rf, int n_rf,
                                       const int *__restrict__ base, int off, int last_turn,
                                       const double *__restrict__ acc,
                                       const double *__restrict__ volt,
                                       const double *__restrict__ bc, int ns, double charge,
                                       double2 *__restrict__ tab, double *__restrict__ params) {
    int t = *base + off;
    if (t < 0) t = 0;
    if (t > last_turn) t = last_turn;
    const double *V = rf + (size_t)t * 3 * n_rf;
    const double *w = V + n_rf;
    const double *phi = w + n_rf;
    const double a = acc[t];
    const double ibw = (double)(ns - 1) / (bc[ns - 1] - bc[0]);
    auto total = [&](int b) {
        double r = 0.0;
        for (int j = 0; j < n_rf; ++j) r = r + V[j] * cephes_sin(w[j] * bc[b] + phi[j]);
        return volt ? r + volt[b] : r;
    };
    const int stride = gridDim.x * blockDim.x;
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < ns - 1; b += stride) {
        const double v0 = total(b), v1 = total(b + 1);
        const double k = charge * (v1 - v0) * ibw;
        tab[b] = make_double2(k, (charge * v0 - bc[b] * k) + a);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        params[0] = bc[0];
        params[1] = ibw;
    }
}

int Engine::set_interpolation(const double *bc_) {
    if (f32) return blmi_set_error(BLMI_ERR_UNSUPPORTED, "interpolation: FP64 particles only");
    if (n_slices < 2 || !bc_)
        return blmi_set_error(BLMI_ERR_ARG, "interpolation: set the slices (>= 2) and bin centres");
    if (ordered())
        return blmi_set_error(BLMI_ERR_UNSUPPORTED, "interpolation with the bin-ordered working set");
    if (!d_table) {
        BLMI_HIP_TRY(hipMalloc(&d_table, sizeof(double) * (2 * (size_t)n_slices + 2)));
        BLMI_HIP_TRY(hipMemset(d_table, 0, sizeof(double) * (2 * (size_t)n_slices + 2)));
    }
    if (!d_zero_acc) {
        BLMI_HIP_TRY(hipMalloc(&d_zero_acc, sizeof(double) * ((size_t)n_turns + 1)));
        BLMI_HIP_TRY(hipMemset(d_zero_acc, 0, sizeof(double) * ((size_t)n_turns + 1)));
    }
    interp_bc = bc_;
    interp_rf = true;
    invalidate_graph();
    return BLMI_OK;
}

int Engine::launch_interp_table(hipStream_t s, int off) {
    const int threads = 256;
    const int blocks = (n_slices - 1 + threads - 1) / threads;
    hipLaunchKernelGGL(interp_rf_table_kernel, dim3(blocks), dim3(threads), 0, s, d_rf, n_rf,
                       d_base, off, n_turns - 1, d_acc, intensity ? volt : nullptr, interp_bc,
                       n_slices, charge, reinterpret_cast<double2 *>(d_table),
                       d_table + 2 * n_slices);
    return blmi_check_launch("interp_rf_table");
}

int Engine::launch_turn_seq(hipStream_t s, int k, const KernelEvents &ei, const KernelEvents &et) {
    int rc = BLMI_OK;
    if (intensity) rc = launch_induced(s, k, ei);
    if (!rc && fb_on && !fb_in_induced()) rc = launch_feedback_turn(s, k);
    if (!rc && interp_rf) rc = launch_interp_table(s, k);
    if (!rc) rc = launch_turn_kernel(s, k, intensity || interp_rf, et);
    if (!rc) rc = launch_reduce(s, k + 1);
    if (!rc && ordered() && ++ord_since == sort_every) {
        ord_since = 0;
        rc = order_sort(s);
    }
    return rc;
}

int Engine::launch_turn_kernel(hipStream_t s, int off, bool interp, const KernelEvents &ev) {
    if (ordered()) return launch_window_turn(s, off, interp, ev);
    if (f32)
        return launch_turn_t<float>(*this, dt32, dE32, d_rf32, d_acc32, d_drift32, cut_left32,
                                    inv_bw32, s, off, interp, ev);
    return launch_turn_t<double>(*this, dt, dE, d_rf, d_acc, d_drift, cut_left, inv_bw, s, off,
                                 interp, ev);
}

// Turns base+off0 .. base+off0+count-1, then (if advance) *base += count and,
// without intensity effects, the publish launch for the last turn.
int Engine::launch_turns(hipStream_t s, int count, bool advance) {
    const KernelEvents none;
    if (multi_turn_ok()) {
        for (int k = 0; k < count;) {
            const int piece = multi_turn_piece(count - k);
            int rc = launch_turns_block(s, k, piece, none);
            if (rc) return rc;
            k += piece;
        }
        return advance ? finish_chunk(s, count) : BLMI_OK;
    }
    for (int k = 0; k < count; ++k) {
        int rc = launch_turn_seq(s, k, none, none);
        if (rc) return rc;
    }
    if (advance) return finish_chunk(s, count);
    return BLMI_OK;
}

int Engine::finish_chunk(hipStream_t s, int count) {
    if (!intensity && n_slices > 0 && d_counts) {
        hipLaunchKernelGGL(publish_kernel, dim3(1), dim3(1024), 0, s, d_base, count, d_counts,
                           d_counts + (size_t)groups * n_slices, groups, d_base + 1, profile,
                           n_slices);
        int rc = blmi_check_launch("publish");
        if (rc) return rc;
    }
    if (f32) {
        // the FP32 caller's profile and induced voltage
        if (profile32 && n_slices > 0)
            hipLaunchKernelGGL(to_float_kernel, dim3((n_slices + 255) / 256), dim3(256), 0, s,
                               (const double *)profile, profile32, n_slices);
        if (volt32 && intensity)
            hipLaunchKernelGGL(to_float_kernel, dim3((n_slices + 255) / 256), dim3(256), 0, s,
                               (const double *)volt, volt32, n_slices);
    }
    if (stats_on) {
        hipLaunchKernelGGL(moments_reduce_kernel, dim3(count), dim3(64), 0, s, d_stats, stats_grid,
                           d_base, (double)n, stat_shift_t, stat_shift_e, d_moments);
        int rc = blmi_check_launch("moments_reduce");
        if (rc) return rc;
    }
    hipLaunchKernelGGL(advance_kernel, dim3(1), dim3(1), 0, s, d_base, count);
    int rc = blmi_check_launch("advance");
    if (rc) return rc;
    // the multi-turn memory of the new turn base, for the caller
    if (n_mem > 0 && mem_user)
        return launch_mtw_publish(d_base, d_mem, d_mem + n_mem, n_mem, mem_user, s);
    return BLMI_OK;
}

int Engine::time_kernels(int turns, hipStream_t user, float *ms_induced, float *ms_turn) {
    // Eager turns; each kernel launched with start/stop events recorded by its
    // own dispatch: average device duration of the induced-voltage kernel and
    // of the particle kernel.
    if (turns <= 0) return blmi_set_error(BLMI_ERR_ARG, "time_kernels: turns must be positive");
    if (host_turn + turns > n_turns)
        return blmi_set_error(BLMI_ERR_ARG, "time_kernels: %d turns from turn %d passes n_turns=%d",
                              turns, host_turn, n_turns);
    int prc = ensure_partials();
    if (!prc) prc = ensure_stats();
    if (prc) return prc;
    const bool ord = ordered();
    if (ord) {
        prc = ensure_order_buffers();
        if (prc) return prc;
    }
    BLMI_HIP_TRY(hipEventRecord(ev_in, user));
    BLMI_HIP_TRY(hipStreamWaitEvent(own_stream, ev_in, 0));
    if (ord) {
        prc = order_in(own_stream);
        if (prc) return prc;
        ord_since = 0;
    }
    if (n_mem > 0 && mem_user) {
        prc = launch_mtw_load(d_base, d_mem, d_mem + n_mem, n_mem, mem_user, own_stream);
        if (prc) return prc;
    }
    std::vector<hipEvent_t> ev(4 * (size_t)turns, nullptr);
    for (auto &x : ev) BLMI_HIP_TRY(hipEventCreate(&x));
    int rc = BLMI_OK;
    // with several plain turns to a launch, launch k's events time its
    // `piece` turns (the other launches' events stay unused: elapsed 0)
    const bool multi = multi_turn_ok();
    int launches = 0;
    for (int k = 0; k < turns && rc == BLMI_OK;) {
        KernelEvents ei{ev[4 * launches], ev[4 * launches + 1]},
            et{ev[4 * launches + 2], ev[4 * launches + 3]};
        const int piece = multi ? multi_turn_piece(turns - k) : 1;
        rc = multi ? launch_turns_block(own_stream, k, piece, et)
                   : launch_turn_seq(own_stream, k, ei, et);
        k += piece;
        ++launches;
    }
    if (rc == BLMI_OK) rc = finish_chunk(own_stream, turns);
    if (rc == BLMI_OK && ord) rc = order_out(own_stream);
    if (rc == BLMI_OK) {
        host_turn += turns;
        BLMI_HIP_TRY(hipStreamSynchronize(own_stream));
        double a = 0, b = 0;
        for (int k = 0; k < launches; ++k) {
            float x = 0, y = 0;
            if (intensity) BLMI_HIP_TRY(hipEventElapsedTime(&x, ev[4 * k], ev[4 * k + 1]));
            BLMI_HIP_TRY(hipEventElapsedTime(&y, ev[4 * k + 2], ev[4 * k + 3]));
            a += x;
            b += y;
        }
        *ms_induced = (float)(a / turns);
        *ms_turn = (float)(b / turns);
    }
    BLMI_HIP_TRY(hipStreamSynchronize(own_stream));
    for (auto &x : ev) (void)hipEventDestroy(x);
    BLMI_HIP_TRY(hipEventRecord(ev_out, own_stream));
    BLMI_HIP_TRY(hipStreamWaitEvent(user, ev_out, 0));
    return rc;
}

int Engine::track(int turns, hipStream_t user, bool use_graph) {
    if (turns <= 0) return BLMI_OK;
    if (f32 ? (!dt32 || !dE32) : (!dt || !dE))
        return blmi_set_error(BLMI_ERR_ARG, "engine has no particle arrays");
    // the per-turn tables hold turns [0, n_turns]; turn t reads entries t and t+1
    if (host_turn + turns > n_turns)
        return blmi_set_error(BLMI_ERR_ARG, "tracking %d turns from turn %d passes n_turns=%d",
                              turns, host_turn, n_turns);
    int prc = ensure_partials();
    if (!prc) prc = ensure_stats();
    if (prc) return prc;
    const bool ord = ordered();
    if (ord) {
        prc = ensure_order_buffers();
        if (prc) return prc;
    }
    // a captured chunk holds an even number of sorts, so it starts and ends
    // on the same ordered buffer
    int chunk = graph_chunk;
    if (ord) chunk = chunk < 2 * sort_every ? 2 * sort_every : chunk - chunk % (2 * sort_every);
    // order the engine stream after the caller's stream
    BLMI_HIP_TRY(hipEventRecord(ev_in, user));
    BLMI_HIP_TRY(hipStreamWaitEvent(own_stream, ev_in, 0));
    if (ord) {
        prc = order_in(own_stream);
        if (prc) return prc;
        ord_since = 0;
    }
    // the caller's multi-turn memory (it may have been changed between calls)
    if (n_mem > 0 && mem_user) {
        prc = launch_mtw_load(d_base, d_mem, d_mem + n_mem, n_mem, mem_user, own_stream);
        if (prc) return prc;
    }
    int done = 0;
    int rc = BLMI_OK;
    if (use_graph && turns >= chunk) {
        if (!graph_exec || graph_turns != chunk) {
            invalidate_graph();
            BLMI_HIP_TRY(hipStreamBeginCapture(own_stream, hipStreamCaptureModeThreadLocal));
            rc = launch_turns(own_stream, chunk, true);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(own_stream, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess)
                return blmi_set_error(BLMI_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
            graph = g;
            BLMI_HIP_TRY(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0));
            graph_turns = chunk;
        }
        while (turns - done >= graph_turns) {
            BLMI_HIP_TRY(hipGraphLaunch(graph_exec, own_stream));
            done += graph_turns;
        }
    }
    if (done < turns) rc = launch_turns(own_stream, turns - done, true);
    if (!rc && ord) rc = order_out(own_stream);
    if (rc) return rc;
    host_turn += turns;
    BLMI_HIP_TRY(hipEventRecord(ev_out, own_stream));
    BLMI_HIP_TRY(hipStreamWaitEvent(user, ev_out, 0));
    return BLMI_OK;
}

}  // namespace blmi
