// Batched ensemble of independent bunches.
//
// Layout in HBM: the bunches' coordinates back to back in one dt and one dE
// array (each bunch starting at an even index, so every workgroup streams
// 16-byte double2 pairs), per-bunch tables indexed by bunch id: RF programs
// [bunch][turn][3][n_rf], cuts, u32 counts [bunch][n_slices], impedance
// [bunch][n_fft/2+1], interpolation tables [bunch][2 n_slices + 2].
//
// A turn is two launches for all bunches:
//   induced: one workgroup per bunch -- counts -> profile (caller's), LDS
//            Stockham rfft, x Z_b, irfft -> V_b (caller's), interpolation
//            table; counts cleared for the particle pass;
//   turn:    the particle pairs of all bunches split evenly over the grid;
//            each workgroup walks its segments (a segment = one bunch's
//            share), loads that bunch's RF program and table, kicks, drifts
//            and bins into an LDS histogram flushed to the bunch's counts
//            at the segment's end.
// Without intensity effects the induced launch is replaced by a publish
// launch after the particle pass (the map order tracker, profile).
// Nothing synchronises workgroups inside a launch; the turn index lives on
// the device (base + offset baked into each launch), so a captured chunk of
// turns replays unchanged.
#include <string.h>

#include <algorithm>
#include <vector>

#include "blmi_common.h"
#include "ensemble.h"
#include "particles.h"
#include "spectral.h"
#include "track_device.h"

namespace blmi {

struct EnsTurnArgs {
    int n_rf;
    double charge;
    const double *rf;          // [bunch][turn][3][n_rf]
    int64_t rf_bunch_stride;
    const double *acc;         // [turn]
    const DriftCoeffs<double> *drift;  // [turn]
    const int *base;
    int off;
    int last_turn;
    const double *table;       // [bunch][tab_stride]: (slope, intercept)[ns], bc0, inv_bw
    int tab_stride;
    const double *cut_left, *inv_bw;  // [bunch]
    unsigned *counts;          // [bunch][n_slices]
    int n_slices;
    const int *seg_ptr;        // [grid + 1]
    const EnsSeg *segs;
    int max_peel;
};

// LDS: the segment's histogram (ns u32) and, with INTERP, its bunch's
// interpolation table ((ns - 1) double2, 16-byte aligned after the counts).
template <int SOLVER, bool INTERP, bool HIST>
__global__ void __launch_bounds__(1024) ensemble_turn_kernel(double *__restrict__ dt,
                                                             double *__restrict__ dE,
                                                             EnsTurnArgs a) {
    extern __shared__ unsigned lds_hist[];
    double2 *ltab = reinterpret_cast<double2 *>(lds_hist + ((a.n_slices + 3) & ~3));
    __shared__ __align__(16) double sin_tab[kCephesTabDoubles];
    const LdsConstDoublePtr stab = (LdsConstDoublePtr)sin_tab;
    cephes_tab_fill((LdsDoublePtr)sin_tab);
    __syncthreads();
    int t = *a.base + a.off;
    if (t < 0) t = 0;
    if (t > a.last_turn) t = a.last_turn;
    const double acc = a.acc[t];
    const DriftCoeffs<double> dc = a.drift[t + 1];
    const int ns = a.n_slices;
    const int s0 = a.seg_ptr[blockIdx.x], s1 = a.seg_ptr[blockIdx.x + 1];
    for (int sg = s0; sg < s1; ++sg) {
        const EnsSeg g = a.segs[sg];
        const int b = g.bunch;
        const double *V = a.rf + (size_t)b * a.rf_bunch_stride + (size_t)t * 3 * a.n_rf;
        const double *w = V + a.n_rf;
        const double *phi = w + a.n_rf;
        const RfProg<double> rfp(V, w, phi, a.n_rf);
        const double2 *tab = nullptr;
        double bc0 = 0.0, ibw = 0.0;
        if (INTERP) {
            const double *tb = a.table + (size_t)b * a.tab_stride;
            tab = reinterpret_cast<const double2 *>(tb);
            bc0 = tb[2 * ns];
            ibw = tb[2 * ns + 1];
        }
        const double cl = a.cut_left[b], hbw = a.inv_bw[b];
        if (HIST || INTERP) {
            for (int k = threadIdx.x; k < ns; k += blockDim.x) {
                if (HIST) lds_hist[k] = 0u;
                if (INTERP && k < ns - 1) ltab[k] = tab[k];
            }
            __syncthreads();
        }
        auto step = [&](double &x, double &e) -> int {
            if (INTERP) {
                // linear_interp_kick.cpp:56-64 (acceleration kick 0)
                const double f = floor((x - bc0) * ibw);
                if (f >= 0.0 && f < (double)(ns - 1)) {
                    const int k = (int)f;
                    const double2 tk = ltab[k];
                    e = e + (x * tk.x + tk.y);
                }
            }
            e = e + rfp.template kick_tab<-1>(x, a.charge, acc, stab);
            x = x + e_drift(dc, e, SOLVER);
            if (HIST) {
                const double f = floor((x - cl) * hbw);
                return (f >= 0.0 && f < (double)ns) ? (int)f : -1;
            }
            return -1;
        };
        // software pipeline: a thread's next pair is loaded while the
        // current one is computed
        double2 xn{}, en{};
        if (g.p0 + threadIdx.x < g.p1) {
            xn = reinterpret_cast<const double2 *>(dt)[g.p0 + threadIdx.x];
            en = reinterpret_cast<const double2 *>(dE)[g.p0 + threadIdx.x];
        }
        for (int64_t p = g.p0 + threadIdx.x; p < g.p1; p += blockDim.x) {
            double2 x = xn, e = en;
            if (p + blockDim.x < g.p1) {
                xn = reinterpret_cast<const double2 *>(dt)[p + blockDim.x];
                en = reinterpret_cast<const double2 *>(dE)[p + blockDim.x];
            }
            int b0 = step(x.x, e.x);
            int b1 = step(x.y, e.y);
            reinterpret_cast<double2 *>(dt)[p] = x;
            reinterpret_cast<double2 *>(dE)[p] = e;
            if (HIST) {
                if (2 * p >= g.end) b0 = -1;
                if (2 * p + 1 >= g.end) b1 = -1;
                lds_count<false>(lds_hist, b0, a.max_peel);
                lds_count<false>(lds_hist, b1, a.max_peel);
            }
        }
        if (HIST || INTERP) __syncthreads();  // every lane done with this segment's LDS
        if (HIST) {
            hist_flush<1>(lds_hist, a.counts + (size_t)b * ns, ns);
            __syncthreads();
        }
    }
}

// Histogram of the current coordinates into the (cleared) counts.
__global__ void __launch_bounds__(1024) ensemble_prime_kernel(const double *__restrict__ dt,
                                                              EnsTurnArgs a) {
    extern __shared__ unsigned lds_hist[];
    const int ns = a.n_slices;
    const int s0 = a.seg_ptr[blockIdx.x], s1 = a.seg_ptr[blockIdx.x + 1];
    for (int sg = s0; sg < s1; ++sg) {
        const EnsSeg g = a.segs[sg];
        const int b = g.bunch;
        const double cl = a.cut_left[b], hbw = a.inv_bw[b];
        for (int k = threadIdx.x; k < ns; k += blockDim.x) lds_hist[k] = 0u;
        __syncthreads();
        for (int64_t p = g.p0 + threadIdx.x; p < g.p1; p += blockDim.x) {
            const double2 x = reinterpret_cast<const double2 *>(dt)[p];
            const double f0 = floor((x.x - cl) * hbw), f1 = floor((x.y - cl) * hbw);
            int b0 = (f0 >= 0.0 && f0 < (double)ns) ? (int)f0 : -1;
            int b1 = (f1 >= 0.0 && f1 < (double)ns) ? (int)f1 : -1;
            if (2 * p >= g.end) b0 = -1;
            if (2 * p + 1 >= g.end) b1 = -1;
            lds_count<false>(lds_hist, b0, 0);
            lds_count<false>(lds_hist, b1, 0);
        }
        __syncthreads();
        hist_flush<1>(lds_hist, a.counts + (size_t)b * ns, ns);
        __syncthreads();
    }
}

// Plain mode: profile[b] = counts[b], counts cleared (one workgroup per bunch).
__global__ void ensemble_publish_kernel(unsigned *__restrict__ counts, double *__restrict__ profiles,
                                        int n_slices) {
    const size_t o = (size_t)blockIdx.x * n_slices;
    for (int k = threadIdx.x; k < n_slices; k += blockDim.x) {
        profiles[o + k] = (double)counts[o + k];
        counts[o + k] = 0u;
    }
}

__global__ void ensemble_advance_kernel(int *base, int k) { *base += k; }

// ---------------------------------------------------------------------------
Ensemble::~Ensemble() { release(); }

void Ensemble::invalidate_graph() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph = nullptr;
    graph_turns = 0;
}

void Ensemble::release() {
    invalidate_graph();
    void *bufs[] = {d_rf, d_acc, d_drift, d_base, d_cut_left, d_inv_bw, d_counts,
                    d_Z, d_table, d_iargs, d_segs, d_seg_ptr};
    for (void *p : bufs) (void)hipFree(p);
    d_rf = d_acc = d_cut_left = d_inv_bw = d_table = nullptr;
    d_drift = nullptr;
    d_base = d_seg_ptr = nullptr;
    d_counts = nullptr;
    d_Z = nullptr;
    d_iargs = nullptr;
    d_segs = nullptr;
    if (own_stream) (void)hipStreamDestroy(own_stream);
    own_stream = nullptr;
    if (ev_in) (void)hipEventDestroy(ev_in);
    if (ev_out) (void)hipEventDestroy(ev_out);
    ev_in = ev_out = nullptr;
}

int Ensemble::init(int n_bunches_, int n_rf_, int n_turns_, int solver_) {
    n_bunches = n_bunches_;
    n_rf = n_rf_;
    n_turns = n_turns_;
    solver = solver_;
    BLMI_HIP_TRY(hipGetDevice(&device));
    BLMI_HIP_TRY(hipMalloc(&d_rf, sizeof(double) * 3 * n_rf * (size_t)(n_turns + 1) * n_bunches));
    BLMI_HIP_TRY(hipMalloc(&d_acc, sizeof(double) * (size_t)(n_turns + 1)));
    BLMI_HIP_TRY(hipMalloc(&d_drift, sizeof(DriftCoeffs<double>) * (size_t)(n_turns + 1)));
    BLMI_HIP_TRY(hipMalloc(&d_base, sizeof(int) * 4));
    BLMI_HIP_TRY(hipMemset(d_base, 0, sizeof(int) * 4));
    BLMI_HIP_TRY(hipStreamCreateWithFlags(&own_stream, hipStreamNonBlocking));
    BLMI_HIP_TRY(hipEventCreateWithFlags(&ev_in, hipEventDisableTiming));
    BLMI_HIP_TRY(hipEventCreateWithFlags(&ev_out, hipEventDisableTiming));
    return BLMI_OK;
}

int64_t Ensemble::total_particles() const {
    int64_t s = 0;
    for (int64_t c : cnt) s += c;
    return s;
}

int Ensemble::set_particles(double *dt_, double *dE_, const int64_t *off_, const int64_t *cnt_) {
    if (((((uintptr_t)dt_) | ((uintptr_t)dE_)) & 15u) != 0)
        return blmi_set_error(BLMI_ERR_ARG, "ensemble: dt/dE must be 16-byte aligned");
    for (int b = 0; b < n_bunches; ++b) {
        if (off_[b] < 0 || (off_[b] & 1) || cnt_[b] < 0)
            return blmi_set_error(BLMI_ERR_ARG, "ensemble: bunch %d offset %lld must be even, count >= 0",
                                  b, (long long)off_[b]);
        if (b > 0 && off_[b] < off_[b - 1] + cnt_[b - 1] + (cnt_[b - 1] & 1))
            return blmi_set_error(BLMI_ERR_ARG, "ensemble: bunch %d overlaps bunch %d (odd bunches "
                                  "need one pad element)", b, b - 1);
    }
    dt = dt_;
    dE = dE_;
    off.assign(off_, off_ + n_bunches);
    cnt.assign(cnt_, cnt_ + n_bunches);
    return build_segments();
}

// Split the particle pairs of all bunches evenly over the grid, in bunch
// order, so a workgroup touches few bunches (one LDS flush per segment).
int Ensemble::build_segments() {
    invalidate_graph();
    if (off.empty()) return BLMI_OK;
    grid = device_cu_count() * wg_per_cu;
    if (grid < 1) grid = 1;
    std::vector<int64_t> pairs(n_bunches);
    int64_t total = 0;
    for (int b = 0; b < n_bunches; ++b) {
        pairs[b] = (cnt[b] + 1) / 2;
        total += pairs[b];
    }
    std::vector<EnsSeg> segs;
    std::vector<int> ptr(grid + 1, 0);
    int b = 0;
    int64_t used = 0;  // pairs of bunch b already assigned
    for (int w = 0; w < grid; ++w) {
        ptr[w] = (int)segs.size();
        int64_t quota = total / grid + (w < total % grid ? 1 : 0);
        while (quota > 0 && b < n_bunches) {
            const int64_t left = pairs[b] - used;
            if (left == 0) {
                ++b;
                used = 0;
                continue;
            }
            const int64_t take = std::min(quota, left);
            EnsSeg s;
            s.bunch = b;
            s.pad = 0;
            s.p0 = off[b] / 2 + used;
            s.p1 = s.p0 + take;
            s.end = off[b] + cnt[b];
            segs.push_back(s);
            used += take;
            quota -= take;
        }
    }
    ptr[grid] = (int)segs.size();
    (void)hipFree(d_segs);
    (void)hipFree(d_seg_ptr);
    d_segs = nullptr;
    d_seg_ptr = nullptr;
    BLMI_HIP_TRY(hipMalloc(&d_segs, sizeof(EnsSeg) * std::max<size_t>(segs.size(), 1)));
    BLMI_HIP_TRY(hipMalloc(&d_seg_ptr, sizeof(int) * (grid + 1)));
    if (!segs.empty())
        BLMI_HIP_TRY(hipMemcpy(d_segs, segs.data(), sizeof(EnsSeg) * segs.size(),
                               hipMemcpyHostToDevice));
    BLMI_HIP_TRY(hipMemcpy(d_seg_ptr, ptr.data(), sizeof(int) * (grid + 1), hipMemcpyHostToDevice));
    return BLMI_OK;
}

int Ensemble::set_programs(const double *rf_host, const double *acc_host, const double *t_rev,
                           double length_ratio, double alpha_order, const double *eta0,
                           const double *eta1, const double *eta2, const double *alpha0,
                           const double *alpha1, const double *alpha2, const double *beta,
                           const double *energy, double charge_) {
    const size_t nt = (size_t)n_turns + 1;
    charge = charge_;
    BLMI_HIP_TRY(hipMemcpy(d_rf, rf_host, sizeof(double) * 3 * n_rf * nt * n_bunches,
                           hipMemcpyHostToDevice));
    BLMI_HIP_TRY(hipMemcpy(d_acc, acc_host, sizeof(double) * nt, hipMemcpyHostToDevice));
    std::vector<DriftCoeffs<double>> dc(nt);
    for (size_t i = 0; i < nt; ++i)
        dc[i] = make_drift_coeffs<double>(solver, t_rev[i], length_ratio, alpha_order, eta0[i],
                                          eta1[i], eta2[i], alpha0[i], alpha1[i], alpha2[i],
                                          beta[i], energy[i]);
    BLMI_HIP_TRY(hipMemcpy(d_drift, dc.data(), sizeof(DriftCoeffs<double>) * nt,
                           hipMemcpyHostToDevice));
    invalidate_graph();
    return BLMI_OK;
}

int Ensemble::set_slices(int n_slices_, const double *cut_left, const double *cut_right,
                         double *profiles_) {
    if ((size_t)n_slices_ * sizeof(unsigned) > 131072)
        return blmi_set_error(BLMI_ERR_UNSUPPORTED, "ensemble: %d slices exceed the LDS histogram",
                              n_slices_);
    n_slices = n_slices_;
    profiles = profiles_;
    std::vector<double> ibw(n_bunches);
    for (int b = 0; b < n_bunches; ++b) ibw[b] = (double)n_slices / (cut_right[b] - cut_left[b]);
    (void)hipFree(d_cut_left);
    (void)hipFree(d_inv_bw);
    (void)hipFree(d_counts);
    BLMI_HIP_TRY(hipMalloc(&d_cut_left, sizeof(double) * n_bunches));
    BLMI_HIP_TRY(hipMalloc(&d_inv_bw, sizeof(double) * n_bunches));
    BLMI_HIP_TRY(hipMemcpy(d_cut_left, cut_left, sizeof(double) * n_bunches, hipMemcpyHostToDevice));
    BLMI_HIP_TRY(hipMemcpy(d_inv_bw, ibw.data(), sizeof(double) * n_bunches, hipMemcpyHostToDevice));
    const size_t cbytes = sizeof(unsigned) * (size_t)n_bunches * n_slices;
    BLMI_HIP_TRY(hipMalloc(&d_counts, cbytes));
    BLMI_HIP_TRY(hipMemset(d_counts, 0, cbytes));
    invalidate_graph();
    return BLMI_OK;
}

int Ensemble::set_induced(int64_t n_fft_, const double *Z_host, const double *scale, double *volt_,
                          const double *bc_) {
    if (n_slices <= 0) return blmi_set_error(BLMI_ERR_ARG, "ensemble: set_slices first");
    // the particle kernel holds a segment's counts and table in LDS
    if (sizeof(unsigned) * ((n_slices + 3) & ~3) + sizeof(double2) * (n_slices - 1) > 150 * 1024)
        return blmi_set_error(BLMI_ERR_UNSUPPORTED,
                              "ensemble: %d slices with intensity effects exceed LDS", n_slices);
    int rc = fft_plan_get(n_fft_, &plan);
    if (rc) return rc;
    if (n_slices > n_fft_)
        return blmi_set_error(BLMI_ERR_ARG, "ensemble: n_slices %d > n_fft %lld", n_slices,
                              (long long)n_fft_);
    n_fft = n_fft_;
    volt = volt_;
    bc = bc_;
    const int64_t nz = n_fft / 2 + 1;
    (void)hipFree(d_Z);
    (void)hipFree(d_table);
    (void)hipFree(d_iargs);
    const int ts = 2 * n_slices + 2;
    BLMI_HIP_TRY(hipMalloc(&d_Z, sizeof(double2) * nz * n_bunches));
    BLMI_HIP_TRY(hipMemcpy(d_Z, Z_host, sizeof(double2) * nz * n_bunches, hipMemcpyHostToDevice));
    BLMI_HIP_TRY(hipMalloc(&d_table, sizeof(double) * ts * (size_t)n_bunches));
    BLMI_HIP_TRY(hipMemset(d_table, 0, sizeof(double) * ts * (size_t)n_bunches));
    std::vector<InducedArgs> ia(n_bunches);
    for (int b = 0; b < n_bunches; ++b) {
        InducedArgs &a = ia[b];
        a = InducedArgs{};
        a.counts0 = d_counts + (size_t)b * n_slices;
        a.groups = 1;
        a.profile_out = profiles + (size_t)b * n_slices;
        a.n_slices = n_slices;
        a.Z = d_Z + (size_t)b * nz;
        a.scale = scale[b];
        a.volt = volt + (size_t)b * n_slices;
        a.n_out = n_slices;
        a.bc = bc + (size_t)b * n_slices;
        a.charge = charge;
        a.tab = reinterpret_cast<double2 *>(d_table + (size_t)b * ts);
        a.tab_params = d_table + (size_t)b * ts + 2 * n_slices;
    }
    BLMI_HIP_TRY(hipMalloc(&d_iargs, sizeof(InducedArgs) * n_bunches));
    BLMI_HIP_TRY(hipMemcpy(d_iargs, ia.data(), sizeof(InducedArgs) * n_bunches,
                           hipMemcpyHostToDevice));
    intensity = true;
    invalidate_graph();
    return BLMI_OK;
}

int Ensemble::set_launch(int block_, int wg_per_cu_, int graph_chunk_, int max_peel_) {
    if (block_ < 64 || block_ > 1024 || (block_ & 63))
        return blmi_set_error(BLMI_ERR_ARG, "ensemble: block must be a multiple of 64 in [64, 1024]");
    if (wg_per_cu_ < 1 || wg_per_cu_ > 64 || graph_chunk_ < 1 || graph_chunk_ > 1000 ||
        max_peel_ < 0 || max_peel_ > 64)
        return blmi_set_error(BLMI_ERR_ARG, "ensemble: bad launch parameters");
    block = block_;
    wg_per_cu = wg_per_cu_;
    graph_chunk = graph_chunk_;
    max_peel = max_peel_;
    return build_segments();
}

int Ensemble::set_turn(int turn) {
    if (turn < 0 || turn >= n_turns)
        return blmi_set_error(BLMI_ERR_ARG, "turn %d outside [0, %d)", turn, n_turns);
    BLMI_HIP_TRY(hipMemcpy(d_base, &turn, sizeof(int), hipMemcpyHostToDevice));
    host_turn = turn;
    if (d_counts)
        BLMI_HIP_TRY(hipMemset(d_counts, 0, sizeof(unsigned) * (size_t)n_bunches * n_slices));
    return BLMI_OK;
}

static EnsTurnArgs turn_args(const Ensemble &e, int off) {
    EnsTurnArgs a{};
    a.n_rf = e.n_rf;
    a.charge = e.charge;
    a.rf = e.d_rf;
    a.rf_bunch_stride = (int64_t)3 * e.n_rf * (e.n_turns + 1);
    a.acc = e.d_acc;
    a.drift = e.d_drift;
    a.base = e.d_base;
    a.off = off;
    a.last_turn = e.n_turns - 1;
    a.table = e.d_table;
    a.tab_stride = 2 * e.n_slices + 2;
    a.cut_left = e.d_cut_left;
    a.inv_bw = e.d_inv_bw;
    a.counts = e.d_counts;
    a.n_slices = e.n_slices;
    a.seg_ptr = e.d_seg_ptr;
    a.segs = e.d_segs;
    a.max_peel = e.max_peel;
    return a;
}

static int ens_lds_attr(const void *fn, size_t lds) {
    if (lds <= 65536) return BLMI_OK;
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess)
        return blmi_set_error(BLMI_ERR_HIP, "ensemble LDS attribute: %s", hipGetErrorString(e));
    return BLMI_OK;
}

int Ensemble::prime(hipStream_t s) {
    if (!intensity || !d_counts) return BLMI_OK;
    if (!d_segs) return blmi_set_error(BLMI_ERR_ARG, "ensemble: no particles");
    int rc = launch_fill<unsigned>(d_counts, (int64_t)n_bunches * n_slices, 0u, s);
    if (rc) return rc;
    const size_t lds = sizeof(unsigned) * n_slices;
    rc = ens_lds_attr((const void *)ensemble_prime_kernel, lds);
    if (rc) return rc;
    hipLaunchKernelGGL(ensemble_prime_kernel, dim3(grid), dim3(block), lds, s, dt,
                       turn_args(*this, 0));
    return blmi_check_launch("ensemble_prime");
}

using EnsFn = void (*)(double *, double *, EnsTurnArgs);

template <int S>
static EnsFn pick_ens(bool interp, bool hist) {
    if (interp) return hist ? ensemble_turn_kernel<S, true, true> : ensemble_turn_kernel<S, true, false>;
    return hist ? ensemble_turn_kernel<S, false, true> : ensemble_turn_kernel<S, false, false>;
}

int Ensemble::launch_turn(hipStream_t s, int k, const KernelEvents &ev) {
    const bool hist = n_slices > 0 && d_counts;
    const bool interp = intensity;
    EnsFn fn = solver == 0 ? pick_ens<0>(interp, hist)
               : solver == 1 ? pick_ens<1>(interp, hist)
                             : pick_ens<2>(interp, hist);
    size_t lds = hist ? sizeof(unsigned) * ((n_slices + 3) & ~3) : 0;
    if (interp) lds = sizeof(unsigned) * ((n_slices + 3) & ~3) + sizeof(double2) * (n_slices - 1);
    int rc = ens_lds_attr((const void *)fn, lds);
    if (rc) return rc;
    launch(fn, dim3(grid), dim3(block), lds, s, ev, dt, dE, turn_args(*this, k));
    return blmi_check_launch("ensemble_turn");
}

int Ensemble::launch_induced(hipStream_t s, const KernelEvents &ev) {
    return launch_induced_lds_batch(d_iargs, n_bunches, n_slices, plan, s, ev);
}

int Ensemble::launch_publish(hipStream_t s) {
    hipLaunchKernelGGL(ensemble_publish_kernel, dim3(n_bunches), dim3(256), 0, s, d_counts,
                       profiles, n_slices);
    return blmi_check_launch("ensemble_publish");
}

int Ensemble::launch_turns(hipStream_t s, int count, bool advance) {
    const KernelEvents none;
    const bool hist = n_slices > 0 && d_counts;
    for (int k = 0; k < count; ++k) {
        int rc = BLMI_OK;
        if (intensity) rc = launch_induced(s, none);
        if (!rc) rc = launch_turn(s, k, none);
        if (!rc && !intensity && hist) rc = launch_publish(s);
        if (rc) return rc;
    }
    if (advance) {
        hipLaunchKernelGGL(ensemble_advance_kernel, dim3(1), dim3(1), 0, s, d_base, count);
        return blmi_check_launch("ensemble_advance");
    }
    return BLMI_OK;
}

int Ensemble::track(int turns, hipStream_t user, bool use_graph) {
    if (turns <= 0) return BLMI_OK;
    if (!d_segs) return blmi_set_error(BLMI_ERR_ARG, "ensemble: no particles");
    if (host_turn + turns > n_turns)
        return blmi_set_error(BLMI_ERR_ARG, "tracking %d turns from turn %d passes n_turns=%d",
                              turns, host_turn, n_turns);
    BLMI_HIP_TRY(hipEventRecord(ev_in, user));
    BLMI_HIP_TRY(hipStreamWaitEvent(own_stream, ev_in, 0));
    int done = 0, rc = BLMI_OK;
    if (use_graph && turns >= graph_chunk) {
        if (!graph_exec || graph_turns != graph_chunk) {
            invalidate_graph();
            BLMI_HIP_TRY(hipStreamBeginCapture(own_stream, hipStreamCaptureModeThreadLocal));
            rc = launch_turns(own_stream, graph_chunk, true);
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
            graph_turns = graph_chunk;
        }
        while (turns - done >= graph_turns) {
            BLMI_HIP_TRY(hipGraphLaunch(graph_exec, own_stream));
            done += graph_turns;
        }
    }
    if (done < turns) rc = launch_turns(own_stream, turns - done, true);
    if (rc) return rc;
    host_turn += turns;
    BLMI_HIP_TRY(hipEventRecord(ev_out, own_stream));
    BLMI_HIP_TRY(hipStreamWaitEvent(user, ev_out, 0));
    return BLMI_OK;
}

int Ensemble::time_turns(int turns, hipStream_t user, float *ms_induced, float *ms_turn) {
    if (turns <= 0) return blmi_set_error(BLMI_ERR_ARG, "time_turns: turns must be positive");
    if (host_turn + turns > n_turns)
        return blmi_set_error(BLMI_ERR_ARG, "time_turns: %d turns from turn %d passes n_turns=%d",
                              turns, host_turn, n_turns);
    BLMI_HIP_TRY(hipEventRecord(ev_in, user));
    BLMI_HIP_TRY(hipStreamWaitEvent(own_stream, ev_in, 0));
    std::vector<hipEvent_t> ev(4 * (size_t)turns, nullptr);
    for (auto &x : ev) BLMI_HIP_TRY(hipEventCreate(&x));
    const bool hist = n_slices > 0 && d_counts;
    int rc = BLMI_OK;
    for (int k = 0; k < turns && rc == BLMI_OK; ++k) {
        KernelEvents ei{ev[4 * k], ev[4 * k + 1]}, et{ev[4 * k + 2], ev[4 * k + 3]};
        if (intensity) rc = launch_induced(own_stream, ei);
        if (rc == BLMI_OK) rc = launch_turn(own_stream, k, et);
        if (rc == BLMI_OK && !intensity && hist) rc = launch_publish(own_stream);
    }
    if (rc == BLMI_OK) {
        hipLaunchKernelGGL(ensemble_advance_kernel, dim3(1), dim3(1), 0, own_stream, d_base, turns);
        rc = blmi_check_launch("ensemble_advance");
    }
    if (rc == BLMI_OK) {
        host_turn += turns;
        BLMI_HIP_TRY(hipStreamSynchronize(own_stream));
        double x = 0, y = 0;
        for (int k = 0; k < turns; ++k) {
            float a = 0, b = 0;
            if (intensity) BLMI_HIP_TRY(hipEventElapsedTime(&a, ev[4 * k], ev[4 * k + 1]));
            BLMI_HIP_TRY(hipEventElapsedTime(&b, ev[4 * k + 2], ev[4 * k + 3]));
            x += a;
            y += b;
        }
        *ms_induced = (float)(x / turns);
        *ms_turn = (float)(y / turns);
    }
    BLMI_HIP_TRY(hipStreamSynchronize(own_stream));
    for (auto &x : ev) (void)hipEventDestroy(x);
    BLMI_HIP_TRY(hipEventRecord(ev_out, own_stream));
    BLMI_HIP_TRY(hipStreamWaitEvent(user, ev_out, 0));
    return rc;
}

// Launch-shape autotuner: every candidate (workgroup size x workgroups per
// CU, within 2048 resident threads per CU) tracks `turns` eager turns from
// the same saved state; the fastest (induced + particle kernel time per
// turn) is kept and the state restored.
int Ensemble::autotune(int turns, hipStream_t user, int *best_block, int *best_wpc, float *best_ms) {
    if (!d_segs) return blmi_set_error(BLMI_ERR_ARG, "ensemble: no particles");
    if (turns < 1 || host_turn + turns > n_turns)
        return blmi_set_error(BLMI_ERR_ARG, "autotune: %d turns from turn %d passes n_turns=%d",
                              turns, host_turn, n_turns);
    BLMI_HIP_TRY(hipStreamSynchronize(user));
    const int64_t span = off.back() + cnt.back() + (cnt.back() & 1);
    const size_t pbytes = sizeof(double) * (size_t)span;
    const size_t cbytes = d_counts ? sizeof(unsigned) * (size_t)n_bunches * n_slices : 0;
    // scratch copies of the state, freed on every return path
    struct Scratch {
        void *p = nullptr;
        ~Scratch() { (void)hipFree(p); }
    } save_buf, csave_buf;
    BLMI_HIP_TRY(hipMalloc(&save_buf.p, 2 * pbytes));
    if (cbytes) BLMI_HIP_TRY(hipMalloc(&csave_buf.p, cbytes));
    double *save = static_cast<double *>(save_buf.p);
    unsigned *csave = static_cast<unsigned *>(csave_buf.p);
    auto snapshot = [&](bool restore) -> int {
        hipMemcpyKind k = hipMemcpyDeviceToDevice;
        if (restore) {
            BLMI_HIP_TRY(hipMemcpy(dt, save, pbytes, k));
            BLMI_HIP_TRY(hipMemcpy(dE, save + span, pbytes, k));
            if (cbytes) BLMI_HIP_TRY(hipMemcpy(d_counts, csave, cbytes, k));
        } else {
            BLMI_HIP_TRY(hipMemcpy(save, dt, pbytes, k));
            BLMI_HIP_TRY(hipMemcpy(save + span, dE, pbytes, k));
            if (cbytes) BLMI_HIP_TRY(hipMemcpy(csave, d_counts, cbytes, k));
        }
        return BLMI_OK;
    };
    const int turn0 = host_turn;
    const int blk0 = block, wpc0 = wg_per_cu;
    int rc = snapshot(false);
    float best = 1e30f;
    int bb = block, bw = wg_per_cu;
    const int blocks[3] = {256, 512, 1024};
    for (int bi = 0; bi < 3 && rc == BLMI_OK; ++bi) {
        for (int w = 1; w * blocks[bi] <= 2048 && rc == BLMI_OK; w *= 2) {
            block = blocks[bi];
            wg_per_cu = w;
            rc = build_segments();
            float mi = 0, mt = 0;
            if (rc == BLMI_OK) rc = time_turns(turns, own_stream, &mi, &mt);
            if (rc == BLMI_OK) {
                BLMI_HIP_TRY(hipStreamSynchronize(own_stream));
                if (mi + mt < best) {
                    best = mi + mt;
                    bb = block;
                    bw = wg_per_cu;
                }
                rc = snapshot(true);
                if (rc == BLMI_OK) {
                    host_turn = turn0;
                    BLMI_HIP_TRY(hipMemcpy(d_base, &turn0, sizeof(int), hipMemcpyHostToDevice));
                }
            }
        }
    }
    block = rc == BLMI_OK ? bb : blk0;
    wg_per_cu = rc == BLMI_OK ? bw : wpc0;
    int rc2 = build_segments();
    if (rc) return rc;
    if (rc2) return rc2;
    *best_block = bb;
    *best_wpc = bw;
    *best_ms = best;
    return BLMI_OK;
}

}  // namespace blmi
