// Fused tracking turn.
//
// One turn of the reference's map is, with intensity effects
// (EX_18_robinson_instability.py:226: profile, total induced voltage, tracker):
//   profile(dt_t) -> V_t -> linear_interp_kick(V_t) -> RF kick(t) -> drift(t+1)
// and without them (EX_01_Acceleration.py:152-153: tracker, profile):
//   RF kick(t) -> drift(t+1) -> profile(dt_{t+1}).
//
// Here a turn is at most two launches:
//   induced(t): staged counts of dt_t -> profile (f64, user visible) -> V_t
//               (user visible) -> interpolation table (slope, intercept)
//   turn(t):    per particle: interp kick, RF kick, drift, bin of dt_{t+1}
//               into a workgroup-private LDS histogram; the histogram is
//               flushed to staging counts with device-scope atomics.
// dt and dE are read once and written once per turn (32 B/particle in FP64).
//
// Every per-turn scalar (RF voltage/frequency/phase, acceleration kick, drift
// coefficients) lives in device tables indexed by a device-side turn counter,
// so a captured graph of K turns replays unchanged for any turn range.  The
// counter is advanced by the last workgroup of the turn kernel to finish
// (ticket counter), after every workgroup has read it.
//
// Staging counts are double buffered by turn parity: turn(t) accumulates
// into counts[(t+1)&1] while induced(t) reads counts[t&1] and clears
// counts[(t+1)&1] beforehand; so no memset node is ever needed.
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "blmi_common.h"
#include "engine.h"
#include "particles.h"
#include "spectral.h"

namespace blmi {

// ---- shared per-particle arithmetic (same order as the single kernels) ----
template <typename T>
__device__ __forceinline__ T e_rf_kick(T t, int n_rf, T charge, const T *__restrict__ V,
                                       const T *__restrict__ w, const T *__restrict__ phi, T acc) {
    T s = T(0);
    for (int j = 0; j < n_rf; ++j) s = s + V[j] * Trig<T>::sin(w[j] * t + phi[j]);
    return charge * s + acc;
}

template <typename T>
__device__ __forceinline__ T e_drift(const DriftCoeffs<T> &c, T e, int solver) {
    if (solver == 0) return c.coeff * e;
    if (solver == 1) {
        T den;
        if (c.alpha_order == 0)
            den = T(1) - c.e0 * e;
        else if (c.alpha_order == 1)
            den = T(1) - c.e0 * e - c.e1 * e * e;
        else
            den = T(1) - c.e0 * e - c.e1 * e * e - c.e2 * e * e * e;
        return c.tl * (T(1) / den - T(1));
    }
    const T delta = sqrt(T(1) + c.ib2 * (e * e * c.ie2 + T(2) * e / c.energy)) - T(1);
    return c.tl * ((T(1) + c.a0 * delta + c.a1 * (delta * delta) + c.a2 * (delta * delta * delta)) *
                       (T(1) + e / c.energy) / (T(1) + delta) -
                   T(1));
}

struct TurnArgs {
    int n_rf;
    double charge;
    const double *rf;                  // [turn][3][n_rf]: V, omega, phi
    const double *acc;                 // [turn]
    const DriftCoeffs<double> *drift;  // [turn]
    int *turn;                         // device turn counter
    int last_turn;                     // n_turns - 1
    unsigned *ticket;                  // zero between launches
    // interpolation kick (intensity effects)
    const double *vk, *fac, *tab_params;
    int n_tab;
    // slicing
    unsigned *counts0, *counts1;
    int n_slices;
    double cut_left, inv_bw;
    // plain mode: last workgroup publishes counts as the f64 profile
    double *profile;
};

template <int SOLVER, bool INTERP, bool HIST, bool LDSH, bool PUBLISH>
__global__ void __launch_bounds__(512) turn_kernel(double *__restrict__ dt, double *__restrict__ dE,
                                                   int64_t n, TurnArgs a) {
    extern __shared__ unsigned lds_hist[];
    int t = *a.turn;
    if (t < 0) t = 0;
    if (t > a.last_turn) t = a.last_turn;  // host checks the range; never read past the tables
    const double *V = a.rf + (size_t)t * 3 * a.n_rf;
    const double *w = V + a.n_rf;
    const double *phi = w + a.n_rf;
    const double acc = a.acc[t];
    const DriftCoeffs<double> dc = a.drift[t + 1];
    unsigned *counts = ((t + 1) & 1) ? a.counts1 : a.counts0;
    if (PUBLISH) counts = a.counts0;
    double bc0 = 0.0, ibw = 0.0;
    if (INTERP) {
        bc0 = a.tab_params[0];
        ibw = a.tab_params[1];
    }
    if (HIST && LDSH) {
        for (int b = threadIdx.x; b < a.n_slices; b += blockDim.x) lds_hist[b] = 0u;
        __syncthreads();
    }
    auto step = [&](double &x, double &e) {
        if (INTERP) {
            const double f = floor((x - bc0) * ibw);
            if (f >= 0.0 && f < (double)a.n_tab) {
                const int k = (int)f;
                e = e + (x * a.vk[k] + a.fac[k]);
            }
        }
        e = e + e_rf_kick(x, a.n_rf, a.charge, V, w, phi, acc);
        x = x + e_drift(dc, e, SOLVER);
        if (HIST) {
            const double f = floor((x - a.cut_left) * a.inv_bw);
            if (f >= 0.0 && f < (double)a.n_slices) {
                if (LDSH)
                    atomicAdd(&lds_hist[(int)f], 1u);
                else
                    atomic_add_agent(&counts[(int)f], 1u);
            }
        }
    };
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t n2 = n / 2;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += stride) {
        double2 x = reinterpret_cast<const double2 *>(dt)[i];
        double2 e = reinterpret_cast<const double2 *>(dE)[i];
        step(x.x, e.x);
        step(x.y, e.y);
        reinterpret_cast<double2 *>(dt)[i] = x;
        reinterpret_cast<double2 *>(dE)[i] = e;
    }
    if ((n & 1) && blockIdx.x == gridDim.x - 1 && threadIdx.x == blockDim.x - 1) {
        double x = dt[n - 1], e = dE[n - 1];
        step(x, e);
        dt[n - 1] = x;
        dE[n - 1] = e;
    }
    if (HIST && LDSH) {
        __syncthreads();
        for (int b = threadIdx.x; b < a.n_slices; b += blockDim.x) {
            const unsigned c = lds_hist[b];
            if (c) atomic_add_agent(&counts[b], c);
        }
    }
    // ticket: the last workgroup to finish advances the turn (and in plain
    // mode publishes the profile and clears the counts)
    __shared__ int is_last;
    __syncthreads();
    if (threadIdx.x == 0) {
        __atomic_thread_fence(__ATOMIC_RELEASE);  // this block's count updates first
        const unsigned prev = __hip_atomic_fetch_add(a.ticket, 1u, __ATOMIC_ACQ_REL,
                                                     __HIP_MEMORY_SCOPE_AGENT);
        is_last = (prev == gridDim.x - 1);
    }
    __syncthreads();
    if (is_last) {
        __atomic_thread_fence(__ATOMIC_ACQUIRE);
        if (PUBLISH) {
            for (int b = threadIdx.x; b < a.n_slices; b += blockDim.x) {
                const unsigned c = __hip_atomic_load(&counts[b], __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT);
                a.profile[b] = (double)c;
                __hip_atomic_store(&counts[b], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        if (threadIdx.x == 0) {
            __hip_atomic_store(a.turn, t + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(a.ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// Histogram of the current dt into counts[turn & 1] (engine priming).
__global__ void prime_hist_kernel(const double *__restrict__ dt, int64_t n, const int *turn,
                                  unsigned *counts0, unsigned *counts1, int n_slices,
                                  double cut_left, double inv_bw) {
    const int t = *turn;
    unsigned *counts = (t & 1) ? counts1 : counts0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double f = floor((dt[i] - cut_left) * inv_bw);
        if (f >= 0.0 && f < (double)n_slices) atomic_add_agent(&counts[(int)f], 1u);
    }
}

// ---------------------------------------------------------------------------
// Engine object
// ---------------------------------------------------------------------------
Engine::~Engine() { release(); }

void Engine::release() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph = nullptr;
    (void)hipFree(d_rf);
    (void)hipFree(d_acc);
    (void)hipFree(d_drift);
    (void)hipFree(d_turn);
    (void)hipFree(d_ticket);
    (void)hipFree(d_counts);
    (void)hipFree(d_table);
    (void)hipFree(d_Z);
    d_rf = d_acc = nullptr;
    d_drift = nullptr;
    d_turn = nullptr;
    d_ticket = nullptr;
    d_counts = nullptr;
    d_table = nullptr;
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
    BLMI_HIP_TRY(hipMalloc(&d_turn, sizeof(int)));
    BLMI_HIP_TRY(hipMalloc(&d_ticket, sizeof(unsigned)));
    BLMI_HIP_TRY(hipMemset(d_turn, 0, sizeof(int)));
    BLMI_HIP_TRY(hipMemset(d_ticket, 0, sizeof(unsigned)));
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

int Engine::set_slices(int n_slices_, double cut_left_, double cut_right_, double *profile_) {
    n_slices = n_slices_;
    cut_left = cut_left_;
    cut_right = cut_right_;
    inv_bw = (double)n_slices / (cut_right - cut_left);  // histogram.cpp:29
    profile = profile_;
    (void)hipFree(d_counts);
    d_counts = nullptr;
    if (n_slices > 0) {
        BLMI_HIP_TRY(hipMalloc(&d_counts, sizeof(unsigned) * 2 * (size_t)n_slices));
        BLMI_HIP_TRY(hipMemset(d_counts, 0, sizeof(unsigned) * 2 * (size_t)n_slices));
    }
    invalidate_graph();
    return BLMI_OK;
}

int Engine::set_induced(int64_t n_fft, const double *Z_host, double scale_, double *volt_,
                        const double *bc_, double charge_) {
    int rc = fft_plan_get(n_fft, &plan);
    if (rc) return rc;
    scale = scale_;
    volt = volt_;
    bc = bc_;
    charge = charge_;
    const size_t nz = (size_t)plan.M + 1;
    (void)hipFree(d_Z);
    (void)hipFree(d_table);
    BLMI_HIP_TRY(hipMalloc(&d_Z, sizeof(double2) * nz));
    BLMI_HIP_TRY(hipMemcpy(d_Z, Z_host, sizeof(double2) * nz, hipMemcpyHostToDevice));
    BLMI_HIP_TRY(hipMalloc(&d_table, sizeof(double) * (2 * (size_t)n_slices + 2)));
    BLMI_HIP_TRY(hipMemset(d_table, 0, sizeof(double) * (2 * (size_t)n_slices + 2)));
    intensity = true;
    invalidate_graph();
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
    BLMI_HIP_TRY(hipMemcpy(d_turn, &turn, sizeof(int), hipMemcpyHostToDevice));
    host_turn = turn;
    return BLMI_OK;
}

int Engine::get_turn(int *turn) {
    BLMI_HIP_TRY(hipMemcpy(turn, d_turn, sizeof(int), hipMemcpyDeviceToHost));
    return BLMI_OK;
}

int Engine::prime(hipStream_t s) {
    if (!intensity || n_slices <= 0) return BLMI_OK;
    int rc = launch_fill<unsigned>(d_counts, 2 * (int64_t)n_slices, 0u, s);
    if (rc) return rc;
    hipLaunchKernelGGL(prime_hist_kernel, dim3(stream_grid(n, 256)), dim3(256), 0, s, dt, n, d_turn,
                       d_counts, d_counts + n_slices, n_slices, cut_left, inv_bw);
    return blmi_check_launch("engine_prime");
}

template <int SOLVER>
static int launch_turn_solver(const Engine &e, const TurnArgs &a, bool interp, bool hist, bool ldsh,
                              bool publish, int grid, int block, size_t lds, hipStream_t s) {
#define BLMI_TURN(I, H, L, P)                                                                   \
    hipLaunchKernelGGL((turn_kernel<SOLVER, I, H, L, P>), dim3(grid), dim3(block), lds, s, e.dt, \
                       e.dE, e.n, a)
    if (interp) {
        if (hist) {
            if (ldsh) BLMI_TURN(true, true, true, false); else BLMI_TURN(true, true, false, false);
        } else {
            BLMI_TURN(true, false, false, false);
        }
    } else {
        if (hist) {
            if (publish) {
                if (ldsh) BLMI_TURN(false, true, true, true); else BLMI_TURN(false, true, false, true);
            } else {
                if (ldsh) BLMI_TURN(false, true, true, false); else BLMI_TURN(false, true, false, false);
            }
        } else {
            BLMI_TURN(false, false, false, false);
        }
    }
#undef BLMI_TURN
    return blmi_check_launch("turn");
}

InducedArgs Engine::induced_args() const {
    // staged counts of dt_t -> profile, V_t and the interpolation table; the
    // kernel picks counts[t&1] from the device turn counter and clears them
    // after reading (turn(t) accumulates into the other buffer)
    InducedArgs ia{};
    ia.turn = d_turn;
    ia.counts0 = d_counts;
    ia.counts1 = d_counts + n_slices;
    ia.profile = profile;
    ia.n_slices = n_slices;
    ia.Z = reinterpret_cast<const double2 *>(d_Z);
    ia.scale = scale;
    ia.volt = volt;
    ia.n_out = n_slices;
    ia.bc = bc;
    ia.charge = charge;
    ia.vk = d_table;
    ia.fac = d_table + n_slices;
    ia.tab_params = d_table + 2 * n_slices;
    return ia;
}

int Engine::launch_turn_kernel(hipStream_t s, bool interp) {
    const bool hist = n_slices > 0 && d_counts != nullptr;
    TurnArgs a{};
    a.n_rf = n_rf;
    a.charge = charge;
    a.rf = d_rf;
    a.acc = d_acc;
    a.drift = d_drift;
    a.turn = d_turn;
    a.last_turn = n_turns - 1;
    a.ticket = d_ticket;
    a.vk = d_table;
    a.fac = d_table ? d_table + n_slices : nullptr;
    a.tab_params = d_table ? d_table + 2 * n_slices : nullptr;
    a.n_tab = n_slices - 1;
    a.counts0 = d_counts;
    a.counts1 = d_counts ? d_counts + n_slices : nullptr;
    a.n_slices = n_slices;
    a.cut_left = cut_left;
    a.inv_bw = inv_bw;
    a.profile = profile;
    const size_t lds_bytes = (size_t)n_slices * sizeof(unsigned);
    const bool ldsh = hist && lds_bytes <= 65536;
    const int block = block_size;
    int grid = device_cu_count() * blocks_per_cu;
    const int64_t need = (n / 2 + block - 1) / block;
    if (grid > need) grid = (int)(need > 0 ? need : 1);
    const size_t lds = ldsh ? lds_bytes : 0;
    const bool publish = hist && !interp;
    switch (solver) {
        case 0: return launch_turn_solver<0>(*this, a, interp, hist, ldsh, publish, grid, block, lds, s);
        case 1: return launch_turn_solver<1>(*this, a, interp, hist, ldsh, publish, grid, block, lds, s);
        default: return launch_turn_solver<2>(*this, a, interp, hist, ldsh, publish, grid, block, lds, s);
    }
}

int Engine::launch_one_turn(hipStream_t s) {
    if (intensity) {
        InducedArgs ia = induced_args();
        int rc = launch_induced_voltage_lds(ia, plan, s);
        if (rc) return rc;
    }
    return launch_turn_kernel(s, intensity);
}

int Engine::time_kernels(int turns, hipStream_t user, float *ms_induced, float *ms_turn) {
    // Eager turns with HIP events around each kernel on the engine's stream:
    // average device duration of the induced-voltage and the particle kernels.
    if (turns <= 0) return blmi_set_error(BLMI_ERR_ARG, "time_kernels: turns must be positive");
    if (host_turn + turns > n_turns)
        return blmi_set_error(BLMI_ERR_ARG, "time_kernels: %d turns from turn %d passes n_turns=%d",
                              turns, host_turn, n_turns);
    BLMI_HIP_TRY(hipEventRecord(ev_in, user));
    BLMI_HIP_TRY(hipStreamWaitEvent(own_stream, ev_in, 0));
    std::vector<hipEvent_t> ev(3 * (size_t)turns, nullptr);
    for (auto &x : ev) BLMI_HIP_TRY(hipEventCreate(&x));
    int rc = BLMI_OK;
    const bool was_intensity = intensity;
    for (int k = 0; k < turns && rc == BLMI_OK; ++k) {
        BLMI_HIP_TRY(hipEventRecord(ev[3 * k], own_stream));
        if (was_intensity) {
            InducedArgs ia = induced_args();
            rc = launch_induced_voltage_lds(ia, plan, own_stream);
            if (rc) break;
        }
        BLMI_HIP_TRY(hipEventRecord(ev[3 * k + 1], own_stream));
        rc = launch_turn_kernel(own_stream, was_intensity);
        BLMI_HIP_TRY(hipEventRecord(ev[3 * k + 2], own_stream));
    }
    if (rc == BLMI_OK) {
        BLMI_HIP_TRY(hipEventSynchronize(ev[3 * (size_t)turns - 1]));
        double a = 0, b = 0;
        for (int k = 0; k < turns; ++k) {
            float x = 0, y = 0;
            BLMI_HIP_TRY(hipEventElapsedTime(&x, ev[3 * k], ev[3 * k + 1]));
            BLMI_HIP_TRY(hipEventElapsedTime(&y, ev[3 * k + 1], ev[3 * k + 2]));
            a += x;
            b += y;
        }
        *ms_induced = (float)(a / turns);
        *ms_turn = (float)(b / turns);
        host_turn += turns;
    }
    for (auto &x : ev) (void)hipEventDestroy(x);
    BLMI_HIP_TRY(hipEventRecord(ev_out, own_stream));
    BLMI_HIP_TRY(hipStreamWaitEvent(user, ev_out, 0));
    return rc;
}

int Engine::track(int turns, hipStream_t user, bool use_graph) {
    if (turns <= 0) return BLMI_OK;
    if (!dt || !dE) return blmi_set_error(BLMI_ERR_ARG, "engine has no particle arrays");
    // the per-turn tables hold turns [0, n_turns]; turn t reads entries t and t+1
    if (host_turn + turns > n_turns)
        return blmi_set_error(BLMI_ERR_ARG, "tracking %d turns from turn %d passes n_turns=%d",
                              turns, host_turn, n_turns);
    // order the engine stream after the caller's stream
    BLMI_HIP_TRY(hipEventRecord(ev_in, user));
    BLMI_HIP_TRY(hipStreamWaitEvent(own_stream, ev_in, 0));
    int done = 0;
    if (use_graph && turns >= graph_chunk) {
        if (!graph_exec || graph_turns != graph_chunk) {
            invalidate_graph();
            BLMI_HIP_TRY(hipStreamBeginCapture(own_stream, hipStreamCaptureModeThreadLocal));
            int rc = BLMI_OK;
            for (int k = 0; k < graph_chunk && rc == BLMI_OK; ++k) rc = launch_one_turn(own_stream);
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
    for (; done < turns; ++done) {
        int rc = launch_one_turn(own_stream);
        if (rc) return rc;
    }
    host_turn += turns;
    BLMI_HIP_TRY(hipEventRecord(ev_out, own_stream));
    BLMI_HIP_TRY(hipStreamWaitEvent(user, ev_out, 0));
    return BLMI_OK;
}

}  // namespace blmi
