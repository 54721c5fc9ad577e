// C ABI of the MI355X tracking library (declared in include/blond_mi355x.h).
//
// Conventions: every array argument is a device pointer; lengths of particle
// arrays are int64; every call takes the HIP stream to run on (NULL = the
// default stream) and returns 0 on success or a BLMI_ERR_* code, with a
// message available from blmi_last_error().  Calls are asynchronous with
// respect to the host unless stated otherwise.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "blmi_common.h"
#include "engine.h"
#include "ensemble.h"
#include "adjoint.h"
#include "particles.h"
#include "spectral.h"
#include "wake.h"
#include "feedback.h"

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local char g_err[512] = "";

int blmi_set_error(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

int blmi_check_launch(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return blmi_set_error(BLMI_ERR_HIP, "%s launch failed: %s", what, hipGetErrorString(e));
    return BLMI_OK;
}

namespace blmi {

// ---------------------------------------------------------------------------
// device properties and library workspace (per device)
// ---------------------------------------------------------------------------
struct DeviceInfo {
    int cus = 256;
    int max_lds = 65536;
};
static std::mutex g_mu;
static std::map<int, DeviceInfo> g_info;

static DeviceInfo &info() {
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_info.find(dev);
    if (it != g_info.end()) return it->second;
    DeviceInfo d;
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
        d.cus = v;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess &&
        v > 0)
        d.max_lds = v;
    return g_info[dev] = d;
}

int device_cu_count() { return info().cus; }
int device_max_lds() { return info().max_lds; }

struct Workspace {
    void *ptr[4] = {nullptr, nullptr, nullptr, nullptr};
    size_t size[4] = {0, 0, 0, 0};
    hipEvent_t done[4] = {nullptr, nullptr, nullptr, nullptr};  // after the last user's work
    hipStream_t last[4] = {nullptr, nullptr, nullptr, nullptr};
    bool used[4] = {false, false, false, false};
};
// held from a lease's construction to its destruction (the launches of one
// call); separate from g_mu, which the launch helpers take for device info
static std::mutex g_ws_mu;
static std::map<int, Workspace> g_ws;

WorkspaceLease::WorkspaceLease(size_t bytes, int slot, hipStream_t s) : slot_(slot), s_(s) {
    g_ws_mu.lock();
    (void)hipGetDevice(&dev_);
    Workspace &w = g_ws[dev_];
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    capturing_ = hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
    if (w.size[slot] < bytes || !w.ptr[slot]) {
        if (w.ptr[slot]) (void)hipFree(w.ptr[slot]);  // synchronises the device
        w.ptr[slot] = nullptr;
        w.size[slot] = 0;
        w.used[slot] = false;
        const size_t want = bytes < 4096 ? 4096 : bytes;
        if (hipMalloc(&w.ptr[slot], want) != hipSuccess) {
            w.ptr[slot] = nullptr;
            return;
        }
        w.size[slot] = want;
    }
    if (!capturing_ && w.used[slot] && w.last[slot] != s) (void)hipStreamWaitEvent(s, w.done[slot], 0);
    p_ = w.ptr[slot];
}

WorkspaceLease::~WorkspaceLease() {
    Workspace &w = g_ws[dev_];
    if (p_ && !capturing_) {
        if (!w.done[slot_]) (void)hipEventCreateWithFlags(&w.done[slot_], hipEventDisableTiming);
        if (w.done[slot_] && hipEventRecord(w.done[slot_], s_) == hipSuccess) {
            w.last[slot_] = s_;
            w.used[slot_] = true;
        }
    }
    g_ws_mu.unlock();
}

// ---------------------------------------------------------------------------
// Fixed interpolation table for blmi_linear_interp_kick_noalloc: the slopes,
// intercepts and (bc0, inv_bw) of up to kNoAllocMaxSlices slices, for
// callers that cannot allocate device memory.  Every device the library runs
// on holds its own copy (1 MiB) until the process ends, at its own address:
// host code uses the address hipGetSymbolAddress returns for the current
// device (never the array's name, which is a host-side stand-in).  Calls on
// one device share the table, so calls that overlap on one device race.
// ---------------------------------------------------------------------------
static constexpr int kNoAllocMaxSlices = 65536;
static constexpr int kNoAllocMaxDevices = 64;
static __device__ __align__(256) double scratch[2 * kNoAllocMaxSlices + 2];
static double *g_scratch_ptr[kNoAllocMaxDevices] = {};

// Address of the current device's table, looked up once per device.
static int scratch_table(double **out) {
    int dev = 0;
    BLMI_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocMaxDevices)
        return blmi_set_error(BLMI_ERR_UNSUPPORTED, "device %d beyond the fixed table's %d", dev,
                              kNoAllocMaxDevices);
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return blmi_set_error(BLMI_ERR_HIP, "fixed interpolation table: %s",
                                  hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[dev];
    return BLMI_OK;
}

}  // namespace blmi

using namespace blmi;

static inline hipStream_t S(void *s) { return reinterpret_cast<hipStream_t>(s); }

#define BLMI_REQUIRE(cond, ...)                                       \
    do {                                                              \
        if (!(cond)) return blmi_set_error(BLMI_ERR_ARG, __VA_ARGS__); \
    } while (0)

extern "C" {

const char *blmi_last_error(void) { return g_err; }

int blmi_version(void) { return 100; }

int blmi_init(int device) {
    BLMI_HIP_TRY(hipSetDevice(device));
    (void)info();
    double *tab = nullptr;
    return scratch_table(&tab);
}

int blmi_device_info(int *cus, int *max_lds) {
    if (cus) *cus = device_cu_count();
    if (max_lds) *max_lds = device_max_lds();
    return BLMI_OK;
}

// ---------------------------------------------------------------- kick ----
int blmi_kick(const double *dt, double *dE, int n_rf, double charge, const double *voltage,
              const double *omega_rf, const double *phi_rf, int64_t n, double acc_kick,
              void *stream) {
    BLMI_REQUIRE(n >= 0 && n_rf >= 0, "kick: negative size");
    BLMI_REQUIRE(n == 0 || (dt && dE), "kick: null particle array");
    BLMI_REQUIRE(n_rf == 0 || (voltage && omega_rf && phi_rf), "kick: null RF array");
    return launch_kick<double>(dt, dE, n, n_rf, charge, voltage, omega_rf, phi_rf, acc_kick,
                               S(stream));
}

int blmi_kick_f32(const float *dt, float *dE, int n_rf, float charge, const float *voltage,
                  const float *omega_rf, const float *phi_rf, int64_t n, float acc_kick,
                  void *stream) {
    BLMI_REQUIRE(n >= 0 && n_rf >= 0, "kick_f32: negative size");
    BLMI_REQUIRE(n == 0 || (dt && dE), "kick_f32: null particle array");
    return launch_kick<float>(dt, dE, n, n_rf, charge, voltage, omega_rf, phi_rf, acc_kick,
                              S(stream));
}

// --------------------------------------------------------------- drift ----
int blmi_drift(double *dt, const double *dE, int solver, double t_rev, double length_ratio,
               double alpha_order, double eta_0, double eta_1, double eta_2, double alpha_0,
               double alpha_1, double alpha_2, double beta, double energy, int64_t n,
               void *stream) {
    BLMI_REQUIRE(n >= 0, "drift: negative size");
    BLMI_REQUIRE(solver >= 0 && solver <= 2, "drift: solver %d not in {0 simple, 1 legacy, 2 exact}",
                 solver);
    BLMI_REQUIRE(n == 0 || (dt && dE), "drift: null particle array");
    DriftCoeffs<double> c = make_drift_coeffs<double>(solver, t_rev, length_ratio, alpha_order,
                                                      eta_0, eta_1, eta_2, alpha_0, alpha_1,
                                                      alpha_2, beta, energy);
    return launch_drift<double>(dt, dE, n, c, S(stream));
}

int blmi_drift_f32(float *dt, const float *dE, int solver, float t_rev, float length_ratio,
                   float alpha_order, float eta_0, float eta_1, float eta_2, float alpha_0,
                   float alpha_1, float alpha_2, float beta, float energy, int64_t n,
                   void *stream) {
    BLMI_REQUIRE(n >= 0, "drift_f32: negative size");
    BLMI_REQUIRE(solver >= 0 && solver <= 2, "drift_f32: bad solver %d", solver);
    BLMI_REQUIRE(n == 0 || (dt && dE), "drift_f32: null particle array");
    DriftCoeffs<float> c = make_drift_coeffs<float>(solver, t_rev, length_ratio, alpha_order,
                                                    eta_0, eta_1, eta_2, alpha_0, alpha_1,
                                                    alpha_2, beta, energy);
    return launch_drift<float>(dt, dE, n, c, S(stream));
}

// ----------------------------------------------------------- histogram ----
int blmi_histogram(const double *dt, double *profile, double cut_left, double cut_right,
                   int n_slices, int64_t n, void *stream) {
    BLMI_REQUIRE(n >= 0 && n_slices >= 0, "histogram: negative size");
    BLMI_REQUIRE(n_slices == 0 || profile, "histogram: null profile");
    BLMI_REQUIRE(n == 0 || dt, "histogram: null dt");
    BLMI_REQUIRE(cut_right > cut_left, "histogram: cut_right must exceed cut_left");
    return launch_histogram<double>(dt, n, profile, cut_left, cut_right, n_slices, S(stream));
}

int blmi_histogram_f32(const float *dt, float *profile, float cut_left, float cut_right,
                       int n_slices, int64_t n, void *stream) {
    BLMI_REQUIRE(n >= 0 && n_slices >= 0, "histogram_f32: negative size");
    BLMI_REQUIRE(cut_right > cut_left, "histogram_f32: cut_right must exceed cut_left");
    return launch_histogram<float>(dt, n, profile, cut_left, cut_right, n_slices, S(stream));
}

int blmi_sparse_histogram(const double *dt, double *profile, const double *cut_left,
                          const double *cut_right, const double *bunch_indexes,
                          int n_slices_bucket, int n_filled_buckets, int n_pattern, int64_t n,
                          void *stream) {
    BLMI_REQUIRE(n >= 0 && n_slices_bucket >= 0 && n_filled_buckets >= 1 && n_pattern >= 1,
                 "sparse_histogram: bad sizes");
    BLMI_REQUIRE(profile && cut_left && cut_right && bunch_indexes && (n == 0 || dt),
                 "sparse_histogram: null array");
    return launch_sparse_histogram<double>(dt, n, profile, cut_left, cut_right, bunch_indexes,
                                           n_slices_bucket, n_filled_buckets, n_pattern,
                                           S(stream));
}

int blmi_sparse_histogram_f32(const float *dt, float *profile, const float *cut_left,
                              const float *cut_right, const float *bunch_indexes,
                              int n_slices_bucket, int n_filled_buckets, int n_pattern, int64_t n,
                              void *stream) {
    BLMI_REQUIRE(n >= 0 && n_slices_bucket >= 0 && n_filled_buckets >= 1 && n_pattern >= 1,
                 "sparse_histogram_f32: bad sizes");
    BLMI_REQUIRE(profile && cut_left && cut_right && bunch_indexes && (n == 0 || dt),
                 "sparse_histogram_f32: null array");
    return launch_sparse_histogram<float>(dt, n, profile, cut_left, cut_right, bunch_indexes,
                                          n_slices_bucket, n_filled_buckets, n_pattern, S(stream));
}

int blmi_smooth_histogram(const double *dt, double *profile, double cut_left, double cut_right,
                          int n_slices, int64_t n, void *stream) {
    BLMI_REQUIRE(n >= 0 && n_slices >= 0, "smooth_histogram: negative size");
    BLMI_REQUIRE(n_slices == 0 || profile, "smooth_histogram: null profile");
    BLMI_REQUIRE(cut_right > cut_left, "smooth_histogram: cut_right must exceed cut_left");
    WorkspaceLease ws(sizeof(long long) * (size_t)(n_slices + 1), 0, S(stream));
    long long *scratch = (long long *)ws.ptr();
    if (!scratch) return blmi_set_error(BLMI_ERR_ALLOC, "smooth_histogram: workspace allocation");
    return launch_smooth_histogram<double>(dt, n, profile, cut_left, cut_right, n_slices, scratch,
                                           S(stream));
}

int blmi_smooth_histogram_f32(const float *dt, float *profile, float cut_left, float cut_right,
                              int n_slices, int64_t n, void *stream) {
    BLMI_REQUIRE(n >= 0 && n_slices >= 0, "smooth_histogram_f32: negative size");
    BLMI_REQUIRE(cut_right > cut_left, "smooth_histogram_f32: cut_right must exceed cut_left");
    WorkspaceLease ws(sizeof(long long) * (size_t)(n_slices + 1), 0, S(stream));
    long long *scratch = (long long *)ws.ptr();
    if (!scratch) return blmi_set_error(BLMI_ERR_ALLOC, "smooth_histogram_f32: workspace allocation");
    return launch_smooth_histogram<float>(dt, n, profile, cut_left, cut_right, n_slices, scratch,
                                          S(stream));
}

// --------------------------------------------------- linear_interp_kick ----
int blmi_linear_interp_kick(const double *dt, double *dE, const double *voltage,
                            const double *bin_centers, double charge, int n_slices, int64_t n,
                            double acc_kick, void *stream) {
    BLMI_REQUIRE(n >= 0, "linear_interp_kick: negative size");
    BLMI_REQUIRE(n_slices >= 2, "linear_interp_kick: needs at least 2 slices (got %d)", n_slices);
    BLMI_REQUIRE(voltage && bin_centers, "linear_interp_kick: null voltage/bin_centers");
    WorkspaceLease ws(sizeof(double) * (2 * (size_t)n_slices + 2), 1, S(stream));
    double *tab = (double *)ws.ptr();
    if (!tab) return blmi_set_error(BLMI_ERR_ALLOC, "linear_interp_kick: workspace allocation");
    double *vk = tab, *fac = tab + n_slices, *par = tab + 2 * n_slices;
    int rc = launch_interp_table<double>(voltage, bin_centers, n_slices, charge, acc_kick, vk, fac,
                                         par, S(stream));
    if (rc) return rc;
    return launch_interp_kick<double>(dt, dE, n, vk, fac, par, n_slices - 1, S(stream));
}

int blmi_linear_interp_kick_f32(const float *dt, float *dE, const float *voltage,
                                const float *bin_centers, float charge, int n_slices, int64_t n,
                                float acc_kick, void *stream) {
    BLMI_REQUIRE(n >= 0, "linear_interp_kick_f32: negative size");
    BLMI_REQUIRE(n_slices >= 2, "linear_interp_kick_f32: needs at least 2 slices");
    WorkspaceLease ws(sizeof(float) * (2 * (size_t)n_slices + 2), 1, S(stream));
    float *tab = (float *)ws.ptr();
    if (!tab) return blmi_set_error(BLMI_ERR_ALLOC, "linear_interp_kick_f32: workspace allocation");
    float *vk = tab, *fac = tab + n_slices, *par = tab + 2 * n_slices;
    int rc = launch_interp_table<float>(voltage, bin_centers, n_slices, charge, acc_kick, vk, fac,
                                        par, S(stream));
    if (rc) return rc;
    return launch_interp_kick<float>(dt, dE, n, vk, fac, par, n_slices - 1, S(stream));
}

// No device allocation: the table lives in the fixed per-device array above
// (at most kNoAllocMaxSlices slices); otherwise as blmi_linear_interp_kick.
// Every table entry the kick reads is written by this call's table launch.
int blmi_linear_interp_kick_noalloc(const double *dt, double *dE, const double *voltage,
                                    const double *bin_centers, double charge, int n_slices,
                                    int64_t n, double acc_kick, void *stream) {
    BLMI_REQUIRE(n >= 0, "linear_interp_kick_noalloc: negative size");
    BLMI_REQUIRE(n_slices >= 2, "linear_interp_kick_noalloc: needs at least 2 slices (got %d)",
                 n_slices);
    BLMI_REQUIRE(n_slices <= kNoAllocMaxSlices,
                 "linear_interp_kick_noalloc: %d slices exceed the fixed table of %d", n_slices,
                 kNoAllocMaxSlices);
    BLMI_REQUIRE(voltage && bin_centers, "linear_interp_kick_noalloc: null voltage/bin_centers");
    double *tab = nullptr;
    int rc = scratch_table(&tab);
    if (rc) return rc;
    double *vk = tab, *fac = tab + n_slices, *par = tab + 2 * n_slices;
    rc = launch_interp_table<double>(voltage, bin_centers, n_slices, charge, acc_kick, vk, fac,
                                     par, S(stream));
    if (rc) return rc;
    return launch_interp_kick<double>(dt, dE, n, vk, fac, par, n_slices - 1, S(stream));
}

// --------------------------------------------------------- rf_volt_comp ----
int blmi_rf_volt_comp(const double *voltage, const double *omega_rf, const double *phi_rf,
                      const double *bin_centers, int n_rf, int n_bins, double *rf_voltage,
                      void *stream) {
    BLMI_REQUIRE(n_rf >= 0 && n_bins >= 0, "rf_volt_comp: negative size");
    return launch_rf_volt_comp<double>(voltage, omega_rf, phi_rf, bin_centers, n_rf, n_bins,
                                       rf_voltage, S(stream));
}

int blmi_rf_volt_comp_f32(const float *voltage, const float *omega_rf, const float *phi_rf,
                          const float *bin_centers, int n_rf, int n_bins, float *rf_voltage,
                          void *stream) {
    BLMI_REQUIRE(n_rf >= 0 && n_bins >= 0, "rf_volt_comp_f32: negative size");
    return launch_rf_volt_comp<float>(voltage, omega_rf, phi_rf, bin_centers, n_rf, n_bins,
                                      rf_voltage, S(stream));
}

// ------------------------------------------------------- fast_resonator ----
int blmi_fast_resonator(double *re, double *im, const double *freq, const double *R_S,
                        const double *Q, const double *f_r, int n_res, int n_freq, void *stream) {
    BLMI_REQUIRE(n_res >= 0 && n_freq >= 0, "fast_resonator: negative size");
    return launch_fast_resonator(re, im, freq, R_S, Q, f_r, n_res, n_freq, S(stream));
}

// ------------------------------------------------------------------ FFT ----
int blmi_fft_max_lds_length(int64_t *n) {
    *n = fft_lds_max_real();
    return BLMI_OK;
}

int blmi_rfft(const double *in, int64_t n_in, double *out, int64_t n, void *stream) {
    BLMI_REQUIRE(n_in >= 0 && n > 0, "rfft: bad sizes");
    FftPlan p;
    int rc = fft_plan_get(n, &p);
    if (rc) return rc;
    return launch_rfft(in, n_in, reinterpret_cast<double2 *>(out), p, S(stream));
}

int blmi_irfft(const double *in, int64_t m_in, double *out, int64_t n, void *stream) {
    BLMI_REQUIRE(m_in >= 0 && n > 0, "irfft: bad sizes");
    FftPlan p;
    int rc = fft_plan_get(n, &p);
    if (rc) return rc;
    return launch_irfft(reinterpret_cast<const double2 *>(in), m_in, out, n, p, S(stream));
}

int blmi_beam_phase(const double *bin_centers, const double *profile, int n_bins, double alpha,
                    double omega_rf, double phi_rf, double bin_size, double *coeff, void *stream) {
    BLMI_REQUIRE(n_bins >= 1 && bin_centers && profile && coeff, "beam_phase: bad arguments");
    return launch_beam_phase(bin_centers, profile, n_bins, alpha, omega_rf, phi_rf, bin_size, false,
                             coeff, S(stream));
}

int blmi_beam_phase_fast(const double *bin_centers, const double *profile, int n_bins,
                         double omega_rf, double phi_rf, double bin_size, double *coeff,
                         void *stream) {
    BLMI_REQUIRE(n_bins >= 1 && bin_centers && profile && coeff, "beam_phase_fast: bad arguments");
    return launch_beam_phase(bin_centers, profile, n_bins, 0.0, omega_rf, phi_rf, bin_size, true,
                             coeff, S(stream));
}

int blmi_strided_mean(const double *x, int64_t n, int64_t stride, double *out, void *stream) {
    BLMI_REQUIRE(x && out && n >= 1 && stride >= 1, "strided_mean: bad arguments");
    return launch_strided_mean(x, n, stride, out, S(stream));
}

int blmi_interp(const double *x, int64_t n, const double *xp, const double *fp, int64_t m,
                double left, double right, double *out, void *stream) {
    BLMI_REQUIRE(n >= 0 && m >= 1, "interp: xp must not be empty (m = %lld)", (long long)m);
    BLMI_REQUIRE(n == 0 || (x && xp && fp && out), "interp: null array");
    return launch_interp(x, n, xp, fp, m, left, right, out, S(stream));
}

int blmi_gradient(const double *f, int64_t n, double h, double *out, void *stream) {
    BLMI_REQUIRE(n >= 2, "gradient: at least 2 values (edge order 1), got %lld", (long long)n);
    BLMI_REQUIRE(f && out, "gradient: null array");
    return launch_gradient(f, n, h, out, S(stream));
}

int blmi_diff(const double *f, int64_t n, double *out, void *stream) {
    BLMI_REQUIRE(n >= 0, "diff: negative length");
    BLMI_REQUIRE(n < 2 || (f && out), "diff: null array");
    return launch_diff(f, n, out, S(stream));
}

int blmi_mtw_shift_add(const double *mem_in, double *mem_out, const double *time, int64_t n_mem,
                       double t_rev, const double *add, int64_t n_add, void *stream) {
    BLMI_REQUIRE(n_mem >= 0 && n_add >= 0 && n_add <= n_mem, "mtw_shift_add: bad lengths");
    BLMI_REQUIRE(n_mem == 0 || (mem_in && mem_out && time), "mtw_shift_add: null array");
    BLMI_REQUIRE(mem_in != mem_out || n_mem == 0, "mtw_shift_add: in-place shift is not supported");
    BLMI_REQUIRE(n_add == 0 || add, "mtw_shift_add: null voltage");
    return launch_mtw_shift_add(mem_in, mem_out, time, n_mem, t_rev, add, n_add, S(stream));
}

int blmi_induced_voltage(const double *profile, int n_slices, const double *Z, int64_t n_fft,
                         double scale, double *volt, int n_out, void *stream) {
    BLMI_REQUIRE(n_slices >= 1 && profile && Z && volt, "induced_voltage: bad arguments");
    BLMI_REQUIRE(n_fft >= 2 && (n_fft & 1) == 0, "induced_voltage: n_fft %lld must be even",
                 (long long)n_fft);
    BLMI_REQUIRE(n_slices <= n_fft && n_out >= 0 && n_out <= n_fft,
                 "induced_voltage: n_slices %d / n_out %d exceed n_fft %lld", n_slices, n_out,
                 (long long)n_fft);
    InducedArgs a{};
    a.src = profile;
    a.n_slices = n_slices;
    a.Z = reinterpret_cast<const double2 *>(Z);
    a.scale = scale;
    a.volt = volt;
    a.n_out = n_out;
    if (n_fft < 8192 && n_fft <= fft_lds_max_real()) {
        FftPlan p;
        int rc = fft_plan_get(n_fft, &p);
        if (rc == BLMI_OK) return launch_induced_voltage_lds(a, p, S(stream));
    }
    // from 8192 points (or beyond one workgroup's LDS): the four-step
    // transform, faster there than the one-workgroup one; its plan holds
    // scratch, so plans are kept per (device, n_fft, stream)
    static std::mutex mu;
    static std::map<std::tuple<int, int64_t, void *>, Fft2Plan> plans;
    int dev = 0;
    BLMI_HIP_TRY(hipGetDevice(&dev));
    Fft2Plan p2;
    {
        std::lock_guard<std::mutex> lk(mu);
        auto key = std::make_tuple(dev, n_fft, stream);
        auto it = plans.find(key);
        if (it == plans.end()) {
            int rc = fft2_plan_create(n_fft, &p2);
            if (rc) {
                // no four-step factorisation: the one-workgroup transform
                // when the length fits its LDS
                FftPlan p;
                if (n_fft <= fft_lds_max_real() && fft_plan_get(n_fft, &p) == BLMI_OK)
                    return launch_induced_voltage_lds(a, p, S(stream));
                return rc;
            }
            plans[key] = p2;
        } else {
            p2 = it->second;
        }
    }
    return launch_induced_fourstep(a, p2, S(stream));
}

// Bunch moments over the alive particles (id != 0; id may be null): out =
// device double[5] {n_alive, mean_dt, sigma_dt, mean_dE, sigma_dE}.
int blmi_beam_statistics(const double *dt, const double *dE, const int64_t *id, int64_t n,
                         double *out, void *stream) {
    BLMI_REQUIRE(n >= 0 && out && (n == 0 || (dt && dE)), "beam_statistics: bad arguments");
    return launch_moments<double>(dt, dE, id, n, out, S(stream));
}

int blmi_beam_statistics_f32(const float *dt, const float *dE, const int64_t *id, int64_t n,
                             double *out, void *stream) {
    BLMI_REQUIRE(n >= 0 && out && (n == 0 || (dt && dE)), "beam_statistics: bad arguments");
    return launch_moments<float>(dt, dE, id, n, out, S(stream));
}

// --------------------------------------------------------------- engine ----
void *blmi_engine_create(int64_t n, int n_rf, int n_turns, int solver) {
    if (n < 0 || n_rf < 0 || n_turns < 1 || solver < 0 || solver > 2) {
        blmi_set_error(BLMI_ERR_ARG, "engine_create: bad arguments");
        return nullptr;
    }
    Engine *e = new Engine();
    if (e->init(n, n_rf, n_turns, solver) != BLMI_OK) {
        delete e;
        return nullptr;
    }
    return e;
}

int blmi_engine_destroy(void *h) {
    delete reinterpret_cast<Engine *>(h);
    return BLMI_OK;
}

int blmi_engine_set_particles(void *h, double *dt, double *dE, int64_t n) {
    Engine *e = reinterpret_cast<Engine *>(h);
    BLMI_REQUIRE(e, "engine: null handle");
    BLMI_REQUIRE(n == e->n, "engine: particle count %lld differs from %lld", (long long)n,
                 (long long)e->n);
    BLMI_REQUIRE(((((uintptr_t)dt) | ((uintptr_t)dE)) & 15u) == 0,
                 "engine: dt and dE must be 16-byte aligned");
    e->dt = dt;
    e->dE = dE;
    e->f32 = false;
    e->invalidate_graph();
    return BLMI_OK;
}

int blmi_engine_set_particles_f32(void *h, float *dt, float *dE, int64_t n) {
    Engine *e = reinterpret_cast<Engine *>(h);
    BLMI_REQUIRE(e, "engine: null handle");
    BLMI_REQUIRE(n == e->n, "engine: particle count %lld differs from %lld", (long long)n,
                 (long long)e->n);
    BLMI_REQUIRE(((((uintptr_t)dt) | ((uintptr_t)dE)) & 7u) == 0,
                 "engine: float dt and dE must be 8-byte aligned");
    return e->set_particles_f32(dt, dE);
}

int blmi_engine_set_slices_f32(void *h, int n_slices, double cut_left, double cut_right,
                               float *profile) {
    Engine *e = reinterpret_cast<Engine *>(h);
    BLMI_REQUIRE(e, "engine: null handle");
    BLMI_REQUIRE(n_slices >= 0 && (n_slices == 0 || profile), "engine: bad slices");
    BLMI_REQUIRE(n_slices == 0 || cut_right > cut_left, "engine: cut_right must exceed cut_left");
    return e->set_slices_f32(n_slices, cut_left, cut_right, profile);
}

int blmi_engine_set_induced_f32(void *h, int64_t n_fft, const double *Z_host, double scale,
                                float *volt, const float *bin_centers) {
    Engine *e = reinterpret_cast<Engine *>(h);
    BLMI_REQUIRE(e, "engine: null handle");
    BLMI_REQUIRE(e->n_slices >= 2 && e->d_prof64, "engine: set float slices (>= 2) first");
    BLMI_REQUIRE(volt && bin_centers && Z_host, "engine: null induced-voltage arrays");
    return e->set_induced_f32(n_fft, Z_host, scale, volt, bin_centers, e->charge);
}

int blmi_engine_set_programs(void *h, const double *rf, const double *acc, const double *t_rev,
                             double length_ratio, double alpha_order, const double *eta0,
                             const double *eta1, const double *eta2, const double *alpha0,
                             const double *alpha1, const double *alpha2, const double *beta,
                             const double *energy, double charge) {
    Engine *e = reinterpret_cast<Engine *>(h);
    BLMI_REQUIRE(e, "engine: null handle");
    e->charge = charge;
    return e->set_programs(rf, acc, t_rev, length_ratio, alpha_order, eta0, eta1, eta2, alpha0,
                           alpha1, alpha2, beta, energy);
}

int blmi_engine_set_slices(void *h, int n_slices, double cut_left, double cut_right,
                           double *profile) {
    Engine *e = reinterpret_cast<Engine *>(h);
    BLMI_REQUIRE(e, "engine: null handle");
    BLMI_REQUIRE(n_slices >= 0 && (n_slices == 0 || profile), "engine: bad slices");
    BLMI_REQUIRE(n_slices == 0 || cut_right > cut_left, "engine: cut_right must exceed cut_left");
    return e->set_slices(n_slices, cut_left, cut_right, profile);
}

int blmi_engine_set_induced(void *h, int64_t n_fft, const double *Z_host, double scale,
                            double *volt, const double *bin_centers) {
    Engine *e = reinterpret_cast<Engine *>(h);
    BLMI_REQUIRE(e, "engine: null handle");
    BLMI_REQUIRE(e->n_slices >= 2, "engine: set slices (>= 2) before the induced voltage");
    BLMI_REQUIRE(volt && bin_centers, "engine: null induced-voltage arrays");
    BLMI_REQUIRE((Z_host && n_fft >= 2) || (!Z_host && n_fft == 0),
                 "engine: Z with n_fft >= 2, or no Z with n_fft = 0 (no convolution)");
    return e->set_induced(n_fft, Z_host, scale, volt, bin_centers, e->charge);
}

int blmi_engine_set_feedback(void *h, int machine, int delay, int first_bin, double alpha,
                             double reference, const double *bin_centers,
                             const double *harmonic_host, const double *omega_rf_d_host,
                             const double *prog_host, const double *on_time_host, int n_on_time,
                             const double *state_host) {
    Engine *e = reinterpret_cast<Engine *>(h);
    BLMI_REQUIRE(e, "engine: null handle");
    return e->set_feedback(machine, delay, first_bin, alpha, reference, bin_centers, harmonic_host,
                           omega_rf_d_host, prog_host, on_time_host, n_on_time, state_host);
}

int blmi_engine_get_feedback(void *h, double *state_host) {
    Engine *e = reinterpret_cast<Engine *>(h);
    BLMI_REQUIRE(e && state_host, "engine: null handle");
    return e->get_feedback_state(state_host);
}

int blmi_engine_get_programs(void *h, int turn0, int turns, double *rf_host) {
    Engine *e = reinterpret_cast<Engine *>(h);
    BLMI_REQUIRE(e && rf_host, "engine: null handle");
    return e->get_programs(turn0, turns, rf_host);
}

int blmi_engine_set_statistics(void *h, int enable, double shift_dt, double shift_dE) {
    Engine *e = reinterpret_cast<Engine *>(h);
    BLMI_REQUIRE(e, "engine: null handle");
    return e->set_statistics(enable != 0, shift_dt, shift_dE);
}

int blmi_engine_get_statistics(void *h, int turn0, int turns, double *moments_host) {
    Engine *e = reinterpret_cast<Engine *>(h);
    BLMI_REQUIRE(e && moments_host, "engine: null handle");
    return e->get_statistics(turn0, turns, moments_host);
}

int blmi_engine_set_wake(void *h, int64_t n_mem, int64_t n_add, const double *time, double *memory,
                         const double *ind_coef_host, int deriv_mode) {
    Engine *e = reinterpret_cast<Engine *>(h);
    BLMI_REQUIRE(e, "engine: null handle");
    return e->set_wake(n_mem, n_add, time, memory, ind_coef_host, deriv_mode);
}

int blmi_engine_set_turn_block(void *h, int max_turns) {
    Engine *e = reinterpret_cast<Engine *>(h);
    BLMI_REQUIRE(e, "engine: null handle");
    BLMI_REQUIRE(max_turns >= 1 && max_turns <= 1000, "engine: turn block in [1, 1000]");
    e->turn_block = max_turns;
    e->invalidate_graph();
    return BLMI_OK;
}

int blmi_engine_set_launch(void *h, int block, int blocks_per_cu, int graph_chunk) {
    Engine *e = reinterpret_cast<Engine *>(h);
    BLMI_REQUIRE(e, "engine: null handle");
    BLMI_REQUIRE(block == 64 || block == 128 || block == 256 || block == 512 || block == 1024,
                 "engine: block size must be 64/128/256/512/1024");
    BLMI_REQUIRE(blocks_per_cu >= 1 && blocks_per_cu <= 64, "engine: blocks_per_cu in [1, 64]");
    BLMI_REQUIRE(graph_chunk >= 1 && graph_chunk <= 1000, "engine: graph_chunk in [1, 1000]");
    e->block_size = block;
    e->blocks_per_cu = blocks_per_cu;
    e->graph_chunk = graph_chunk;
    e->invalidate_graph();
    return BLMI_OK;
}

int blmi_engine_set_turn(void *h, int turn) {
    Engine *e = reinterpret_cast<Engine *>(h);
    BLMI_REQUIRE(e, "engine: null handle");
    return e->set_turn(turn);
}

int blmi_engine_get_turn(void *h, int *turn) {
    Engine *e = reinterpret_cast<Engine *>(h);
    BLMI_REQUIRE(e && turn, "engine: null handle");
    return e->get_turn(turn);
}

int blmi_engine_prime(void *h, void *stream) {
    Engine *e = reinterpret_cast<Engine *>(h);
    BLMI_REQUIRE(e, "engine: null handle");
    return e->prime(S(stream));
}

int blmi_engine_track(void *h, int turns, int use_graph, void *stream) {
    Engine *e = reinterpret_cast<Engine *>(h);
    BLMI_REQUIRE(e, "engine: null handle");
    BLMI_REQUIRE(turns >= 0, "engine: negative turn count");
    return e->track(turns, S(stream), use_graph != 0);
}

int blmi_engine_set_interpolation(void *h, const double *bin_centers) {
    Engine *e = reinterpret_cast<Engine *>(h);
    BLMI_REQUIRE(e, "engine: null handle");
    return e->set_interpolation(bin_centers);
}

int blmi_engine_set_method(void *h, int method) {
    Engine *e = reinterpret_cast<Engine *>(h);
    BLMI_REQUIRE(e, "engine: null handle");
    return e->set_method(method);
}

int blmi_engine_get_method(void *h, int *method) {
    Engine *e = reinterpret_cast<Engine *>(h);
    BLMI_REQUIRE(e && method, "engine: null handle");
    *method = e->intensity ? e->method : -1;
    return BLMI_OK;
}

int blmi_engine_set_hist(void *h, int groups, int max_peel) {
    Engine *e = reinterpret_cast<Engine *>(h);
    BLMI_REQUIRE(e, "engine: null handle");
    return e->set_hist(groups, max_peel);
}

int blmi_engine_set_order(void *h, int sort_every) {
    Engine *e = reinterpret_cast<Engine *>(h);
    BLMI_REQUIRE(e, "engine: null handle");
    BLMI_REQUIRE(sort_every >= 0 && sort_every <= 1000, "engine: sort_every in [0, 1000]");
    e->sort_every = sort_every;
    e->ord_since = 0;
    e->invalidate_graph();
    return BLMI_OK;
}

int blmi_engine_time_kernels(void *h, int turns, float *ms_induced, float *ms_turn,
                             void *stream) {
    Engine *e = reinterpret_cast<Engine *>(h);
    BLMI_REQUIRE(e && ms_induced && ms_turn, "engine: null argument");
    return e->time_kernels(turns, S(stream), ms_induced, ms_turn);
}

// ------------------------------------------------------------- ensemble ----
void *blmi_ensemble_create(int n_bunches, int n_rf, int n_turns, int solver) {
    if (n_bunches < 1 || n_rf < 0 || n_turns < 1 || solver < 0 || solver > 2) {
        blmi_set_error(BLMI_ERR_ARG, "ensemble_create: bad arguments");
        return nullptr;
    }
    Ensemble *e = new Ensemble();
    if (e->init(n_bunches, n_rf, n_turns, solver) != BLMI_OK) {
        delete e;
        return nullptr;
    }
    return e;
}

int blmi_ensemble_destroy(void *h) {
    delete reinterpret_cast<Ensemble *>(h);
    return BLMI_OK;
}

int blmi_ensemble_set_particles(void *h, double *dt, double *dE, const int64_t *offsets,
                                const int64_t *counts) {
    Ensemble *e = reinterpret_cast<Ensemble *>(h);
    BLMI_REQUIRE(e, "ensemble: null handle");
    BLMI_REQUIRE(dt && dE && offsets && counts, "ensemble: null particle arrays");
    return e->set_particles(dt, dE, offsets, counts);
}

int blmi_ensemble_set_programs(void *h, const double *rf, const double *acc, const double *t_rev,
                               double length_ratio, double alpha_order, const double *eta0,
                               const double *eta1, const double *eta2, const double *alpha0,
                               const double *alpha1, const double *alpha2, const double *beta,
                               const double *energy, double charge) {
    Ensemble *e = reinterpret_cast<Ensemble *>(h);
    BLMI_REQUIRE(e, "ensemble: null handle");
    return e->set_programs(rf, acc, t_rev, length_ratio, alpha_order, eta0, eta1, eta2, alpha0,
                           alpha1, alpha2, beta, energy, charge);
}

int blmi_ensemble_set_slices(void *h, int n_slices, const double *cut_left,
                             const double *cut_right, double *profiles) {
    Ensemble *e = reinterpret_cast<Ensemble *>(h);
    BLMI_REQUIRE(e, "ensemble: null handle");
    BLMI_REQUIRE(n_slices >= 2 && cut_left && cut_right && profiles, "ensemble: bad slices");
    for (int b = 0; b < e->n_bunches; ++b)
        BLMI_REQUIRE(cut_right[b] > cut_left[b], "ensemble: bunch %d: cut_right <= cut_left", b);
    return e->set_slices(n_slices, cut_left, cut_right, profiles);
}

int blmi_ensemble_set_induced(void *h, int64_t n_fft, const double *Z_host, const double *scale,
                              double *volt, const double *bin_centers) {
    Ensemble *e = reinterpret_cast<Ensemble *>(h);
    BLMI_REQUIRE(e, "ensemble: null handle");
    BLMI_REQUIRE(Z_host && scale && volt && bin_centers, "ensemble: null induced-voltage arrays");
    return e->set_induced(n_fft, Z_host, scale, volt, bin_centers);
}

int blmi_ensemble_set_launch(void *h, int block, int wg_per_cu, int graph_chunk, int max_peel) {
    Ensemble *e = reinterpret_cast<Ensemble *>(h);
    BLMI_REQUIRE(e, "ensemble: null handle");
    return e->set_launch(block, wg_per_cu, graph_chunk, max_peel);
}

int blmi_ensemble_get_launch(void *h, int *block, int *wg_per_cu) {
    Ensemble *e = reinterpret_cast<Ensemble *>(h);
    BLMI_REQUIRE(e && block && wg_per_cu, "ensemble: null argument");
    *block = e->block;
    *wg_per_cu = e->wg_per_cu;
    return BLMI_OK;
}

int blmi_ensemble_set_turn(void *h, int turn) {
    Ensemble *e = reinterpret_cast<Ensemble *>(h);
    BLMI_REQUIRE(e, "ensemble: null handle");
    return e->set_turn(turn);
}

int blmi_ensemble_prime(void *h, void *stream) {
    Ensemble *e = reinterpret_cast<Ensemble *>(h);
    BLMI_REQUIRE(e, "ensemble: null handle");
    return e->prime(S(stream));
}

int blmi_ensemble_track(void *h, int turns, int use_graph, void *stream) {
    Ensemble *e = reinterpret_cast<Ensemble *>(h);
    BLMI_REQUIRE(e, "ensemble: null handle");
    BLMI_REQUIRE(turns >= 0, "ensemble: negative turn count");
    return e->track(turns, S(stream), use_graph != 0);
}

int blmi_ensemble_time_turns(void *h, int turns, float *ms_induced, float *ms_turn, void *stream) {
    Ensemble *e = reinterpret_cast<Ensemble *>(h);
    BLMI_REQUIRE(e && ms_induced && ms_turn, "ensemble: null argument");
    return e->time_turns(turns, S(stream), ms_induced, ms_turn);
}

int blmi_ensemble_autotune(void *h, int turns, int *block, int *wg_per_cu, float *ms_per_turn,
                           void *stream) {
    Ensemble *e = reinterpret_cast<Ensemble *>(h);
    BLMI_REQUIRE(e && block && wg_per_cu && ms_per_turn, "ensemble: null argument");
    return e->autotune(turns, S(stream), block, wg_per_cu, ms_per_turn);
}

// ------------------------------------------------------------- adjoint -----
void *blmi_adjoint_create(int64_t n, int n_rf, int n_turns, int n_slices, int checkpoint_every) {
    if (n < 1 || n_rf < 1 || n_rf > 4 || n_turns < 1 || n_slices < 2 ||
        n_slices > kAdjMaxSlices || checkpoint_every < 1) {
        blmi_set_error(BLMI_ERR_ARG, "adjoint_create: bad arguments (n_rf 1..4, 2..%d slices)",
                       kAdjMaxSlices);
        return nullptr;
    }
    Adjoint *a = new Adjoint();
    if (a->init(n, n_rf, n_turns, n_slices, checkpoint_every) != BLMI_OK) {
        delete a;
        return nullptr;
    }
    return a;
}

int blmi_adjoint_destroy(void *h) {
    delete reinterpret_cast<Adjoint *>(h);
    return BLMI_OK;
}

int blmi_adjoint_set_particles(void *h, const double *dt, const double *dE) {
    Adjoint *a = reinterpret_cast<Adjoint *>(h);
    BLMI_REQUIRE(a && dt && dE, "adjoint: null argument");
    return a->set_particles(dt, dE);
}

int blmi_adjoint_set_programs(void *h, const double *rf, const double *acc, const double *t_rev,
                              double length_ratio, double alpha_order, const double *eta0,
                              const double *eta1, const double *eta2, const double *alpha0,
                              const double *alpha1, const double *alpha2, const double *beta,
                              const double *energy, double charge) {
    Adjoint *a = reinterpret_cast<Adjoint *>(h);
    BLMI_REQUIRE(a, "adjoint: null handle");
    return a->set_programs(rf, acc, t_rev, length_ratio, alpha_order, eta0, eta1, eta2, alpha0,
                           alpha1, alpha2, beta, energy, charge);
}

int blmi_adjoint_set_slices(void *h, double cut_left, double cut_right) {
    Adjoint *a = reinterpret_cast<Adjoint *>(h);
    BLMI_REQUIRE(a, "adjoint: null handle");
    BLMI_REQUIRE(cut_right > cut_left, "adjoint: cut_right must exceed cut_left");
    return a->set_slices(cut_left, cut_right);
}

int blmi_adjoint_set_map(void *h, int map, const double *bin_centers) {
    Adjoint *a = reinterpret_cast<Adjoint *>(h);
    BLMI_REQUIRE(a, "adjoint: null handle");
    return a->set_map(map, bin_centers);
}

int blmi_adjoint_set_wake(void *h, int64_t n_fft, const double *g, double scale) {
    Adjoint *a = reinterpret_cast<Adjoint *>(h);
    BLMI_REQUIRE(a && g, "adjoint: null argument");
    return a->set_wake(n_fft, g, scale);
}

int blmi_adjoint_forward(void *h, int turns, double *profile_out, void *stream) {
    Adjoint *a = reinterpret_cast<Adjoint *>(h);
    BLMI_REQUIRE(a && profile_out, "adjoint: null argument");
    return a->forward(turns, profile_out, S(stream));
}

int blmi_adjoint_backward(void *h, const double *dprofile, double *grad_rf, double *corr,
                          void *stream) {
    Adjoint *a = reinterpret_cast<Adjoint *>(h);
    BLMI_REQUIRE(a && dprofile && grad_rf && corr, "adjoint: null argument");
    return a->backward(dprofile, grad_rf, corr, S(stream));
}

}  // extern "C"
