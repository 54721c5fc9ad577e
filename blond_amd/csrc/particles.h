// Internal launcher declarations of the MI355X tracking library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace blmi {

template <typename T> struct Vec2;
template <> struct Vec2<double> { using type = double2; };
template <> struct Vec2<float> { using type = float2; };

// Scalars of one drift call, precomputed on the host in the reference's
// order (drift.cpp:30-61).
template <typename T> struct DriftCoeffs {
    int solver;       // 0 simple, 1 legacy, 2 exact
    int alpha_order;  // 0, 1, 2
    T tl;             // t_rev * length_ratio
    T coeff;          // T*eta0/(beta^2 E)      (simple)
    T e0, e1, e2;     // eta_i / (beta^2 E)^(i+1) (legacy)
    T ib2, ie2;       // 1/beta^2, 1/E^2         (exact)
    T energy;
    T a0, a1, a2;
};

int device_cu_count();
int device_max_lds();
int stream_grid(int64_t work_items, int block);

template <typename T>
int launch_kick(const T *dt, T *dE, int64_t n, int n_rf, T charge, const T *V, const T *w,
                const T *phi, T acc, hipStream_t s);
template <typename T>
DriftCoeffs<T> make_drift_coeffs(int solver, T t_rev, T length_ratio, T alpha_order, T eta_0,
                                 T eta_1, T eta_2, T alpha_0, T alpha_1, T alpha_2, T beta,
                                 T energy);
template <typename T>
int launch_drift(T *dt, const T *dE, int64_t n, const DriftCoeffs<T> &c, hipStream_t s);
template <typename T>
int launch_interp_table(const T *volt, const T *bc, int n_slices, T charge, T acc, T *vk,
                        T *fac, T *params, hipStream_t s);
template <typename T>
int launch_interp_kick(const T *dt, T *dE, int64_t n, const T *vk, const T *fac,
                       const T *params, int n_tab, hipStream_t s);
template <typename T>
int launch_rf_volt_comp(const T *V, const T *w, const T *phi, const T *bc, int n_rf, int n_bins,
                        T *out, hipStream_t s);

// slices.hip
template <typename T> int launch_fill(T *p, int64_t n, T v, hipStream_t s);
// sparse_histogram.cpp:22-87 (profile: [n_filled][n_slices_bucket], overwritten)
template <typename T>
int launch_sparse_histogram(const T *dt, int64_t n, T *profile, const T *cut_left,
                            const T *cut_right, const T *bunch_indexes, int n_slices_bucket,
                            int n_filled, int n_pattern, hipStream_t s);
template <typename T>
int launch_histogram(const T *dt, int64_t n, T *profile, T cut_left, T cut_right, int n_slices,
                     hipStream_t s);
template <typename T>
int launch_smooth_histogram(const T *dt, int64_t n, T *profile, T cut_left, T cut_right,
                            int n_slices, long long *scratch, hipStream_t s);

// bunch moments: res[5] = {n_alive, mean_dt, sigma_dt, mean_dE, sigma_dE} (device)
template <typename T>
int launch_moments(const T *dt, const T *dE, const int64_t *id, int64_t n, double *res,
                   hipStream_t s);

// A slot of the library's per-device workspace (capi.hip: device scratch
// owned by the library, grown on demand), held for one call's launches on
// stream s.  The constructor orders s after the slot's previous user when
// that was another stream, and the destructor records s as the last user,
// so calls on different streams that share a slot run one after the other
// instead of racing.  Under stream capture the ordering is left to the
// caller (no event is waited on or recorded).  ptr() is null when the
// allocation failed.
class WorkspaceLease {
  public:
    WorkspaceLease(size_t bytes, int slot, hipStream_t s);
    ~WorkspaceLease();
    WorkspaceLease(const WorkspaceLease &) = delete;
    WorkspaceLease &operator=(const WorkspaceLease &) = delete;
    void *ptr() const { return p_; }

  private:
    void *p_ = nullptr;
    int dev_ = 0, slot_ = 0;
    hipStream_t s_ = nullptr;
    bool capturing_ = false;
};

}  // namespace blmi
