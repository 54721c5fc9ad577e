/*
 * CPU oracle for the BLonD per-turn tracking path -- TEST INFRASTRUCTURE ONLY.
 *
 * This file is the checker, not the product: only tests/, the smoke check and
 * the CPU-baseline leg of bench.py may load it.  The MI355X library never
 * links or calls it.
 *
 * It restates, in plain C, the arithmetic of BLonD's C++ backend for the
 * functions on the tracking path, with the same C symbol names and argument
 * order, so that it is also a valid LIBBLOND drop-in for the reference's
 * ctypes layer (blond/utils/butils_wrap_cpp.py:33-44).  Built with
 * -ffp-contract=off and without fast-math, every function here reproduces the
 * reference's strict build bit for bit (see tests/test_oracle_vs_reference.py).
 *
 * Parity pins (what checks each function):
 *   kick                 golden vectors from the reference build; closed form
 *                        V*sin() at libm accuracy; sin against libm.
 *   drift (3 solvers)    golden vectors; closed-form drifts of
 *                        unittests/trackers/test_drift.py:40-96.
 *   histogram            golden vectors; reference counts of
 *                        unittests/beam_profile/test_beam_profile_object.py:169-204;
 *                        brute-force integer count on small inputs.
 *   smooth_histogram     golden vectors; reference values of
 *                        test_beam_profile_object.py:206-230; total weight.
 *   linear_interp_kick   golden vectors; numpy.interp closed form.
 *   rf_volt_comp         golden vectors; closed form.
 *   fast_resonator_real_imag  golden vectors; complex closed form R/(1+jQ(f/fr-fr/f)).
 *   sparse_histogram     the reference build (bit for bit); the uniform
 *                        histogram over the whole train (the reference's
 *                        unittests/beam_profile/test_sparse_profile.py).
 *   beam_phase, beam_phase_fast  the reference build run on one thread (bit
 *                        for bit); the closed form tan(w t + phi) of a one-bin
 *                        profile; cos and exp against libm.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------
 * Sine: the Cephes method (S. L. Moshier, Cephes Math Library, sin.c), in
 * the arrangement BLonD's vendored VDT uses (cpp_routines/sincos.h:72-160):
 * reduce |x| to [-pi/4, pi/4] with an even octant index j and a three-part
 * Cody-Waite subtraction of j*pi/4, evaluate both the sine and the cosine
 * minimax polynomials of Cephes on the reduced argument, then choose and
 * sign by octant.  The polynomial coefficients are Cephes's published
 * sincof[] / coscof[] tables and the pi/4 split is Cephes's DP1..DP3.
 * ---------------------------------------------------------------------- */
static const double ORC_PIO4_INV = 4.0 / 3.14159265358979323846; /* 4/M_PI */
static const double ORC_DP1 = 7.853981554508209228515625E-1;
static const double ORC_DP2 = 7.94662735614792836714E-9;
static const double ORC_DP3 = 3.06161699786838294307E-17;
static const double ORC_SIN[6] = {
    1.58962301576546568060E-10, -2.50507477628578072866E-8,
    2.75573136213857245213E-6,  -1.98412698295895385996E-4,
    8.33333333332211858878E-3,  -1.66666666666666307295E-1};
static const double ORC_COS[6] = {
    -1.13585365213876817300E-11, 2.08757008419747316778E-9,
    -2.75573141792967388112E-7,  2.48015872888517045348E-5,
    -1.38888888888730564116E-3,  4.16666666666665929218E-2};

static double oracle_poly6(const double *c, double z) {
    double p = c[0];
    for (int k = 1; k < 6; ++k) {
        p = p * z;
        p = p + c[k];
    }
    return p;
}

double oracle_sin(double x) {
    double ax = fabs(x);
    int oct = (int)(ORC_PIO4_INV * ax);
    oct = (oct + 1) & ~1;
    double y = (double)oct;
    double r = ((ax - y * ORC_DP1) - y * ORC_DP2) - y * ORC_DP3;
    double zz = r * r;
    double s = r + r * zz * oracle_poly6(ORC_SIN, zz);
    double c = 1.0 - zz * 0.5 + zz * zz * oracle_poly6(ORC_COS, zz);
    /* octant bookkeeping: bit 2 of oct flips the sine, bit 1 of (oct-2)
       selects the cosine polynomial */
    int flip = oct & 4;
    int use_cos = ((oct - 2) & 2) == 0;
    double out = use_cos ? c : s;
    if (flip) out = -out;
    if (x < 0.0) out = -out;
    return out;
}

/* float variant (sincos.h:166-233): Cephes sinf polynomials */
float oracle_sinf(float x) {
    const float PIO4_INV_F = (float)(4.0 / 3.14159265358979323846);
    float ax = fabsf(x);
    int oct = (int)(PIO4_INV_F * ax);
    oct = (oct + 1) & ~1;
    float y = (float)oct;
    float r = ((ax - y * 0.78515625f) - y * 2.4187564849853515625e-4f) -
              y * 3.77489497744594108e-8f;
    float z = r * r;
    float s = (((-1.9515295891E-4f * z + 8.3321608736E-3f) * z -
                1.6666654611E-1f) * z * r) + r;
    float c = ((2.443315711809948E-005f * z - 1.388731625493765E-003f) * z +
               4.166664568298827E-002f) * z * z - 0.5f * z + 1.0f;
    int flip = oct & 4;
    int use_cos = ((oct - 2) & 2) == 0;
    float out = use_cos ? c : s;
    if (flip) out = -out;
    if (x < 0.0f) out = -out;
    return out;
}

/* sparse_histogram.cpp:22-87: bucket from the first filled bucket's left
 * cut, its row from bunch_indexes (-1: empty), bin by truncation; right cut
 * of the last bucket inclusive.  Particles the reference would look up past
 * bunch_indexes (n_pattern entries) or count outside the rows are skipped. */
void sparse_histogram_n(const double *input, double *output, const double *cut_left,
                        const double *cut_right, const double *bunch_indexes,
                        int n_slices_bucket, int n_filled, int n, int n_pattern) {
    const double inv_bucket = 1.0 / (cut_right[0] - cut_left[0]);
    const double inv_bw = inv_bucket * (double)n_slices_bucket;
    const int total = n_filled * n_slices_bucket;
    for (int i = 0; i < total; ++i) output[i] = 0.0;
    for (int j = 0; j < n; ++j) {
        const double a = input[j];
        if (a < cut_left[0] || a > cut_right[n_filled - 1] || a != a) continue;
        const int fb = (int)((a - cut_left[0]) * inv_bucket);
        if (fb >= n_pattern) continue;
        const int ib = (int)bunch_indexes[fb];
        if (ib == -1) continue;
        const int bin = ib * n_slices_bucket + (int)((a - cut_left[ib]) * inv_bw);
        if (bin < 0 || bin >= total) continue;
        output[bin] += 1.0;
    }
}

/* Cosine with the sine's reduction and polynomials (VDT fast_sincos,
 * cpp_routines/sincos.h:132-160, cos.h:22): the other polynomial of the
 * octant pair, negated when bit 2 of (j - 2) is clear. */
double oracle_cos(double x) {
    double ax = fabs(x);
    int oct = (int)(ORC_PIO4_INV * ax);
    oct = (oct + 1) & ~1;
    double y = (double)oct;
    double r = ((ax - y * ORC_DP1) - y * ORC_DP2) - y * ORC_DP3;
    double zz = r * r;
    double s = r + r * zz * oracle_poly6(ORC_SIN, zz);
    double c = 1.0 - zz * 0.5 + zz * zz * oracle_poly6(ORC_COS, zz);
    int use_sin = ((oct - 2) & 2) == 0;
    double out = use_sin ? s : c;
    if (((oct - 2) & 4) == 0) out = -out;
    return out;
}

/* Exponential: Cephes exp.c as arranged in VDT (cpp_routines/exp.h:70-113):
 * n = trunc(y) - signbit((float)y) with y = x log2(e) + 1/2, x - n C1 - n C2,
 * 1 + 2 x P / (Q - x P) with Cephes's P(x^2), Q(x^2), times 2^n. */
double oracle_exp(double x0) {
    double x = x0;
    const double t = 1.4426950408889634073599 * x + 0.5;
    float tf = (float)t;
    uint32_t tb;
    memcpy(&tb, &tf, sizeof tb);
    double px = (double)((int32_t)t - (int32_t)(tb >> 31));
    const int32_t n = (int32_t)px;
    x -= px * 6.93145751953125E-1;
    x -= px * 1.42860682030941723212E-6;
    const double xx = x * x;
    px = 1.26177193074810590878E-4;
    px *= xx;
    px += 3.02994407707441961300E-2;
    px *= xx;
    px += 9.99999999999999999910E-1;
    px *= x;
    double qx = 3.00198505138664455042E-6;
    qx *= xx;
    qx += 2.52448340349684104192E-3;
    qx *= xx;
    qx += 2.27265548208155028766E-1;
    qx *= xx;
    qx += 2.00000000000000000009E0;
    x = px / (qx - px);
    x = 1.0 + 2.0 * x;
    uint64_t bits = (((uint64_t)(int64_t)n) + 1023) << 52;
    double scale;
    memcpy(&scale, &bits, sizeof scale);
    x *= scale;
    if (x0 > 708.0) x = INFINITY;
    if (x0 < -708.0) x = 0.0;
    return x;
}

/* trapz_const_delta (blondmath.cpp:496-512) as its OpenMP reduction runs on
 * one thread: the interior summed into a zero-initialised private copy,
 * which is then added to (f[0] + f[n-1]) / 2. */
static double oracle_trapz_const_delta(const double *f, double dx, int n) {
    double psum = (f[0] + f[n - 1]) / 2.;
    double priv = 0.0;
    for (int i = 1; i < n - 1; ++i) priv += f[i];
    psum = psum + priv;
    return dx * psum;
}

/* beam_phase.cpp:6-39 */
double beam_phase(const double *bin_centers, const double *profile, double alpha,
                  double omega_rf, double phi_rf, double bin_size, int n_bins) {
    if (n_bins < 1) return NAN;
    double *a1 = (double *)malloc(sizeof(double) * (size_t)n_bins);
    double *a2 = (double *)malloc(sizeof(double) * (size_t)n_bins);
    for (int i = 0; i < n_bins; ++i) {
        const double base = oracle_exp(alpha * bin_centers[i]) * profile[i];
        const double a = omega_rf * bin_centers[i] + phi_rf;
        a1[i] = base * oracle_sin(a);
        a2[i] = base * oracle_cos(a);
    }
    const double s = oracle_trapz_const_delta(a1, bin_size, n_bins);
    const double c = oracle_trapz_const_delta(a2, bin_size, n_bins);
    free(a1);
    free(a2);
    return s / c;
}

/* beam_phase.cpp:42-65 */
double beam_phase_fast(const double *bin_centers, const double *profile, double omega_rf,
                       double phi_rf, double bin_size, int n_bins) {
    if (n_bins < 1) return NAN;
    double *a1 = (double *)malloc(sizeof(double) * (size_t)n_bins);
    double *a2 = (double *)malloc(sizeof(double) * (size_t)n_bins);
    for (int i = 0; i < n_bins; ++i) {
        const double a = omega_rf * bin_centers[i] + phi_rf;
        a1[i] = profile[i] * oracle_sin(a);
        a2[i] = profile[i] * oracle_cos(a);
    }
    const double s = oracle_trapz_const_delta(a1, bin_size, n_bins);
    const double c = oracle_trapz_const_delta(a2, bin_size, n_bins);
    free(a1);
    free(a2);
    return s / c;
}

/* kick.cpp:16-35 : dE[i] += charge * sum_j V_j sin(w_j dt_i + phi_j) + acc */
void kick(const double *dt, double *dE, int n_rf, double charge,
          const double *voltage, const double *omega_rf, const double *phi_rf,
          int n, double acc_kick) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        const double t = dt[i];
        double acc = 0.0;
        for (int j = 0; j < n_rf; ++j)
            acc = acc + voltage[j] * oracle_sin(omega_rf[j] * t + phi_rf[j]);
        dE[i] = dE[i] + (charge * acc + acc_kick);
    }
}

/* kick.cpp:37-51 : rf_voltage[b] += sum_j V_j sin(w_j bc_b + phi_j), harmonic-major */
void rf_volt_comp(const double *voltage, const double *omega_rf,
                  const double *phi_rf, const double *bin_centers, int n_rf,
                  int n_bins, double *rf_voltage) {
    for (int j = 0; j < n_rf; ++j)
        for (int b = 0; b < n_bins; ++b)
            rf_voltage[b] = rf_voltage[b] +
                            voltage[j] * oracle_sin(omega_rf[j] * bin_centers[b] + phi_rf[j]);
}

/* drift.cpp:19-78 : solver 0 simple, 1 legacy, 2 exact */
void drift(double *dt, const double *dE, int solver, double t_rev,
           double length_ratio, double alpha_order, double eta_0, double eta_1,
           double eta_2, double alpha_0, double alpha_1, double alpha_2,
           double beta, double energy, int n) {
    const double T = t_rev * length_ratio;
    if (solver == 0) {
        const double coeff = T * eta_0 / (beta * beta * energy);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) dt[i] = dt[i] + coeff * dE[i];
    } else if (solver == 1) {
        const double k = 1.0 / (beta * beta * energy);
        const double e0 = eta_0 * k;
        const double e1 = eta_1 * k * k;
        const double e2 = eta_2 * k * k * k;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
            const double d = dE[i];
            double den;
            if (alpha_order == 0)
                den = 1.0 - e0 * d;
            else if (alpha_order == 1)
                den = 1.0 - e0 * d - e1 * d * d;
            else
                den = 1.0 - e0 * d - e1 * d * d - e2 * d * d * d;
            dt[i] = dt[i] + T * (1.0 / den - 1.0);
        }
    } else {
        const double ib2 = 1 / (beta * beta);
        const double ie2 = 1 / (energy * energy);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
            const double d = dE[i];
            const double delta = sqrt(1.0 + ib2 * (d * d * ie2 + 2.0 * d / energy)) - 1.0;
            dt[i] = dt[i] + T * ((1.0 + alpha_0 * delta + alpha_1 * (delta * delta) +
                                  alpha_2 * (delta * delta * delta)) *
                                     (1.0 + d / energy) / (1.0 + delta) -
                                 1.0);
        }
    }
}

/* histogram.cpp:22-73 : nearest-bin counts, index staged through float,
   right edge exclusive, output overwritten */
void histogram(const double *input, double *output, double cut_left,
               double cut_right, int n_slices, int n) {
    const double inv_bw = n_slices / (cut_right - cut_left);
    int *cnt = (int *)calloc((size_t)n_slices, sizeof(int));
    /* integer counts: per-thread tables summed in any order give the same
       result as the serial loop */
#pragma omp parallel
    {
        int *mine = (int *)calloc((size_t)n_slices, sizeof(int));
#pragma omp for schedule(static)
        for (int i = 0; i < n; ++i) {
            float fb = (float)floor((input[i] - cut_left) * inv_bw);
            /* the reference converts with (int); values outside int range are
               out of the histogram either way */
            if (!(fb >= 0.0f) || fb >= (float)n_slices) continue;
            int b = (int)fb;
            if (b < 0 || b >= n_slices) continue;
            mine[b] += 1;
        }
#pragma omp critical
        for (int b = 0; b < n_slices; ++b) cnt[b] += mine[b];
        free(mine);
    }
    for (int b = 0; b < n_slices; ++b) output[b] = (double)cnt[b];
    free(cnt);
}

/* histogram.cpp:75-132 : two-bin weights with the reference's quirk --
   d<=0.5 puts 0.5+d on the bin BELOW, d>0.5 puts 0.5-d (negative) on the own
   bin and 0.5+d on the bin above; particles in the outer half-bins dropped */
void smooth_histogram(const double *input, double *output, double cut_left,
                      double cut_right, int n_slices, int n) {
    const double inv_bw = n_slices / (cut_right - cut_left);
    const double bw = (cut_right - cut_left) / n_slices;
    const double lo = cut_left + bw * 0.5;
    const double hi = cut_right - bw * 0.5;
    for (int b = 0; b < n_slices; ++b) output[b] = 0.0;
    for (int i = 0; i < n; ++i) {
        const double a = input[i];
        if (a < lo || a > hi) continue;
        const double f = (a - cut_left) * inv_bw;
        const int k = (int)f;
        const double d = f - (double)k;
        const int k2 = (d > 0.5) ? (int)(f + 1.0) : (int)(f - 1.0);
        output[k] = output[k] + 0.5 - d;
        output[k2] = output[k2] + 0.5 + d;
    }
}

/* linear_interp_kick.cpp:21-71 : table (slope, intercept) per bin, gather kick */
void linear_interp_kick(const double *dt, double *dE, const double *voltage,
                        const double *bin_centers, double charge, int n_slices,
                        int n, double acc_kick) {
    const double inv_bw = (n_slices - 1) / (bin_centers[n_slices - 1] - bin_centers[0]);
    double *vk = (double *)malloc(sizeof(double) * (size_t)(n_slices > 1 ? n_slices - 1 : 1));
    double *fc = (double *)malloc(sizeof(double) * (size_t)(n_slices > 1 ? n_slices - 1 : 1));
    for (int i = 0; i < n_slices - 1; ++i) {
        vk[i] = charge * (voltage[i + 1] - voltage[i]) * inv_bw;
        fc[i] = (charge * voltage[i] - bin_centers[i] * vk[i]) + acc_kick;
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        const double f = floor((dt[i] - bin_centers[0]) * inv_bw);
        /* the reference casts to unsigned and keeps k < n_slices-1: negative
           and NaN positions get no kick */
        if (!(f >= 0.0) || f >= (double)(n_slices - 1)) continue;
        const int k = (int)f;
        dE[i] = dE[i] + (dt[i] * vk[k] + fc[k]);
    }
    free(vk);
    free(fc);
}

/* fast_resonator.cpp:19-73 : resonator impedance accumulated for f[1:] */
void fast_resonator_real_imag(double *re, double *im, const double *freq,
                              const double *R, const double *Q,
                              const double *fr, int n_res, int n_f) {
    for (int r = 0; r < n_res; ++r) {
        const double q2 = Q[r] * Q[r];
        for (int k = 1; k < n_f; ++k) {
            const double u = freq[k] / fr[r] - fr[r] / freq[k];
            re[k] = re[k] + R[r] / (1.0 + q2 * u * u);
            im[k] = im[k] - R[r] * (Q[r] * u) / (1.0 + q2 * u * u);
        }
    }
}

/* ------------------------------------------------------------------------
 * float32 variants (the reference's libblond_single.so, -DUSEFLOAT):
 * same arithmetic with real_t = float and vdt::fast_sinf.
 * ---------------------------------------------------------------------- */
void kick_f32(const float *dt, float *dE, int n_rf, float charge,
              const float *voltage, const float *omega_rf, const float *phi_rf,
              int n, float acc_kick) {
    for (int i = 0; i < n; ++i) {
        const float t = dt[i];
        float acc = 0.0f;
        for (int j = 0; j < n_rf; ++j)
            acc = acc + voltage[j] * oracle_sinf(omega_rf[j] * t + phi_rf[j]);
        dE[i] = dE[i] + (charge * acc + acc_kick);
    }
}

void drift_simple_f32(float *dt, const float *dE, float t_rev, float length_ratio,
                      float eta_0, float beta, float energy, int n) {
    const float T = t_rev * length_ratio;
    const float coeff = T * eta_0 / (beta * beta * energy);
    for (int i = 0; i < n; ++i) dt[i] = dt[i] + coeff * dE[i];
}

void histogram_f32(const float *input, float *output, float cut_left,
                   float cut_right, int n_slices, int n) {
    const float inv_bw = n_slices / (cut_right - cut_left);
    int *cnt = (int *)calloc((size_t)n_slices, sizeof(int));
    for (int i = 0; i < n; ++i) {
        float fb = floorf((input[i] - cut_left) * inv_bw);
        if (!(fb >= 0.0f) || fb >= (float)n_slices) continue;
        cnt[(int)fb] += 1;
    }
    for (int b = 0; b < n_slices; ++b) output[b] = (float)cnt[b];
    free(cnt);
}

void linear_interp_kick_f32(const float *dt, float *dE, const float *voltage,
                            const float *bin_centers, float charge, int n_slices,
                            int n, float acc_kick) {
    const float inv_bw = (n_slices - 1) / (bin_centers[n_slices - 1] - bin_centers[0]);
    float *vk = (float *)malloc(sizeof(float) * (size_t)(n_slices > 1 ? n_slices - 1 : 1));
    float *fc = (float *)malloc(sizeof(float) * (size_t)(n_slices > 1 ? n_slices - 1 : 1));
    for (int i = 0; i < n_slices - 1; ++i) {
        vk[i] = charge * (voltage[i + 1] - voltage[i]) * inv_bw;
        fc[i] = (charge * voltage[i] - bin_centers[i] * vk[i]) + acc_kick;
    }
    for (int i = 0; i < n; ++i) {
        const float f = floorf((dt[i] - bin_centers[0]) * inv_bw);
        if (!(f >= 0.0f) || f >= (float)(n_slices - 1)) continue;
        const int k = (int)f;
        dE[i] = dE[i] + (dt[i] * vk[k] + fc[k]);
    }
    free(vk);
    free(fc);
}
