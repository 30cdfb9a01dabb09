// Device-side building blocks shared by the three stages of the radiation hot path.
// Physics formulas follow the reference files cited at each function; the code is
// written for gfx950 (64-wide wavefronts, FP64 VALU).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/pbagpu.h"

namespace pbagpu {

// CGS constants (src/utilitites/utils.h:18-60)
constexpr double kC = 2.99792458e10;
constexpr double kPi = 3.141592653589793;
constexpr double kMp = 1.6726e-24;
constexpr double kMe = 9.1094e-28;
constexpr double kMppme = kMp + kMe;
constexpr double kQe = 4.803204e-10;
constexpr double kSigmaT = 6.6524e-25;
constexpr double kH = 6.6260755e-27;
constexpr double kMec2 = 8.187105649650028e-07;
constexpr double kCgs2mJy = 1.e26;

#define PBA_HD __host__ __device__ __forceinline__

PBA_HD double beta_from_gamma(double g) { return sqrt((g - 1.) * (g + 1.)) / g; }
PBA_HD double mom_from_gamma(double g) { return g * beta_from_gamma(g); }
PBA_HD double beta_from_mom(double m) { return m / sqrt(m * m + 1.); }
PBA_HD double gamma_from_mom(double m) { return sqrt(m * m + 1.); }
// EOSadi Nava13 adiabatic index (blastwave_components.h:944-946)
PBA_HD double adi_nava(double g) { return (4.0 + 1.0 / g) / 3.0; }
// EQS::rho2t / GammaSh (blastwave_components.h:270-279)
PBA_HD double rho2t(double g, double adi, double rho) { return (adi * g + 1.) / (adi - 1.) * rho; }
PBA_HD double gamma_shock(double g, double adi) {
    return (g + 1.0) * (adi * (g - 1.0) + 1.0) / (adi * (2.0 - adi) * (g - 1.0) + 2.0);
}

// bisection index search (interpolators.h:13-35)
template <class Arr>
__device__ __forceinline__ int find_index(double x, const Arr &arr, int N) {
    if (x <= arr(0)) return 0;
    if (x >= arr(N - 1)) return N - 2;
    unsigned a = 0, b = (unsigned)N - 1;
    while (b - a > 1u) {
        unsigned i = (b + a) >> 1;
        if (arr(i) > x) b = i; else a = i;
    }
    return (int)a;
}

__device__ __forceinline__ double interp_lin(double x, double xa, double xb, double ya, double yb) {
    return ya + (yb - ya) * (x - xa) / (xb - xa);
}
// log-log segment interpolation (interpolators.h:114-133)
__device__ __forceinline__ double interp_log(double x, double xa, double xb, double ya, double yb) {
    return ya * pow(yb / ya, log(x / xa) / log(xb / xa));
}

// computeIntensity with the smooth optical-depth rule (microphysics.h:11-58)
__device__ __forceinline__ double intensity_smooth(double em, double dtau) {
    if (dtau == 0) return em;
    const double abs_fac = dtau > 0.0 ? -expm1(-dtau) / dtau : expm1(dtau) / dtau;
    return em * abs_fac;
}

// computeFluxDensity (microphysics.h:799-867) for method_ne = useNe, method_shock_vel = sameAsBW,
// method_tau = smooth: the luminosity of one shock from its comoving emissivity and the optical
// depth along the line of sight.  `bad` is set where the reference stops (negative or non-finite
// intensity or luminosity).
__device__ __forceinline__ double shock_luminosity(double em_prime, double dtau, double Gamma, double beta,
                                                   double delta_D, double mu, double ne, double omc, bool &bad) {
    const double em_lab = em_prime / (delta_D * delta_D);
    const double intensity = intensity_smooth(em_lab, dtau);
    if (intensity < 0 || !isfinite(intensity)) bad = true;
    const double lum = intensity * ne / (2. * kPi * omc) / Gamma / (1.0 - mu * beta);
    if (!(lum >= 0) || !isfinite(lum)) bad = true;
    return lum;
}

// bilinear interpolation of a [step][frequency] table at (freq, r) inside the cell
// (ia, ib) x (ia_nu, ib_nu) (Interp2d::InterpolateBilinear, interpolators.h:724-808)
__device__ __forceinline__ double table_bilinear(const double *tab, int nf, int ia, int ib, int ia_nu, int ib_nu,
                                                 double x1, double x2, double y1, double y2, double freq, double r) {
    const double q11 = tab[(size_t)ia * nf + ia_nu], q12 = tab[(size_t)ib * nf + ia_nu];
    const double q21 = tab[(size_t)ia * nf + ib_nu], q22 = tab[(size_t)ib * nf + ib_nu];
    const double x2x1 = x2 - x1, y2y1 = y2 - y1, x2x = x2 - freq, y2y = y2 - r, yy1 = r - y1, xx1 = freq - x1;
    return 1.0 / (x2x1 * y2y1) * (q11 * x2x * y2y + q21 * xx1 * y2y + q12 * x2x * yy1 + q22 * xx1 * yy1);
}

// Optical depth of the shell with synchrotron self-absorption (computeFluxDensity,
// microphysics.h:806-832): absorption interpolated from the table, boosted to the lab frame,
// over the shell thickness of Johannesson et al. (2006, eq. 33).
__device__ __forceinline__ double absorption_depth(const double *aall, int nf, int ia, int ib, int ia_nu,
                                                       int ib_nu, double x1, double x2, double y1, double y2,
                                                       double freq, double r, double Gamma, double beta,
                                                       double delta_D, double mu, double ne, double n_prime,
                                                       double omc) {
    const double abs_prime = table_bilinear(aall, nf, ia, ib, ia_nu, ib_nu, x1, x2, y1, y2, freq, r);
    const double abs_lab = abs_prime * delta_D;
    const double m2 = ne * kMp;
    const double dr = m2 / (2. * kPi * kMp * r * r * omc * Gamma * n_prime);
    return abs_lab * dr * (1. - mu * beta) / (mu - beta);
}

// log grid by repeated multiplication (utils.h:91-105)
__device__ __forceinline__ double logspace_at(double start, double ratio, int i) {
    double v = start;
    for (int k = 0; k < i; k++) v *= ratio;
    return v;
}

// Wave-level helpers (64 lanes).
__device__ __forceinline__ double wave_shfl_down(double v, int d) { return __shfl_down(v, d, 64); }
__device__ __forceinline__ double wave_bcast(double v, int lane) { return __shfl(v, lane, 64); }

}  // namespace pbagpu
