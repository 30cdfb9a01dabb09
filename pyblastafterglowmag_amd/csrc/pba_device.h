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
