/*
 * CPU oracle of the per-blastwave radiation hot path — TEST INFRASTRUCTURE ONLY.
 *
 * A plain, scalar, FP64 restatement of the reference algorithm, used to check
 * the GPU path in tests/ and __graft_entry__.smoke(), and as the CPU baseline leg of
 * bench.py.  Nothing in the product (pyblastafterglowmag_amd/, include/) links
 * or calls it.  Each function cites the reference file:line it restates.
 */
#ifndef PBA_ORACLE_COMMON_H
#define PBA_ORACLE_COMMON_H

#include <cmath>
#include <cstddef>
#include <vector>
#include <algorithm>
#include "../include/pbagpu.h"

namespace orc {

// CGS constants (src/utilitites/utils.h:18-60)
constexpr double c = 2.99792458e10;
constexpr double pi = 3.141592653589793;
constexpr double mp = 1.6726e-24;
constexpr double me = 9.1094e-28;
constexpr double mppme = mp + me;
constexpr double qe = 4.803204e-10;
constexpr double sigmaT = 6.6524e-25;
constexpr double h = 6.6260755e-27;
constexpr double mec2 = 8.187105649650028e-07;
constexpr double cgs2mJy = 1.e26;
const double h_by_mec2 = h / me / c / c;

// log grid by repeated multiplication, last point < stop (utils.h:91-105)
inline std::vector<double> logspace(double start, double stop, int num) {
    std::vector<double> v(num);
    double cur = std::pow(10.0, start);
    double ratio = std::pow(10.0, (stop - start) / num);
    for (int i = 0; i < num; i++) { v[i] = cur; cur *= ratio; }
    return v;
}

// Chang-Cooper electron grid e_i = step^i, step = exp(ln(gam2)/n) (base_state.h:229-252)
struct EleGrid {
    int n = 0;
    std::vector<double> e, half_e, dg, dgbar; // dg has n+1 entries
    void build(double gam2, int n_) {
        n = n_;
        e.assign(n, 0.); half_e.assign(n - 1, 0.); dg.assign(n + 1, 0.); dgbar.assign(n, 0.);
        double step = std::exp(std::log(gam2) / (double)n);
        double spo = step + 1.0;
        for (int i = 0; i < n; i++) {
            e[i] = std::pow(step, (double)i);
            if (i < n - 1) half_e[i] = .5 * e[i] * spo;
        }
        for (int i = 0; i < n - 1; i++) dg[i + 1] = e[i + 1] - e[i];
        dg[0] = e[0] * (1 - 1.0 / step);
        dg[n] = e[n - 1] * (step - 1);
        for (int i = 0; i < n; i++) dgbar[i] = (dg[i] + dg[i + 1]) / 2.;
    }
};

inline std::vector<double> freq_grid(const pba_shock_params &s) {
    return logspace(std::log10(s.freq1), std::log10(s.freq2), s.nfreq);
}

// EQS helpers (src/blastwave/blastwave_components.h:13-463)
inline double BetaFromGamma(double g) { return std::sqrt((g - 1.) * (g + 1.)) / g; }
inline double MomFromGamma(double g) { return g * BetaFromGamma(g); }
inline double BetaFromMom(double m) { return m / std::sqrt(m * m + 1.); }
inline double GammaFromMom(double m) { return std::sqrt(m * m + 1.); }
inline double rho2t(double G, double adi, double rho) { return (adi * G + 1.) / (adi - 1.) * rho; }
inline double GammaSh(double G, double adi) {
    return (G + 1.0) * (adi * (G - 1.0) + 1.0) / (adi * (2.0 - adi) * (G - 1.0) + 2.0);
}
inline double adiNava(double G) { return (4.0 + 1.0 / G) / 3.0; } // EOSadi::AdiabaticIndexNava (:944-946)

// interpolation helpers (src/utilitites/interpolators.h:13-133)
inline size_t findIndex(double x, const double *arr, size_t N) {
    if (x <= arr[0]) return 0;
    else if (x >= arr[N - 1]) return N - 2;
    unsigned int i = ((unsigned int)N) >> 1;
    unsigned int a = 0, b = (unsigned int)N - 1;
    while (b - a > 1u) {
        i = (b + a) >> 1;
        if (arr[i] > x) b = i; else a = i;
    }
    return a;
}
inline double interpSegLin(size_t a, size_t b, double x, const double *X, const double *Y) {
    double xa = X[a], xb = X[b], ya = Y[a], yb = Y[b];
    return ya + (yb - ya) * (x - xa) / (xb - xa);
}
inline double interpSegLog(size_t a, size_t b, double x, const double *X, const double *Y) {
    double xa = X[a], xb = X[b], ya = Y[a], yb = Y[b];
    return ya * std::pow(yb / ya, std::log(x / xa) / std::log(xb / xa));
}

} // namespace orc
#endif
