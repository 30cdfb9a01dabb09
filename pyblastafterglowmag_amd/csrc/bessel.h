// Modified Bessel functions K_{1/3}(x) and K_{4/3}(x) for the Bessel synchrotron kernel
// (SynKernel::synchBessel, src/microphysics/kernels.h:396-412, which calls std::cyl_bessel_k for
// both orders).
//
// Method: N. M. Temme, J. Comput. Phys. 19 (1975) 324, in the form of the bessik scheme of
// Press et al., Numerical Recipes (2nd ed., section 6.7), which libstdc++'s cyl_bessel_k also
// follows.  With nu = mu + n and |mu| <= 1/2, K_mu and K_{mu+1} come from Temme's series for x < 2
// and from Steed's evaluation of the continued fraction CF2 for x >= 2; K_{mu+n} then follows by
// upward recurrence.  Orders 1/3 and 4/3 share mu = 1/3, so one evaluation gives both:
// K_{1/3} = K_mu and K_{4/3} = K_{mu+1}.  (libstdc++ forms mu = nu - n in double, which for
// nu = 4/3 lands one ulp below 1/3; the difference is of that order.)  The ratio I'_nu / I_nu
// (CF1) is not needed for K and is not evaluated.
//
// Callable on the host and on the device: the electron-stage kernel evaluates it per kernel
// entry, and pbagpu_bessel_k_thirds evaluates the same code on the host for checks.
#pragma once
#include <hip/hip_runtime.h>
#include <float.h>
#include <math.h>

namespace pbagpu {

// Constants of mu = 1/3 taken from the host's libm once (the device does not recompute
// Gamma(1 +- mu) or sin(pi mu)): the coefficients of Temme's series.
struct BesselThird {
    double mu;     // 1/3
    double fact;   // pi mu / sin(pi mu)
    double gam1;   // (1/Gamma(1-mu) - 1/Gamma(1+mu)) / (2 mu)
    double gam2;   // (1/Gamma(1-mu) + 1/Gamma(1+mu)) / 2
    double gampl;  // 1/Gamma(1+mu)
    double gammi;  // 1/Gamma(1-mu)
};

inline BesselThird bessel_third_constants() {
    BesselThird b;
    b.mu = 1. / 3.;
    const double pimu = 3.14159265358979323846 * b.mu;
    b.fact = pimu / sin(pimu);
    b.gampl = 1. / tgamma(1. + b.mu);
    b.gammi = 1. / tgamma(1. - b.mu);
    b.gam1 = (b.gammi - b.gampl) / (2. * b.mu);
    b.gam2 = (b.gammi + b.gampl) / 2.;
    return b;
}

// K_{1/3}(x) and K_{4/3}(x) for 0 < x; both are 0 in double precision for x >= 745.
__host__ __device__ inline void bessel_k_thirds(double x, const BesselThird &b, double &k13, double &k43) {
    k13 = k43 = 0.;
    if (!(x < 745.)) {   // NaN propagates; beyond 745 both underflow to 0
        if (x != x) k13 = k43 = x;
        return;
    }
    const int max_iter = 15000;
    const double eps = DBL_EPSILON, mu = b.mu, mu2 = mu * mu, xi = 1. / x;
    double kmu, kmu1;
    if (x < 2.) {
        // Temme's series: K_mu = sum_k c_k f_k, K_{mu+1} = (2/x) sum_k c_k (p_k - k f_k)
        const double x2 = x / 2.;
        double d = -log(x2);
        double e = mu * d;
        const double sinhc = fabs(e) < eps ? 1. : sinh(e) / e;
        double f = b.fact * (b.gam1 * cosh(e) + b.gam2 * sinhc * d);
        double sum = f;
        e = exp(e);
        double p = e / (2. * b.gampl);
        double q = 1. / (2. * e * b.gammi);
        double c = 1.;
        d = x2 * x2;
        double sum1 = p;
        for (int k = 1; k <= max_iter; k++) {
            f = (k * f + p + q) / (k * k - mu2);
            c *= d / k;
            p /= k - mu;
            q /= k + mu;
            const double del = c * f;
            sum += del;
            sum1 += c * (p - k * f);
            if (fabs(del) < eps * fabs(sum)) break;
        }
        kmu = sum;
        kmu1 = sum1 * (2. * xi);
    } else {
        // Steed's algorithm for CF2 with Temme's normalisation sum S
        double bb = 2. * (1. + x);
        double d = 1. / bb;
        double delh = d, h = d;
        double q1 = 0., q2 = 1.;
        const double a1 = 0.25 - mu2;
        double q = a1, c = a1, a = -a1;
        double s = 1. + q * delh;
        for (int k = 2; k <= max_iter; k++) {
            a -= 2 * (k - 1);
            c = -a * c / k;
            const double qn = (q1 - bb * q2) / a;
            q1 = q2;
            q2 = qn;
            q += c * qn;
            bb += 2.;
            d = 1. / (bb + a * d);
            delh = (bb * d - 1.) * delh;
            h += delh;
            const double dels = q * delh;
            s += dels;
            if (fabs(dels / s) < eps) break;
        }
        h = a1 * h;
        kmu = sqrt(3.14159265358979323846 / (2. * x)) * exp(-x) / s;
        kmu1 = kmu * (mu + x + 0.5 - h) * xi;
    }
    k13 = kmu;
    k43 = kmu1;
}

// synchBessel's emissivity kernel without its prefactor: with t = nu / nu_crit,
// 2 t^2 (K43 K13 - 3/5 t (K43^2 - K13^2)).
__host__ __device__ inline double bessel_synch_shape(double t, const BesselThird &b) {
    double k13, k43;
    bessel_k_thirds(t, b, k13, k43);
    return 2. * t * t * (k43 * k13 - 3. / 5 * t * (k43 * k43 - k13 * k13));
}

}  // namespace pbagpu
