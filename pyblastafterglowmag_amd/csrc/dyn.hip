// Stage 1 — blastwave dynamics on the GPU.
//
// One workgroup (one 64-lane wavefront) integrates one run: lane i owns
// blastwave i of the run.  All blastwaves of a run form one ODE system with one
// shared adaptive step, as in EvolveODEsystem (src/model_evolve.h:72, 294): the
// DOP853 error norm is a sum over every component of every blastwave, so the
// per-lane partial norms are combined in lane order through LDS and every lane
// then takes the same step decision.  After every output interval each lane
// writes its blastwave's derived quantities and injection-spectrum limits
// (one column of the [PBA_NV][nt] table).
//
// Follows: BlastWave::setInitConditions / rhs_fs / rhs_fsrs / insertSolution /
// applyUnits / addOtherVars (src/blastwave/blastwave.h:73-756, 1515-2116),
// ElectronAndRadiation::computeRadiationLoss (src/microphysics/microphysics.h:749-795),
// computeEleSpectrumLimits (src/microphysics/microphysics_base.h:575-966),
// dop853_integrator::integrate (src/utilitites/ode_solvers.h:1289-1980),
// EvolveODEsystem::advanceTimeSubStep (src/model_evolve.h:439-566).
// The reference evaluates Gamma43 and dM3/dR in x87 long double; FP64 is used here.
#include "pba_device.h"
#include "timing.h"
#include "runtime.h"

namespace pbagpu {
namespace dyn {

enum QS { iR, iRsh, itt, itcomov, iGamma, iEint2, iEint3, itheta, iRrsh, iErad2, iErad3, iEsh2, iEsh3,
          iEad2, iEad3, iM2, iM3, ideltaR4, NA };
constexpr int NEQ_REF = 23;  // components per blastwave in the reference state vector
// Accepted+rejected DOP853 steps allowed inside one output interval before the run is flagged
// PBAGPU_S_NMAX.  The stiff start of a reverse-shock run (h ~ 1e-10 s) takes ~2.4e5 steps in the
// first interval of a 5-layer Gaussian jet, so the budget sits well above that.
constexpr int kMaxStepsPerInterval = 10000000;

struct MState { double acc, tcomov; };

struct Shock { double e0, etop, nu0_43; };

struct BW {
    double E0, M0, R0, Gamma0, beta0, mom0, theta_b0, theta_core, theta_max, rho_def, drhodr0;
    double ncells;     // cells sharing this blastwave: 1 (adaptive EATS) or the total cell count (piece-wise)
    double tan_ratio;  // tan(theta_b0 / 2) / tan(theta_core / 2) of the spreading rate
    double Rw0, Wtt0;  // passive components of the reference state vector (enter the step-size estimate)
    int ilayer, nlayers;
    int fsrs, do_rs, do_rs_rad, shutOff, end_evolution, end_spreading;
    int prev_idx_x, i_end_r, status;
    MState fs_st, rs_st;
};

__device__ double x86_int_abs(double v) {
    // int abs(int) applied to a double on x86 (the initial-step estimator of the
    // reference calls the int overload): out-of-range values give INT_MIN.
    if (!(v > -2147483649.0 && v < 2147483648.0)) return -2147483648.0;
    int iv = (int)v;
    return (double)(iv < 0 ? -iv : iv);
}

__device__ double dthetadr_our(const BW &b, double Gamma, double Gsh, double R, double adi, double theta) {
    double Q0 = 2.0, Q = sqrt(2.0) * 3.0;
    double mom = mom_from_gamma(Gsh);
    double vperp = sqrt(adi * (adi - 1.0) * (Gamma - 1.0) / (1.0 + adi * (Gamma - 1.0))) * kC;
    if ((theta < 0.5 * kPi) && (Q0 * mom * b.theta_core < 1)) {
        double fac = 1.;
        if (mom * b.theta_core * Q >= 1.0) fac = Q * (1. - Q0 * mom * b.theta_core) / (Q - Q0);
        if (b.theta_b0 < b.theta_core) fac *= b.tan_ratio;   // tan(theta_b0/2) / tan(theta_core/2), set once
        double beta = beta_from_gamma(Gsh);
        return vperp / R / Gsh / beta / kC * fac;
    }
    return 0.;
}

// Integrated WSPN99 synchrotron power (analytic_models.h:889-911, 1051-1078).  Every power x^a
// is exp(a ln x) with each logarithm taken once (3 logarithms and 8 exponentials instead of 10
// calls of pow); nu0_43 = nu0^(4/3) is a constant of the run.
__device__ __forceinline__ double wspn99_ptot(double p, double gm, double gc, double gM, double B, double nprime, double nu0_43) {
    double f = (3.0 / (4.0 * kPi)) * (kQe * B) / (kMe * kC);
    double nu_m, nu_c, p_max;
    if (gm < gc) {
        double XpS = 0.06 + 0.28 * p;
        nu_m = XpS * gm * gm * f;
        nu_c = XpS * gc * gc * f;
        p_max = 11.17 * (p - 1.0) / (3.0 * p - 1.0) * (0.54 + 0.08 * p) * kQe * kQe * kQe * nprime * B / (kMe * kC * kC);
    } else {
        double XpF = 0.455 + 0.08 * p;
        double phipF = 1.89 - 0.935 * p + 0.17 * (p * p);
        nu_m = XpF * gm * gm * f;
        nu_c = XpF * gc * gc * f;
        p_max = 2.234 * phipF * kQe * kQe * kQe * nprime * B / (kMe * kC * kC);
    }
    double nu_M = 3.0 / (4.0 * kPi) * gM * gM * kQe * B / (kMe * kC);
    const double lm = log(nu_m), lc = log(nu_c), lM = log(nu_M);
    double t0, t1, t2;
    if (nu_m <= nu_c) {
        t0 = (3. / 4.) * exp(-(1. / 3.) * lm) * (exp((4. / 3.) * lm) - nu0_43);
        t1 = -1. / (3. - p) * (2 * exp(-(1. / 2. - p / 2.) * lm)) *
             (exp((3. / 2. - p / 2.) * lm) - exp((3. / 2. - p / 2.) * lc));
        t2 = exp((-(p - 1.) / 2.) * (lc - lm)) * 2. / (p - 2.) * (nu_c * exp((p / 2.) * lM) - nu_M * exp((p / 2.) * lc)) *
             exp((-p / 2.) * lM);
    } else {
        t0 = (3. / 4.) * exp(-(1. / 3.) * lc) * (exp((4. / 3.) * lc) - nu0_43);
        t1 = -2. / sqrt(1. / nu_c) * (sqrt(nu_c) - sqrt(nu_m));
        t2 = 1. / sqrt(nu_m / nu_c) * 2. / (p - 2.) * (nu_m * exp((p / 2.) * lM) - nu_M * exp((p / 2.) * lm)) *
             exp((-p / 2.) * lM);
    }
    return (t0 + t1 + t2) * p_max;
}

// computeRadiationLoss (microphysics.h:749-795); cth = cos(theta) from the caller
__device__ __forceinline__ double radiation_loss(const pba_shock_params &s, const Shock &sh, MState &st, double dmdR,
                                                 double dEshdR, double R, double Gamma, double Gsh, double adi, double m,
                                                 double Eint, double rho, double cth, double ncells) {
    if (dmdR == 0 || dEshdR == 0 || m == 0 || rho == 0) return 0.;
    double rho2 = rho2t(Gamma, adi, rho);
    double U_p = (Eint == 0 || m == 0) ? 0. : Eint / (m / rho2);
    double betaSh = beta_from_gamma(Gsh), n_prime = rho2 / kMp;
    double omc = (1.0 - cth / ncells);
    double thickness = m / (2.0 * kPi * omc * rho2 * Gamma * R * R);
    double p_tot = 0.;
    if (!(betaSh < s.beta_min)) {
        double B = sqrt(8. * kPi * (U_p * s.eps_b));
        double gM = sqrt(6.0 * kPi * kQe / kSigmaT / B);
        if (gM > 0.99 * sh.etop) gM = 0.99 * sh.etop;
        double gm = (s.p - 2.) / (s.p - 1.) * (U_p * s.eps_e) / (n_prime * kMe * kC * kC);
        if (gm < 2.) {
            double sir = (s.p - 2.0) / (s.p - 1.0) * s.eps_e * U_p / (n_prime * kMe * kC * kC);
            st.acc = sir > 1. ? 1. : sir;
            gm = gm < 1. ? 1. : gm;
        }
        if (gm < sh.e0) gm = sh.e0;
        double gc = (6. * kPi * kMe * kC) / (kSigmaT * B * B * st.tcomov);
        if (gc <= 1.) gc = 1.001;
        if (gc > gM) gc = gM;
        p_tot = wspn99_ptot(s.p, gm, gc, gM, B, dmdR / kMp, sh.nu0_43);
    }
    p_tot *= st.acc;
    p_tot = fmax(p_tot, 0.);
    double dErad2dR = p_tot * (thickness * Gsh / kC);
    double eps_rad = dErad2dR / (dEshdR * s.eps_e);
    if (eps_rad < 0. || !isfinite(eps_rad)) eps_rad = 0.;
    return eps_rad;
}

// Ambient density and its radial derivative at radius R (RhoISM::evaluateRhoDrhoDrDefault,
// src/blastwave/blastwave_components.h:846-893): uniform for n_ism > 0, else constant inside r_ej,
// A0 R^-s between r_ej and r_ism, constant beyond.  rho > 1 g/cm^3 is where the reference exits.
__device__ __forceinline__ void ism_density(const pba_run_params &P, double R, double &rho, double &drhodr,
                                            int &status) {
    if (P.n_ism > 0.0) {
        rho = P.n_ism * kMppme;
        drhodr = 0.0;
    } else {
        if (R < P.ism_r_ej) { rho = P.ism_A0 * pow(P.ism_r_ej, -P.ism_s) * kMppme; drhodr = 0.0; }
        if ((P.ism_r_ej <= R) && (R < P.ism_r_ism)) {
            rho = P.ism_A0 * pow(R, -P.ism_s) * kMppme;
            drhodr = rho * (-P.ism_s / R);
        }
        if (P.ism_r_ism <= R) { rho = P.ism_A0 * pow(P.ism_r_ism, -P.ism_s) * kMppme; drhodr = 0.0; }
    }
    if (rho < 0 || rho > 1.) status |= PBAGPU_S_NONFINITE;
}

__device__ void rhs_fs(const pba_run_params &P, const Shock &shfs, BW &b, const double *Y, double *o) {
    double R = Y[iR], Gamma = Y[iGamma], Eint2 = Y[iEint2], theta = Y[itheta], M2 = Y[iM2];
    if (Gamma < 1 || M2 < 0 || theta < 0) b.status |= PBAGPU_S_NONFINITE;
    double rho, drhodr;
    ism_density(P, R, rho, drhodr, b.status);
    rho /= b.M0;
    drhodr /= b.M0;
    double beta = beta_from_gamma(Gamma);
    double adi = adi_nava(Gamma);
    double Gsh = gamma_shock(Gamma, adi);
    double dRdt = beta * kC;
    double dthetadr = 0.0;
    if ((theta < b.theta_max) && !b.end_spreading && Gamma * beta < fmax(P.mom_when_start_spread, 0. * b.mom0))
        dthetadr = dthetadr_our(b, Gamma, Gsh, R, adi, theta);
    double sth, cth;
    sincos(theta, &sth, &cth);
    double dM2dR = 2. * kPi * rho * ((1. - cth) + (1. / 3.) * sth * R * dthetadr) * R * R / b.ncells;
    double GE = (adi * Gamma * Gamma - adi + 1.0) / Gamma;
    double dGE = (adi * Gamma * Gamma + adi - 1.0) / (Gamma * Gamma);
    double num = (Gamma - 1.) * (GE + 1.) * dM2dR - GE * (adi - 1.) * Eint2 * (dM2dR / M2 - drhodr / rho);
    double den = (1. + M2) + Eint2 * dGE + GE * (adi - 1.) * Eint2 / Gamma;
    double dGdR = -num / den;
    double dEsh2dR = (Gamma - 1.0) * dM2dR;
    double dlnV2dR = dM2dR / M2 - drhodr / rho - dGdR / Gamma;
    double dEad2dR = P.adiab_loss ? -(adi - 1.0) * Eint2 * dlnV2dR : 0.0;
    double eps_rad = P.fs.epsilon_e_rad;
    if (P.fs.epsilon_e_rad < 0)
        eps_rad = radiation_loss(P.fs, shfs, b.fs_st, dM2dR * b.M0, dEsh2dR * b.M0 * kC * kC, R, Gamma, Gsh, adi,
                                 M2 * b.M0, Eint2 * b.M0 * kC * kC, rho * b.M0, cth, b.ncells);
    double dErad2dR = eps_rad * P.fs.eps_e * dEsh2dR;
    double dEint2dR = dEsh2dR + dEad2dR - dErad2dR;
    double mom = mom_from_gamma(Gamma);
    double bm = beta_from_mom(mom);
    double dttdr = 1. / (kC * bm) * sqrt(1. + R * R * dthetadr * dthetadr) - (1. / kC);
    if (!isfinite(dGdR) || !isfinite(dlnV2dR) || !isfinite(dEint2dR) || fabs(dRdt * dGdR) > b.Gamma0)
        b.status |= PBAGPU_S_NONFINITE;
    o[iR] = dRdt;
    o[iRsh] = beta_from_gamma(Gsh) * kC;
    o[itt] = dRdt * dttdr;
    o[itcomov] = dRdt * (1.0 / mom / kC);
    o[iGamma] = dRdt * dGdR;
    o[iEint2] = dRdt * dEint2dR;
    o[itheta] = dRdt * dthetadr;
    o[iErad2] = dRdt * dErad2dR;
    o[iEsh2] = dRdt * dEsh2dR;
    o[iEad2] = dRdt * dEad2dR;
    o[iM2] = dRdt * dM2dR;
    o[iEint3] = 0.; o[iRrsh] = 0.; o[iErad3] = 0.; o[iEsh3] = 0.; o[iEad3] = 0.; o[iM3] = 0.; o[ideltaR4] = 0.;
}

__device__ void rhs_fsrs(const pba_run_params &P, const Shock &shfs, BW &b, const double *Y, double *o) {
    double R = Y[iR], Gamma = Y[iGamma], Eint2 = Y[iEint2], Eint3 = Y[iEint3], theta = Y[itheta];
    double M2 = Y[iM2], M3 = Y[iM3], deltaR4 = Y[ideltaR4];
    if (Gamma < 1.) Gamma = (b.prev_idx_x == 0) ? b.Gamma0 : 1. + 1.e-5;
    if (Eint3 < 0) Eint3 = 0.0;
    if (M3 < 0 && b.prev_idx_x == 0) { M3 = 0; Eint3 = 0; deltaR4 = 0; }
    double beta = beta_from_gamma(Gamma);
    double G0 = b.Gamma0;
    double G43 = Gamma * G0 * (1.0 / (G0 * G0) + 1.0 / (Gamma * Gamma) - 1.0 / (Gamma * Gamma) / (G0 * G0)) /
                 (1.0 + beta * b.beta0);
    if (G43 < 1.) G43 = 1.;
    double adi3 = adi_nava(G43);
    double GRsh = gamma_shock(G43, adi3);
    double rho, drhodr;
    ism_density(P, R, rho, drhodr, b.status);
    rho /= b.M0;
    drhodr /= b.M0;
    double dRdt = beta * kC;
    double adi = adi_nava(Gamma);
    double Gsh = gamma_shock(Gamma, adi);
    double dthetadr = 0.0;
    if ((theta < b.theta_max) && !b.end_spreading && Gamma * beta < fmax(P.mom_when_start_spread, 0. * b.mom0))
        dthetadr = dthetadr_our(b, Gamma, Gsh, R, adi, theta);
    double sth, cth;
    sincos(theta, &sth, &cth);
    double dM2dR = 2. * kPi * rho * ((1. - cth) + (1. / 3.) * sth * R * dthetadr) * R * R / b.ncells;
    double ddR4 = 0., dM3dR = 0., dlnrho4dR = 0.;
    if (!b.shutOff && Gamma < 0.9999 * G0) {
        double alpha_of = P.tprompt * b.beta0 * kC;
        double b0 = b.beta0;
        ddR4 = (1.0 / (beta * beta * beta * beta) - 1.0 / (b0 * b0 * b0 * b0)) / (1.0 / beta + 1.0 / b0) /
               (1.0 / (beta * beta) + 1.0 / (b0 * b0)) * b0;
        if (P.exponential_rho4) {
            dM3dR = (1.0 / alpha_of) * ddR4 * exp(-deltaR4 / alpha_of);
            dlnrho4dR = -2.0 / R - ddR4 / alpha_of;
        } else {
            dM3dR = (1.0 / alpha_of) * ddR4;
            dlnrho4dR = -2.0 / R;
        }
    }
    double dlnrhodr = drhodr / rho;
    double dGdR;
    {
        double dg43 = G0 - (G0 * G0 * Gamma - Gamma) / sqrt(Gamma * Gamma * G0 * G0 - G0 * G0 - Gamma * Gamma + 1.0);
        double GE3 = (adi3 * Gamma * Gamma - adi3 + 1.0) / Gamma;
        double dGE3 = adi3 * (1.0 + 1.0 / (Gamma * Gamma)) - dg43 / 3.0 / G43 * G43 * (Gamma - 1.0 / Gamma) -
                      1.0 / (Gamma * Gamma);
        double dGE = (adi * Gamma * Gamma + adi - 1.0) / (Gamma * Gamma);
        double GE = (adi * Gamma * Gamma - adi + 1.0) / Gamma;
        double f_2 = GE * (adi - 1.0) * Eint2 / Gamma;
        double h_2 = GE * (adi - 1.0) * Eint2 * (dM2dR / M2 - dlnrhodr);
        double fh3 = GE3 * (adi3 - 1.0) * Eint3;
        double f_3 = fh3 / G43 * dg43;
        double h_3 = (Eint3 != 0) ? fh3 * (dM3dR / M3 - dlnrho4dR) : 0.0;
        dGdR = -((Gamma - 1.0) * (GE + 1.0) * dM2dR + (Gamma - G0 + GE3 * (G43 - 1.0)) * dM3dR - h_2 - h_3) /
               ((M2 + M3) + Eint2 * dGE + Eint3 * dGE3 + f_2 + f_3);
    }
    if (dGdR > 0 && Gamma > P.rs_gamma0_frac_no_exceed * G0) dGdR = 0.0;
    double dEsh2dR = (Gamma - 1.0) * dM2dR;
    double dEsh3dR = (G43 - 1.0) * dM3dR;
    double dlnV2dR = dM2dR / M2 - drhodr / rho - dGdR / Gamma;
    double dEad2dR = P.adiab_loss ? -(adi - 1.0) * Eint2 * dlnV2dR : 0.0;
    double dEad3dR = 0.;
    if (!b.shutOff && M3 > 0.0) {
        double dlnV3dR = dM3dR / M3 - dlnrho4dR - dGdR / Gamma;
        if (P.adiab_loss_rs) dEad3dR = -(adi3 - 1) * Eint3 * dlnV3dR;
    }
    double eps_rad = P.fs.epsilon_e_rad;
    if (P.fs.epsilon_e_rad < 0)
        eps_rad = radiation_loss(P.fs, shfs, b.fs_st, dM2dR * b.M0, dEsh2dR * b.M0 * kC * kC, R, Gamma, Gsh, adi,
                                 M2 * b.M0, Eint2 * b.M0 * kC * kC, rho * b.M0, cth, b.ncells);
    double dErad2dR = eps_rad * P.fs.eps_e * dEsh2dR;
    // the reverse-shock loss is evaluated with rho4 = 0 and returns 0 (blastwave.h:1991-1997)
    double eps_rad_rs = P.rs.epsilon_e_rad < 0 ? 0. : P.rs.epsilon_e_rad;
    double dErad3dR = eps_rad_rs * P.rs.eps_e * dEsh3dR;
    double dEint2dR = dEsh2dR + dEad2dR - dErad2dR;
    double dEint3dR = dEsh3dR + dEad3dR - dErad3dR;
    double mom = Gamma * beta;
    double bm = beta_from_mom(mom);
    double dttdr = 1. / (kC * bm) * sqrt(1. + R * R * dthetadr * dthetadr) - (1. / kC);
    if (!isfinite(dGdR) || dM2dR < 0. || !isfinite(dlnV2dR)) b.status |= PBAGPU_S_NONFINITE;
    o[iR] = dRdt;
    o[iRsh] = beta_from_gamma(Gsh) * kC;
    o[itt] = dRdt * dttdr;
    o[itcomov] = dRdt * (1.0 / beta / Gamma / kC);
    o[iGamma] = dRdt * dGdR;
    o[iRrsh] = beta_from_gamma(GRsh) * kC;
    o[iEint2] = dRdt * dEint2dR;
    o[iEint3] = dRdt * dEint3dR;
    o[itheta] = dRdt * dthetadr;
    o[iErad2] = dRdt * dErad2dR;
    o[iErad3] = dRdt * dErad3dR;
    o[iEsh2] = dRdt * dEsh2dR;
    o[iEsh3] = dRdt * dEsh3dR;
    o[iEad2] = dRdt * dEad2dR;
    o[iEad3] = dRdt * dEad3dR;
    o[iM2] = dRdt * dM2dR;
    o[iM3] = dRdt * dM3dR;
    o[ideltaR4] = dRdt * ddR4;
}

// Kind of the blastwaves of a launch: the jet's (rhs_fs / rhs_fsrs), or kilonova ejecta evolved by
// rhs_fs_dense with dt_obs/dR in its radial or (method_spread set) spreading form.
enum Kind { KIND_JET = 0, KIND_EJECTA = 1, KIND_EJECTA_SPREAD = 2 };

// The upstream state that an ejecta blastwave's last RHS call left in its density object
// (set_standard_ism, src/blastwave/blastwave.h:1282-1295); addOtherVars stores it (:539-560).
struct Upstream { double rho, drhodr, grel; int seen, pad_; };

// Ejecta blastwave in a standard ISM: setStandardISM + BlastWave::rhs_fs_dense
// (src/blastwave/blastwave.h:769-781, 2120-2497; src/model_evolve.h:1627-1651) with method_dgdr =
// our, method_dmdr = usingdthdr and no energy injection, nuclear heating or luminosity loss: the
// upstream is at rest (GammaRho = 1), GammaRel = Gamma, dtheta/dR = 0, and Rsh is not evolved.
__device__ void rhs_fs_dense(const pba_run_params &P, BW &b, Upstream &up, bool spread, const double *Y, double *o) {
    double Gamma = Y[iGamma];
    const double R = Y[iR], Eint2 = Y[iEint2], theta = Y[itheta], M2 = Y[iM2];
    if (Gamma < 1.) { Gamma = 1. + 1.e-5; b.status |= PBAGPU_S_NONFINITE; }
    const double beta = beta_from_gamma(Y[iGamma]);
    if (!isfinite(R) || !isfinite(Gamma) || M2 < 0. || !isfinite(M2) || !isfinite(Eint2) || Eint2 < 0)
        b.status |= PBAGPU_S_NONFINITE;
    double rho_c, drhodr_c;
    ism_density(P, R, rho_c, drhodr_c, b.status);
    const double GRel = Gamma;   // GammaRel, clamped with Gamma (setStandardISM)
    up.rho = rho_c; up.drhodr = drhodr_c; up.grel = GRel; up.seen = 1;
    const double rho = rho_c / b.M0, drhodr = drhodr_c / b.M0;
    const double dRdt = beta * kC;
    const double adi = adi_nava(GRel);
    const double dthetadr = 0.0;
    const double dM2dR = 2. * kPi * rho * ((1. - cos(theta)) + (1. / 3.) * sin(theta) * R * dthetadr) * R * R / b.ncells;
    // energy injection, nuclear heating and luminosity terms: all zero here
    const double Dop = Gamma / (1. - beta * cos(0.));
    const double dEinjdR = 0. / (b.M0 * kC * kC) / b.ncells / dRdt;
    const double dEinj_dop = dEinjdR / Dop / Dop;
    const double dEnucdR = 0. / (b.M0 * kC * kC) / dRdt, dElumdR = 0. / (b.M0 * kC * kC) / dRdt;
    const double GE = (adi * GRel * GRel - adi + 1.0) / GRel;
    const double dGE = (adi * GRel * GRel + adi - 1.0) / (GRel * GRel);
    const double num1 = (Gamma - 1. + GE * (GRel - 1.)) * dM2dR;
    const double num2 = -GE * (adi - 1.) * Eint2 * (dM2dR / M2 - drhodr / rho - 0. / 1.);
    const double num3 = -(1. + GE / Gamma * 1.) * dEinjdR;
    const double den = (1. + M2) + Eint2 * dGE + GE * (adi - 1.) * Eint2 * 1. / GRel;
    const double dGdR = -1. * (num1 + num2 + num3) / den;
    const double dlnV2dR = dM2dR / M2 - drhodr / rho - (1. / GRel) * 1. * dGdR + 0. / 1.;
    const double dEad2dR = P.adiab_loss ? -(adi - 1.0) * Eint2 * dlnV2dR : 0.0;
    const double dEsh2dR = (GRel - 1.0) * dM2dR;   // a subsonic upstream would scale it by 1e-10; c_s = 0 here
    const double dErad2dR = P.fs.epsilon_e_rad * dEsh2dR;
    const double dEint2dR = dEsh2dR + dEad2dR - dErad2dR + dEnucdR - dElumdR + dEinj_dop;
    // EQS::evalElapsedTime (src/blastwave/blastwave_components.h:165-175)
    const double mom = Gamma * beta;
    const double bm = beta_from_mom(mom);
    double dttdr;
    if (spread) dttdr = 1. / (kC * bm) * sqrt(1. + R * R * dthetadr * dthetadr) - (1. / kC);
    else dttdr = 1. / (kC * (1. + mom * mom * mom / sqrt(1. + mom * mom)) * (1. + bm));
    if (!isfinite(dRdt) || !isfinite(dGdR) || dM2dR < 0. || !isfinite(dlnV2dR)) b.status |= PBAGPU_S_NONFINITE;
    o[iR] = dRdt;
    o[itt] = dRdt * dttdr;
    o[itcomov] = dRdt * (1.0 / beta / Gamma / kC);
    o[iGamma] = dRdt * dGdR;
    o[iEint2] = dRdt * dEint2dR;
    o[itheta] = dRdt * dthetadr;
    o[iErad2] = dRdt * dErad2dR;
    o[iEsh2] = dRdt * dEsh2dR;
    o[iEad2] = dRdt * dEad2dR;
    o[iM2] = dRdt * dM2dR;
}

template <int KIND>
__device__ __forceinline__ void rhs(const pba_run_params &P, const Shock &shfs, BW &b, Upstream *up, bool active,
                                    const double *Y, double *o) {
    if (!active || b.end_evolution) return;  // derivative entries of inactive blastwaves keep their (zero) values
    if constexpr (KIND != KIND_JET) rhs_fs_dense(P, b, *up, KIND == KIND_EJECTA_SPREAD, Y, o);
    else if (b.fsrs) rhs_fsrs(P, shfs, b, Y, o);
    else rhs_fs(P, shfs, b, Y, o);
}

struct Lim { double B, gm, gM, gc, Theta, z_cool, nprime, acc; bool ok; };

__device__ Lim ele_limits(const pba_shock_params &s, const Shock &sh, MState &st, double U_p, double Gsh, double t_e,
                          double tcomov, double n_prime) {
    Lim L;
    st.tcomov = tcomov;
    double betaSh = beta_from_gamma(Gsh);
    L.ok = !(betaSh < s.beta_min);
    if (!L.ok) return L;
    double B = sqrt(8. * kPi * (U_p * s.eps_b));
    double gM = sqrt(6.0 * kPi * kQe / kSigmaT / B);
    if (gM > 0.99 * sh.etop) gM = 0.99 * sh.etop;
    double gm1 = (s.p - 2.) / (s.p - 1.) * (U_p * s.eps_e) / (n_prime * kMe * kC * kC);
    double Theta0 = s.eps_t * (9.0 * 0.62 * kMp / (32.0 * 1.18 * kMe)) * betaSh * betaSh;
    double Theta = (5.0 * Theta0 - 6.0 + sqrt(25.0 * Theta0 * Theta0 + 180.0 * Theta0 + 36.0)) / 30.0;
    double z_cool = (6.0 * kPi * kMe * kC / (kSigmaT * B * B * t_e)) / Theta;
    // bisection for gamma_min (rootfinders.h:17-58 with microphysics_base.h:792-806)
    double gamma_ave = s.eps_e * U_p / (n_prime * kMe * kC * kC);
    double p = s.p;
    double gMa = pow(gM, -p + 2), gMb = pow(gM, -p + 1);
    auto f = [&](double x) { return (p - 1) / (p - 2) * (gMa - pow(x, -p + 2)) / (gMb - pow(x, -p + 1)) - gamma_ave; };
    int status = 0;
    double gm;
    {
        double xa = 1., xb = 1.e8, fa = f(xa), fb = f(xb);
        if (fa * fb > 0) { status = -1; gm = 0.; }
        else if (fa == 0) { status = -1; gm = xa; }
        else if (fb == 0) { status = -1; gm = xb; }
        else {
            double dm = xb - xa;
            int i;
            gm = 0.;
            for (i = 0; i < 100; i++) {
                dm *= 0.5;
                double xm = xa + dm;
                double fm = f(xm);
                if (fm * fa >= 0) xa = xm;
                if (fm == 0 || fabs(dm) < .001 * fabs(xm)) { gm = xm; break; }
            }
            if (i == 100) { status = -1; gm = xa; }
        }
    }
    if (status < 0 && gm1 > 1.5) gm = gm1;
    if (gm < 2. || status < 0) {
        double sir = (p - 2.0) / (p - 1.0) * s.eps_e * U_p / (n_prime * kMe * kC * kC);
        st.acc = sir > 1. ? 1. : sir;
        gm = gm < 1. ? 1. : gm;
    }
    if (gm < sh.e0) gm = sh.e0;
    double gc = (6. * kPi * kMe * kC) / (kSigmaT * B * B * tcomov);
    if (gc <= 1.) gc = 1.001;
    if (gc > gM) gc = gM;
    L.B = B; L.gm = gm; L.gM = gM; L.gc = gc; L.Theta = Theta; L.z_cool = z_cool; L.nprime = n_prime; L.acc = st.acc;
    return L;
}

__device__ double ctheta_of(double theta, int ilayer, int nlayers) {
    if (ilayer <= 0) return 0.;
    double fac0 = (double)ilayer / (double)nlayers, fac1 = (double)(ilayer + 1) / (double)nlayers;
    double xx2 = 2. * asin(fac1 * sin(theta / 2.)), xx1 = 2. * asin(fac0 * sin(theta / 2.));
    return 0.5 * (xx1 + xx2);
}

template <int KIND>
__device__ __forceinline__ void post_step(const pba_run_params &P, const Shock &shfs, const Shock &shrs, BW &b,
                                          const Upstream *up, const double *Yc, int it, double tb, double *T, int nt) {
    if (b.end_evolution) return;
    auto W = [&](int v, double x) { T[(size_t)v * nt + it] = x; };
    auto Rd = [&](int v, int t) { return T[(size_t)v * nt + t]; };
    const double u = b.M0 * kC * kC;
    double M2 = Yc[iM2] * b.M0, Eint2 = Yc[iEint2] * u;
    double G = Yc[iGamma];
    W(PV_tburst, tb); W(PV_R, Yc[iR]); W(PV_Rsh, Yc[iRsh]); W(PV_tt, Yc[itt]);
    W(PV_mom, mom_from_gamma(G)); W(PV_Gamma, G); double beta = beta_from_gamma(G); W(PV_beta, beta);
    W(PV_theta, Yc[itheta]); W(PV_M2, M2); W(PV_tcomov, Yc[itcomov]); W(PV_Ead2, Yc[iEad2] * u);
    W(PV_Eint2, Eint2); W(PV_Esh2, Yc[iEsh2] * u); W(PV_Erad2, Yc[iErad2] * u);
    double M3 = Yc[iM3] * b.M0, Eint3 = Yc[iEint3] * u;
    if (b.fsrs) {
        W(PV_Rrsh, Yc[iRrsh]); W(PV_Eint3, Eint3); W(PV_Ead3, Yc[iEad3] * u); W(PV_Erad3, Yc[iErad3] * u);
        W(PV_Esh3, Yc[iEsh3] * u); W(PV_M3, M3); W(PV_deltaR4, Yc[ideltaR4]);
    }
    double R = Yc[iR], theta = Yc[itheta];
    if (R < 1. || G < 1.) b.status |= PBAGPU_S_NONFINITE;
    W(PV_ctheta, ctheta_of(theta, b.ilayer, b.nlayers));
    // addOtherVars stores the density set at the blastwave's start (m_rho_, blastwave.h:92-101,
    // 539-541), and the downstream quantities are formed from it; an ejecta blastwave's density
    // object holds what its last RHS call evaluated, and GammaRel (unset, -1, before the first
    // call: then Gamma is stored, :544-560)
    double rho = b.rho_def, drhodr = b.drhodr0;
    if constexpr (KIND != KIND_JET) {
        if (up->seen) { rho = up->rho; drhodr = up->drhodr; }
        const double grel = up->seen ? up->grel : -1.;
        W(PV_GammaRel, grel < 1. ? G : grel);
    }
    W(PV_rho, rho); W(PV_drhodr, drhodr);
    double adi = adi_nava(G);
    double rho2 = rho2t(G, adi, rho);
    double Gsh = gamma_shock(G, adi);
    W(PV_adi, adi); W(PV_rho2, rho2); W(PV_GammaFsh, Gsh);
    double omc = (1. - cos(theta) / b.ncells);
    W(PV_thickness, M2 / (2. * kPi * kMp * R * R * omc * G * (rho2 / kMp)));
    double U_p = (Eint2 == 0 || M2 == 0) ? 0. : Eint2 / (M2 / rho2);
    W(PV_U_p, U_p);
    double GRsh = 0., U_p3 = 0., rho3 = 0.;
    if (it > 0 && b.do_rs && !b.shutOff && b.fsrs) {
        double lbeta = beta_from_mom(mom_from_gamma(G)), G0 = b.Gamma0;
        double G43 = G * G0 * (1.0 / (G0 * G0) + 1.0 / (G * G) - 1.0 / (G * G) / (G0 * G0)) / (1.0 + lbeta * b.beta0);
        double alpha_of = P.tprompt * b.beta0 * kC;
        double rho4f = b.M0 / (2.0 * alpha_of * kPi * (1.0 - cos(b.theta_b0))) / (R * R);
        double rho4 = P.exponential_rho4 ? rho4f * exp(-Yc[ideltaR4] / alpha_of) : rho4f;
        double adi3 = adi_nava(G43);
        rho3 = rho2t(G, adi3, rho4);
        GRsh = gamma_shock(G43, adi3);
        W(PV_Gamma43, G43); W(PV_rho4, rho4); W(PV_adi3, adi3); W(PV_rho3, rho3); W(PV_GammaRsh, GRsh);
        W(PV_thickness_rs, M3 / (2. * kPi * kMp * R * R * omc * G43 * (rho3 / kMp)));
        U_p3 = (Eint3 == 0 || M3 == 0) ? 0. : Eint3 / (M3 / rho3);
        W(PV_U_p3, U_p3);
    }
    if (adi < 1. || rho2 < 0. || U_p < 0 || Gsh < 1.) b.status |= PBAGPU_S_NONFINITE;
    int flags = 0;
    bool valid = b.R0 >= 0;
    if (valid && beta < P.fs.beta_min) {
        if (b.i_end_r > it - 1) b.i_end_r = it - 1;
        valid = false;
    }
    if constexpr (KIND != KIND_JET) {   // subsonic upstream (sound speed 0 here): no shock (blastwave_radiation.h:918-928)
        if (valid && 0. >= beta) {
            W(PV_B, 0.);
            if (b.i_end_r > it - 1) b.i_end_r = it - 1;
            valid = false;
        }
    }
    if (valid && (U_p <= 0 || rho2 <= 0)) valid = false;
    if (valid && R < 1) valid = false;
    if (valid) {
        bool rs = b.fsrs && b.do_rs_rad && it > 0 && Rd(PV_thickness_rs, it - 1) > 0 && GRsh > 0 && U_p3 > 0;
        if (rs) {
            Lim L = ele_limits(P.rs, shrs, b.rs_st, U_p3, GRsh, Yc[itt], Yc[itcomov], rho3 / kMp);
            if (L.ok) {
                W(PV_B3, L.B); W(PV_gm_rs, L.gm); W(PV_gM_rs, L.gM); W(PV_gc_rs, L.gc); W(PV_Theta_rs, L.Theta);
                W(PV_z_cool_rs, L.z_cool); W(PV_nprime_rs, L.nprime); W(PV_accel_frac_rs, L.acc);
            }
            flags |= 2;
        }
        Lim L = ele_limits(P.fs, shfs, b.fs_st, U_p, Gsh, Yc[itt], Yc[itcomov], rho2 / kMp);
        if (L.ok) {
            W(PV_B, L.B); W(PV_gm, L.gm); W(PV_gM, L.gM); W(PV_gc, L.gc); W(PV_Theta, L.Theta);
            W(PV_z_cool, L.z_cool); W(PV_nprime, L.nprime); W(PV_accel_frac, L.acc);
        }
        flags |= 1;
    }
    W(PV_mflags, (double)flags);
}

// DOP853 tableau (Hairer & Wanner's published dop853 coefficients)
namespace tab {
__constant__ double c_[12] = {0., 0.526001519587677318785587544488E-01, 0.789002279381515978178381316732E-01,
                              0.118350341907227396726757197510E+00, 0.281649658092772603273242802490E+00,
                              0.333333333333333333333333333333E+00, 0.25E+00, 0.307692307692307692307692307692E+00,
                              0.651282051282051282051282051282E+00, 0.6E+00, 0.857142857142857142857142857142E+00, 1.0};
}
constexpr double a21 = 5.26001519587677318785587544488E-2, a31 = 1.97250569845378994544595329183E-2,
                 a32 = 5.91751709536136983633785987549E-2, a41 = 2.95875854768068491816892993775E-2,
                 a43 = 8.87627564304205475450678981324E-2, a51 = 2.41365134159266685502369798665E-1,
                 a53 = -8.84549479328286085344864962717E-1, a54 = 9.24834003261792003115737966543E-1,
                 a61 = 3.7037037037037037037037037037E-2, a64 = 1.70828608729473871279604482173E-1,
                 a65 = 1.25467687566822425016691814123E-1, a71 = 3.7109375E-2,
                 a74 = 1.70252211019544039314978060272E-1, a75 = 6.02165389804559606850219397283E-2,
                 a76 = -1.7578125E-2, a81 = 3.70920001185047927108779319836E-2,
                 a84 = 1.70383925712239993810214054705E-1, a85 = 1.07262030446373284651809199168E-1,
                 a86 = -1.53194377486244017527936158236E-2, a87 = 8.27378916381402288758473766002E-3,
                 a91 = 6.24110958716075717114429577812E-1, a94 = -3.36089262944694129406857109825,
                 a95 = -8.68219346841726006818189891453E-1, a96 = 2.75920996994467083049415600797E+1,
                 a97 = 2.01540675504778934086186788979E+1, a98 = -4.34898841810699588477366255144E+1,
                 aA1 = 4.77662536438264365890433908527E-1, aA4 = -2.48811461997166764192642586468,
                 aA5 = -5.90290826836842996371446475743E-1, aA6 = 2.12300514481811942347288949897E+1,
                 aA7 = 1.52792336328824235832596922938E+1, aA8 = -3.32882109689848629194453265587E+1,
                 aA9 = -2.03312017085086261358222928593E-2, aB1 = -9.3714243008598732571704021658E-1,
                 aB4 = 5.18637242884406370830023853209, aB5 = 1.09143734899672957818500254654,
                 aB6 = -8.14978701074692612513997267357, aB7 = -1.85200656599969598641566180701E+1,
                 aB8 = 2.27394870993505042818970056734E+1, aB9 = 2.49360555267965238987089396762,
                 aBA = -3.0467644718982195003823669022, aC1 = 2.27331014751653820792359768449,
                 aC4 = -1.05344954667372501984066689879E+1, aC5 = -2.00087205822486249909675718444,
                 aC6 = -1.79589318631187989172765950534E+1, aC7 = 2.79488845294199600508499808837E+1,
                 aC8 = -2.85899827713502369474065508674, aC9 = -8.87285693353062954433549289258,
                 aCA = 1.23605671757943030647266201528E+1, aCB = 6.43392746015763530355970484046E-1;
constexpr double b1 = 5.42937341165687622380535766363E-2, b6 = 4.45031289275240888144113950566,
                 b7 = 1.89151789931450038304281599044, b8 = -5.8012039600105847814672114227,
                 b9 = 3.1116436695781989440891606237E-1, bA = -1.52160949662516078556178806805E-1,
                 bB = 2.01365400804030348374776537501E-1, bC = 4.47106157277725905176885569043E-2;
constexpr double er1 = 0.1312004499419488073250102996E-01, er6 = -0.1225156446376204440720569753E+01,
                 er7 = -0.4957589496572501915214079952E+00, er8 = 0.1664377182454986536961530415E+01,
                 er9 = -0.3503288487499736816886487290E+00, erA = 0.3341791187130174790297318841E+00,
                 erB = 0.8192320648511571246570742613E-01, erC = -0.2235530786388629525884427845E-01;
constexpr double bhh1 = 0.244094488188976377952755905512E+00, bhh9 = 0.733846688281611857341361741547E+00,
                 bhhC = 0.220588235294117647058823529412E-01;

// Stage vectors live in LDS, KS slots of [NA][bw_per_run] doubles.  Slot of each k_i:
//   k1 -> 0; k2, k3, k5 -> 1 (each is dead once the next of them is formed); k4, k12 -> 2;
//   k6 .. k11 -> 3 .. 8.  The RK4 integrator keeps its running increment in slot 3.
constexpr int KS = 9;
// Stage s = 2 .. 12 of one DOP853 step: x_s = Y + h * sum_j a_sj k_j over the non-zero a_sj,
// added in the tableau's order; the stage's derivative goes to slot st_out[s].
__constant__ int st_n[13] = {0, 0, 1, 2, 2, 3, 3, 4, 5, 6, 7, 8, 9};
__constant__ int st_out[13] = {0, 0, 1, 1, 2, 1, 3, 4, 5, 6, 7, 8, 2};
__constant__ int st_slot[13][9] = {
    {0}, {0}, {0}, {0, 1}, {0, 1}, {0, 1, 2}, {0, 2, 1}, {0, 2, 1, 3}, {0, 2, 1, 3, 4}, {0, 2, 1, 3, 4, 5},
    {0, 2, 1, 3, 4, 5, 6}, {0, 2, 1, 3, 4, 5, 6, 7}, {0, 2, 1, 3, 4, 5, 6, 7, 8}};
__constant__ double st_a[13][9] = {
    {0}, {0}, {a21}, {a31, a32}, {a41, a43}, {a51, a53, a54}, {a61, a64, a65}, {a71, a74, a75, a76},
    {a81, a84, a85, a86, a87}, {a91, a94, a95, a96, a97, a98}, {aA1, aA4, aA5, aA6, aA7, aA8, aA9},
    {aB1, aB4, aB5, aB6, aB7, aB8, aB9, aBA}, {aC1, aC4, aC5, aC6, aC7, aC8, aC9, aCA, aCB}};

// Sums of two values per lane over every run's lanes [base, base + n), added in lane order (each
// lane gets its run's sums).  Called by the whole wavefront.
__device__ __forceinline__ void run_sums(double &v1, double &v2, double *red, int base, int n) {
    const int lane = threadIdx.x;
    __syncthreads();
    red[lane] = v1;
    red[64 + lane] = v2;
    __syncthreads();
    double s1 = 0., s2 = 0.;
    for (int i = 0; i < n; i++) { s1 += red[base + i]; s2 += red[64 + base + i]; }
    v1 = s1;
    v2 = s2;
}

// Program of one output interval: every pass of the loop evaluates the right-hand side exactly
// once (one call site, so one copy of the RHS in the kernel's code) and then moves the integrator
// to the state that needs the next evaluation.
enum Phase { PH_K1I, PH_H2, PH_STAGE, PH_FSAL, PH_RK1, PH_RK2, PH_RK3, PH_RK4 };

// One wavefront integrates rpw runs: lane l owns blastwave l mod bw_per_run of run
// blockIdx.x * rpw + l / bw_per_run (rpw * bw_per_run <= 64).  Each run keeps its own output
// interval, step size and integrator phase in its lanes, so the runs of a wavefront need not step
// together; every pass of the loop evaluates one right-hand side on every lane whose run is still
// integrating, and the loop ends when the last of them has done its grid.  The run-wide norms of
// the step control are sums over the run's lanes in lane order, as with one run per wavefront, so
// a run's arithmetic does not depend on its neighbours.
// KG = false: one run per wavefront (rpw = 1), stage vectors in LDS as [KS][NA][bw_per_run].
// KG = true: stage vectors in the global work array kg, [wavefront][KS][NA][64] (a wavefront of
// three 21-blastwave runs would need 82 KB of LDS, one wavefront per CU).
// nt is the table's step stride; run r integrates nt_run[r] steps (all nt when nt_run is null) on
// the time grid tgrid + r * tg_stride (tg_stride 0: one grid for every run).
template <bool KG, int KIND = KIND_JET>
__global__ void __launch_bounds__(64) dyn_kernel(const pba_run_params *__restrict__ params,
                                                 const pba_layer_init *__restrict__ init, int n_runs, int bw_per_run,
                                                 int rpw, int nt, const int *__restrict__ nt_run,
                                                 const double *__restrict__ tgrid_all, int tg_stride,
                                                 double *__restrict__ table, pba_bw_result *__restrict__ res,
                                                 double *__restrict__ kg) {
    // LDS: (KG = false: [KS][NA][bw_per_run] stage vectors,) the state Y as [NA][ys], 128 doubles for
    // sums, then one BW per lane slot (ys slots), then (ejecta) one Upstream per lane slot
    extern __shared__ double dyn_lds[];
    const int ys = KG ? 64 : bw_per_run;
    double *ylds = KG ? dyn_lds : dyn_lds + (size_t)KS * NA * bw_per_run;
    double *red = ylds + (size_t)NA * ys;
    BW *bws = reinterpret_cast<BW *>(red + 128);
    const int lane = threadIdx.x;
    const int g = lane / bw_per_run, li = lane - g * bw_per_run;
    const int run0 = blockIdx.x * rpw;
    const int run = run0 + g;
    const bool active = g < rpw && run < n_runs;           // the lane holds a blastwave
    const int runc = KG ? (active ? run : run0) : run0;     // one run per wavefront: uniform
    const int base = active ? g * bw_per_run : 0;          // first lane of the lane's run
    double *kw = KG ? kg + (size_t)blockIdx.x * KS * NA * 64 : nullptr;
    // the ODE state of the lane's blastwave (in LDS: registers are the loop's scarcest resource)
    auto Yv = [&](int c) -> double & { return ylds[c * ys + lane]; };
    auto K = [&](int slot, int c) -> double & {
        if constexpr (KG) return kw[(slot * NA + c) * 64 + lane];
        else return dyn_lds[(slot * NA + c) * bw_per_run + lane];
    };
    const int ntr = nt_run ? nt_run[runc] : nt;
    const double *__restrict__ tgrid = tgrid_all + (size_t)runc * tg_stride;
    // one run per wavefront: a uniform copy (scalar registers); packed: read where used (a
    // lane-varying copy of the 400-byte struct would live in scratch)
    const pba_run_params Pu = params[run0];
    const pba_run_params &P = KG ? params[runc] : Pu;
    Shock shfs, shrs;
    {
        double step = exp(log(P.fs.gam2) / (double)P.fs.ngam);
        shfs.e0 = pow(step, 0.); shfs.etop = pow(step, (double)(P.fs.ngam - 1));
        double stepr = exp(log(P.rs.gam2) / (double)P.rs.ngam);
        shrs.e0 = pow(stepr, 0.); shrs.etop = pow(stepr, (double)(P.rs.ngam - 1));
        shfs.nu0_43 = shrs.nu0_43 = pow(1e6, 4. / 3.);   // computeRadiationLoss's nu0 = 1e6 Hz
    }
    // the blastwave's constants and flags (in LDS; padding lanes alias slot 0 and never touch it)
    BW &b = bws[active ? lane : 0];
    if (active) b = BW{};
    Upstream *up = nullptr;
    if constexpr (KIND != KIND_JET) {
        up = reinterpret_cast<Upstream *>(bws + ys) + (active ? lane : 0);
        if (active) *up = Upstream{0., 0., -1., 0, 0};
    }
    double x[NA];
    for (int k = 0; k < NA; k++) x[k] = 0.;
    if (active) {
        for (int k = 0; k < NA; k++) Yv(k) = 0.;
        for (int sl = 0; sl < KS; sl++)
            for (int k = 0; k < NA; k++) K(sl, k) = 0.;
    }
    const size_t bw_global = (size_t)run * bw_per_run + li;
    {   // clear the tables of this wavefront's runs, all lanes striding over contiguous memory
        const int nrb = min(rpw, n_runs - run0);
        double *Tr = table + (size_t)run0 * bw_per_run * PBA_NV * nt;
        const size_t n = (size_t)nrb * bw_per_run * PBA_NV * nt;
        for (size_t i = lane; i < n; i += 64) Tr[i] = 0.;
        __syncthreads();
    }
    double *T = active ? table + bw_global * PBA_NV * nt : nullptr;
    if (active) {
        const pba_layer_init id = init[bw_global];
        // this blastwave's own electron grids (the grid top clamps gamma_max, computeGammaMax,
        // microphysics_base.h:824-837, which enters the radiative-loss term)
        if (id.ngam_fs > 0) {
            const double step = exp(log(P.fs.gam2) / (double)id.ngam_fs);
            shfs.e0 = pow(step, 0.); shfs.etop = pow(step, (double)(id.ngam_fs - 1));
        }
        if (id.ngam_rs > 0) {
            const double step = exp(log(P.rs.gam2) / (double)id.ngam_rs);
            shrs.e0 = pow(step, 0.); shrs.etop = pow(step, (double)(id.ngam_rs - 1));
        }
        b.E0 = id.E0; b.M0 = id.M0; b.mom0 = id.mom0;
        b.Gamma0 = gamma_from_mom(id.mom0);
        b.R0 = beta_from_mom(id.mom0) * kC * tgrid[0];
        b.theta_b0 = id.theta_b0; b.ncells = id.ncells; b.theta_core = id.theta_core; b.theta_max = P.theta_max;
        b.tan_ratio = tan(0.5 * b.theta_b0) / tan(0.5 * b.theta_core);
        b.ilayer = id.ilayer; b.nlayers = id.nlayers;
        b.fsrs = P.fsrs; b.do_rs = P.fsrs && P.do_rs; b.do_rs_rad = P.fsrs && P.do_rs_radiation;
        ism_density(P, b.R0, b.rho_def, b.drhodr0, b.status);   // at R0 (blastwave.h:92)
        b.fs_st.acc = 1.; b.fs_st.tcomov = -1.; b.rs_st.acc = 1.; b.rs_st.tcomov = -1.;
        b.i_end_r = ntr;
        if (b.fsrs) {
            double min_g0 = b.do_rs ? P.min_gamma0_for_rs : 0.;
            if (b.Gamma0 < min_g0) { b.do_rs = 0; b.do_rs_rad = 0; b.shutOff = 1; b.fsrs = 0; }
        }
        if (b.M0 == 0. && b.E0 == 0.) b.end_evolution = 1;
        else {
            b.beta0 = beta_from_gamma(b.Gamma0);
            double M20 = (2.0 / 3.0) * kPi * (1.0 - cos(b.theta_b0)) * b.rho_def * pow(b.R0, 3) / b.ncells;
            if (b.M0 <= 0.) b.M0 = b.E0 / (b.Gamma0 * kC * kC);
            Yv(iR) = b.R0;
            double b0m = beta_from_mom(b.mom0);
            // EQS::init_elapsed_time (blastwave_components.h:141-163): the spreading form for the jet
            // and for ejecta with method_spread set, else the radial one
            if constexpr (KIND == KIND_EJECTA) {
                const double xx = 1. + b.mom0 * b.mom0 * b.mom0 / (sqrt(1. + b.mom0 * b.mom0));
                Yv(itt) = b.R0 / (kC * xx * (1.0 + b0m));
            } else {
                Yv(itt) = b.R0 / (kC * b0m) * sqrt(1. + b.R0 * b.R0 * 0. * 0.) - b.R0 / kC;
            }
            Yv(itcomov) = b.R0 / (b.beta0 * b.Gamma0 * kC);
            Yv(iEint2) = (b.Gamma0 - 1.) * M20 / b.M0 + 0. / b.M0 / kC / kC;
            Yv(iGamma) = b.Gamma0;
            Yv(itheta) = b.theta_b0;
            Yv(iM2) = M20 / b.M0;
            if (b.fsrs && P.init_deltaR4) Yv(ideltaR4) = kC * P.tprompt * b.beta0;
            b.Rw0 = .5 * Yv(iR);
            b.Wtt0 = Yv(itt);
        }
        double yc[NA];
        for (int k = 0; k < NA; k++) yc[k] = Yv(k);
        post_step<KIND>(P, shfs, shrs, b, up, yc, 0, tgrid[0], T, nt);
    }
    const double rtol = P.rtol, atol = P.rtol;
    const double nsz = (double)(NEQ_REF * bw_per_run);
    double time = P.tb0;
    int nsteps = 0, nrhs = 0;
    // the run's state inside its current output interval it
    int it = 1;
    bool live = active && ntr > 1;
    double timeN = 0., hmax = 0., hrk = 0.;
    int ph = PH_K1I, st = 0, outslot = 0, guard = 0;
    double h = 0., h1 = 0., dnf = 0.;
    bool reject = false, last = false;
    // x_s for stage s, from the stage vectors
    auto combo = [&](int s) {
        const int n = st_n[s];
        for (int k = 0; k < NA; k++) x[k] = st_a[s][0] * K(st_slot[s][0], k);
        for (int j = 1; j < n; j++) {
            const double a = st_a[s][j];
            const int sl = st_slot[s][j];
            for (int k = 0; k < NA; k++) x[k] += a * K(sl, k);
        }
        for (int k = 0; k < NA; k++) x[k] = Yv(k) + h * x[k];
    };
    // start an attempt of one DOP853 step of size h (ode_solvers.h:1434-1479)
    auto begin_attempt = [&]() {
        if ((time + 1.01 * h - timeN) > 0.0) { h = timeN - time; last = true; }
        st = 2;
        combo(2);
        outslot = st_out[2];
        ph = PH_STAGE;
    };
    auto start_interval = [&]() {
        timeN = time + (tgrid[it] - tgrid[it - 1]);
        hmax = fabs(timeN - time);
        hrk = tgrid[it] - tgrid[it - 1];
        ph = P.integrator == 1 ? PH_RK1 : PH_K1I;
        st = 0; outslot = 0; guard = 0;
        h = 0.; h1 = 0.; dnf = 0.;
        reject = false; last = false;
        for (int k = 0; k < NA; k++) x[k] = Yv(k);
    };
    if (live) start_interval();
    while (__ballot(live) != 0ull) {
        double o[NA];
        for (int k = 0; k < NA; k++) o[k] = 0.;
        rhs<KIND>(P, shfs, b, up, live, x, o);   // the one right-hand-side evaluation of the loop
        if (live) nrhs++;
        // per-lane terms of the run-wide norms this pass needs (initial step, error estimate)
        double c1 = 0., c2 = 0.;
        bool need = false;
        if (live) {
            if (ph == PH_K1I) {
                // k1 = f(Y), then the initial step (ode_solvers.h:1596-1628)
                need = true;
                for (int k = 0; k < NA; k++) {
                    K(0, k) = o[k];
                    double sk = atol + rtol * fabs(Yv(k));
                    c1 += (o[k] / sk) * (o[k] / sk);
                    c2 += (Yv(k) / sk) * (Yv(k) / sk);
                }
                double skw = atol + rtol * fabs(b.Rw0);
                c2 += (b.Rw0 / skw) * (b.Rw0 / skw);
                double skt = atol + rtol * fabs(b.Wtt0);
                c2 += (b.Wtt0 / skt) * (b.Wtt0 / skt);
            } else if (ph == PH_H2) {
                need = true;
                for (int k = 0; k < NA; k++) {
                    double sk = atol + rtol * x86_int_abs(Yv(k));
                    double z = (o[k] - K(0, k)) / sk;
                    c1 += z * z;
                }
            } else if (ph == PH_STAGE) {
                for (int k = 0; k < NA; k++) K(outslot, k) = o[k];
                if (st == 12) {
                    // all 12 stages done: 8th-order solution and the error estimate
                    need = true;
                    for (int k = 0; k < NA; k++) {
                        const double k1 = K(0, k), k6 = K(3, k), k7 = K(4, k), k8 = K(5, k), k9 = K(6, k),
                                     kA = K(7, k), kB = K(8, k), kC = K(2, k);
                        double slope = b1 * k1 + b6 * k6 + b7 * k7 + b8 * k8 + b9 * k9 + bA * kA + bB * kB + bC * kC;
                        double yn = Yv(k) + h * slope;
                        double sk = atol + rtol * fmax(fabs(Yv(k)), fabs(yn));
                        double e1 = er1 * k1 + er6 * k6 + er7 * k7 + er8 * k8 + er9 * k9 + erA * kA + erB * kB +
                                    erC * kC;
                        double e2 = slope - bhh1 * k1 - bhh9 * k9 - bhhC * kC;
                        c1 += (e1 / sk) * (e1 / sk);
                        c2 += (e2 / sk) * (e2 / sk);
                        x[k] = yn;
                    }
                }
            }
        }
        if (__ballot(need) != 0ull) run_sums(c1, c2, red, base, bw_per_run);
        bool iv_done = false;
        if (live) {
            if (ph == PH_K1I) {
                dnf = c1;
                const double dny = c2;
                h1 = (dnf <= 1e-10 || dny <= 1e-10) ? 1.0e-6 : sqrt(dny / dnf) * 0.01;
                h1 = fmin(h1, hmax);
                for (int k = 0; k < NA; k++) x[k] = Yv(k) + h1 * o[k];
                ph = PH_H2;
            } else if (ph == PH_H2) {
                const double n2 = sqrt(c1) / h1;
                double der12 = fmax(n2, sqrt(dnf));
                double h2 = (der12 <= 1e-15) ? fmax(1e-6, fabs(h1) * 1e-3) : pow(0.01 / der12, 1.0 / 8);
                h = fmin(fmin(100 * h1, h2), hmax);
                begin_attempt();
            } else if (ph == PH_STAGE) {
                if (st < 12) {
                    st++;
                    combo(st);
                    outslot = st_out[st];
                } else {
                    const double e1s = c1, e2s = c2;
                    double deno = e1s + 0.01 * e2s;
                    if (deno <= 0.0) deno = 1.0;
                    double err = fabs(h) * e1s / sqrt(nsz * deno);
                    double fac11 = pow(err, 1.0 / 8.0);
                    if (err > 1.0) {
                        h /= fmin(1.0 / 0.333, fac11 / 0.9);
                        reject = true;
                        last = false;
                        if (guard == kMaxStepsPerInterval - 1) { b.status |= PBAGPU_S_NMAX; iv_done = true; }
                        else { guard++; begin_attempt(); }
                    } else {
                        double fac = fac11 / (1.0 * 0.9);
                        fac = fac < 1.0 / 6.0 ? 1.0 / 6.0 : (fac > 1.0 / 0.333 ? 1.0 / 0.333 : fac);
                        h1 = h / fac;   // the next step size, before the limits applied after acceptance
                        ph = PH_FSAL;   // f(y_new) becomes k1 of the next step
                    }
                }
            } else if (ph == PH_FSAL) {
                for (int k = 0; k < NA; k++) { K(0, k) = o[k]; Yv(k) = x[k]; }
                time += h;
                nsteps++;
                if (last) iv_done = true;
                else {
                    double hnew = h1;
                    if (fabs(hnew) > hmax) hnew = hmax;
                    if (reject) hnew = fmin(fabs(hnew), fabs(h));
                    reject = false;
                    if (guard == kMaxStepsPerInterval - 1) { b.status |= PBAGPU_S_NMAX; iv_done = true; }
                    else { guard++; h = hnew; begin_attempt(); }
                }
            } else if (ph == PH_RK1) {
                // RK4: one classical step over the output interval (rk4_integrator, ode_solvers.h:319-366),
                // every blastwave on its own (the step has no error control that couples them);
                // the running increment is kept in slot 3
                for (int k = 0; k < NA; k++) { K(3, k) = (1.0 / 6.0) * hrk * o[k]; x[k] = Yv(k) + 0.5 * hrk * o[k]; }
                ph = PH_RK2;
            } else if (ph == PH_RK2 || ph == PH_RK3) {
                const double c = ph == PH_RK2 ? 0.5 : 1.0;
                for (int k = 0; k < NA; k++) { K(3, k) += (2.0 / 6.0) * hrk * o[k]; x[k] = Yv(k) + c * hrk * o[k]; }
                ph = ph == PH_RK2 ? PH_RK3 : PH_RK4;
            } else {   // PH_RK4
                for (int k = 0; k < NA; k++) {
                    const double a = K(3, k) + (1.0 / 6.0) * hrk * o[k];
                    Yv(k) += a;
                }
                time += hrk;
                nsteps++;
                iv_done = true;
            }
        }
        if (iv_done) {
            // end of output interval it (advanceTimeSubStep's checks, then insertSolution)
            if (!b.end_evolution) {
                double gg = Yv(iGamma), be = beta_from_gamma(gg), E2 = Yv(iEint2);
                if ((E2 <= 0.) || !isfinite(E2) || (gg < 1.) || (be < 0) || (fabs(be) < 1.e-6) || !isfinite(be))
                    b.status |= PBAGPU_S_TERMINATED;
                if (!b.end_spreading && Yv(itheta) > 0.99 * b.theta_max) b.end_spreading = 1;
                if (!(b.shutOff || !b.do_rs)) {
                    double alpha_of = P.tprompt * b.beta0 * kC;
                    double rho4f = b.M0 / (2.0 * alpha_of * kPi * (1.0 - cos(b.theta_b0))) / (Yv(iR) * Yv(iR));
                    double rho4 = P.exponential_rho4 ? rho4f * exp(-Yv(ideltaR4) / alpha_of) : rho4f;
                    if ((Yv(iGamma) < b.Gamma0) && (rho4 < P.rs_shutoff_rho || Yv(iM3) >= 1.)) b.shutOff = 1;
                }
                if (gg < 1. || Yv(itheta) < b.theta_b0 * 0.99 || Yv(itt) <= 0.) b.status |= PBAGPU_S_NONFINITE;
            }
            b.prev_idx_x = it;
            double yc[NA];   // a copy, so that Y itself stays in registers
            for (int k = 0; k < NA; k++) yc[k] = Yv(k);
            post_step<KIND>(P, shfs, shrs, b, up, yc, it, tgrid[it], T, nt);
            it++;
            if (it < ntr) start_interval();
            else live = false;
        }
    }
    if (active) {
        pba_bw_result r;
        r.status = b.status; r.i_end_r = b.i_end_r; r.fsrs = b.fsrs; r.n_steps = nsteps; r.n_rhs = nrhs; r.pad_ = 0;
        res[bw_global] = r;
    }
}

}  // namespace dyn
}  // namespace pbagpu

// kind: pbagpu::dyn::Kind of every blastwave of the launch
static int dyn_evolve_impl(const pba_run_params *params, int n_runs, int bw_per_run, const pba_layer_init *init,
                           int nt, const int *nt_run, const double *d_tgrid, int tg_stride, double *d_table,
                           pba_bw_result *res, void *stream, int kind = pbagpu::dyn::KIND_JET) {
    if (!params || !init || !d_tgrid || !d_table || !res || n_runs < 1 || bw_per_run < 1 || nt < 2)
        return PBAGPU_E_ARG;
    if (bw_per_run > 64) {
        // one wavefront per run: the blastwaves of a run share one adaptive step through LDS
        pbagpu::set_error_msg(kind == pbagpu::dyn::KIND_JET
                                  ? "dyn_evolve: at most 64 blastwaves per run"
                                  : "dyn_evolve_ejecta: at most 64 blastwaves per run (one wavefront per run); "
                                    "larger ejecta need a kernel that spreads one run over several wavefronts");
        return PBAGPU_E_UNSUPPORTED;
    }
    hipStream_t s = (hipStream_t)stream;
    if (kind != pbagpu::dyn::KIND_JET)
        for (int r = 0; r < n_runs; r++)
            if (params[r].fsrs) {
                pbagpu::set_error_msg("dyn_evolve_ejecta: ejecta blastwaves have no reverse shock (fsrs must be 0)");
                return PBAGPU_E_ARG;
            }
    for (int r = 0; r < n_runs; r++)
        if (params[r].integrator != 0 && params[r].integrator != 1) {
            pbagpu::set_error_msg("pba_run_params: integrator must be 0 (DOP853E) or 1 (RK4)");
            return PBAGPU_E_UNSUPPORTED;
        }
    if (nt_run)
        for (int r = 0; r < n_runs; r++)
            if (nt_run[r] < 2 || nt_run[r] > nt) {
                pbagpu::set_error_msg("dyn_evolve_ragged: every run needs 2 <= nt_run <= nt_max steps");
                return PBAGPU_E_ARG;
            }
    const size_t nbw = (size_t)n_runs * bw_per_run;
    for (size_t i = 0; i < nbw; i++) {
        // ncells < 1 is an error of the reference too (blastwave.h:175-181)
        if (!(init[i].ncells >= 1.0) || !std::isfinite(init[i].ncells) || !(init[i].theta_b0 > 0.0)) {
            pbagpu::set_error_msg("pba_layer_init: ncells must be >= 1 and theta_b0 > 0");
            return PBAGPU_E_ARG;
        }
        if (init[i].ngam_fs < 0 || init[i].ngam_fs == 1 || init[i].ngam_rs < 0 || init[i].ngam_rs == 1) {
            pbagpu::set_error_msg("pba_layer_init: ngam_fs / ngam_rs must be 0 (the run's) or >= 2");
            return PBAGPU_E_ARG;
        }
    }
    auto *dP = (pba_run_params *)pbagpu::workspace(pbagpu::WS_DYN_PARAMS, sizeof(pba_run_params) * n_runs, s);
    auto *dI = (pba_layer_init *)pbagpu::workspace(pbagpu::WS_DYN_INIT, sizeof(pba_layer_init) * nbw, s);
    auto *dR = (pba_bw_result *)pbagpu::workspace(pbagpu::WS_DYN_RES, sizeof(pba_bw_result) * nbw, s);
    int *dN = nt_run ? (int *)pbagpu::workspace(pbagpu::WS_DYN_NT, sizeof(int) * n_runs, s) : nullptr;
    if (!dP || !dI || !dR || (nt_run && !dN)) return PBAGPU_E_HIP;
    int rc = PBAGPU_OK;
    PBA_TRY(rc, hipMemcpyAsync(dP, params, sizeof(pba_run_params) * n_runs, hipMemcpyHostToDevice, s));
    PBA_TRY(rc, hipMemcpyAsync(dI, init, sizeof(pba_layer_init) * nbw, hipMemcpyHostToDevice, s));
    if (nt_run) PBA_TRY(rc, hipMemcpyAsync(dN, nt_run, sizeof(int) * n_runs, hipMemcpyHostToDevice, s));
    // Runs per wavefront.  A wavefront's pass time grows with the runs it carries (on a full chip,
    // three 21-blastwave runs per wavefront take 2.7 times one run's time), so runs are packed only
    // as far as they outnumber the SIMDs the stream may use (one wavefront per SIMD: the kernel
    // needs all 512 registers of a lane); the option overrides this with a fixed count.
    const long long opt_rpw = pbagpu::option(PBAGPU_OPT_DYN_RUNS_PER_WAVE);
    const int rpw_max = 64 / bw_per_run;
    int rpw;
    if (kind != pbagpu::dyn::KIND_JET) rpw = 1;   // ejecta: one run per wavefront
    else if (opt_rpw > 0) rpw = (int)std::min<long long>(opt_rpw, rpw_max);
    else {
        const int simds = 4 * std::max(1, pbagpu::stream_cu_count(s));
        rpw = std::min(rpw_max, std::max(1, (n_runs + simds - 1) / simds));
    }
    if (rpw > n_runs) rpw = n_runs;
    const int nblk = (n_runs + rpw - 1) / rpw;
    using pbagpu::dyn::KS;
    using pbagpu::dyn::NA;
    double *dK = nullptr;
    if (rpw > 1) {
        dK = (double *)pbagpu::workspace(pbagpu::WS_DYN_K, sizeof(double) * (size_t)nblk * KS * NA * 64, s);
        if (!dK) return PBAGPU_E_HIP;
    }
    pbagpu::KernelTimer timer(PBAGPU_STAGE_DYN);
    if (rc == PBAGPU_OK) {
        timer.start(s);
        if (rpw > 1) {
            hipLaunchKernelGGL(pbagpu::dyn::dyn_kernel<true>, dim3(nblk), dim3(64),
                               sizeof(double) * (NA * 64 + 128) + sizeof(pbagpu::dyn::BW) * 64, s, dP, dI,
                               n_runs, bw_per_run, rpw, nt, (const int *)dN, d_tgrid, tg_stride, d_table, dR, dK);
        } else if (kind == pbagpu::dyn::KIND_JET) {
            const size_t lds = sizeof(double) * ((size_t)(KS + 1) * NA * bw_per_run + 128) +
                               sizeof(pbagpu::dyn::BW) * bw_per_run;
            (void)hipFuncSetAttribute((const void *)pbagpu::dyn::dyn_kernel<false>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            hipLaunchKernelGGL(pbagpu::dyn::dyn_kernel<false>, dim3(nblk), dim3(64), lds, s, dP, dI, n_runs,
                               bw_per_run, 1, nt, (const int *)dN, d_tgrid, tg_stride, d_table, dR, (double *)nullptr);
        } else {
            // the jet's LDS layout plus one Upstream record per blastwave
            const size_t lds = sizeof(double) * ((size_t)(KS + 1) * NA * bw_per_run + 128) +
                               (sizeof(pbagpu::dyn::BW) + sizeof(pbagpu::dyn::Upstream)) * bw_per_run;
            auto kern = kind == pbagpu::dyn::KIND_EJECTA_SPREAD
                            ? pbagpu::dyn::dyn_kernel<false, pbagpu::dyn::KIND_EJECTA_SPREAD>
                            : pbagpu::dyn::dyn_kernel<false, pbagpu::dyn::KIND_EJECTA>;
            (void)hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            hipLaunchKernelGGL(kern, dim3(nblk), dim3(64), lds, s, dP, dI, n_runs, bw_per_run, 1, nt,
                               (const int *)dN, d_tgrid, tg_stride, d_table, dR, (double *)nullptr);
        }
        PBA_TRY(rc, hipGetLastError());
        timer.stop(s);
    }
    if (rc == PBAGPU_OK) PBA_TRY(rc, hipMemcpyAsync(res, dR, sizeof(pba_bw_result) * nbw, hipMemcpyDeviceToHost, s));
    PBA_TRY(rc, hipStreamSynchronize(s));
    timer.collect();
    return rc;
}

extern "C" int pbagpu_dyn_evolve(const pba_run_params *params, int n_runs, int bw_per_run, const pba_layer_init *init,
                                 int nt, const double *d_tgrid, double *d_table, pba_bw_result *res, void *stream) {
    return dyn_evolve_impl(params, n_runs, bw_per_run, init, nt, nullptr, d_tgrid, 0, d_table, res, stream);
}

extern "C" int pbagpu_dyn_evolve_ragged(const pba_run_params *params, int n_runs, int bw_per_run,
                                        const pba_layer_init *init, int nt_max, const int *nt_run,
                                        const double *d_tgrid, double *d_table, pba_bw_result *res, void *stream) {
    if (!nt_run) return PBAGPU_E_ARG;
    return dyn_evolve_impl(params, n_runs, bw_per_run, init, nt_max, nt_run, d_tgrid, nt_max, d_table, res, stream);
}

extern "C" int pbagpu_dyn_evolve_ejecta(const pba_run_params *params, int n_runs, int bw_per_run,
                                        const pba_layer_init *init, int spread, int nt, const double *d_tgrid,
                                        double *d_table, pba_bw_result *res, void *stream) {
    return dyn_evolve_impl(params, n_runs, bw_per_run, init, nt, nullptr, d_tgrid, 0, d_table, res, stream,
                           spread ? pbagpu::dyn::KIND_EJECTA_SPREAD : pbagpu::dyn::KIND_EJECTA);
}
