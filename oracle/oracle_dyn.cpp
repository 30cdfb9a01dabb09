/*
 * Oracle, stage 1: blastwave dynamics (TEST INFRASTRUCTURE ONLY — see oracle_common.h).
 *
 * Restates, for the option set the workloads use (adaptive layers, uniform ISM,
 * Nava13 EOS, "our" lateral spreading limited by MomValAndFrac, dM/dR from
 * dtheta/dR, "our" dGamma/dR, radiative losses inside the RHS, numeric
 * electrons with useU_b / useB / useNumeric / useTcomov / use_Sironi limits):
 *   - initial conditions      BlastWave::setInitConditions   (blastwave.h:73-241)
 *   - forward-shock RHS       BlastWave::rhs_fs              (blastwave.h:1515-1768)
 *   - forward+reverse RHS     BlastWave::rhs_fsrs            (blastwave.h:1771-2116)
 *   - kilonova ejecta RHS     BlastWave::setStandardISM + rhs_fs_dense (blastwave.h:769-781,
 *                             1282-1295, 2120-2497), bw_type = fs_dense, entry orc_dyn_evolve_ejecta
 *   - radiative loss          ElectronAndRadiation::computeRadiationLoss / computePtot
 *                             (microphysics.h:204-220, 749-795; analytic_models.h:889-911, 1051-1078)
 *   - embedded DOP853         dop853_integrator::integrate / _integrate8 / _determine_initial_step
 *                             (ode_solvers.h:1289-1980), one shared step for all blastwaves of a run
 *   - per-output-step logic   EvolveODEsystem::advanceTimeSubStep / processSolution... (model_evolve.h:345-566)
 *   - derived quantities      BlastWave::insertSolution / applyUnits / addOtherVars (blastwave.h:270-756)
 *   - injection limits        computeInjectionElectronDistributionLimits{FS,RS}
 *                             (blastwave_radiation.h:881-1107; microphysics_base.h:575-966)
 */
#include "oracle_common.h"
#include <cstdio>
#include <cstring>
#include <climits>
#include <cstdlib>

namespace orc {

// state-vector layout of one blastwave (SOL::QS, blastwave_pars.h:308-316)
enum QS { iR, iRsh, itt, itcomov, iGamma, iEint2, iEint3, itheta, iRrsh, iErad2, iErad3, iEsh2, iEsh3,
          iEad2, iEad3, iM2, iM3, ideltaR4, iRw, iWmom, iWenb, iWepwn, iWtt, NEQ };

// Hidden state of one ElectronAndRadiation object that survives between calls:
// accel_frac is only rewritten when gamma_min < 2 (microphysics_base.h:887-920), and
// tcomov is set by the per-step microphysics and read by the RHS radiative loss
// through computeGammaCool (microphysics_base.h:938-966, microphysics.h:757).
struct MState { double acc = 1.0; double tcomov = -1.0; };

struct Shock {
    const pba_shock_params *sp;
    double e0, etop;   // first and last Chang-Cooper grid points
};

struct BW {
    double E0, M0, R0, Gamma0, beta0, mom0, theta_a, theta_b0, ncells, ctheta0, theta_w, theta_core;
    double theta_max, Rd, rho_def, drhodr0;
    int ilayer, nlayers;
    bool fsrs, do_rs, do_rs_rad, shutOff, end_evolution, end_spreading;
    // ejecta blastwave (bw_type = fs_dense): evolved by rhs_fs_dense in a standard ISM; `spread`
    // is method_spread != None.  The upstream state that the RHS leaves in the blastwave's density
    // object (RhoISM m_rho_, m_drhodr_, m_GammaRel) is what addOtherVars stores afterwards.
    bool dense = false, spread = true, rhs_seen = false;
    double rho_rhs = 0., drhodr_rhs = 0., grel_rhs = -1.;
    int prev_idx_x;
    int i_end_r;
    MState fs_st, rs_st;
    int status;
    Shock fs_sh, rs_sh;   // this blastwave's electron grids (the run's, or pba_layer_init ngam_fs / ngam_rs)
};

struct Run {
    const pba_run_params *P;
    Shock fs, rs;
    std::vector<BW> bws;
    int n_rhs = 0;
};

// x86 behaviour of int abs(int) applied to a double (the unqualified abs() in
// dop853_integrator::_determine_initial_step, ode_solvers.h:1614, resolves to
// the int overload; out-of-range conversions give INT_MIN and abs(INT_MIN) == INT_MIN).
static double x86_int_abs(double v) {
    if (!(v > -2147483649.0 && v < 2147483648.0)) return -2147483648.0;
    int iv = (int)v;
    return (double)(iv < 0 ? -iv : iv);
}

// EjectaID2::ctheta (ejecta_id.h:81-117)
static double ctheta_of(double theta, int ilayer, int nlayers) {
    double ct = 0.;
    if (ilayer > 0) {
        double fac0 = (double)ilayer / (double)nlayers;
        double fac1 = (double)(ilayer + 1) / (double)nlayers;
        double x2 = fac1 * std::sin(theta / 2.);
        double xx2 = 2. * std::asin(x2);
        double x1 = fac0 * std::sin(theta / 2.);
        double xx1 = 2. * std::asin(x1);
        ct = 0.5 * (xx1 + xx2);
    }
    return ct;
}

// LatSpread::dthetadr_our (blastwave_components.h:1060-1080)
static double dthetadr_our(const BW &b, double Gamma, double GammaShk, double R, double adi, double theta) {
    double Q0 = 2.0;
    double Q = std::sqrt(2.0) * 3.0;
    double mom = MomFromGamma(GammaShk);
    double vperp = std::sqrt(adi * (adi - 1.0) * (Gamma - 1.0) / (1.0 + adi * (Gamma - 1.0))) * c;
    if ((theta < 0.5 * pi) && (Q0 * mom * b.theta_core < 1)) {
        double fac = 1.;
        if (mom * b.theta_core * Q >= 1.0) fac = Q * (1. - Q0 * mom * b.theta_core) / (Q - Q0);
        if (b.theta_b0 < b.theta_core) fac *= std::tan(0.5 * b.theta_b0) / std::tan(0.5 * b.theta_core);
        double beta = BetaFromGamma(GammaShk);
        return vperp / R / GammaShk / beta / c * fac;
    }
    return 0.;
}

// EQS::dmdr (usingdthdR) (blastwave_components.h:59-61)
// RhoISM::evaluateRhoDrhoDrDefault (blastwave_components.h:846-893): the ambient density and its
// radial derivative at R; uniform for n_ism > 0, else the step-function (wind) profile of A0, s,
// r_ej, r_ism.  The reference exits for rho > 1 (:886-890).
static void ism_density(const pba_run_params &P, double R, double &rho, double &drhodr, int &status) {
    if (P.n_ism > 0.0) {
        rho = P.n_ism * mppme;
        drhodr = 0.0;
    } else {
        if (R < P.ism_r_ej) { rho = P.ism_A0 * std::pow(P.ism_r_ej, -P.ism_s) * mppme; drhodr = 0.0; }
        if ((P.ism_r_ej <= R) && (R < P.ism_r_ism)) {
            rho = P.ism_A0 * std::pow(R, -P.ism_s) * mppme;
            drhodr = rho * (-P.ism_s / R);
        }
        if (P.ism_r_ism <= R) { rho = P.ism_A0 * std::pow(P.ism_r_ism, -P.ism_s) * mppme; drhodr = 0.0; }
    }
    if (rho < 0 || rho > 1.) status |= PBAGPU_S_NONFINITE;
}

static double dmdr(double R, double dthetadR, double theta, double rho) {
    return 2. * pi * rho * ((1. - std::cos(theta)) + (1. / 3.) * std::sin(theta) * R * dthetadR) * R * R;
}

// EQS::evalElapsedTime (blastwave_components.h:110-120)
static double elapsed_dttdr(double R, double mom, double dthetadr) {
    double beta = BetaFromMom(mom);
    return 1. / (c * beta) * std::sqrt(1. + R * R * dthetadr * dthetadr) - (1. / c);
}

// SynchrotronAnalytic::computeAnalyticSynchJOH06_limits + computeIntegratedAnalyticSynchWSPN99
// (analytic_models.h:889-911, 1051-1078)
static double wspn99_ptot(double p, double gm, double gc, double gM, double B, double nprime, double nu0) {
    double gamToNuFactor = (3.0 / (4.0 * pi)) * (qe * B) / (me * c);
    double nu_m, nu_c, p_max;
    if (gm < gc) {
        double XpS = 0.06 + 0.28 * p;
        nu_m = XpS * gm * gm * gamToNuFactor;
        nu_c = XpS * gc * gc * gamToNuFactor;
        p_max = 11.17 * (p - 1.0) / (3.0 * p - 1.0) * (0.54 + 0.08 * p) * qe * qe * qe * nprime * B / (me * c * c);
    } else {
        double XpF = 0.455 + 0.08 * p;
        double phipF = 1.89 - 0.935 * p + 0.17 * (p * p);
        nu_m = XpF * gm * gm * gamToNuFactor;
        nu_c = XpF * gc * gc * gamToNuFactor;
        double _phip = 2.234 * phipF;
        p_max = _phip * qe * qe * qe * nprime * B / (me * c * c);
    }
    double nu_M = 3.0 / (4.0 * pi) * gM * gM * qe * B / (me * c);
    double t0, t1, t2, ptot;
    if (nu_m <= nu_c) {
        t0 = (3. / 4.) * std::pow(1. / nu_m, 1. / 3.) * (std::pow(nu_m, 4. / 3.) - std::pow(nu0, 4. / 3.));
        t1 = -1. / (3. - p) * (2 * std::pow(1. / nu_m, 1. / 2. - p / 2.)) *
             (std::pow(nu_m, 3. / 2. - p / 2.) - std::pow(nu_c, 3. / 2. - p / 2.));
        t2 = std::pow(nu_c / nu_m, -(p - 1.) / 2.) * 2. / (p - 2.) *
             (nu_c * std::pow(nu_M, p / 2.) - nu_M * std::pow(nu_c, p / 2.)) * std::pow(nu_c * nu_M / nu_c, -p / 2.);
        ptot = t0 + t1 + t2;
    } else {
        t0 = (3. / 4.) * std::pow(1. / nu_c, 1. / 3.) * (std::pow(nu_c, 4. / 3.) - std::pow(nu0, 4. / 3.));
        t1 = -2. / std::sqrt(1. / nu_c) * (std::sqrt(nu_c) - std::sqrt(nu_m));
        t2 = 1. / std::sqrt(nu_m / nu_c) * 2. / (p - 2.) *
             (nu_m * std::pow(nu_M, p / 2.) - nu_M * std::pow(nu_m, p / 2.)) * std::pow(nu_m * nu_M / nu_m, -p / 2.);
        ptot = t0 + t1 + t2;
    }
    return ptot * p_max;
}

// ElectronAndRadiation::computeRadiationLoss (microphysics.h:749-795) with computePtot (:204-220)
static double radiation_loss(const Shock &sh, MState &st, double dmdR, double dEshdR, double R, double Gamma,
                             double GammaShk, double adi, double m, double Eint, double rho, double theta,
                             double ncells) {
    const pba_shock_params &s = *sh.sp;
    if (dmdR == 0 || dEshdR == 0 || m == 0 || rho == 0) return 0.;
    double rho2 = rho2t(Gamma, adi, rho);
    double U_p = (Eint == 0 || m == 0) ? 0. : Eint / (m / rho2);
    double eprime = U_p, betaSh = BetaFromGamma(GammaShk), n_prime = rho2 / mp;
    double one_min_cos = (std::cos(0.) - std::cos(theta) / ncells);
    double thickness = m / (2.0 * pi * one_min_cos * rho2 * Gamma * R * R);
    double p_tot = 0.;
    if (!(betaSh < s.beta_min)) {
        double B = std::sqrt(8. * pi * (eprime * s.eps_b));
        double gM = std::sqrt(6.0 * pi * qe / sigmaT / B);
        if (gM > 0.99 * sh.etop) gM = 0.99 * sh.etop;
        double gm = (s.p - 2.) / (s.p - 1.) * (eprime * s.eps_e) / (n_prime * me * c * c);
        if (gm < 2.) {
            double sironi = (s.p - 2.0) / (s.p - 1.0) * s.eps_e * eprime / (n_prime * me * c * c);
            st.acc = sironi > 1. ? 1 : sironi;
            gm = gm < 1. ? 1 : gm;
        }
        if (gm < sh.e0) gm = sh.e0;
        double gc = ((6. * pi * me * c) / (sigmaT * B * B * (st.tcomov)));
        if (gc <= 1.) gc = 1.001;
        if (gc > gM) gc = gM;
        p_tot = wspn99_ptot(s.p, gm, gc, gM, B, dmdR / mp, 1e6);
    }
    p_tot *= st.acc;
    p_tot = std::max(p_tot, 0.);
    double drcomov = thickness * GammaShk;
    double dt_esc = drcomov / c;
    double dErad2dR = p_tot * dt_esc;
    double eps_rad = dErad2dR / (dEshdR * s.eps_e);
    if (eps_rad < 0. || !std::isfinite(eps_rad)) eps_rad = 0.;
    return eps_rad;
}

// Margalit21::Theta_fun (analytic_models.h:56-60)
static double Theta_fun(double beta, double mu, double mu_e, double epsT) {
    double Theta0 = epsT * (9.0 * mu * mp / (32.0 * mu_e * me)) * beta * beta;
    return (5.0 * Theta0 - 6.0 + std::sqrt(25.0 * Theta0 * Theta0 + 180.0 * Theta0 + 36.0)) / 30.0;
}

struct Lim { double B, gm, gM, gc, Theta, z_cool, nprime, acc; bool ok; };

// computeEleSpectrumLimits with useU_b / useB / useNumeric / useTcomov / use_Sironi
// (microphysics_base.h:575-652, 792-966; rootfinders.h:17-58)
static Lim ele_limits(const Shock &sh, MState &st, double U_p, double GammaShk, double t_e, double tcomov,
                      double n_prime) {
    const pba_shock_params &s = *sh.sp;
    Lim L{};
    st.tcomov = tcomov;
    double eprime = U_p, betaSh = BetaFromGamma(GammaShk);
    if (betaSh < s.beta_min) { L.ok = false; return L; }
    L.ok = true;
    double B = std::sqrt(8. * pi * (eprime * s.eps_b));
    double gM = std::sqrt(6.0 * pi * qe / sigmaT / B);
    if (gM > 0.99 * sh.etop) gM = 0.99 * sh.etop;
    double gm1 = (s.p - 2.) / (s.p - 1.) * (eprime * s.eps_e) / (n_prime * me * c * c);
    double Theta = Theta_fun(betaSh, 0.62, 1.18, s.eps_t);
    double z_cool = (6.0 * pi * me * c / (sigmaT * B * B * t_e)) / Theta;
    // Bisect(gammaMinFunc, 1, 1e8, xtol=0, rtol=1e-3, 100 iterations)
    double gamma_ave = s.eps_e * eprime / (n_prime * me * c * c);
    auto f = [&](double x) {
        return (s.p - 1) / (s.p - 2) * (std::pow(gM, -s.p + 2) - std::pow(x, -s.p + 2)) /
               (std::pow(gM, -s.p + 1) - std::pow(x, -s.p + 1)) - gamma_ave;
    };
    int status = 0;
    double gm;
    {
        double xa = 1., xb = 1.e8;
        double fa = f(xa), fb = f(xb);
        if (fa * fb > 0) { status = -1; gm = 0.; }
        else if (fa == 0) { status = -1; gm = xa; }
        else if (fb == 0) { status = -1; gm = xb; }
        else {
            double dm = xb - xa, xm = xa;
            int i;
            gm = 0.;
            for (i = 0; i < 100; i++) {
                dm *= 0.5;
                xm = xa + dm;
                double fm = f(xm);
                if (fm * fa >= 0) xa = xm;
                if (fm == 0 || std::fabs(dm) < 0. + .001 * std::fabs(xm)) { gm = xm; break; }
            }
            if (i == 100) { status = -1; gm = xa; }
        }
    }
    if (status < 0 && gm1 > 1.5) gm = gm1;
    if (gm < 2. || status < 0) {
        double sironi = (s.p - 2.0) / (s.p - 1.0) * s.eps_e * eprime / (n_prime * me * c * c);
        st.acc = sironi > 1. ? 1 : sironi;
        gm = gm < 1. ? 1 : gm;
    }
    if (gm < sh.e0) gm = sh.e0;
    double gc = ((6. * pi * me * c) / (sigmaT * B * B * (tcomov)));
    if (gc <= 1.) gc = 1.001;
    if (gc > gM) gc = gM;
    L.B = B; L.gm = gm; L.gM = gM; L.gc = gc; L.Theta = Theta; L.z_cool = z_cool; L.nprime = n_prime;
    L.acc = st.acc;
    return L;
}

static void rhs_fs(Run &run, BW &b, const double *Y, double *out) {
    const pba_run_params &P = *run.P;
    double R = Y[iR], Gamma = Y[iGamma], Eint2 = Y[iEint2], theta = Y[itheta], M2 = Y[iM2];
    if (Gamma < 1 || M2 < 0 || theta < 0) b.status |= PBAGPU_S_NONFINITE;
    double rho = 0., drhodr = 0.;
    ism_density(P, R, rho, drhodr, b.status);
    rho /= b.M0;
    drhodr /= b.M0;
    double beta = BetaFromGamma(Gamma);
    double adi = adiNava(Gamma);
    double Gsh = GammaSh(Gamma, adi);
    double dRshdt = (double)((long double)BetaFromGamma(Gsh) * c);
    double dRdt = beta * c;
    double dthetadr = 0.0;
    if ((theta < b.theta_max) && (!b.end_spreading)) {
        if (Gamma * beta < std::max(P.mom_when_start_spread, 0. * b.mom0))
            dthetadr = dthetadr_our(b, Gamma, Gsh, R, adi, theta);
    }
    double dM2dR = dmdr(R, dthetadr, theta, rho) / b.ncells;
    double GammaEff = (adi * Gamma * Gamma - adi + 1.0) / Gamma;
    double dGammaEffdGamma = (adi * Gamma * Gamma + adi - 1.0) / (Gamma * Gamma);
    double num = ((Gamma - 1.) * (GammaEff + 1.) * dM2dR - GammaEff * (adi - 1.) * Eint2 * (dM2dR / M2 - drhodr / rho));
    double denum = ((1. + M2) + Eint2 * dGammaEffdGamma + GammaEff * (adi - 1.) * Eint2 / Gamma);
    double dGammadR = -num / denum;
    double dtcomov_dR = 1.0 / MomFromGamma(Gamma) / c;
    double dEsh2dR = (Gamma - 1.0) * dM2dR;
    double dlnV2dR = dM2dR / M2 - drhodr / rho - dGammadR / Gamma;
    double dEad2dR = 0.0;
    if (P.adiab_loss) dEad2dR = -(adi - 1.0) * Eint2 * dlnV2dR;
    double eps_rad = P.fs.epsilon_e_rad;
    if (P.fs.epsilon_e_rad < 0)
        eps_rad = radiation_loss(b.fs_sh, b.fs_st, dM2dR * b.M0, dEsh2dR * b.M0 * c * c, R, Gamma, Gsh, adi,
                                 M2 * b.M0, Eint2 * b.M0 * c * c, rho * b.M0, theta, b.ncells);
    double dErad2dR = eps_rad * P.fs.eps_e * dEsh2dR;
    double dEint2dR = dEsh2dR + dEad2dR - dErad2dR;
    double dttdr = elapsed_dttdr(R, MomFromGamma(Gamma), dthetadr);
    if (!std::isfinite(dRdt) || !std::isfinite(dGammadR) || !std::isfinite(dlnV2dR) || !std::isfinite(dEint2dR))
        b.status |= PBAGPU_S_NONFINITE;
    if (std::fabs(dRdt * dGammadR) > b.Gamma0) b.status |= PBAGPU_S_NONFINITE;
    out[iR] = dRdt;
    out[iRsh] = dRshdt;
    out[itt] = dRdt * dttdr;
    out[itcomov] = dRdt * dtcomov_dR;
    out[iGamma] = dRdt * dGammadR;
    out[iEint2] = dRdt * dEint2dR;
    out[itheta] = dRdt * dthetadr;
    out[iErad2] = dRdt * dErad2dR;
    out[iEsh2] = dRdt * dEsh2dR;
    out[iEad2] = dRdt * dEad2dR;
    out[iM2] = dRdt * dM2dR;
}

static void rhs_fsrs(Run &run, BW &b, const double *Y, double *out) {
    const pba_run_params &P = *run.P;
    double R = Y[iR], Gamma = Y[iGamma], Eint2 = Y[iEint2], Eint3 = Y[iEint3], theta = Y[itheta];
    double M2 = Y[iM2], M3 = Y[iM3], deltaR4 = Y[ideltaR4];
    if (Gamma < 1.) Gamma = (b.prev_idx_x == 0) ? b.Gamma0 : 1. + 1.e-5;
    if (Eint3 < 0) Eint3 = 0.0;
    if (M3 < 0 && b.prev_idx_x == 0) { M3 = 0; Eint3 = 0; deltaR4 = 0; }
    double beta = BetaFromGamma(Gamma);
    // EQS::get_Gamma43 (blastwave_components.h:327-340), long double as in the reference
    long double lbeta = beta, lbeta0 = b.beta0;
    long double Gamma43 = Gamma * b.Gamma0 * (1.0 / (b.Gamma0 * b.Gamma0) + 1.0 / (Gamma * Gamma) -
                                              1.0 / (Gamma * Gamma) / (b.Gamma0 * b.Gamma0)) / (1.0 + lbeta * lbeta0);
    if (Gamma43 < 1.l) Gamma43 = 1.l;
    double adi3 = adiNava((double)Gamma43);
    double GammaRsh = GammaSh((double)Gamma43, adi3);
    double dRshRshdt = (double)((long double)BetaFromGamma(GammaRsh) * c);
    double rho = 0., drhodr = 0.;
    ism_density(P, R, rho, drhodr, b.status);
    rho /= b.M0;
    drhodr /= b.M0;
    double dRdt = beta * c;
    double adi = adiNava(Gamma);
    double Gsh = GammaSh(Gamma, adi);
    double dRshdt = (double)((long double)BetaFromGamma(Gsh) * c);
    double dthetadr = 0.0;
    if ((theta < b.theta_max) && (!b.end_spreading)) {
        if (Gamma * beta < std::max(P.mom_when_start_spread, 0. * b.mom0))
            dthetadr = dthetadr_our(b, Gamma, Gsh, R, adi, theta);
    }
    double dM2dR = dmdr(R, dthetadr, theta, rho) / b.ncells;
    double ddeltaR4dR = 0.;
    long double dM3dR = 0.;
    double dlnrho4dR = 0.;
    if ((!b.shutOff) && (Gamma < 0.9999 * b.Gamma0)) {
        double alpha_of = P.tprompt * b.beta0 * c;
        ddeltaR4dR = (1.0 / (beta * beta * beta * beta) - 1.0 / (b.beta0 * b.beta0 * b.beta0 * b.beta0)) /
                     (1.0 / beta + 1.0 / b.beta0) / (1.0 / (beta * beta) + 1.0 / (b.beta0 * b.beta0)) * b.beta0;
        long double sf = -deltaR4 / alpha_of;
        if (P.exponential_rho4) {
            dM3dR = (1.0 / alpha_of) * ddeltaR4dR * std::exp(sf);
            dlnrho4dR = -2.0 / R - ddeltaR4dR / alpha_of;
        } else {
            dM3dR = (1.0 / alpha_of) * ddeltaR4dR;
            dlnrho4dR = -2.0 / R;
        }
    }
    double dlnrhodr = drhodr / rho;
    // EQS::get_dGammadR_fs_rs (blastwave_components.h:381-411)
    double dGammadR;
    {
        double G0 = b.Gamma0;
        long double dg43 = G0 - (G0 * G0 * Gamma - Gamma) / std::sqrt(Gamma * Gamma * G0 * G0 - G0 * G0 - Gamma * Gamma + 1.0);
        long double GammaEff3 = (adi3 * Gamma * Gamma - adi3 + 1.0) / Gamma;
        // returned as double by get_dGammaEff3dGamma, then held in a long double (blastwave_components.h:416)
        long double dGammaEff3dGamma = (double)(adi3 * (1.0 + 1.0 / (Gamma * Gamma)) -
                                                dg43 / 3.0 / Gamma43 * Gamma43 * (Gamma - 1.0 / Gamma) - 1.0 / (Gamma * Gamma));
        double dGammaEffdGamma = (adi * Gamma * Gamma + adi - 1.0) / (Gamma * Gamma);
        double GammaEff = (adi * Gamma * Gamma - adi + 1.0) / Gamma;
        double f_2 = GammaEff * (adi - 1.0) * Eint2 / Gamma;
        double h_2 = GammaEff * (adi - 1.0) * Eint2 * (dM2dR / M2 - dlnrhodr);
        long double fh3 = GammaEff3 * (adi3 - 1.0) * Eint3;
        long double f_3 = fh3 / (Gamma43) * dg43;
        long double h_3 = 0.0;
        if (Eint3 != 0) h_3 = fh3 * (dM3dR / M3 - dlnrho4dR);
        long double d = -((Gamma - 1.0) * (GammaEff + 1.0) * dM2dR + (Gamma - G0 + GammaEff3 * (Gamma43 - 1.0)) * dM3dR -
                          h_2 - h_3) /
                        ((M2 + M3) + Eint2 * dGammaEffdGamma + Eint3 * dGammaEff3dGamma + f_2 + f_3);
        dGammadR = (double)d;
    }
    if (dGammadR > 0 && Gamma > P.rs_gamma0_frac_no_exceed * b.Gamma0) dGammadR = 0.0;
    double dEsh2dR = (Gamma - 1.0) * dM2dR;
    long double dEsh3dR = (Gamma43 - 1.0) * dM3dR;
    double dEad2dR = 0.0;
    double dlnV2dR = dM2dR / M2 - drhodr / rho - dGammadR / Gamma;
    if (P.adiab_loss) dEad2dR = -(adi - 1.0) * Eint2 * dlnV2dR;
    double dEad3dR = 0.;
    if ((!b.shutOff) && (M3 > 0.0)) {
        double dlnV3dR = (double)dM3dR / M3 - dlnrho4dR - dGammadR / Gamma;
        if (P.adiab_loss_rs) dEad3dR = -(adi3 - 1) * Eint3 * dlnV3dR;
    }
    double eps_rad = P.fs.epsilon_e_rad;
    if (P.fs.epsilon_e_rad < 0)
        eps_rad = radiation_loss(b.fs_sh, b.fs_st, dM2dR * b.M0, dEsh2dR * b.M0 * c * c, R, Gamma, Gsh, adi,
                                 M2 * b.M0, Eint2 * b.M0 * c * c, rho * b.M0, theta, b.ncells);
    double dErad2dR = eps_rad * P.fs.eps_e * dEsh2dR;
    // the reverse-shock loss is called with rho4 = 0 and therefore returns 0 (blastwave.h:1991-1997)
    double eps_rad_rs = P.rs.epsilon_e_rad < 0 ? 0. : P.rs.epsilon_e_rad;
    double dErad3dR = eps_rad_rs * P.rs.eps_e * (double)dEsh3dR;
    double dEint2dR = dEsh2dR + dEad2dR - dErad2dR;
    double dEint3dR = (double)(dEsh3dR + dEad3dR - dErad3dR);
    double dtcomov_dR = 1.0 / beta / Gamma / c;
    double dttdr = elapsed_dttdr(R, Gamma * beta, dthetadr);
    if (!std::isfinite(dRdt) || !std::isfinite(dGammadR) || dM2dR < 0. || !std::isfinite(dlnV2dR))
        b.status |= PBAGPU_S_NONFINITE;
    out[iR] = dRdt;
    out[iRsh] = dRshdt;
    out[itt] = dRdt * dttdr;
    out[itcomov] = dRdt * dtcomov_dR;
    out[iGamma] = dRdt * dGammadR;
    out[iRrsh] = dRshRshdt;
    out[iEint2] = dRdt * dEint2dR;
    out[iEint3] = dRdt * dEint3dR;
    out[itheta] = dRdt * dthetadr;
    out[iErad2] = dRdt * dErad2dR;
    out[iErad3] = dRdt * dErad3dR;
    out[iEsh2] = dRdt * dEsh2dR;
    out[iEsh3] = (double)(dRdt * dEsh3dR);
    out[iEad2] = dRdt * dEad2dR;
    out[iEad3] = dRdt * dEad3dR;
    out[iM2] = dRdt * dM2dR;
    out[iM3] = (double)(dRdt * dM3dR);
    out[ideltaR4] = dRdt * ddeltaR4dR;
}

// EQS::evalElapsedTime (blastwave_components.h:165-175): dt_obs/dR, the spreading form or the
// radial one
static double elapsed_dttdr_radial(double R, double mom, double dthetadr, bool spread) {
    double xx = 1. + mom * mom * mom / (std::sqrt(1. + mom * mom));
    double beta = BetaFromMom(mom);
    if (spread) return 1. / (c * beta) * std::sqrt(1. + R * R * dthetadr * dthetadr) - (1. / c);
    return 1. / (c * xx * (1. + beta));
}

// Ejecta blastwave in a standard ISM: BlastWave::setStandardISM / set_standard_ism
// (blastwave.h:769-781, 1282-1295), then BlastWave::rhs_fs_dense (blastwave.h:2120-2497) with
// method_dgdr = our, method_dmdr = usingdthdr, no energy injection, nuclear heating or
// luminosity loss (dEinjdt = dEnuc = dElum = 0: no magnetar, do_nucinj = no, no collisions).
// The upstream is at rest (GammaRho = 1, dGammaRho/dR = 0, sound speed 0) and GammaRel = Gamma,
// dGammaRel/dGamma = 1; the zero terms are kept so that every sum rounds as the reference's.
// dtheta/dR = 0 and Rsh is not evolved (its derivative is never written).
static void rhs_fs_dense(Run &run, BW &b, const double *Y, double *out) {
    const pba_run_params &P = *run.P;
    double Gamma = Y[iGamma];
    double R = Y[iR], Eint2 = Y[iEint2], theta = Y[itheta], M2 = Y[iM2];
    if (Gamma < 1.) { Gamma = 1. + 1.e-5; b.status |= PBAGPU_S_NONFINITE; }
    double beta = BetaFromGamma(Y[iGamma]);
    if (!std::isfinite(R) || !std::isfinite(Gamma) || M2 < 0. || !std::isfinite(M2) || !std::isfinite(Eint2) ||
        Eint2 < 0)
        b.status |= PBAGPU_S_NONFINITE;
    // set_standard_ism: evaluateRhoDrhoDrDefault at (R, ctheta); ctheta only matters for profiles
    // that depend on angle, which the default one does not.  GammaRel is the blastwave's Gamma,
    // clamped as in the RHS (setStandardISM, blastwave.h:771-775).
    double rho_c = 0., drhodr_c = 0.;
    ism_density(P, R, rho_c, drhodr_c, b.status);
    const double GammaRho = 1., GammaRel = Gamma, dGammaRelDGamma = 1., dGammaRhodR = 0., cs_cbm = 0.;
    b.rho_rhs = rho_c; b.drhodr_rhs = drhodr_c; b.grel_rhs = GammaRel; b.rhs_seen = true;
    double rho = rho_c / b.M0;
    double drhodr = drhodr_c / b.M0;
    double dRdt = beta * c;
    double Gamma_ = GammaRel;
    double adi = adiNava(Gamma_);
    double dthetadr = 0.0;
    double dM2dR = dmdr(R, dthetadr, theta, rho) / b.ncells;
    if (dM2dR < 0.) b.status |= PBAGPU_S_NONFINITE;
    // energy sources, all zero here
    const double xi_inj = 1.;
    double dEinjdt = 0. / (b.M0 * c * c) / b.ncells;
    double dEinjdR = dEinjdt / dRdt;
    double Doppler = Gamma / (1. - beta * std::cos(0.));
    double dEingdR_abs = dEinjdR;
    double dEingdR_abs_dop = dEingdR_abs / Doppler / Doppler;
    double dEnucdR = 0. / (b.M0 * c * c) / dRdt;
    double dElumdR = 0. / (b.M0 * c * c) / dRdt;
    // dGamma/dR, method "our"
    double GammaEff = (adi * Gamma_ * Gamma_ - adi + 1.0) / Gamma_;
    double dGammaEffdGamma = (adi * Gamma_ * Gamma_ + adi - 1.0) / (Gamma_ * Gamma_);
    double num1 = (Gamma - GammaRho + GammaEff * (GammaRel - 1.)) * dM2dR;
    double num2 = -GammaEff * (adi - 1.) * Eint2 * (dM2dR / M2 - drhodr / rho - dGammaRhodR / GammaRho);
    double num3 = -(1. + GammaEff / Gamma * xi_inj) * dEingdR_abs;
    double denum1 = (1. + M2);
    double denum2 = Eint2 * dGammaEffdGamma;
    double denom3 = GammaEff * (adi - 1.) * Eint2 * dGammaRelDGamma / GammaRel;
    double dGammadR = -1. * (num1 + num2 + num3) / (denum1 + denum2 + denom3);
    double dlnV2dR = dM2dR / M2 - drhodr / rho - (1. / GammaRel) * dGammaRelDGamma * dGammadR + dGammaRhodR / GammaRho;
    double dEad2dR = 0.0;
    if (P.adiab_loss) dEad2dR = -(adi - 1.0) * Eint2 * dlnV2dR;
    double dEsh2dR = (GammaRel - 1.0) * dM2dR;
    if (cs_cbm > BetaFromGamma(GammaRel)) dEsh2dR *= 1e-10;   // subsonic upstream: no shock
    double dErad2dR = P.fs.epsilon_e_rad * dEsh2dR;
    double dEint2dR = dEsh2dR + dEad2dR - dErad2dR + dEnucdR - dElumdR + dEingdR_abs_dop;
    double dtcomov_dR = 1.0 / beta / Gamma / c;
    double dttdr = elapsed_dttdr_radial(R, Gamma * beta, dthetadr, b.spread);
    if (!std::isfinite(dRdt) || !std::isfinite(dGammadR) || dM2dR < 0. || !std::isfinite(dlnV2dR))
        b.status |= PBAGPU_S_NONFINITE;
    out[iR] = dRdt;
    out[itt] = dRdt * dttdr;
    out[itcomov] = dRdt * dtcomov_dR;
    out[iGamma] = dRdt * dGammadR;
    out[iEint2] = dRdt * dEint2dR;
    out[itheta] = dRdt * dthetadr;
    out[iErad2] = dRdt * dErad2dR;
    out[iEsh2] = dRdt * dEsh2dR;
    out[iEad2] = dRdt * dEad2dR;
    out[iM2] = dRdt * dM2dR;
}

// RHS_comb over all blastwaves of one run (model_evolve.h:1587-1689): jet layers by bw_type,
// ejecta blastwaves always through setStandardISM + rhs_fs_dense (:1631-1651)
static void rhs_all(Run &run, double x, const double *Y, double *out) {
    (void)x;
    run.n_rhs++;
    for (size_t i = 0; i < run.bws.size(); i++) {
        BW &b = run.bws[i];
        if (b.end_evolution) continue;
        if (b.dense) rhs_fs_dense(run, b, Y + i * NEQ, out + i * NEQ);
        else if (b.fsrs) rhs_fsrs(run, b, Y + i * NEQ, out + i * NEQ);
        else rhs_fs(run, b, Y + i * NEQ, out + i * NEQ);
    }
}

// DOP853 coefficients: Hairer & Wanner's published dop853 tableau (as used at ode_solvers.h:1311-1433)
namespace dp {
const double c2 = 0.526001519587677318785587544488E-01, c3 = 0.789002279381515978178381316732E-01,
             c4 = 0.118350341907227396726757197510E+00, c5 = 0.281649658092772603273242802490E+00,
             c6 = 0.333333333333333333333333333333E+00, c7 = 0.25E+00, c8 = 0.307692307692307692307692307692E+00,
             c9 = 0.651282051282051282051282051282E+00, cA = 0.6E+00, cB = 0.857142857142857142857142857142E+00,
             cC = 1.0;
const double a21 = 5.26001519587677318785587544488E-2, a31 = 1.97250569845378994544595329183E-2,
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
const double b1 = 5.42937341165687622380535766363E-2, b6 = 4.45031289275240888144113950566,
             b7 = 1.89151789931450038304281599044, b8 = -5.8012039600105847814672114227,
             b9 = 3.1116436695781989440891606237E-1, bA = -1.52160949662516078556178806805E-1,
             bB = 2.01365400804030348374776537501E-1, bC = 4.47106157277725905176885569043E-2;
const double er1 = 0.1312004499419488073250102996E-01, er6 = -0.1225156446376204440720569753E+01,
             er7 = -0.4957589496572501915214079952E+00, er8 = 0.1664377182454986536961530415E+01,
             er9 = -0.3503288487499736816886487290E+00, erA = 0.3341791187130174790297318841E+00,
             erB = 0.8192320648511571246570742613E-01, erC = -0.2235530786388629525884427845E-01;
const double bhh1 = 0.244094488188976377952755905512E+00, bhh9 = 0.733846688281611857341361741547E+00,
             bhhC = 0.220588235294117647058823529412E-01;
} // namespace dp

struct Dop {
    size_t n;
    std::vector<double> x, k1, k2, k4, k6, k7, k8, k9, kA, kB, kD; // k3,k5 alias k2; kC aliases k4
    explicit Dop(size_t n_) : n(n_), x(n_, 0.), k1(n_, 0.), k2(n_, 0.), k4(n_, 0.), k6(n_, 0.), k7(n_, 0.),
                              k8(n_, 0.), k9(n_, 0.), kA(n_, 0.), kB(n_, 0.), kD(n_, 0.) {}
};

// dop853_integrator::_integrate8 (ode_solvers.h:1434-1479)
static void integrate8(Run &run, Dop &d, double time, const double *value, double h, double atol, double rtol,
                       double &err_out, double &stf_out) {
    using namespace dp;
    size_t n = d.n;
    double *x = d.x.data(), *k1 = d.k1.data(), *k2 = d.k2.data(), *k4 = d.k4.data(), *k6 = d.k6.data();
    double *k7 = d.k7.data(), *k8 = d.k8.data(), *k9 = d.k9.data(), *kA = d.kA.data(), *kB = d.kB.data();
    double *k3 = k2, *k5 = k2, *kC = k4;
    for (size_t i = 0; i < n; i++) x[i] = value[i] + h * (a21 * k1[i]);
    rhs_all(run, time + c2 * h, x, k2);
    for (size_t i = 0; i < n; i++) x[i] = value[i] + h * (a31 * k1[i] + a32 * k2[i]);
    rhs_all(run, time + c3 * h, x, k3);
    for (size_t i = 0; i < n; i++) x[i] = value[i] + h * (a41 * k1[i] + a43 * k3[i]);
    rhs_all(run, time + c4 * h, x, k4);
    for (size_t i = 0; i < n; i++) x[i] = value[i] + h * (a51 * k1[i] + a53 * k3[i] + a54 * k4[i]);
    rhs_all(run, time + c5 * h, x, k5);
    for (size_t i = 0; i < n; i++) x[i] = value[i] + h * (a61 * k1[i] + a64 * k4[i] + a65 * k5[i]);
    rhs_all(run, time + c6 * h, x, k6);
    for (size_t i = 0; i < n; i++) x[i] = value[i] + h * (a71 * k1[i] + a74 * k4[i] + a75 * k5[i] + a76 * k6[i]);
    rhs_all(run, time + c7 * h, x, k7);
    for (size_t i = 0; i < n; i++)
        x[i] = value[i] + h * (a81 * k1[i] + a84 * k4[i] + a85 * k5[i] + a86 * k6[i] + a87 * k7[i]);
    rhs_all(run, time + c8 * h, x, k8);
    for (size_t i = 0; i < n; i++)
        x[i] = value[i] + h * (a91 * k1[i] + a94 * k4[i] + a95 * k5[i] + a96 * k6[i] + a97 * k7[i] + a98 * k8[i]);
    rhs_all(run, time + c9 * h, x, k9);
    for (size_t i = 0; i < n; i++)
        x[i] = value[i] + h * (aA1 * k1[i] + aA4 * k4[i] + aA5 * k5[i] + aA6 * k6[i] + aA7 * k7[i] + aA8 * k8[i] +
                               aA9 * k9[i]);
    rhs_all(run, time + cA * h, x, kA);
    for (size_t i = 0; i < n; i++)
        x[i] = value[i] + h * (aB1 * k1[i] + aB4 * k4[i] + aB5 * k5[i] + aB6 * k6[i] + aB7 * k7[i] + aB8 * k8[i] +
                               aB9 * k9[i] + aBA * kA[i]);
    rhs_all(run, time + cB * h, x, kB);
    for (size_t i = 0; i < n; i++)
        x[i] = value[i] + h * (aC1 * k1[i] + aC4 * k4[i] + aC5 * k5[i] + aC6 * k6[i] + aC7 * k7[i] + aC8 * k8[i] +
                               aC9 * k9[i] + aCA * kA[i] + aCB * kB[i]);
    rhs_all(run, time + cC * h, x, kC);
    double err1 = 0.0, err2 = 0.0, err3 = 0.0;
    for (size_t i = 0; i < n; i++) {
        double slope = b1 * k1[i] + b6 * k6[i] + b7 * k7[i] + b8 * k8[i] + b9 * k9[i] + bA * kA[i] + bB * kB[i] +
                       bC * kC[i];
        double _y = value[i] + h * slope;
        double sk = atol + rtol * std::max(std::abs(value[i]), std::abs(_y));
        double e1 = er1 * k1[i] + er6 * k6[i] + er7 * k7[i] + er8 * k8[i] + er9 * k9[i] + erA * kA[i] + erB * kB[i] +
                    erC * kC[i];
        double e2 = slope - bhh1 * k1[i] - bhh9 * k9[i] - bhhC * kC[i];
        err1 += (e1 / sk) * (e1 / sk);
        err2 += (e2 / sk) * (e2 / sk);
        err3 += slope * slope;
        x[i] = _y;
    }
    double deno = err1 + 0.01 * err2;
    if (deno <= 0.0) deno = 1.0;
    err_out = std::abs(h) * err1 / std::sqrt(n * deno);
    stf_out = h * h * err3;
}

// dop853_integrator::_determine_initial_step (ode_solvers.h:1596-1628), k1 already evaluated
static double initial_step(Run &run, Dop &d, double time, const double *value, double atol, double rtol,
                           double hmax) {
    size_t n = d.n;
    double dnf = 0.0, dny = 0.0;
    for (size_t i = 0; i < n; i++) {
        double sk = atol + rtol * std::abs(value[i]);
        dnf += (d.k1[i] / sk) * (d.k1[i] / sk);
        dny += (value[i] / sk) * (value[i] / sk);
    }
    double h1;
    if (dnf <= 1e-10 || dny <= 1e-10) h1 = 1.0e-6;
    else h1 = std::sqrt(dny / dnf) * 0.01;
    h1 = std::min(h1, hmax);
    for (size_t i = 0; i < n; i++) d.x[i] = value[i] + h1 * d.k1[i];
    rhs_all(run, time + h1, d.x.data(), d.k2.data());
    double norm2 = 0.0;
    for (size_t i = 0; i < n; i++) {
        double sk = atol + rtol * x86_int_abs(value[i]);
        double z = (d.k2[i] - d.k1[i]) / sk;
        norm2 += z * z;
    }
    norm2 = std::sqrt(norm2) / h1;
    double der12 = std::max(norm2, std::sqrt(dnf));
    double h2;
    if (der12 <= 1e-15) h2 = std::max(1e-6, std::abs(h1) * 1e-3);
    else h2 = std::pow(0.01 / der12, 1.0 / 8);
    return std::min(std::min(100 * h1, h2), hmax);
}

// dop853_integrator::integrate (ode_solvers.h:1793-1912): time advanced in place to timeN
static int dop853(Run &run, Dop &d, double &time, double *value, double timeN, double rtol, double atol,
                  int &nsteps) {
    size_t n = d.n;
    const double safe = 0.9, facc1 = 1.0 / 0.333, facc2 = 1.0 / 6.000, expo1 = 1.0 / 8.0;
    double hmax = std::abs(timeN - time);
    rhs_all(run, time, value, d.k1.data());
    double h = initial_step(run, d, time, value, atol, rtol, hmax);
    double facold = 1e-4;
    bool reject = false, last = false;
    long guard = 0;
    for (;;) {
        if (++guard > 10000000) return -1;
        if ((time + 1.01 * h - timeN) > 0.0) { h = timeN - time; last = true; }
        double err, stf;
        integrate8(run, d, time, value, h, atol, rtol, err, stf);
        double fac11 = std::pow(err, expo1);
        double fac = fac11 / (1.0 * safe);
        fac = fac < facc2 ? facc2 : (fac > facc1 ? facc1 : fac);
        double hnew = h / fac;
        if (err > 1.0) {
            h /= std::min(facc1, fac11 / safe);
            reject = true;
            last = false;
        } else {
            facold = std::max(err, 1e-4);
            (void)facold;
            rhs_all(run, time + h, d.x.data(), d.kD.data());
            for (size_t i = 0; i < n; i++) { d.k1[i] = d.kD[i]; value[i] = d.x[i]; }
            time += h;
            nsteps++;
            if (last) return 0;
            if (std::abs(hnew) > hmax) hnew = hmax;
            if (reject) hnew = std::min(std::abs(hnew), std::abs(h));
            reject = false;
            h = hnew;
        }
    }
}

static void bw_init(Run &run, BW &b, const pba_layer_init &id, double t0_grid, double *Y) {
    const pba_run_params &P = *run.P;
    std::memset(Y, 0, sizeof(double) * NEQ);
    b.E0 = id.E0; b.M0 = id.M0; b.mom0 = id.mom0;
    b.Gamma0 = GammaFromMom(id.mom0);
    b.beta0 = BetaFromMom(id.mom0);
    b.R0 = BetaFromMom(id.mom0) * c * t0_grid;      // ejecta_id.h:333-337
    b.theta_a = 0.; b.theta_b0 = id.theta_b0; b.ncells = id.ncells;   // blastwave_base.h:304-312
    b.ctheta0 = id.ctheta0; b.theta_w = id.theta_wing; b.theta_core = id.theta_core;
    b.theta_max = P.theta_max;
    b.ilayer = id.ilayer; b.nlayers = id.nlayers;
    b.fsrs = !b.dense && P.fsrs != 0;
    b.do_rs = b.fsrs && P.do_rs != 0; b.do_rs_rad = b.fsrs && P.do_rs_radiation != 0;
    b.shutOff = false; b.end_evolution = false; b.end_spreading = false;
    b.prev_idx_x = 0; b.status = 0;
    // density at the start radius (evaluateRhoDrhoDrDefault(R0), blastwave.h:92-101): the RHS
    // re-evaluates it at every R, the stored m_rho_ keeps this value
    b.rho_def = 0.; b.drhodr0 = 0.;
    ism_density(P, b.R0, b.rho_def, b.drhodr0, b.status);
    // setParamsRS: too slow layers evolve without the reverse shock (blastwave_base.h:437-445)
    if (b.fsrs) {
        double min_g0 = b.do_rs ? P.min_gamma0_for_rs : 0.;
        if (b.Gamma0 < min_g0) { b.do_rs = false; b.do_rs_rad = false; b.shutOff = true; b.fsrs = false; }
    }
    if (b.M0 == 0. && b.E0 == 0.) { b.end_evolution = true; return; }
    b.beta0 = BetaFromGamma(b.Gamma0);
    double rho = b.rho_def;
    double M20 = (2.0 / 3.0) * pi * (std::cos(b.theta_a) - std::cos(b.theta_b0)) * rho * std::pow(b.R0, 3) / b.ncells;
    if (b.M0 <= 0.) b.M0 = b.E0 / (b.Gamma0 * c * c);
    b.Rd = std::pow(3. / (4. * pi) * 1. / (c * c * mp) * b.E0 / (rho / mp * b.Gamma0 * b.Gamma0), 1. / 3.);
    Y[iR] = b.R0;
    // EQS::init_elapsed_time (blastwave_components.h:141-163): the jet spreads ("our"); an
    // ejecta blastwave takes the radial form unless method_spread is set
    if (b.spread) {
        double beta0m = BetaFromMom(b.mom0);
        double dThetadr0 = 0.0;
        Y[itt] = b.R0 / (c * beta0m) * (std::sqrt(1. + b.R0 * b.R0 * dThetadr0 * dThetadr0)) - b.R0 / c;
    } else {
        double xx = 1. + b.mom0 * b.mom0 * b.mom0 / (std::sqrt(1. + b.mom0 * b.mom0));
        double beta0m = BetaFromMom(b.mom0);
        Y[itt] = b.R0 / (c * xx * (1.0 + beta0m));
    }
    Y[itcomov] = b.R0 / (b.beta0 * b.Gamma0 * c);
    Y[iEint2] = (b.Gamma0 - 1.) * M20 / b.M0;
    Y[iEint2] += 0. / b.M0 / c / c;
    Y[iGamma] = b.Gamma0;
    Y[itheta] = b.theta_b0;
    Y[iM2] = M20 / b.M0;
    Y[iRw] = .5 * Y[iR];
    Y[iWtt] = Y[itt];
    if (b.fsrs) {
        Y[iEint3] = 0.;
        if (P.init_deltaR4) Y[ideltaR4] = c * P.tprompt * b.beta0;
    }
}

// insertSolution + applyUnits + addOtherVars + limits for one output step
static void post_step(Run &run, BW &b, const double *Ycode, int it, double tb, double *T, int nt) {
    const pba_run_params &P = *run.P;
    auto V = [&](int v) -> double & { return T[(size_t)v * nt + it]; };
    if (b.end_evolution) return;
    double Y[NEQ];
    for (int k = 0; k < NEQ; k++) Y[k] = Ycode[k];
    // applyUnits (blastwave.h:352-368)
    Y[iM2] *= b.M0; Y[iEint2] *= (b.M0 * c * c); Y[iEad2] *= (b.M0 * c * c); Y[iEsh2] *= (b.M0 * c * c);
    Y[iErad2] *= (b.M0 * c * c); Y[iM3] *= b.M0; Y[iEint3] *= (b.M0 * c * c); Y[iEad3] *= (b.M0 * c * c);
    Y[iEsh3] *= (b.M0 * c * c); Y[iErad3] *= (b.M0 * c * c);
    // insertSolution (blastwave.h:270-340)
    V(PV_tburst) = tb;
    V(PV_R) = Y[iR]; V(PV_Rsh) = Y[iRsh]; V(PV_tt) = Y[itt];
    V(PV_mom) = MomFromGamma(Y[iGamma]); V(PV_Gamma) = Y[iGamma]; V(PV_beta) = BetaFromGamma(Y[iGamma]);
    V(PV_theta) = Y[itheta]; V(PV_M2) = Y[iM2]; V(PV_tcomov) = Y[itcomov]; V(PV_Ead2) = Y[iEad2];
    V(PV_Eint2) = Y[iEint2]; V(PV_Esh2) = Y[iEsh2]; V(PV_Erad2) = Y[iErad2];
    if (b.fsrs) {
        V(PV_Rrsh) = Y[iRrsh]; V(PV_Eint3) = Y[iEint3]; V(PV_Ead3) = Y[iEad3]; V(PV_Erad3) = Y[iErad3];
        V(PV_Esh3) = Y[iEsh3]; V(PV_M3) = Y[iM3]; V(PV_deltaR4) = Y[ideltaR4];
    }
    if (Y[iR] < 1. || V(PV_Gamma) < 1.) b.status |= PBAGPU_S_NONFINITE;
    // addOtherVars (blastwave.h:525-640)
    V(PV_ctheta) = ctheta_of(V(PV_theta), b.ilayer, b.nlayers);
    V(PV_rho) = b.rho_def;
    V(PV_drhodr) = b.drhodr0;
    if (b.dense) {
        // the density object as the last RHS call left it (set_standard_ism); before the first
        // call, setInitConditions' values and GammaRel unset (-1), stored as Gamma (:544-560)
        if (b.rhs_seen) { V(PV_rho) = b.rho_rhs; V(PV_drhodr) = b.drhodr_rhs; }
        double grel = b.rhs_seen ? b.grel_rhs : -1.;
        V(PV_GammaRel) = grel < 1. ? V(PV_Gamma) : grel;
    }
    V(PV_adi) = adiNava(V(PV_Gamma));
    V(PV_rho2) = rho2t(V(PV_Gamma), V(PV_adi), V(PV_rho));
    V(PV_GammaFsh) = GammaSh(V(PV_Gamma), V(PV_adi));
    {
        double nprime = V(PV_rho2) / mp;
        V(PV_thickness) = V(PV_M2) / (2. * pi * mp * V(PV_R) * V(PV_R) * (1. - std::cos(V(PV_theta)) / b.ncells) *
                                      V(PV_Gamma) * nprime);
        double m2 = V(PV_M2), E = V(PV_Eint2);
        V(PV_U_p) = (E == 0 || m2 == 0) ? 0. : E / (m2 / V(PV_rho2));
    }
    if ((it > 0) && b.do_rs && (!b.shutOff) && b.fsrs) {
        long double lbeta = BetaFromMom(V(PV_mom)), lbeta0 = b.beta0;
        double G = V(PV_Gamma), G0 = b.Gamma0;
        long double G43 = G * G0 * (1.0 / (G0 * G0) + 1.0 / (G * G) - 1.0 / (G * G) / (G0 * G0)) / (1.0 + lbeta * lbeta0);
        V(PV_Gamma43) = (double)G43;
        double alpha_of = P.tprompt * b.beta0 * c;
        double rho4_fac_1 = b.M0 / (2.0 * alpha_of * pi * (std::cos(b.theta_a) - std::cos(b.theta_b0)));
        double rho4_fac = rho4_fac_1 / (V(PV_R) * V(PV_R));
        V(PV_rho4) = P.exponential_rho4 ? rho4_fac * std::exp(-V(PV_deltaR4) / alpha_of) : rho4_fac;
        V(PV_adi3) = adiNava(V(PV_Gamma43));
        V(PV_rho3) = rho2t(V(PV_Gamma), V(PV_adi3), V(PV_rho4));
        V(PV_GammaRsh) = GammaSh(V(PV_Gamma43), V(PV_adi3));
        double nprime3 = V(PV_rho3) / mp;
        V(PV_thickness_rs) = V(PV_M3) / (2. * pi * mp * V(PV_R) * V(PV_R) * (1. - std::cos(V(PV_theta)) / b.ncells) *
                                         V(PV_Gamma43) * nprime3);
        double m3 = V(PV_M3), E3 = V(PV_Eint3);
        V(PV_U_p3) = (E3 == 0 || m3 == 0) ? 0. : E3 / (m3 / V(PV_rho3));
    }
    if ((V(PV_adi) < 1.) || (V(PV_R) < 1.) || (V(PV_rho2) < 0.) || (V(PV_U_p) < 0) || (V(PV_GammaFsh) < 1.))
        b.status |= PBAGPU_S_NONFINITE;
    // computeMicrophsysics limits part (blastwave_radiation.h:881-1107, 1185-1198)
    int flags = 0;
    bool valid = true;
    if (b.R0 < 0) valid = false;
    if (valid && V(PV_beta) < P.fs.beta_min) {
        if (b.i_end_r > it - 1) b.i_end_r = it - 1;
        valid = false;
    }
    // ejecta: a subsonic upstream (sound speed 0 in a standard ISM) has no shock (:918-928); the
    // velocity test below uses GammaRel, which equals Gamma here
    if (valid && b.dense && (0. >= V(PV_beta))) {
        V(PV_B) = 0.;
        if (b.i_end_r > it - 1) b.i_end_r = it - 1;
        valid = false;
    }
    if (valid && ((V(PV_U_p) <= 0) || (V(PV_rho2) <= 0))) valid = false;
    if (valid && ((V(PV_R) < 1) || (V(PV_beta) < P.fs.beta_min))) valid = false;
    if (valid) {
        bool rs = b.fsrs && b.do_rs_rad && (it > 0) && (T[(size_t)PV_thickness_rs * nt + it - 1] > 0) &&
                  (V(PV_GammaRsh) > 0) && (V(PV_U_p3) > 0);
        if (rs) {
            Lim L = ele_limits(b.rs_sh, b.rs_st, V(PV_U_p3), V(PV_GammaRsh), V(PV_tt), V(PV_tcomov), V(PV_rho3) / mp);
            if (L.ok) {
                V(PV_B3) = L.B; V(PV_gm_rs) = L.gm; V(PV_gM_rs) = L.gM; V(PV_gc_rs) = L.gc;
                V(PV_Theta_rs) = L.Theta; V(PV_z_cool_rs) = L.z_cool; V(PV_nprime_rs) = L.nprime;
                V(PV_accel_frac_rs) = L.acc;
            }
            flags |= 2;
        }
        Lim L = ele_limits(b.fs_sh, b.fs_st, V(PV_U_p), V(PV_GammaFsh), V(PV_tt), V(PV_tcomov), V(PV_rho2) / mp);
        if (L.ok) {
            V(PV_B) = L.B; V(PV_gm) = L.gm; V(PV_gM) = L.gM; V(PV_gc) = L.gc; V(PV_Theta) = L.Theta;
            V(PV_z_cool) = L.z_cool; V(PV_nprime) = L.nprime; V(PV_accel_frac) = L.acc;
        }
        flags |= 1;
    }
    V(PV_mflags) = flags;
}

// ejecta: 0 for jet layers; 1 for ejecta blastwaves (bw_type = fs_dense), 2 for ejecta blastwaves
// with method_spread set (the radial or the spreading form of dt_obs/dR)
static void evolve_run(const pba_run_params *P, const pba_layer_init *init, int nbw, int nt, const double *tgrid,
                       double *table, pba_bw_result *res, int ejecta = 0) {
    Run run;
    run.P = P;
    EleGrid gfs, grs;
    gfs.build(P->fs.gam2, P->fs.ngam);
    run.fs = Shock{&P->fs, gfs.e[0], gfs.e[P->fs.ngam - 1]};
    if (P->fsrs) {
        grs.build(P->rs.gam2, P->rs.ngam);
        run.rs = Shock{&P->rs, grs.e[0], grs.e[P->rs.ngam - 1]};
    } else run.rs = run.fs;
    run.bws.resize(nbw);
    size_t neq = (size_t)nbw * NEQ;
    std::vector<double> Y(neq, 0.);
    for (int i = 0; i < nbw; i++) {
        // a blastwave with its own electron-grid size: the grid's first and last points for the
        // gamma_max clamp (computeGammaMax, microphysics_base.h:824-837)
        BW &bw = run.bws[i];
        bw.fs_sh = run.fs;
        bw.rs_sh = run.rs;
        if (init[i].ngam_fs > 0) {
            EleGrid g;
            g.build(P->fs.gam2, init[i].ngam_fs);
            bw.fs_sh = Shock{&P->fs, g.e[0], g.e[init[i].ngam_fs - 1]};
        }
        if (init[i].ngam_rs > 0 && P->fsrs) {
            EleGrid g;
            g.build(P->rs.gam2, init[i].ngam_rs);
            bw.rs_sh = Shock{&P->rs, g.e[0], g.e[init[i].ngam_rs - 1]};
        }
        bw.dense = ejecta != 0;
        bw.spread = ejecta != 1;
        bw_init(run, run.bws[i], init[i], tgrid[0], &Y[(size_t)i * NEQ]);
        run.bws[i].i_end_r = nt;
        std::fill(table + (size_t)i * PBA_NV * nt, table + (size_t)(i + 1) * PBA_NV * nt, 0.);
    }
    int nsteps = 0;
    // step 0: insertSolution(0) / addOtherVariables(0) / computeMicrophysics(0) (model_evolve.h:276-283)
    for (int i = 0; i < nbw; i++)
        post_step(run, run.bws[i], &Y[(size_t)i * NEQ], 0, tgrid[0], table + (size_t)i * PBA_NV * nt, nt);
    Dop d(neq);
    std::vector<double> kn(neq, 0.), xn(neq, 0.), dl(neq, 0.);
    double x = P->tb0;
    for (int it = 1; it < nt; it++) {
        double dt = tgrid[it] - tgrid[it - 1];
        int rc = 0;
        if (P->integrator == 1) {
            // IntegratorStatic<rk4_integrator>: one classical RK4 step over the interval (ode_solvers.h:319-366)
            const double h = dt;
            rhs_all(run, x, Y.data(), kn.data());
            for (size_t i = 0; i < neq; i++) { dl[i] = (1.0 / 6.0) * h * kn[i]; xn[i] = Y[i] + 0.5 * h * kn[i]; }
            rhs_all(run, x + 0.5 * h, xn.data(), kn.data());
            for (size_t i = 0; i < neq; i++) { dl[i] += (2.0 / 6.0) * h * kn[i]; xn[i] = Y[i] + 0.5 * h * kn[i]; }
            rhs_all(run, x + 0.5 * h, xn.data(), kn.data());
            for (size_t i = 0; i < neq; i++) { dl[i] += (2.0 / 6.0) * h * kn[i]; xn[i] = Y[i] + h * kn[i]; }
            rhs_all(run, x + h, xn.data(), kn.data());
            for (size_t i = 0; i < neq; i++) {
                const double a = dl[i] + (1.0 / 6.0) * h * kn[i];
                Y[i] += a;
            }
            x += h;
            nsteps++;
        } else {
            rc = dop853(run, d, x, Y.data(), x + dt, P->rtol, P->rtol, nsteps);
        }
        for (int i = 0; i < nbw; i++) {
            BW &b = run.bws[i];
            if (rc != 0) b.status |= PBAGPU_S_NMAX;
            const double *y = &Y[(size_t)i * NEQ];
            if (b.end_evolution) continue;
            // isToTerminate (blastwave.h:420-471)
            double g = y[iGamma], be = BetaFromGamma(g), E2 = y[iEint2];
            if ((E2 <= 0.) || !std::isfinite(E2) || (g < 1.) || (be < 0) || (std::fabs(be) < 1.e-6) || !std::isfinite(be))
                b.status |= PBAGPU_S_TERMINATED;
        }
        for (int i = 0; i < nbw; i++) {
            BW &b = run.bws[i];
            const double *y = &Y[(size_t)i * NEQ];
            // isToStopLateralExpansion (blastwave.h:473-485)
            if (!b.end_spreading && y[itheta] > 0.99 * b.theta_max) b.end_spreading = true;
        }
        for (int i = 0; i < nbw; i++) {
            BW &b = run.bws[i];
            const double *y = &Y[(size_t)i * NEQ];
            // isToStopReverseShock (blastwave.h:487-510)
            if (!(b.shutOff || !b.do_rs)) {
                double alpha_of = P->tprompt * b.beta0 * c;
                double rho4_fac_1 = b.M0 / (2.0 * alpha_of * pi * (std::cos(b.theta_a) - std::cos(b.theta_b0)));
                double rho4_fac = rho4_fac_1 / (y[iR] * y[iR]);
                double rho4 = P->exponential_rho4 ? rho4_fac * std::exp(-y[ideltaR4] / alpha_of) : rho4_fac;
                if ((y[iGamma] < b.Gamma0) && (rho4 < P->rs_shutoff_rho || y[iM3] >= 1.)) b.shutOff = true;
            }
        }
        for (int i = 0; i < nbw; i++) {
            BW &b = run.bws[i];
            const double *y = &Y[(size_t)i * NEQ];
            // isSolutionOk (blastwave.h:512-560)
            if (!b.end_evolution) {
                double g = y[iGamma];
                if (g < 1. || BetaFromGamma(g) < 0 || y[itheta] < b.theta_b0 * 0.99 || y[itt] <= 0.)
                    b.status |= PBAGPU_S_NONFINITE;
            }
            b.prev_idx_x = it;
        }
        for (int i = 0; i < nbw; i++)
            post_step(run, run.bws[i], &Y[(size_t)i * NEQ], it, tgrid[it], table + (size_t)i * PBA_NV * nt, nt);
    }
    for (int i = 0; i < nbw; i++) {
        res[i].status = run.bws[i].status;
        res[i].i_end_r = run.bws[i].i_end_r;
        res[i].fsrs = run.bws[i].fsrs ? 1 : 0;
        res[i].n_steps = nsteps;
        res[i].n_rhs = run.n_rhs;
    }
}

} // namespace orc

extern "C" int orc_dyn_evolve(const pba_run_params *params, int n_runs, int bw_per_run, const pba_layer_init *init,
                              int nt, const double *tgrid, double *table, pba_bw_result *res) {
    if (!params || !init || !tgrid || !table || !res || n_runs < 1 || bw_per_run < 1 || nt < 2) return PBAGPU_E_ARG;
#pragma omp parallel for schedule(dynamic)   // runs are independent
    for (int r = 0; r < n_runs; r++)
        orc::evolve_run(&params[r], init + (size_t)r * bw_per_run, bw_per_run, nt, tgrid,
                        table + (size_t)r * bw_per_run * PBA_NV * nt, res + (size_t)r * bw_per_run);
    return PBAGPU_OK;
}

// The ejecta of a kilonova run: every blastwave evolved as BlastWave::rhs_fs_dense in a standard
// ISM (model_evolve.h:1627-1651), ordered as the reference's state vector, layer-major
// (for each layer, its velocity shells; model_evolve.h:225-234).  spread: method_spread != None.
extern "C" int orc_dyn_evolve_ejecta(const pba_run_params *params, int n_runs, int bw_per_run,
                                     const pba_layer_init *init, int spread, int nt, const double *tgrid,
                                     double *table, pba_bw_result *res) {
    if (!params || !init || !tgrid || !table || !res || n_runs < 1 || bw_per_run < 1 || nt < 2) return PBAGPU_E_ARG;
    for (int r = 0; r < n_runs; r++)
        if (params[r].fsrs) return PBAGPU_E_ARG;   // ejecta blastwaves have no reverse shock
#pragma omp parallel for schedule(dynamic)
    for (int r = 0; r < n_runs; r++)
        orc::evolve_run(&params[r], init + (size_t)r * bw_per_run, bw_per_run, nt, tgrid,
                        table + (size_t)r * bw_per_run * PBA_NV * nt, res + (size_t)r * bw_per_run,
                        spread ? 2 : 1);
    return PBAGPU_OK;
}
