/*
 * Oracle, stage 3: equal-time-arrival-surface light curve (TEST INFRASTRUCTURE ONLY).
 *
 * Adaptive method, per layer and (t_obs, nu_obs) pair, primary jet + counter-jet:
 *   EATS::evalLightCurve / evalFluxDensA / parsPars / integrate_theta_phi /
 *   phi_integrand / costheta_integrand / integrand / find_jet_edge (src/eats.h:103-983)
 *   with the integrand callback fluxDensAdaptiveWithComov
 *   (src/blastwave/blastwave_radiation.h:75-141) ->
 *   ElectronAndRadiation::fluxDens / computeFluxDensity / computeIntensity
 *   (src/microphysics/microphysics.h:11-58, 799-867, 1185-1217),
 *   interpolation helpers (src/utilitites/interpolators.h:13-133, 724-808) and
 *   observer angles (src/image.h:361-378).
 * Quadrature: the CADRE adaptive Romberg scheme with its binary max-heap of
 *   intervals, m9_adapt / cadreInitInterval / cadreProcessInterval /
 *   cadreSplitInterval / mesh9* (src/utilitites/quadratures.h:655-770, 1613-2115),
 *   restated with the same interval store so that the summation order matches.
 */
#include "oracle_common.h"
#include <cstring>
#include <cstdio>
#include <limits>

namespace orc {

struct Interval9 {
    double a, b, I, err, fa, fll, fl, flr, fm, frl, fr, frr, fb;
    int refinement;
};

struct Mesh9 {
    std::vector<Interval9> heap;
    size_t N = 0;
    size_t maxN = 0;
    void insert(const Interval9 &i) {
        if (N >= heap.size()) heap.resize(heap.empty() ? 4 : heap.size() * 2);
        heap[N] = i;
        N++;
        if (N > maxN) maxN = N;
        size_t cI = N - 1, p = (cI - 1) / 2;
        while (cI != 0 && heap[p].err < heap[cI].err) {
            std::swap(heap[p], heap[cI]);
            cI = p;
            p = (cI - 1) / 2;
        }
    }
    void extract(Interval9 &worst) {
        worst = heap[0];
        heap[0] = heap[N - 1];
        N--;
        size_t p = 0, c1 = 1, c2 = 2;
        while (c1 < N) {
            size_t cI = c1;
            double e = heap[c1].err;
            if (c2 < N && heap[c2].err > e) { cI = c2; e = heap[c2].err; }
            if (e <= heap[p].err) break;
            std::swap(heap[p], heap[cI]);
            p = cI;
            c1 = 2 * p + 1;
            c2 = c1 + 1;
        }
    }
    double total_I() const { double s = 0.; for (size_t i = 0; i < N; i++) s += heap[i].I; return s; }
    double total_err() const { double s = 0.; for (size_t i = 0; i < N; i++) s += heap[i].err; return s; }
};

template <class F> static int cadre_process(F &f, Interval9 *i) {
    const double scaleTol = 0.2;
    int count = 0;
    if (i->refinement < 4) {
        double fa = i->fa, fm = i->fm, fb = i->fb, fl, fr;
        if (i->refinement == 0) {
            fl = f(0.75 * i->a + 0.25 * i->b); count++;
            fr = f(0.25 * i->a + 0.75 * i->b); count++;
            i->fl = fl; i->fr = fr;
        } else {
            fl = i->fl; fr = i->fr;
            i->refinement = 0;
        }
        double hh = 0.25 * (i->b - i->a);
        double T0 = 2 * hh * (fa + fb);
        double T1 = 0.5 * T0 + 2 * hh * fm;
        double T2 = 0.5 * T1 + hh * (fl + fr);
        double R0 = (T1 - T0) / (T2 - T1);
        if (R0 > 4 - scaleTol && R0 < 4 + scaleTol) {
            double fll = f(0.875 * i->a + 0.125 * i->b); count++;
            double flr = f(0.625 * i->a + 0.375 * i->b); count++;
            double frl = f(0.375 * i->a + 0.625 * i->b); count++;
            double frr = f(0.125 * i->a + 0.875 * i->b); count++;
            i->fll = fll; i->flr = flr; i->frl = frl; i->frr = frr;
            i->refinement = 1;
            double T3 = 0.5 * T2 + 0.5 * hh * (fll + flr + frl + frr);
            double R1 = (T2 - T1) / (T3 - T2);
            if (R1 > 4 - scaleTol && R1 < 4 + scaleTol) {
                double R11 = (4 * T1 - T0) / 3.0;
                double R21 = (4 * T2 - T1) / 3.0;
                double R22 = (16 * R21 - R11) / 15.0;
                i->fa = T3;
                i->fll = (4 * T3 - T2) / 3.0;
                i->fl = (16 * i->fll - R21) / 15.0;
                i->flr = (64 * i->fl - R22) / 63.0;
                i->I = i->flr;
                i->err = std::fabs((i->fl - R22) / 63.0);
                i->refinement = 4;
                return count;
            } else {
                double errt = (T3 - T2) / 3.0;
                i->err = std::fabs(errt);
                i->I = T3 + errt;
                return count;
            }
        } else if (R0 >= 4 + scaleTol || R0 < 2 || R0 != R0) {
            double err10 = std::fabs(T1 - T0), err21 = std::fabs(T2 - T1);
            i->err = err10 > err21 ? err10 : err21;
            i->I = T2;
            return count;
        } else {
            double err = -(T2 - T1) * (T2 - T1) / (T2 - 2 * T1 + T0);
            i->err = std::fabs(err);
            if (std::fabs(T2 + err) < 1.0e-14 * std::fabs(T2)) i->I = 1.0e-14 * T2;
            else i->I = T2 + err;
            return count;
        }
    }
    if (i->refinement < 0) return 0;
    int nint = 1;
    for (int n = 0; n < i->refinement; n++) nint *= 2;
    double hh = (i->b - i->a) / nint;
    double sum = 0;
    for (int n = 1; n < nint; n += 2) { sum += f(i->a + hh * n); count++; }
    double T = 0.5 * i->fa + hh * sum;
    double Rnm = i->fa, Rnmp = i->fll;
    i->fa = T;
    i->fll = (4.0 * i->fa - Rnm) / 3.0;
    if (i->refinement == 1) { i->I = i->fll; i->err = std::fabs((i->fa - Rnm) / 3.0); i->refinement += 1; return count; }
    Rnm = Rnmp; Rnmp = i->fl;
    i->fl = (16 * i->fll - Rnm) / 15.0;
    if (i->refinement == 2) { i->I = i->fl; i->err = std::fabs((i->fll - Rnm) / 15.0); i->refinement += 1; return count; }
    Rnm = Rnmp; Rnmp = i->flr;
    i->flr = (64 * i->fl - Rnm) / 63.0;
    if (i->refinement == 3) { i->I = i->flr; i->err = std::fabs((i->fl - Rnm) / 63.0); i->refinement += 1; return count; }
    Rnm = Rnmp; Rnmp = i->fm;
    i->fm = (256 * i->flr - Rnm) / 255.0;
    if (i->refinement == 4) { i->I = i->fm; i->err = std::fabs((i->flr - Rnm) / 255.0); i->refinement += 1; return count; }
    Rnm = Rnmp; Rnmp = i->frl;
    i->frl = (1024 * i->fm - Rnm) / 1023.0;
    if (i->refinement == 5) { i->I = i->frl; i->err = std::fabs((i->fm - Rnm) / 1023.0); i->refinement += 1; return count; }
    Rnm = Rnmp; Rnmp = i->fr;
    i->fr = (4096 * i->frl - Rnm) / 4095.0;
    if (i->refinement == 6) { i->I = i->fr; i->err = std::fabs((i->frl - Rnm) / 4095.0); i->refinement += 1; return count; }
    Rnm = Rnmp; Rnmp = i->frr;
    i->frr = (16384 * i->fr - Rnm) / 16383.0;
    if (i->refinement == 7) { i->I = i->frr; i->err = std::fabs((i->fr - Rnm) / 16383.0); i->refinement += 1; return count; }
    Rnm = Rnmp; Rnmp = i->fb;
    i->fb = (65536 * i->frr - Rnm) / 65535.0;
    if (i->refinement == 8) { i->I = i->fb; i->err = std::fabs((i->frr - Rnm) / 65535.0); i->refinement += 1; return count; }
    Rnm = Rnmp;
    i->I = (262144 * i->fb - Rnm) / 262143.0;
    i->err = std::fabs((i->fb - Rnm) / 262143.0);
    i->refinement += 1;
    return count;
}

template <class F> static int cadre_split(F &f, Interval9 *i0, Interval9 *i1, Interval9 *i2) {
    double xm = 0.5 * (i0->a + i0->b);
    i1->a = i0->a; i1->b = xm; i2->a = xm; i2->b = i0->b;
    i1->fa = i0->fa; i1->fl = i0->fll; i1->fm = i0->fl; i1->fr = i0->flr; i1->fb = i0->fm;
    i1->refinement = i0->refinement;
    i2->fa = i0->fm; i2->fl = i0->frl; i2->fm = i0->fr; i2->fr = i0->frr; i2->fb = i0->fb;
    i2->refinement = i0->refinement;
    int n = 0;
    n += cadre_process(f, i1);
    n += cadre_process(f, i2);
    return n;
}

struct CadreStats { long evals = 0; size_t max_heap = 0; };
static size_t g_max_heap_phi = 0, g_max_heap_theta = 0;
static long g_evals = 0;

// m9_adapt with the CADRE interval rules (quadratures.h:1969-2115)
template <class F> static double cadre_adapt(F &f, double xa, double xb, int Nmax, double atol, double rtol,
                                             CadreStats *st) {
    Mesh9 m;
    Interval9 i{};
    i.a = xa; i.b = xb; i.refinement = 0;
    i.fa = f(i.a);
    i.fb = f(i.b);
    i.fm = f(0.5 * (i.a + i.b));
    int n = 3;
    n += cadre_process(f, &i);
    m.insert(i);
    double I = i.I, err = i.err;
    int num_intervals = 1, num_iterations = 0, last_check = 0;
    while (n < Nmax && err > atol + rtol * std::fabs(I)) {
        m.extract(i);
        if (i.refinement >= 4) {
            double oldI = i.I, olderr = i.err;
            n += cadre_process(f, &i);
            m.insert(i);
            err += i.err - olderr;
            I += i.I - oldI;
        } else {
            Interval9 i1{}, i2{};
            n += cadre_split(f, &i, &i1, &i2);
            m.insert(i1);
            m.insert(i2);
            num_intervals++;
            err += i1.err + i2.err - i.err;
            I += i1.I + i2.I - i.I;
        }
        if (num_iterations == 2 * last_check) {
            err = m.total_err();
            I = m.total_I();
            last_check = num_intervals;
        }
        num_iterations++;
    }
    I = m.total_I();
    if (st) { st->evals += n; st->max_heap = std::max(st->max_heap, m.maxN); }
    return I;
}

static inline double obs_angle(double theta, double phi, double alpha, bool cj) {
    if (!cj) return std::sin(alpha) * (std::sin(theta) * std::sin(phi)) + std::cos(alpha) * std::cos(theta);
    return std::sin(alpha) * (std::sin(pi - theta) * std::sin(phi)) + std::cos(alpha) * std::cos(pi - theta);
}

struct LayerCtx {
    const pba_run_params *P;
    const double *T; int nt; int i_end_r;
    const double *jfs, *jrs;           // [nt][nf]
    const double *afs = nullptr, *ars = nullptr;   // absorption (use_ssa) or null
    std::vector<double> nu_fs, nu_rs;
    bool rs;
    double theta_c_l, theta_c_h;
    std::vector<double> mu;           // [nt]
    double t_obs, nu_obs;
    bool cj;
    int status = 0;
    long nevals = 0;
    const double *V(int v) const { return T + (size_t)v * nt; }
};

// EATS::find_jet_edge (eats.h:103-173)
static double find_jet_edge(const LayerCtx &L, double phi, double th_con_hi) {
    const double *a_thj = L.V(PV_theta);
    int N = L.i_end_r;
    double mu = obs_angle(th_con_hi, phi, L.P->theta_obs, L.cj);
    size_t ia = findIndex(mu, L.mu.data(), N);
    if (a_thj[ia] <= th_con_hi && th_con_hi <= a_thj[ia + 1]) return th_con_hi;
    double theta_a, theta_b;
    if (th_con_hi < a_thj[ia]) { theta_a = th_con_hi; theta_b = 0.5 * M_PI; }
    else { theta_a = 0.0; theta_b = th_con_hi; }
    int i = 0;
    while (theta_b - theta_a > 1.0e-5 && i < 100) {
        double tc = 0.5 * (theta_a + theta_b);
        mu = obs_angle(tc, phi, L.P->theta_obs, L.cj);
        ia = findIndex(mu, L.mu.data(), N);
        if (tc < a_thj[ia]) theta_a = tc; else theta_b = tc;
        i++;
    }
    return theta_a;
}

// computeIntensity, smooth optical-depth rule (microphysics.h:11-58)
static inline double intensity_smooth(double em, double dtau) {
    if (dtau == 0) return em;
    double abs_fac;
    if (dtau > 0.0) abs_fac = -std::expm1(-dtau) / dtau;
    else abs_fac = std::expm1(dtau) / dtau;
    return em * abs_fac;
}

// ElectronAndRadiation::fluxDens + computeFluxDensity (useNe, sameAsBW, smooth, no SSA)
static double flux_dens_shock(LayerCtx &L, const double *jall, const double *aall, const std::vector<double> &nu,
                              double freq, size_t ia, size_t ib, double Gamma, double mu, double r, double ne,
                              double n_prime, double theta, double ncells = 1.) {
    int nf = (int)nu.size();
    if (freq >= nu[nf - 1] || freq <= nu[0]) { L.status |= PBAGPU_S_NU_OUT_OF_GRID; return 0.; }
    size_t ia_nu = findIndex(freq, nu.data(), nf), ib_nu = ia_nu + 1;
    const double *R = L.V(PV_R);
    double x1 = nu[ia_nu], x2 = nu[ib_nu], y1 = R[ia], y2 = R[ib];
    double q11 = jall[ia * nf + ia_nu], q12 = jall[ib * nf + ia_nu];
    double q21 = jall[ia * nf + ib_nu], q22 = jall[ib * nf + ib_nu];
    double x2x1 = x2 - x1, y2y1 = y2 - y1, x2x = x2 - freq, y2y = y2 - r, yy1 = r - y1, xx1 = freq - x1;
    double em_prime = 1.0 / (x2x1 * y2y1) * (q11 * x2x * y2y + q21 * xx1 * y2y + q12 * x2x * yy1 + q22 * xx1 * yy1);
    double abs_prime = 0.;  // without synchrotron self-absorption the absorption table is identically 0
    if (aall) {
        double a11 = aall[ia * nf + ia_nu], a12 = aall[ib * nf + ia_nu];
        double a21 = aall[ia * nf + ib_nu], a22 = aall[ib * nf + ib_nu];
        abs_prime = 1.0 / (x2x1 * y2y1) * (a11 * x2x * y2y + a21 * xx1 * y2y + a12 * x2x * yy1 + a22 * xx1 * yy1);
    }
    double beta = BetaFromGamma(Gamma);
    double a = 1.0 - beta * mu;
    double delta_D = Gamma * a;
    double em_lab = em_prime / (delta_D * delta_D);
    double abs_lab = abs_prime * delta_D;
    double beta_shock = BetaFromGamma(Gamma);
    double m2 = ne * mp;
    double dr = m2 / (2. * pi * mp * r * r * (1. - std::cos(theta) / ncells) * Gamma * n_prime);
    double dtau = (mu == beta_shock) ? 1.0e100 : abs_lab * dr * (1. - mu * beta_shock) / (mu - beta_shock);
    double intensity = intensity_smooth(em_lab, dtau);
    if (intensity < 0 || !std::isfinite(intensity)) L.status |= PBAGPU_S_NONFINITE;
    double ashock = (1.0 - mu * beta_shock);
    double lum = intensity * ne / (2. * pi * (1. - std::cos(theta) / ncells)) / Gamma / ashock;
    if (lum < 0 || !std::isfinite(lum)) L.status |= PBAGPU_S_NONFINITE;
    return lum;
}

// EATS::integrand + fluxDensAdaptiveWithComov
static double integrand(LayerCtx &L, double cos_theta, double phi) {
    const double *R = L.V(PV_R), *tb = L.V(PV_tburst);
    int N = L.i_end_r;
    if (R[0] == 0.0 && R[N - 1] == 0.0) return 0.0;
    double a_theta = 2.0 * std::asin(std::sqrt(0.5 * (1.0 - cos_theta)));
    double mu = obs_angle(a_theta, phi, L.P->theta_obs, L.cj);
    if (mu < L.mu[0]) return 0.;
    if (mu > L.mu[N - 1]) return 0.;
    size_t ia = findIndex(mu, L.mu.data(), N), ib = ia + 1;
    double t_e = interpSegLin(ia, ib, mu, L.mu.data(), tb);
    if (mu > L.mu[N - 1] || L.mu[0] >= mu) t_e = -1;  // check_emission_time (eats.h:14-27)
    if ((t_e < 0.0) || !std::isfinite(t_e)) return 0.;
    double r = interpSegLog(ia, ib, t_e, tb, R);
    if ((r <= 0.0) || !std::isfinite(r)) return 0.;
    double Gamma = interpSegLog(ia, ib, r, R, L.V(PV_Gamma));
    double beta = interpSegLog(ia, ib, r, R, L.V(PV_beta));
    double ne = interpSegLog(ia, ib, r, R, L.V(PV_M2)) / mp;
    double n_prime = interpSegLog(ia, ib, r, R, L.V(PV_rho2)) / mp;
    double acc_fac = interpSegLog(ia, ib, r, R, L.V(PV_accel_frac));
    (void)acc_fac;
    double a = 1.0 - beta * mu;
    double delta_D = Gamma * a;
    double nuprime = (1.0 + L.P->z) * L.nu_obs * delta_D;
    double theta = interpSegLog(ia, ib, r, R, L.V(PV_theta));
    double flux = flux_dens_shock(L, L.jfs, L.afs, L.nu_fs, nuprime, ia, ib, Gamma, mu, r, ne, n_prime, theta);
    if (L.rs) {
        const double *th_rs = L.V(PV_thickness_rs);
        if (th_rs[ia] > 0 && th_rs[ib] > 0) {
            double ne_rs = interpSegLog(ia, ib, r, R, L.V(PV_M3)) / mp;
            double n_prime_rs = interpSegLog(ia, ib, r, R, L.V(PV_rho3)) / mp;
            flux += flux_dens_shock(L, L.jrs, L.ars, L.nu_rs, nuprime, ia, ib, Gamma, mu, r, ne_rs, n_prime_rs, theta);
        }
    }
    return flux;
}

static double integrate_theta_phi(LayerCtx &L, CadreStats *st) {
    const pba_run_params &P = *L.P;
    double Fcoeff = cgs2mJy / (4 * M_PI * P.d_l * P.d_l);
    double atol = 0.;  // evalLightCurve: out[it] * rtol / nlayers with out[] starting at 0 (eats.h:808-819)
    double theta_1c = L.theta_c_h, theta_0c = L.theta_c_l;
    auto phi_integrand = [&](double a_phi) -> double {
        double th_1 = find_jet_edge(L, a_phi, theta_1c);
        double frac = theta_0c / theta_1c;
        double th_0 = frac * th_1;
        double theta_0 = th_0, theta_1 = th_1;
        if (theta_0 > 0.5 * M_PI) theta_0 = 0.5 * M_PI;
        if (theta_1 > 0.5 * M_PI) theta_1 = 0.5 * M_PI;
        double s0 = std::sin(0.5 * theta_0), s1 = std::sin(0.5 * theta_1);
        double omc0 = 2.0 * s0 * s0, omc1 = 2.0 * s1 * s1;
        auto ct = [&](double aomct) -> double {
            L.nevals++;
            return integrand(L, 1 - aomct, a_phi);
        };
        CadreStats sti;
        double result = cadre_adapt(ct, omc0, omc1, P.nmax_theta, atol / (2. * Fcoeff * M_PI), P.rtol_theta, &sti);
        if (st) { st->evals += sti.evals; }
        if (sti.max_heap > g_max_heap_theta) g_max_heap_theta = sti.max_heap;
        g_evals += sti.evals;
        if (result != result || result < 0.0) return 0.;
        return result;
    };
    CadreStats sto;
    double r = 2 * Fcoeff * cadre_adapt(phi_integrand, 0.0, M_PI, P.nmax_phi, atol / (2. * Fcoeff * M_PI), P.rtol_phi, &sto);
    if (sto.max_heap > g_max_heap_phi) g_max_heap_phi = sto.max_heap;
    return r;
}

static void pars_mu(LayerCtx &L, double t_obs) {
    const double *tb = L.V(PV_tburst), *R = L.V(PV_R);
    L.mu.assign(L.nt, 0.);
    for (int i = 0; i < L.i_end_r; i++) L.mu[i] = (tb[i] - t_obs / (1.0 + L.P->z)) / R[i] * c;
}

} // namespace orc

extern "C" int orc_eats_lightcurve(const pba_run_params *params, int n_runs, const pba_eats_layer *layers, int n_layers,
                                   const double *table, int nt, const double *jall, const double *aall, int nfreq,
                                   int n_pairs,
                                   const double *tobs, const double *nuobs, double *out, int *status) {
    if (!params || !layers || !table || !jall || !out || !status || n_layers < 1) return PBAGPU_E_ARG;
#pragma omp parallel for schedule(dynamic)
    for (int l = 0; l < n_layers; l++) {
        const pba_eats_layer &Ly = layers[l];
        orc::LayerCtx L;
        L.P = &params[Ly.run];
        L.T = table + (size_t)Ly.bw * PBA_NV * nt;
        L.nt = nt;
        L.i_end_r = Ly.i_end_r;
        L.jfs = jall + (size_t)Ly.pop_fs * nt * nfreq;
        L.rs = Ly.pop_rs >= 0;
        L.jrs = L.rs ? jall + (size_t)Ly.pop_rs * nt * nfreq : nullptr;
        L.afs = (aall && L.P->fs.ssa) ? aall + (size_t)Ly.pop_fs * nt * nfreq : nullptr;
        L.ars = (aall && L.P->rs.ssa && L.rs) ? aall + (size_t)Ly.pop_rs * nt * nfreq : nullptr;
        L.nu_fs = orc::freq_grid(L.P->fs);
        L.nu_rs = orc::freq_grid(L.P->rs);
        L.theta_c_l = Ly.theta_c_l;
        L.theta_c_h = Ly.theta_c_h;
        orc::CadreStats st;
        for (int k = 0; k < n_pairs; k++) {
            double fl = 0.;
            L.t_obs = tobs[k];
            L.nu_obs = nuobs[k];
            const double *tb = L.V(PV_tburst);
            if (L.i_end_r > 0 && (L.t_obs < tb[0] || L.t_obs > tb[L.i_end_r - 1])) { L.status |= PBAGPU_S_NONFINITE; }
            if (L.i_end_r > 0) {
                L.cj = false;
                orc::pars_mu(L, L.t_obs);
                fl += orc::integrate_theta_phi(L, &st);
                if (L.P->counter_jet) {
                    L.cj = true;
                    orc::pars_mu(L, L.t_obs);
                    fl += orc::integrate_theta_phi(L, &st);
                }
            }
            out[(size_t)l * n_pairs + k] = 0. + fl;
        }
        status[l] = L.status;
    }
    return PBAGPU_OK;
}

namespace orc {

// findClosestIndex (src/utilitites/interpolators.h:37-68): the closest of the initial guess and
// the entries a bisection of the whole vector visits; an exact hit returns at once.
static int find_closest_index(double target, const std::vector<double> &v, int guess) {
    int size = (int)v.size(), low = 0, high = size - 1, closest = guess;
    if (closest < 0 || closest >= size) return -1;
    double best = std::abs(v[closest] - target);
    while (low <= high) {
        int mid = low + (high - low) / 2;
        double d = std::abs(v[mid] - target);
        if (d < best) { best = d; closest = mid; }
        if (v[mid] == target) return mid;
        else if (v[mid] < target) low = mid + 1;
        else high = mid - 1;
    }
    return closest;
}

struct PwCtx {
    const pba_run_params *P;
    const double *T; int nt; int i_end_r;
    const double *jfs;                 // [nt][nf] forward-shock emissivity
    const double *afs = nullptr;       // [nt][nf] forward-shock absorption (use_ssa) or null
    std::vector<double> nu_fs;
    bool rs;
    double ctheta0, ncells;
    int cil;
    std::vector<double> ttobs;         // [nt], entries past i_end_r stay at DBL_MAX
    int status = 0;
    const double *V(int v) const { return T + (size_t)v * nt; }
};

// fluxDens for the piece-wise method (microphysics.h:1185-1217): theta = the cell angle and the
// jet's total cell count, forward-shock emissivity for both shocks (blastwave_radiation.h:43-66).
static double flux_dens_shock_pw(PwCtx &L, double freq, size_t ia, size_t ib, double Gamma, double mu, double r,
                                 double ne, double n_prime) {
    LayerCtx S;
    S.P = L.P; S.T = L.T; S.nt = L.nt; S.i_end_r = L.i_end_r;
    double f = flux_dens_shock(S, L.jfs, L.afs, L.nu_fs, freq, ia, ib, Gamma, mu, r, ne, n_prime, L.ctheta0,
                               L.ncells);
    L.status |= S.status;
    return f;
}

// EATS::evalSkyMapPW, one jet (src/eats.h:1064-1213), with fluxDensPieceWiseWithComov
// (src/blastwave/blastwave_radiation.h:14-72) as the cell flux.
// Raw image arrays of one (t_obs, nu_obs): IMG::m_names order, 2 ncells entries each.
struct PwImage {
    double *out = nullptr;   // [PBA_IMG_NV][2 * ncells]
    size_t ncells = 0, offset = 0;
    bool remove_mu = false;
};

// imageXXs / imageYYs / imageXXsCJ / imageYYsCJ (src/image.h:379-392)
static void image_xy(bool cj, double theta, double phi, double alpha, double &x, double &y) {
    if (!cj) {
        x = -1.0 * std::cos(alpha) * (std::sin(theta) * std::sin(phi)) + std::sin(alpha) * std::cos(theta);
        y = +1. * std::sin(theta) * std::cos(phi);
    } else {
        x = -1.0 * std::cos(alpha) * std::sin(pi - theta) * std::sin(phi) + std::sin(alpha) * std::cos(pi - theta);
        y = +1. * std::sin(pi - theta) * std::cos(phi);
    }
}

static double sky_map_pw(PwCtx &L, double t_obs, double nu_obs, bool cj, const PwImage *im = nullptr) {
    const pba_run_params &P = *L.P;
    const double *R = L.V(PV_R), *tt = L.V(PV_tt), *G = L.V(PV_Gamma);
    const int N = L.i_end_r;
    if (R[0] == 0. && G[0] == 0.) return 0.;
    if (tt[0] == 0 && N < L.nt && tt[N] == 0) L.status |= PBAGPU_S_NONFINITE;
    const double dmax = std::numeric_limits<double>::max();
    std::vector<double> mu(L.cil, dmax), ta(L.cil, 0.), tb(L.cil, 0.), r(L.cil, 0.);
    std::vector<size_t> ia(L.cil, 0), ib(L.cil, 0);
    for (int i = 0; i < L.cil; i++) {
        double cphi = (double)i * 2.0 * M_PI / (double)L.cil;
        mu[i] = obs_angle(L.ctheta0, cphi, P.theta_obs, cj);
        for (int k = 0; k < N; k++) L.ttobs[k] = tt[k] + R[k] / c * (1.0 - mu[i]);
        if (t_obs < L.ttobs[0]) { L.status |= PBAGPU_S_NONFINITE; mu[i] = dmax; continue; }
        if (t_obs > L.ttobs[N - 1]) { mu[i] = dmax; continue; }
        int guess = i > 0 ? (int)ia[i - 1] : (int)(L.ttobs.size() / 2);
        ia[i] = (size_t)find_closest_index(t_obs, L.ttobs, guess);
        if (ia[i] >= (size_t)(N - 1)) { mu[i] = dmax; continue; }
        ib[i] = ia[i] + 1;
        ta[i] = L.ttobs[ia[i]];
        tb[i] = L.ttobs[ib[i]];
        r[i] = interpSegLog(ia[i], ib[i], t_obs, L.ttobs.data(), R);
        if (r[i] <= 0.0 || !std::isfinite(r[i])) { mu[i] = dmax; continue; }
    }
    double tot = 0.;
    for (int i = 0; i < L.cil; i++) {
        if (mu[i] == dmax) continue;
        size_t a = ia[i], b = ib[i];
        auto lg = [&](int v) {
            const double *Y = L.V(v);
            return Y[a] * std::pow(Y[b] / Y[a], std::log(t_obs / ta[i]) / std::log(tb[i] / ta[i]));
        };
        double Gamma = lg(PV_Gamma), beta = lg(PV_beta);
        double nprime = lg(PV_rho2) / mp, nprotons = lg(PV_M2) / mp;
        double nuprime = (1.0 + P.z) * nu_obs * (Gamma * (1.0 - beta * mu[i]));
        double f = flux_dens_shock_pw(L, nuprime, a, b, Gamma, mu[i], r[i], nprotons, nprime);
        if (L.rs) {
            const double *th_rs = L.V(PV_thickness_rs);
            if (th_rs[a] > 0 && th_rs[b] > 0) {
                double nprime_rs = lg(PV_rho3) / mp, ne_rs = lg(PV_M3) / mp;
                f += flux_dens_shock_pw(L, nuprime, a, b, Gamma, mu[i], r[i], ne_rs, nprime_rs);
            }
        }
        f *= (1.0 + P.z) / (2.0 * P.d_l * P.d_l);
        if (!std::isfinite(f) || f < 0) L.status |= PBAGPU_S_NONFINITE;
        if (im && im->out) {   // eats.h:1163-1184
            const double rad2mas = 206264806.247;
            // fluxDensPieceWiseWithComov's ctheta (interpSegLin with the cell's ta, tb, interpolators.h:107-113)
            const double *cth = L.V(PV_ctheta);
            double ctheta = cth[a] + (cth[b] - cth[a]) * (t_obs - ta[i]) / (tb[i] - ta[i]);
            double phi_cell = (double)i * 2.0 * M_PI / (double)L.cil, x, y;
            image_xy(cj, ctheta, phi_cell, P.theta_obs, x, y);
            const size_t vs = 2 * im->ncells, j = (cj ? im->ncells : 0) + im->offset + i;
            double *o = im->out;
            o[PBA_IMG_intensity * vs + j] = f / (r[i] * r[i]) * cgs2mJy;
            o[PBA_IMG_xrs * vs + j] = r[i] * x;
            o[PBA_IMG_yrs * vs + j] = r[i] * y;
            o[PBA_IMG_r * vs + j] = r[i];
            o[PBA_IMG_ctheta * vs + j] = L.ctheta0;
            o[PBA_IMG_cphi * vs + j] = phi_cell;
            o[PBA_IMG_mu * vs + j] = mu[i];
            if (!im->remove_mu) o[PBA_IMG_intensity * vs + j] /= std::abs(mu[i]);
            o[PBA_IMG_xrs * vs + j] *= (rad2mas / P.d_l);
            o[PBA_IMG_yrs * vs + j] *= (rad2mas / P.d_l);
            o[PBA_IMG_intensity * vs + j] *= (P.d_l * P.d_l / rad2mas / rad2mas);
        }
        tot += f;
    }
    return tot * cgs2mJy;
}

} // namespace orc

extern "C" int orc_eats_lightcurve_pw(const pba_run_params *params, int n_runs, const pba_eats_pw_layer *layers,
                                      int n_layers, const double *table, int nt, const double *jall,
                                      const double *aall, int nfreq,
                                      int n_pairs, const double *tobs, const double *nuobs, double *out, int *status) {
    (void)n_runs;
    if (!params || !layers || !table || !jall || !out || !status || n_layers < 1) return PBAGPU_E_ARG;
#pragma omp parallel for schedule(dynamic)
    for (int l = 0; l < n_layers; l++) {
        const pba_eats_pw_layer &Ly = layers[l];
        orc::PwCtx L;
        L.P = &params[Ly.run];
        L.T = table + (size_t)Ly.bw * PBA_NV * nt;
        L.nt = nt;
        L.i_end_r = Ly.i_end_r;
        L.jfs = jall + (size_t)Ly.pop_fs * nt * nfreq;
        L.afs = (aall && L.P->fs.ssa) ? aall + (size_t)Ly.pop_fs * nt * nfreq : nullptr;
        L.nu_fs = orc::freq_grid(L.P->fs);
        L.rs = Ly.rs != 0;
        L.ctheta0 = Ly.ctheta0;
        L.ncells = Ly.ncells;
        L.cil = Ly.cil;
        L.ttobs.assign(nt, std::numeric_limits<double>::max());
        const double *R = L.V(PV_R), *tb = L.V(PV_tburst);
        for (int k = 0; k < n_pairs; k++) {
            double fl = 0.;
            if (L.i_end_r > 0) {   // evalEjectaLightCurves skips blastwaves that never evolved
                if (R[0] > 0 && (tobs[k] < tb[0] || tobs[k] > tb[L.i_end_r - 1])) L.status |= PBAGPU_S_NONFINITE;
                double pj = orc::sky_map_pw(L, tobs[k], nuobs[k], false), cj = 0.;
                if (L.P->counter_jet) cj = orc::sky_map_pw(L, tobs[k], nuobs[k], true);
                fl = pj + cj;
            }
            out[(size_t)l * n_pairs + k] = fl;
        }
        status[l] = L.status;
    }
    return PBAGPU_OK;
}

/* Piece-wise raw sky maps of one run's layers (run 0): image [n_images][PBA_IMG_NV][2 ncells] and the
 * per-layer flux [n_layers][n_images]; Ejecta::computeEjectaSkyMapPW's call sequence
 * (src/ejecta/ejecta_base.h:471-495), the normalisation after it is host logic. */
extern "C" int orc_eats_skymap_pw(const pba_run_params *params, const pba_eats_pw_layer *layers, int n_layers,
                                  const double *table, int nt, const double *jall, const double *aall, int nfreq,
                                  int n_images,
                                  const double *tobs, const double *nuobs, int remove_mu, double *image,
                                  double *flux, int *status) {
    if (!params || !layers || !table || !jall || !image || !flux || !status || n_layers < 1) return PBAGPU_E_ARG;
    size_t ncells = 0;
    for (int l = 0; l < n_layers; l++) ncells += (size_t)layers[l].cil;
    std::fill(image, image + (size_t)n_images * PBA_IMG_NV * 2 * ncells, 0.);
    size_t offset = 0;
    for (int l = 0; l < n_layers; l++) {
        const pba_eats_pw_layer &Ly = layers[l];
        orc::PwCtx L;
        L.P = &params[0];
        L.T = table + (size_t)Ly.bw * PBA_NV * nt;
        L.nt = nt;
        L.i_end_r = Ly.i_end_r;
        L.jfs = jall + (size_t)Ly.pop_fs * nt * nfreq;
        L.afs = (aall && L.P->fs.ssa) ? aall + (size_t)Ly.pop_fs * nt * nfreq : nullptr;
        L.nu_fs = orc::freq_grid(L.P->fs);
        L.rs = Ly.rs != 0;
        L.ctheta0 = Ly.ctheta0;
        L.ncells = Ly.ncells;
        L.cil = Ly.cil;
        L.ttobs.assign(nt, std::numeric_limits<double>::max());
        const double *R = L.V(PV_R), *tb = L.V(PV_tburst);
        for (int k = 0; k < n_images; k++) {
            double fl = 0.;
            orc::PwImage im;
            im.out = image + (size_t)k * PBA_IMG_NV * 2 * ncells;
            im.ncells = ncells;
            im.offset = offset;
            im.remove_mu = remove_mu != 0;
            if (L.i_end_r > 0) {
                if (R[0] > 0 && (tobs[k] < tb[0] || tobs[k] > tb[L.i_end_r - 1])) L.status |= PBAGPU_S_NONFINITE;
                double pj = orc::sky_map_pw(L, tobs[k], nuobs[k], false, &im), cj = 0.;
                if (L.P->counter_jet) cj = orc::sky_map_pw(L, tobs[k], nuobs[k], true, &im);
                fl = pj + cj;
            }
            flux[(size_t)l * n_images + k] = fl;
        }
        status[l] = L.status;
        offset += (size_t)Ly.cil;
    }
    return PBAGPU_OK;
}

/* Largest interval stores and integrand count seen by the last calls (single-threaded use). */
extern "C" void orc_eats_stats(long *max_heap_phi, long *max_heap_theta, long *evals, int reset) {
    if (max_heap_phi) *max_heap_phi = (long)orc::g_max_heap_phi;
    if (max_heap_theta) *max_heap_theta = (long)orc::g_max_heap_theta;
    if (evals) *evals = orc::g_evals;
    if (reset) { orc::g_max_heap_phi = 0; orc::g_max_heap_theta = 0; orc::g_evals = 0; }
}

/* Stand-alone CADRE on built-in test integrands (closed-form checks in tests). */
extern "C" double orc_cadre_test(int which, double a, double b, int nmax, double atol, double rtol, long *nevals,
                                 long *max_heap) {
    orc::CadreStats st;
    auto f = [&](double x) -> double {
        switch (which) {
            case 0: return std::sin(x);
            case 1: return std::exp(-x * x);
            case 2: return 1.0 / (1e-3 + x * x);
            case 3: return std::sqrt(std::fabs(x));
            default: return x * x * x;
        }
    };
    double I = orc::cadre_adapt(f, a, b, nmax, atol, rtol, &st);
    if (nevals) *nevals = st.evals;
    if (max_heap) *max_heap = (long)st.max_heap;
    return I;
}
