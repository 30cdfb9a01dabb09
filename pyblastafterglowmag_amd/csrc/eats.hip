// Stage 3 — equal-time-arrival-surface (EATS) light curves on the GPU.
//
// One lane per (layer, t_obs, nu_obs) pair; lanes of a wavefront take consecutive
// pairs of one layer (same t_obs, neighbouring nu_obs), so they share the surface
// geometry and read the same rows of the layer's tables.  Each lane runs the
// reference's nested adaptive quadrature: CADRE over phi in [0, pi], and for every
// phi sample CADRE over 1 - cos(theta) between the layer edges found by
// find_jet_edge, for the primary jet and the counter-jet (src/eats.h:103-856).
//
// The CADRE interval store keeps the reference's binary max-heap order
// (src/utilitites/quadratures.h:655-770, 1613-2115) so that the sequence of
// integrand samples and the final summation order are those of the reference.
// Heap entries are (error, slot) pairs; the interval records live in a per-lane
// pool in private memory, so a heap swap moves 12 bytes instead of 112.
//
// Integrand: EATS::integrand + fluxDensAdaptiveWithComov + ElectronAndRadiation::
// fluxDens / computeFluxDensity (src/eats.h:175-458, src/blastwave/blastwave_radiation.h:75-141,
// src/microphysics/microphysics.h:799-867, 1185-1217); the observer-angle
// table mu[i] = (t_b[i] - t_obs/(1+z)) c / R[i] is evaluated on the fly during the
// index search instead of being stored.
#include "pba_device.h"
#include "timing.h"
#include "runtime.h"
#include <algorithm>
#include <vector>

namespace pbagpu {
namespace eats {

constexpr int WAVE = 64;

constexpr int HI = 256;  // inner (theta) interval store (reference cases need <= ~160)
constexpr int HO = 256;  // outer (phi) interval store (off-axis cases need <= ~190)

struct Iv {
    double a, b, I, err, fa, fll, fl, flr, fm, frl, fr, frr, fb;
    int refinement;
};

// Interval store of one lane: pool, heap keys and heap slots live in a lane-private,
// contiguous block of global memory (one Iv is 112 contiguous bytes, so reading or
// writing an interval touches one or two cache lines of that lane only).
template <int H>
struct Store {
    Iv *pool;
    double *herr;
    int *hslot;
    int N, used, overflow;
    static constexpr size_t bytes() { return (size_t)H * (sizeof(Iv) + sizeof(double) + sizeof(int)); }
    __device__ void bind(char *base) {
        pool = (Iv *)base;
        herr = (double *)(base + (size_t)H * sizeof(Iv));
        hslot = (int *)(base + (size_t)H * (sizeof(Iv) + sizeof(double)));
    }
    __device__ void reset() { N = 0; used = 0; overflow = 0; }
    __device__ int alloc() {
        if (used >= H) { overflow = 1; return H - 1; }
        return used++;
    }
    __device__ void push(int slot) {
        if (N >= H) { overflow = 1; return; }
        double e = pool[slot].err;
        int c = N++;
        herr[c] = e; hslot[c] = slot;
        while (c != 0) {
            int p = (c - 1) / 2;
            if (!(herr[p] < herr[c])) break;
            double te = herr[p]; herr[p] = herr[c]; herr[c] = te;
            int ts = hslot[p]; hslot[p] = hslot[c]; hslot[c] = ts;
            c = p;
        }
    }
    __device__ int pop() {
        int top = hslot[0];
        herr[0] = herr[N - 1]; hslot[0] = hslot[N - 1];
        N--;
        int p = 0, c1 = 1, c2 = 2;
        while (c1 < N) {
            int c = c1;
            double e = herr[c1];
            if (c2 < N && herr[c2] > e) { c = c2; e = herr[c2]; }
            if (e <= herr[p]) break;
            double te = herr[p]; herr[p] = herr[c]; herr[c] = te;
            int ts = hslot[p]; hslot[p] = hslot[c]; hslot[c] = ts;
            p = c; c1 = 2 * p + 1; c2 = c1 + 1;
        }
        return top;
    }
    __device__ double total_I() const { double s = 0.; for (int i = 0; i < N; i++) s += pool[hslot[i]].I; return s; }
    __device__ double total_err() const { double s = 0.; for (int i = 0; i < N; i++) s += pool[hslot[i]].err; return s; }
};

// Resumable CADRE: the reference's cadre_adapt / cadreProcessInterval / cadreSplitInterval
// (src/utilitites/quadratures.h:1613-2115, as used by src/eats.h:460-684) rewritten as a
// state machine that hands every integrand sample to its caller.  The caller evaluates the
// integrand for all lanes of a wavefront at one place in the code, so lanes whose adaptive
// paths differ still sample the integrand together; the bookkeeping between samples
// (heap, Romberg updates) is short.  Samples are requested in the reference's order and
// the arithmetic of every update is the reference's, so the result is the same number.
template <int H>
struct Cadre {
    Store<H> S;
    double atol, rtol;
    int Nmax;
    int n, num_intervals, num_iterations, last_check;
    double I, err;
    int pc;        // main program
    int ppc;       // interval-processing program
    int ret;       // main state to resume after processing
    int cur;       // slot being processed
    int sl, s2;    // slots of the interval popped / created by a split
    int cnt;       // samples taken while processing `cur`
    double oldI, olderr;
    double fl, fr, fll, flr, frl, frr, sum, hh;
    int nint, nidx;
    double x;      // requested abscissa
    double result;

    enum { M_FA, M_FB, M_FM, M_AFTER_INIT, M_LOOP, M_AFTER_REFINE, M_AFTER_SPLIT1, M_AFTER_SPLIT2, M_DONE };
    enum { P_START, P_FL, P_FR, P_R0, P_FLL, P_FLR, P_FRL, P_FRR, P_SUM, P_DONE };

    // Begin the integral over [xa, xb]; always requests a first sample (x).
    __device__ void start(double xa, double xb, int nmax, double at, double rt) {
        S.reset();
        Nmax = nmax; atol = at; rtol = rt;
        int s0 = S.alloc();
        Iv &i0 = S.pool[s0];
        i0.a = xa; i0.b = xb; i0.I = 0; i0.err = 0;
        i0.fll = 0; i0.flr = 0; i0.frl = 0; i0.frr = 0; i0.fl = 0; i0.fr = 0; i0.refinement = 0;
        cur = s0;
        pc = M_FA;
        x = xa;
    }

    // Interval processing (cadreProcessInterval).  Returns true when it needs a sample (x set);
    // false when `cur` is processed (cnt holds the number of samples it took).
    __device__ bool process(double v, bool have_v) {
        const double scaleTol = 0.2;
        Iv &i = S.pool[cur];
        for (;;) {
            switch (ppc) {
            case P_START:
                cnt = 0;
                if (i.refinement < 4) {
                    if (i.refinement == 0) { ppc = P_FL; x = 0.75 * i.a + 0.25 * i.b; return true; }
                    fl = i.fl; fr = i.fr;
                    i.refinement = 0;
                    ppc = P_R0;
                } else {
                    nint = 1 << i.refinement;
                    hh = (i.b - i.a) / nint;
                    sum = 0;
                    nidx = 1;
                    if (nidx < nint) { ppc = P_SUM; x = i.a + hh * nidx; return true; }
                    ppc = P_DONE;   // not reached for refinement >= 4
                }
                break;
            case P_FL:
                fl = v; cnt++;
                ppc = P_FR; x = 0.25 * i.a + 0.75 * i.b; return true;
            case P_FR:
                fr = v; cnt++;
                i.fl = fl; i.fr = fr;
                ppc = P_R0;
                break;
            case P_R0: {
                double fa = i.fa, fm = i.fm, fb = i.fb;
                double h4 = 0.25 * (i.b - i.a);
                double T0 = 2 * h4 * (fa + fb);
                double T1 = 0.5 * T0 + 2 * h4 * fm;
                double T2 = 0.5 * T1 + h4 * (fl + fr);
                double R0 = (T1 - T0) / (T2 - T1);
                if (R0 > 4 - scaleTol && R0 < 4 + scaleTol) {
                    ppc = P_FLL; x = 0.875 * i.a + 0.125 * i.b; return true;
                } else if (R0 >= 4 + scaleTol || R0 < 2 || R0 != R0) {
                    double e10 = fabs(T1 - T0), e21 = fabs(T2 - T1);
                    i.err = e10 > e21 ? e10 : e21;
                    i.I = T2;
                } else {
                    double e = -(T2 - T1) * (T2 - T1) / (T2 - 2 * T1 + T0);
                    i.err = fabs(e);
                    i.I = (fabs(T2 + e) < 1.0e-14 * fabs(T2)) ? 1.0e-14 * T2 : T2 + e;
                }
                ppc = P_DONE;
                break;
            }
            case P_FLL: fll = v; cnt++; ppc = P_FLR; x = 0.625 * i.a + 0.375 * i.b; return true;
            case P_FLR: flr = v; cnt++; ppc = P_FRL; x = 0.375 * i.a + 0.625 * i.b; return true;
            case P_FRL: frl = v; cnt++; ppc = P_FRR; x = 0.125 * i.a + 0.875 * i.b; return true;
            case P_FRR: {
                frr = v; cnt++;
                double fa = i.fa, fm = i.fm, fb = i.fb;
                double h4 = 0.25 * (i.b - i.a);
                double T0 = 2 * h4 * (fa + fb);
                double T1 = 0.5 * T0 + 2 * h4 * fm;
                double T2 = 0.5 * T1 + h4 * (fl + fr);
                i.fll = fll; i.flr = flr; i.frl = frl; i.frr = frr;
                i.refinement = 1;
                double T3 = 0.5 * T2 + 0.5 * h4 * (fll + flr + frl + frr);
                double R1 = (T2 - T1) / (T3 - T2);
                if (R1 > 4 - scaleTol && R1 < 4 + scaleTol) {
                    double R11 = (4 * T1 - T0) / 3.0;
                    double R21 = (4 * T2 - T1) / 3.0;
                    double R22 = (16 * R21 - R11) / 15.0;
                    i.fa = T3;
                    i.fll = (4 * T3 - T2) / 3.0;
                    i.fl = (16 * i.fll - R21) / 15.0;
                    i.flr = (64 * i.fl - R22) / 63.0;
                    i.I = i.flr;
                    i.err = fabs((i.fl - R22) / 63.0);
                    i.refinement = 4;
                } else {
                    double errt = (T3 - T2) / 3.0;
                    i.err = fabs(errt);
                    i.I = T3 + errt;
                }
                ppc = P_DONE;
                break;
            }
            case P_SUM:
                sum += v; cnt++;
                nidx += 2;
                if (nidx < nint) { x = i.a + hh * nidx; return true; }
                {
                    double T = 0.5 * i.fa + hh * sum;
                    double Rnm = i.fa, Rnmp = i.fll;
                    i.fa = T;
                    i.fll = (4.0 * i.fa - Rnm) / 3.0;
                    if (i.refinement == 1) { i.I = i.fll; i.err = fabs((i.fa - Rnm) / 3.0); i.refinement++; ppc = P_DONE; break; }
                    Rnm = Rnmp; Rnmp = i.fl;
                    i.fl = (16 * i.fll - Rnm) / 15.0;
                    if (i.refinement == 2) { i.I = i.fl; i.err = fabs((i.fll - Rnm) / 15.0); i.refinement++; ppc = P_DONE; break; }
                    Rnm = Rnmp; Rnmp = i.flr;
                    i.flr = (64 * i.fl - Rnm) / 63.0;
                    if (i.refinement == 3) { i.I = i.flr; i.err = fabs((i.fl - Rnm) / 63.0); i.refinement++; ppc = P_DONE; break; }
                    Rnm = Rnmp; Rnmp = i.fm;
                    i.fm = (256 * i.flr - Rnm) / 255.0;
                    if (i.refinement == 4) { i.I = i.fm; i.err = fabs((i.flr - Rnm) / 255.0); i.refinement++; ppc = P_DONE; break; }
                    Rnm = Rnmp; Rnmp = i.frl;
                    i.frl = (1024 * i.fm - Rnm) / 1023.0;
                    if (i.refinement == 5) { i.I = i.frl; i.err = fabs((i.fm - Rnm) / 1023.0); i.refinement++; ppc = P_DONE; break; }
                    Rnm = Rnmp; Rnmp = i.fr;
                    i.fr = (4096 * i.frl - Rnm) / 4095.0;
                    if (i.refinement == 6) { i.I = i.fr; i.err = fabs((i.frl - Rnm) / 4095.0); i.refinement++; ppc = P_DONE; break; }
                    Rnm = Rnmp; Rnmp = i.frr;
                    i.frr = (16384 * i.fr - Rnm) / 16383.0;
                    if (i.refinement == 7) { i.I = i.frr; i.err = fabs((i.fr - Rnm) / 16383.0); i.refinement++; ppc = P_DONE; break; }
                    Rnm = Rnmp; Rnmp = i.fb;
                    i.fb = (65536 * i.frr - Rnm) / 65535.0;
                    if (i.refinement == 8) { i.I = i.fb; i.err = fabs((i.frr - Rnm) / 65535.0); i.refinement++; ppc = P_DONE; break; }
                    Rnm = Rnmp;
                    i.I = (262144 * i.fb - Rnm) / 262143.0;
                    i.err = fabs((i.fb - Rnm) / 262143.0);
                    i.refinement++;
                    ppc = P_DONE;
                }
                break;
            default:  // P_DONE
                return false;
            }
        }
    }

    // Feed the sample requested last; returns true with the next request in x, or false
    // when the integral is complete (result).
    __device__ bool feed(double v) {
        bool have = true;
        for (;;) {
            switch (pc) {
            case M_FA: S.pool[cur].fa = v; pc = M_FB; x = S.pool[cur].b; return true;
            case M_FB: S.pool[cur].fb = v; pc = M_FM; x = 0.5 * (S.pool[cur].a + S.pool[cur].b); return true;
            case M_FM:
                S.pool[cur].fm = v;
                n = 3;
                ppc = P_START; ret = M_AFTER_INIT; pc = -1; have = false;
                break;
            case -1:  // processing `cur`
                if (process(v, have)) return true;
                n += cnt;
                pc = ret;
                break;
            case M_AFTER_INIT: {
                S.push(cur);
                I = S.pool[cur].I; err = S.pool[cur].err;
                num_intervals = 1; num_iterations = 0; last_check = 0;
                pc = M_LOOP;
                break;
            }
            case M_LOOP: {
                if (!(n < Nmax && err > atol + rtol * fabs(I) && !S.overflow)) {
                    result = S.total_I();
                    pc = M_DONE;
                    return false;
                }
                sl = S.pop();
                Iv &iv = S.pool[sl];
                if (iv.refinement >= 4) {
                    oldI = iv.I; olderr = iv.err;
                    cur = sl; ppc = P_START; ret = M_AFTER_REFINE; pc = -1; have = false;
                } else {
                    Iv p = iv;
                    s2 = S.alloc();
                    Iv &i1 = S.pool[sl];
                    Iv &i2 = S.pool[s2];
                    double xm = 0.5 * (p.a + p.b);
                    i1.a = p.a; i1.b = xm; i1.fa = p.fa; i1.fl = p.fll; i1.fm = p.fl; i1.fr = p.flr; i1.fb = p.fm;
                    i1.refinement = p.refinement;
                    i2.a = xm; i2.b = p.b; i2.fa = p.fm; i2.fl = p.frl; i2.fm = p.fr; i2.fr = p.frr; i2.fb = p.fb;
                    i2.refinement = p.refinement;
                    oldI = p.I; olderr = p.err;
                    cur = sl; ppc = P_START; ret = M_AFTER_SPLIT1; pc = -1; have = false;
                }
                break;
            }
            case M_AFTER_REFINE: {
                S.push(sl);
                err += S.pool[sl].err - olderr;
                I += S.pool[sl].I - oldI;
                pc = M_LOOP + 100;   // iteration end
                break;
            }
            case M_AFTER_SPLIT1:
                cur = s2; ppc = P_START; ret = M_AFTER_SPLIT2; pc = -1; have = false;
                break;
            case M_AFTER_SPLIT2: {
                S.push(sl);
                S.push(s2);
                num_intervals++;
                err += S.pool[sl].err + S.pool[s2].err - olderr;
                I += S.pool[sl].I + S.pool[s2].I - oldI;
                pc = M_LOOP + 100;
                break;
            }
            case M_LOOP + 100:  // end of one iteration of the adaptive loop
                if (num_iterations == 2 * last_check) {
                    err = S.total_err();
                    I = S.total_I();
                    last_check = num_intervals;
                }
                num_iterations++;
                pc = M_LOOP;
                break;
            default:
                return false;
            }
        }
    }
};

struct LayerArgs {
    const double *T;    // table [PBA_NV][nt]
    const double *jfs;  // [nt][nf]
    const double *jrs;  // or null
    const double *afs;  // absorption [nt][nf] (use_ssa) or null
    const double *ars;  // or null
    const double *nufs, *nurs;
    double theta_c_l, theta_c_h;
    double z, theta_obs, d_l, rtol_theta, rtol_phi;
    int nmax_theta, nmax_phi, counter_jet, i_end_r;
    int nt, nf;
};

// Table columns the integrand reads, staged in LDS once per workgroup (one layer per
// workgroup): every index search and interpolation then reads LDS instead of L2.
enum Slot { S_tb, S_R, S_theta, S_Gamma, S_beta, S_M2, S_rho2, S_thrs, S_M3, S_rho3, NSLOT };
__device__ __forceinline__ constexpr int slot_var(int s) {
    return s == S_tb ? PV_tburst : s == S_R ? PV_R : s == S_theta ? PV_theta : s == S_Gamma ? PV_Gamma
         : s == S_beta ? PV_beta : s == S_M2 ? PV_M2 : s == S_rho2 ? PV_rho2 : s == S_thrs ? PV_thickness_rs
         : s == S_M3 ? PV_M3 : PV_rho3;
}

extern __shared__ double lds_tab[];

// Per-lane context of one (layer, t_obs, nu_obs) pair; the layer's scalars are copied in so
// that the integrand reads registers, not the layer descriptor.
struct Ctx {
    const double *jfs, *jrs;
    const double *afs, *ars;   // absorption tables or null
    double z;
    int nt, nf, i_end_r;
    double tz;          // t_obs / (1 + z)
    double nu_obs;
    double sa, ca;      // sin, cos of the observer angle
    double mu0, muN;    // mu at the first and last usable step for this t_obs
    double sphi;        // sin(phi) of the current phi sample
    bool cj;
    int status;
    __device__ double V(int s, int i) const { return lds_tab[s * nt + i]; }
    __device__ double mu_at(int i) const { return (V(S_tb, i) - tz) / V(S_R, i) * kC; }
    // ImageExtend obsAngle / obsAngleCJ (src/image.h:361-378) with sin(phi) of the current sample
    __device__ double obs_angle(double theta) const {
        if (!cj) return sa * (sin(theta) * sphi) + ca * cos(theta);
        return sa * (sin(kPi - theta) * sphi) + ca * cos(kPi - theta);
    }
    // the same from sin and cos of theta (cj: sin(pi - theta) = sin theta, cos(pi - theta) = -cos theta)
    __device__ double obs_angle_sc(double st, double ct) const {
        if (!cj) return sa * (st * sphi) + ca * ct;
        return sa * (st * sphi) - ca * ct;
    }
    __device__ int mu_index(double mu) const {
        auto arr = [&](int i) { return mu_at(i); };
        return find_index(mu, arr, i_end_r);
    }
};

// find_jet_edge (src/eats.h:103-173) as a resumable bisection: one probe per call of step(),
// so that lanes searching an edge do not hold up lanes sampling the integrand.
struct EdgeSearch {
    double ta, tb;
    int i;
    // first probe at the layer's upper edge; returns true if that already is the answer
    __device__ bool start(const Ctx &C, double th_con_hi) {
        double mu = C.obs_angle(th_con_hi);
        int ia = C.mu_index(mu);
        double a0 = C.V(S_theta, ia), a1 = C.V(S_theta, ia + 1);
        if (a0 <= th_con_hi && th_con_hi <= a1) { ta = th_con_hi; return true; }
        if (th_con_hi < a0) { ta = th_con_hi; tb = 0.5 * M_PI; }
        else { ta = 0.0; tb = th_con_hi; }
        i = 0;
        return !(tb - ta > 1.0e-5 && i < 100);
    }
    // one bisection probe; returns true when done (edge in ta)
    __device__ bool step(const Ctx &C) {
        double tc = 0.5 * (ta + tb);
        double mu = C.obs_angle(tc);
        int ia = C.mu_index(mu);
        if (tc < C.V(S_theta, ia)) ta = tc; else tb = tc;
        i++;
        return !(tb - ta > 1.0e-5 && i < 100);
    }
};

// nu_off: offset of the shock's frequency grid in the staged LDS block
// SSA: absorption tables may be present (a separate instantiation, so that runs without
// synchrotron self-absorption do not carry the absorption code's registers).
template <bool SSA>
__device__ __forceinline__ double shock_flux(Ctx &C, const double *jall, const double *aall, int nu_off, double freq,
                                             int ia, int ib, double Gamma, double mu, double r, double ne,
                                             double n_prime, double theta) {
    const int nf = C.nf;
    auto nu = [&](int i) { return lds_tab[nu_off + i]; };
    if (freq >= nu(nf - 1) || freq <= nu(0)) { C.status |= PBAGPU_S_NU_OUT_OF_GRID; return 0.; }
    const int ia_nu = find_index(freq, nu, nf), ib_nu = ia_nu + 1;
    const double x1 = nu(ia_nu), x2 = nu(ib_nu), y1 = C.V(S_R, ia), y2 = C.V(S_R, ib);
    const double em_prime = table_bilinear(jall, nf, ia, ib, ia_nu, ib_nu, x1, x2, y1, y2, freq, r);
    const double beta = beta_from_gamma(Gamma);
    const double delta_D = Gamma * (1.0 - beta * mu);
    const double omc = (1. - cos(theta) / 1.);
    // no absorption table: the optical depth is 0 unless mu equals the shock velocity
    const double dtau = (mu == beta) ? 1.0e100
                      : (SSA && aall) ? absorption_depth(aall, nf, ia, ib, ia_nu, ib_nu, x1, x2, y1, y2, freq, r, Gamma,
                                                         beta, delta_D, mu, ne, n_prime, omc)
                                      : 0.;
    bool bad = false;
    const double lum = shock_luminosity(em_prime, dtau, Gamma, beta, delta_D, mu, ne, omc, bad);
    if (bad) C.status |= PBAGPU_S_NONFINITE;
    return lum;
}

// Out of line on purpose: inlined at its one call site, the kernel needs more registers than
// the two waves per SIMD it runs at and spills almost twice as much.
template <bool SSA>
__device__ __noinline__ double integrand(Ctx &C, double cos_theta) {
    const int N = C.i_end_r;
    if (C.V(S_R, 0) == 0.0 && C.V(S_R, N - 1) == 0.0) return 0.0;
    // theta = 2 asin(s), s = sqrt((1 - cos_theta)/2): sin(theta) = 2 s sqrt(1 - s^2), cos(theta) = 1 - 2 s^2
    // (the same angle the reference builds with asin and then takes sin/cos of)
    const double sh = sqrt(0.5 * (1.0 - cos_theta));
    const double st = 2.0 * sh * sqrt(1.0 - sh * sh), ct = 1.0 - 2.0 * sh * sh;
    double mu = C.obs_angle_sc(st, ct);
    const double mu0 = C.mu0, muN = C.muN;
    if (mu < mu0 || mu > muN) return 0.;
    if (mu0 >= mu) return 0.;
    int ia = C.mu_index(mu), ib = ia + 1;
    double mua = C.mu_at(ia), mub = C.mu_at(ib);
    double tba = C.V(S_tb, ia), tbb = C.V(S_tb, ib);
    double t_e = interp_lin(mu, mua, mub, tba, tbb);
    if (t_e < 0.0 || !isfinite(t_e)) return 0.;
    double Ra = C.V(S_R, ia), Rb = C.V(S_R, ib);
    // interpSegLog: r = Ra (Rb/Ra)^w, w = ln(t_e/t_a) / ln(t_b/t_a); every state variable below is
    // interpolated in R with the exponent ln(r/Ra)/ln(Rb/Ra), which is w itself
    const double w = log(t_e / tba) / log(tbb / tba);
    double r = Ra * pow(Rb / Ra, w);
    if (r <= 0.0 || !isfinite(r)) return 0.;
    // shared log-interpolation weight in R for all state variables
    const double wr = w;
    auto li = [&](int v) { double ya = C.V(v, ia), yb = C.V(v, ib); return ya * pow(yb / ya, wr); };
    double Gamma = li(S_Gamma);
    double beta = li(S_beta);
    double ne = li(S_M2) / kMp;
    double n_prime = li(S_rho2) / kMp;
    double nuprime = (1.0 + C.z) * C.nu_obs * (Gamma * (1.0 - beta * mu));
    double theta = li(S_theta);
    double flux = shock_flux<SSA>(C, C.jfs, C.afs, NSLOT * C.nt, nuprime, ia, ib, Gamma, mu, r, ne, n_prime, theta);
    if (C.jrs && C.V(S_thrs, ia) > 0 && C.V(S_thrs, ib) > 0) {
        double ne_rs = li(S_M3) / kMp;
        double np_rs = li(S_rho3) / kMp;
        flux += shock_flux<SSA>(C, C.jrs, C.ars, NSLOT * C.nt + C.nf, nuprime, ia, ib, Gamma, mu, r, ne_rs, np_rs, theta);
    }
    return flux;
}

constexpr int EATS_WG = 256;
constexpr size_t LANE_WS = ((Store<HO>::bytes() + Store<HI>::bytes()) + 255) / 256 * 256;

extern __shared__ double lds_tab[];   // [NSLOT][nt] table columns, then the two frequency grids

// Persistent kernel: each workgroup stages one layer's table columns in LDS and its lanes take
// that layer's (t_obs, nu_obs) pairs one at a time from the layer's global counter; a
// workgroup changes layer only when the layer has no pairs left (see the loop below).
// Per lane, the nested adaptive quadrature of one pair (evalFluxDensA -> integrate_theta_phi,
// src/eats.h:587-684, 824-856: CADRE over phi in [0, pi], CADRE over 1 - cos(theta) between
// the layer edges of find_jet_edge for every phi sample, primary jet then counter-jet) runs as
// two resumable CADRE machines; every pass of the lane loop takes exactly one integrand
// sample, at the single call site below.
template <bool SSA>
__global__ void __launch_bounds__(EATS_WG, 2) eats_kernel(const LayerArgs *__restrict__ layers, int n_layers, int n_pairs,
                                                       int *__restrict__ pair_next,
                                                       const double *__restrict__ tobs,
                                                       const double *__restrict__ nuobs, double *__restrict__ out,
                                                       int *__restrict__ status, char *__restrict__ ws,
                                                       unsigned long long *__restrict__ n_samples) {
    __shared__ int s_layer;
    char *lane_ws = ws + ((size_t)blockIdx.x * EATS_WG + threadIdx.x) * LANE_WS;
    Cadre<HO> Mo;
    Cadre<HI> Mi;
    Mo.S.bind(lane_ws);
    Mi.S.bind(lane_ws + Store<HO>::bytes());
    int cur_layer = -1;
    // Workgroups start spread over the layers; each keeps its layer until that layer's pairs are
    // all taken, then joins the next layer (cyclically) that still has pairs.  Lanes take pairs
    // one at a time from the layer's counter, so they wait for each other only when their
    // workgroup changes layer (the staged table is shared).
    int probe = (int)(((long)blockIdx.x * n_layers) / gridDim.x);
    for (;;) {
        __syncthreads();  // every lane is done with the previous layer and its staged table
        if (threadIdx.x < WAVE) {
            int found = -1;
            for (int base = 0; base < n_layers && found < 0; base += WAVE) {
                const int t = base + threadIdx.x;
                int lt = probe + t;
                if (lt >= n_layers) lt -= n_layers;
                const bool open = t < n_layers && __atomic_load_n(&pair_next[lt], __ATOMIC_RELAXED) < n_pairs;
                const unsigned long long m = __ballot(open);
                if (m) {
                    int f = probe + base + __ffsll((long long)m) - 1;
                    found = f >= n_layers ? f - n_layers : f;
                }
            }
            if (threadIdx.x == 0) s_layer = found;
        }
        __syncthreads();
        const int l = s_layer;
        if (l < 0) break;
        probe = l;
        const LayerArgs &L = layers[l];
        const int nt = L.nt, nf = L.nf;
        if (l != cur_layer) {
            for (int sl = 0; sl < NSLOT; sl++) {
                const double *col = L.T + (size_t)slot_var(sl) * nt;
                for (int i = threadIdx.x; i < nt; i += blockDim.x) lds_tab[sl * nt + i] = col[i];
            }
            for (int i = threadIdx.x; i < nf; i += blockDim.x) {
                lds_tab[NSLOT * nt + i] = L.nufs[i];
                lds_tab[NSLOT * nt + nf + i] = L.nurs[i];
            }
            cur_layer = l;
        }
        __syncthreads();
        const double Fcoeff = kCgs2mJy / (4 * M_PI * L.d_l * L.d_l);
        const double th1c = L.theta_c_h, th0c = L.theta_c_l;
        Ctx C;
        C.jfs = L.jfs; C.jrs = L.jrs; C.afs = L.afs; C.ars = L.ars; C.z = L.z;
        C.nt = nt; C.nf = nf; C.i_end_r = L.i_end_r;
        C.sa = sin(L.theta_obs);
        C.ca = cos(L.theta_obs);
        C.status = 0;
        C.sphi = 0.;
        EdgeSearch E;
        enum { NEED_PAIR, NEED_INNER, EDGE, EVAL, DONE };
        int mode = NEED_PAIR, k = 0, st_acc = 0;
        unsigned long long samples = 0;
        double flux = 0.;
        for (;;) {
            // bookkeeping until this lane needs an edge probe or an integrand sample
            while (mode == NEED_PAIR || mode == NEED_INNER) {
                if (mode == NEED_PAIR) {
                    k = atomicAdd(&pair_next[l], 1);
                    if (k >= n_pairs) { mode = DONE; break; }
                    C.status = 0;
                    const double t_obs = tobs[k];
                    C.tz = t_obs / (1.0 + L.z);
                    C.nu_obs = nuobs[k];
                    C.cj = false;
                    flux = 0.;
                    if (L.i_end_r == 1) C.status |= PBAGPU_S_NONFINITE;  // one usable step: no surface
                    if (L.i_end_r > 1) {
                        if (t_obs < C.V(S_tb, 0) || t_obs > C.V(S_tb, L.i_end_r - 1)) C.status |= PBAGPU_S_NONFINITE;
                        C.mu0 = C.mu_at(0);
                        C.muN = C.mu_at(L.i_end_r - 1);
                        Mo.start(0.0, M_PI, L.nmax_phi, 0., L.rtol_phi);
                        mode = NEED_INNER;
                    } else {
                        out[(size_t)l * n_pairs + k] = 0. + flux;
                        st_acc |= C.status;
                    }
                } else {
                    // a new phi sample (phi_integrand, src/eats.h:460-585): find the layer edge first
                    C.sphi = sin(Mo.x);
                    mode = E.start(C, th1c) ? EVAL + 100 : EDGE;
                    break;
                }
            }
            if (mode == DONE) break;
            if (mode == EDGE) {
                if (E.step(C)) mode = EVAL + 100;
            }
            if (mode == EVAL + 100) {
                // edge known: start the inner quadrature over 1 - cos(theta)
                double th_1 = E.ta;
                double th_0 = (th0c / th1c) * th_1;
                double t0 = th_0 > 0.5 * M_PI ? 0.5 * M_PI : th_0;
                double t1 = th_1 > 0.5 * M_PI ? 0.5 * M_PI : th_1;
                double s0 = sin(0.5 * t0), s1 = sin(0.5 * t1);
                Mi.start(2.0 * s0 * s0, 2.0 * s1 * s1, L.nmax_theta, 0., L.rtol_theta);
                mode = EVAL;
                continue;
            }
            if (mode != EVAL) continue;
            // the one integrand call of the loop
            const double v = integrand<SSA>(C, 1 - Mi.x);
            samples++;
            if (Mi.feed(v)) continue;
            double res = Mi.result;
            if (Mi.S.overflow) C.status |= PBAGPU_S_HEAP_FULL;
            if (res != res || res < 0.0) res = 0.;
            if (Mo.feed(res)) { mode = NEED_INNER; continue; }
            if (Mo.S.overflow) C.status |= PBAGPU_S_HEAP_FULL;
            flux += 2 * Fcoeff * Mo.result;
            if (L.counter_jet && !C.cj) {
                C.cj = true;
                Mo.start(0.0, M_PI, L.nmax_phi, 0., L.rtol_phi);
                mode = NEED_INNER;
            } else {
                out[(size_t)l * n_pairs + k] = 0. + flux;
                st_acc |= C.status;
                mode = NEED_PAIR;
            }
        }
        if (st_acc) atomicOr(&status[l], st_acc);
        if (n_samples) atomicAdd(n_samples, samples);
    }
}

}  // namespace eats
}  // namespace pbagpu

extern "C" int pbagpu_eats_lightcurve(const pba_run_params *params, int n_runs, const pba_eats_layer *layers,
                                      int n_layers, const double *d_table, int nt, const double *d_jall,
                                      const double *d_aall, int nfreq, int n_pairs, const double *d_tobs,
                                      const double *d_nuobs, double *d_out, int *d_status, void *stream) {
    using namespace pbagpu::eats;
    if (!params || !layers || !d_table || !d_jall || !d_out || !d_status || n_runs < 1 || n_layers < 1 ||
        n_pairs < 1 || nt < 2 || nfreq < 2)
        return PBAGPU_E_ARG;
    if (sizeof(double) * ((size_t)NSLOT * nt + 2 * (size_t)nfreq) > 160 * 1024) {
        pbagpu::set_error_msg("eats: time grid too long for the LDS-staged table (nt <= 2000)");
        return PBAGPU_E_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    // frequency grids of every (run, shock), built on the host like MakeLogspaceVec (utils.h:91-105)
    std::vector<double> hnu((size_t)n_runs * 2 * nfreq);
    for (int r = 0; r < n_runs; r++)
        for (int sh = 0; sh < 2; sh++) {
            const pba_shock_params &sp = sh ? params[r].rs : params[r].fs;
            if (sp.nfreq != nfreq) return PBAGPU_E_ARG;
            double cur = std::pow(10.0, std::log10(sp.freq1));
            double ratio = std::pow(10.0, (std::log10(sp.freq2) - std::log10(sp.freq1)) / sp.nfreq);
            for (int i = 0; i < nfreq; i++) { hnu[((size_t)r * 2 + sh) * nfreq + i] = cur; cur *= ratio; }
        }
    for (int l = 0; l < n_layers; l++) {
        const pba_eats_layer &Y = layers[l];
        if (Y.run < 0 || Y.run >= n_runs || Y.i_end_r > nt || Y.i_end_r < 0 || Y.pop_fs < 0 || Y.pop_rs < -1 ||
            Y.bw < 0) {
            pbagpu::set_error_msg("eats_lightcurve: bad layer descriptor (run, bw, pop_fs >= 0, pop_rs >= -1, "
                                  "0 <= i_end_r <= nt)");
            return PBAGPU_E_ARG;
        }
    }
    auto *dnu = (double *)pbagpu::workspace(pbagpu::WS_EATS_NU, sizeof(double) * hnu.size(), s);
    auto *dL = (LayerArgs *)pbagpu::workspace(pbagpu::WS_EATS_LAYERS, sizeof(LayerArgs) * n_layers, s);
    if (!dnu || !dL) return PBAGPU_E_HIP;
    std::vector<LayerArgs> hl(n_layers);
    for (int l = 0; l < n_layers; l++) {
        const pba_eats_layer &Y = layers[l];
        const pba_run_params &P = params[Y.run];
        LayerArgs &A = hl[l];
        A.T = d_table + (size_t)Y.bw * PBA_NV * nt;
        A.jfs = d_jall + (size_t)Y.pop_fs * nt * nfreq;
        A.jrs = Y.pop_rs >= 0 ? d_jall + (size_t)Y.pop_rs * nt * nfreq : nullptr;
        A.afs = (d_aall && P.fs.ssa) ? d_aall + (size_t)Y.pop_fs * nt * nfreq : nullptr;
        A.ars = (d_aall && P.rs.ssa && Y.pop_rs >= 0) ? d_aall + (size_t)Y.pop_rs * nt * nfreq : nullptr;
        A.nufs = dnu + ((size_t)Y.run * 2) * nfreq;
        A.nurs = dnu + ((size_t)Y.run * 2 + 1) * nfreq;
        A.theta_c_l = Y.theta_c_l; A.theta_c_h = Y.theta_c_h;
        A.z = P.z; A.theta_obs = P.theta_obs; A.d_l = P.d_l; A.rtol_theta = P.rtol_theta; A.rtol_phi = P.rtol_phi;
        A.nmax_theta = P.nmax_theta; A.nmax_phi = P.nmax_phi; A.counter_jet = P.counter_jet; A.i_end_r = Y.i_end_r;
        A.nt = nt; A.nf = nfreq;
    }
    int rc = PBAGPU_OK;
    PBA_TRY(rc, hipMemcpyAsync(dnu, hnu.data(), sizeof(double) * hnu.size(), hipMemcpyHostToDevice, s));
    PBA_TRY(rc, hipMemcpyAsync(dL, hl.data(), sizeof(LayerArgs) * n_layers, hipMemcpyHostToDevice, s));
    PBA_TRY(rc, hipMemsetAsync(d_status, 0, sizeof(int) * n_layers, s));
    pbagpu::KernelTimer timer(PBAGPU_STAGE_EATS);
    if (rc == PBAGPU_OK) {
        const size_t lds = sizeof(double) * ((size_t)NSLOT * nt + 2 * (size_t)nfreq);
        const int n_cu = pbagpu::stream_cu_count(s);   // the stream's CU mask bounds the persistent grid
        const long n_items = ((long)n_pairs + EATS_WG - 1) / EATS_WG * n_layers;   // workgroup-sized pieces
        bool ssa = false;   // any absorption table: the instantiation with the absorption code
        for (const LayerArgs &A : hl) ssa = ssa || A.afs || A.ars;
        const void *kfn = ssa ? (const void *)eats_kernel<true> : (const void *)eats_kernel<false>;
        (void)hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        int per_cu = 1;
        PBA_TRY(rc, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, EATS_WG, lds));
        const int n_wg = (int)std::min<long>(n_items, (long)std::max(n_cu, 1) * std::max(per_cu, 1));
        char *ws = (char *)pbagpu::workspace(pbagpu::WS_EATS_STORE, (size_t)n_wg * EATS_WG * LANE_WS, s);
        int *counter = (int *)pbagpu::workspace(pbagpu::WS_EATS_COUNTER, sizeof(int) * n_layers, s);
        if (!ws || !counter) rc = PBAGPU_E_HIP;
        if (rc == PBAGPU_OK) PBA_TRY(rc, hipMemsetAsync(counter, 0, sizeof(int) * n_layers, s));
        if (rc == PBAGPU_OK) {
            (void)hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            unsigned long long *d_samples = nullptr;
            if (pbagpu::timing_enabled()) {
                d_samples = (unsigned long long *)pbagpu::workspace(pbagpu::WS_EATS_SAMPLES, sizeof(unsigned long long), s);
                if (d_samples) PBA_TRY(rc, hipMemsetAsync(d_samples, 0, sizeof(unsigned long long), s));
            }
            timer.start(s);
            if (ssa)
                hipLaunchKernelGGL(eats_kernel<true>, dim3((unsigned)n_wg), dim3(EATS_WG), lds, s, dL, n_layers, n_pairs,
                                   counter, d_tobs, d_nuobs, d_out, d_status, ws, d_samples);
            else
                hipLaunchKernelGGL(eats_kernel<false>, dim3((unsigned)n_wg), dim3(EATS_WG), lds, s, dL, n_layers, n_pairs,
                                   counter, d_tobs, d_nuobs, d_out, d_status, ws, d_samples);
            PBA_TRY(rc, hipGetLastError());
            timer.stop(s);
            if (d_samples && rc == PBAGPU_OK) {
                unsigned long long hs = 0;
                PBA_TRY(rc, hipMemcpyAsync(&hs, d_samples, sizeof(hs), hipMemcpyDeviceToHost, s));
                PBA_TRY(rc, hipStreamSynchronize(s));
                pbagpu::timing_add_units(PBAGPU_STAGE_EATS, (double)hs);
            }
        }
    }
    PBA_TRY(rc, hipStreamSynchronize(s));
    timer.collect();
    return rc;
}
