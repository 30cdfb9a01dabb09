// Stage 2 — electron kinetics (Chang-Cooper) and comoving synchrotron + SSC emissivity.
//
// One 64-lane wavefront per electron population (forward or reverse shock of one
// blastwave), stepping through all output steps.  Lane L owns the contiguous
// Lorentz-factor bins [L*m, L*m+m), m = ceil(ngam/64) <= 8.
//
// Chang-Cooper substep (ChangCooper::setupVectors / solve / tridiagonal_solver,
// src/microphysics/numeric_model.h:19-43, 539-708): with zero dispersion the
// sub-diagonal vanishes, the forward sweep is skipped by the reference, and the
// update is the first-order linear recurrence f_j = d_j/b_j - (c_j/b_j) f_{j+1}
// taken from the top bin down.  Each affine map x -> d_j/b_j - (c_j/b_j) x is
// composed per lane, the lane maps are combined by a 6-round suffix scan over the
// wavefront, and each lane then applies its maps: O(m + log 64) dependent steps
// instead of O(ngam).  Heating/cooling terms and the tridiagonal coefficients are
// constant over the substeps of one step and are built once per step.
//
// Emission (computeRadiationNumericMixed, src/microphysics/microphysics.h:492-599):
// lanes own frequencies; the CSYN synchrotron kernel (Aharonian et al. 2010, eq. D7,
// as SynKernel::cycSynch, src/microphysics/kernels.h:414-437) is evaluated on the fly
// and contracted against f(gamma) held in LDS; the SSC contraction
// (computeSSCSpectrum, numeric_model.h:247-285) reads the precomputed
// Blumenthal-Gould / Nava kernel table (SSCKernel::evalSSCkernel, kernels.h:49-117).
#include "pba_device.h"
#include "bessel.h"
#include "timing.h"
#include "runtime.h"
#include "mlp_internal.h"
#include <vector>

namespace pbagpu {
namespace ele {

constexpr int MAXG = 512;      // max Lorentz-factor bins
constexpr int MAXF = 256;      // max frequency bins
constexpr int MPL = MAXG / 64; // bins per lane

// SSCKernel::sscNava (kernels.h:95-117)
__device__ double ssc_nava(double nu, double gam, double nu_t, double h_by_mec2) {
    double en = nu * h_by_mec2, et = nu_t * h_by_mec2;
    if (et / (4. * gam * gam) < en && en <= et) return (en / et) - (1. / (4. * gam * gam));
    if (et < en && en < (4. * gam * gam * et / (1. + 4. * gam * et))) {
        double q = en / (4. * gam * et * (gam - en));
        double t = 4. * gam * et * q;
        return 2. * q * log(q) + (1. + 2. * q) * (1. - q) + .5 * (1. - q) * (t * t) / (1. + t);
    }
    return 0.;
}

// CSYN kernel (kernels.h:414-437); pref = sqrt(3) qe^3 / (me c^2)
__device__ __forceinline__ double cyc_synch(double B, double nu, double gam, double pref) {
    if (gam > 2.) {
        double x = nu / ((3. * kQe * B * gam * gam) / (4. * kPi * kMe * kC));
        double x13 = cbrt(x);
        double x23 = x13 * x13;
        double x43 = x23 * x23;
        double t1n = 1.808 * x13, t1d = sqrt(1 + 3.4 * x23);
        double t2n = 1. + 2.21 * x23 + 0.347 * x43;
        double t2d = 1. + 1.353 * x23 + 0.217 * x43;
        return t1n / t1d * t2n / t2d * exp(-x) * (pref * B);
    }
    double x = nu / ((3. * kQe * B) / (2. * kPi * kMe * kC));
    double psq = gam * gam - 1.;
    return (2. * psq) / (1. + 3. * psq) * exp((2. * (1. - x)) / (1. + 3. * psq)) * (pref * B);
}

// synchDermer (kernels.h:451-473): the Aharonian et al. (2010) fit of the CSYN branch at every
// Lorentz factor (no cyclotron branch), prefactor sqrt(3) qe^3 B / (me c c)
__device__ __forceinline__ double dermer_synch(double B, double nu, double gam) {
    double x = nu / ((3. * kQe * B * gam * gam) / (4. * kPi * kMe * kC));
    double x13 = cbrt(x);
    double x23 = x13 * x13;
    double x43 = x23 * x23;
    double t1n = 1.808 * x13, t1d = sqrt(1 + 3.4 * x23);
    double t2n = 1. + 2.21 * x23 + 0.347 * x43;
    double t2d = 1. + 1.353 * x23 + 0.217 * x43;
    return t1n / t1d * t2n / t2d * exp(-x) * (sqrt(3.) * (kQe * kQe * kQe) * B / (kMe * kC * kC));
}

// synchBessel (kernels.h:396-412): t = nu / nu_crit(B, gam); 2 t^2 (K43 K13 - 3/5 t (K43^2 - K13^2))
// times sqrt(3) qe^3 B / (me c c), with K_{1/3}, K_{4/3} from bessel.h.  Out of line: its series
// loops would otherwise take registers from the default kernels' hot loop.
__device__ __noinline__ double bessel_synch(double B, double nu, double gam, BesselThird b) {
    const double t = nu / ((3. * kQe * B * gam * gam) / (4. * kPi * kMe * kC));
    return bessel_synch_shape(t, b) * (sqrt(3.) * (kQe * kQe * kQe) * B / (kMe * kC * kC));
}

// method_synchrotron = GSL (SynKernel::synchGSL, kernels.h:211-342): the single-pitch-angle
// synchrotron function F(x) = x int_x^inf K_5/3(t) dt at x = (h nu / me c^2) / (1.5 B / 4.14e13
// gam^2), times sqrt(3) qe^3 B / (me c^2), as the reference forms them.  F itself is evaluated
// here from the integral representation K_nu(t) = int_0^inf exp(-t cosh u) cosh(nu u) du
// (DLMF 10.32.9): integrating over t >= x gives
//     F(x) = x int_0^inf cosh(5u/3) / cosh(u) exp(-x cosh u) du,
// a positive integrand analytic in the strip |Im u| < pi/2, so the trapezoidal rule in u converges
// geometrically.  Step h = min(0.14, 0.75 / sqrt(x)) (the integrand's width shrinks as 1/sqrt(x) for
// large x); the sum stops once x (cosh u - 1) > 46 + 2u/3 (terms below e^-46 of the first).
// Below x = 1e-25 the leading term 4 pi / (sqrt(3) Gamma(1/3)) (x/2)^(1/3) is used (next term
// relative ~x^(2/3)); above x = 745, e^-x underflows and F = 0.
PBA_HD double synchrotron_F(double x) {
    if (!(x > 0.)) return 0.;
    if (x < 1e-25) return 4.0 * kPi / (sqrt(3.0) * 2.6789385347077476337) * cbrt(0.5 * x);  // Gamma(1/3)
    if (x > 745.) return 0.;
    const double h = fmin(0.14, 0.75 / sqrt(x));
    double s = 0.5 * exp(-x);
    for (int k = 1; k < 4096; k++) {
        const double u = k * h;
        const double eu = exp(u), e53 = exp((5.0 / 3.0) * u);
        const double ch = 0.5 * (eu + 1.0 / eu);
        s += 0.5 * (e53 + 1.0 / e53) / ch * exp(-x * ch);
        if (x * (ch - 1.0) > 46.0 + (2.0 / 3.0) * u) break;
    }
    return x * h * s;
}

__device__ __noinline__ double gsl_synch(double B, double nu, double gam) {
    const double ec = 1.5 * B / 4.14e13;
    const double x = nu * (kH / kMe / kC / kC) / (ec * gam * gam);
    return synchrotron_F(x) * (sqrt(3.) * pow(kQe, 3.) * B / kMec2);
}

// Kernel table K[i][j][k] and its frequency integral KI[j][k] (kernels.h:49-94).
__global__ void ssc_table_kernel(const double *__restrict__ nu, const double *__restrict__ e, int nf, int ng,
                                 double h_by_mec2, double *__restrict__ K) {
    size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t n = (size_t)nf * ng * nf;
    if (idx >= n) return;
    int k = (int)(idx % nf);
    int j = (int)((idx / nf) % ng);
    int i = (int)(idx / ((size_t)nf * ng));
    K[idx] = ssc_nava(nu[i], e[j], nu[k], h_by_mec2);
}

__global__ void ssc_integral_kernel(const double *__restrict__ nu, const double *__restrict__ K, int nf, int ng,
                                    double *__restrict__ KI) {
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= ng * nf) return;
    int j = idx / nf, k = idx % nf;
    double res = 0;
    for (int i = 0; i < nf - 1; i++) res += K[((size_t)i * ng + j) * nf + k] * nu[i + 1] * (nu[i + 1] - nu[i]);
    KI[idx] = res;
}

struct PopDesc {
    const double *T;      // dynamics table of the blastwave
    const double *K;      // SSC kernel table or null
    const double *KI;
    const double *e;      // Lorentz-factor grid [ng]
    const double *half_e; // [ng-1]
    const double *dg;     // delta_grid [ng+1]
    const double *dgbar;  // delta_grid_bar [ng]
    const double *nu;     // frequency grid [nf]
    double *jall;         // [nt][nf]
    double *aall;         // [nt][nf] absorption (use_ssa) or null
    double *fall;         // [nt][ng] or null
    // spectra of the dense output ([nt][nf], gam_dot [nt][ng]) or null (pba_spectra)
    double *sp_synj, *sp_syni, *sp_sscj, *sp_totf, *sp_toti, *sp_gds, *sp_gda, *sp_gdc;
    double p, eps_e;
    BesselThird bes;      // Temme-series constants of the Bessel kernel (syn = 2)
    int shock, ng, nf, max_substeps, adi_loss, ssc, ssa, syn;
};

__device__ __forceinline__ void power_law(const double *e, int j, double p, double gm, double gc, double gM,
                                          double n_inj, double &N) {
    double gamma = e[j];
    if (gm > gc) {
        double K_f = n_inj / (pow(gm, 1. - p) * (1. / gc - 1. / gm) + 1. / p * (pow(gm, -p) - pow(gM, -p)));
        if ((gc <= gamma) && (gamma < gm)) N = K_f * pow(gm, 1. - p) * pow(gamma, -2.);
        else if ((gm < gamma) && (gamma <= gM)) N = K_f * pow(gamma, -p - 1.);
    } else {
        double K_s = n_inj / ((1. / (1. - p)) * (pow(gc, 1. - p) - pow(gm, 1. - p)) - gc / p * (pow(gM, -p) - pow(gc, -p)));
        if ((gm <= gamma) && (gamma < gc)) N = K_s * pow(gamma, -p);
        else if ((gc <= gamma) && (gamma <= gM)) N = K_s * gc * pow(gamma, -p - 1.);
    }
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// Per-population state between the per-step launches (device workspace).
struct PopState {
    double *f;     // [n_pop][ngs]  electron spectrum (a population's first ng entries)
    double *ft;    // [ngs][n_pop]  the same, transposed for the Compton contraction
    double *tf;    // [n_pop][nf]   photon density of the previous step (total_rad.f)
    double *wt;    // [nf][n_pop]   Compton photon weights of this step
    double *jsyn;  // [n_pop][nf]
    double *jssc;  // [n_pop][nf][SSC_JB] partial sums over blocks of Lorentz-factor bins
    double *scal;  // [n_pop][4]    dr_comov, vol, n_ele, fac
    double *dfde;  // [n_pop][ng]   d(f/gamma^2)/dgamma of the last computed step (computeSSA's
                   //               ele.dfde: bins where f = 0 keep their earlier value)
    int *flag;     // [n_pop][3]    initialised, active at this step, status bits
    double *gdot;  // [n_pop][3][ngs] cooling rates (syn, adi, ssc) of the last Chang-Cooper step, or
                   //               null (only kept for the spectra of the dense output)
    int ngs;       // row stride of the per-population electron arrays (f, dfde, gdot) and of the
                   // f_all / gam_dot outputs: the largest grid of the call
};

// Step `it` of every population, part 1: injection / Chang-Cooper update and the
// synchrotron emissivity.  One 64-lane wavefront per population.
__global__ void __launch_bounds__(64) ele_step_kernel(const PopDesc *__restrict__ pops, int n_pop, int it, int nt,
                                                      PopState S) {
    __shared__ double s_f[MAXG];
    __shared__ double s_tf[MAXF];   // photon density of the previous step (total_rad.f)
    __shared__ double s_w[MAXF];    // tot_f[k] / nu[k] * (nu[k+1] - nu[k])
    __shared__ double s_dge[MAXG];  // e[j+1] - e[j]
    __shared__ double s_wa[MAXG];   // e_j^2 (e_{j+1} - e_j) d(f/gamma^2)/dgamma (synchrotron self-absorption)
    const PopDesc D = pops[blockIdx.x];
    const int lane = threadIdx.x;
    const int ng = D.ng, nf = D.nf;
    const int m = (ng + 63) / 64;
    const int j0 = lane * m;
    const double p = D.p;
    const double pref = sqrt(3.) * pow(kQe, 3.) * (1.0 / kMec2);
    const int bit = D.shock == 0 ? 1 : 2;
    const int vB = D.shock ? PV_B3 : PV_B, vgm = D.shock ? PV_gm_rs : PV_gm, vgM = D.shock ? PV_gM_rs : PV_gM;
    const int vgc = D.shock ? PV_gc_rs : PV_gc, vacc = D.shock ? PV_accel_frac_rs : PV_accel_frac;
    const int vM = D.shock ? PV_M3 : PV_M2, vrho = D.shock ? PV_rho3 : PV_rho2;
    const int vth = D.shock ? PV_thickness_rs : PV_thickness, vGsh = D.shock ? PV_GammaRsh : PV_GammaFsh;
    auto V = [&](int v, int t) { return D.T[(size_t)v * nt + t]; };

    const int q0 = blockIdx.x;
    int *FL = S.flag + 3 * q0;
    double *SC = S.scal + 4 * q0;
    {
        int flags = (int)V(PV_mflags, it);
        if (!(flags & bit)) { if (lane == 0) FL[1] = 0; return; }
    }
    double f[MPL], src[MPL], bb[MPL], cp[MPL];
    for (int r = 0; r < MPL; r++) {
        int j = j0 + r;
        f[r] = (r < m && j < ng) ? S.f[(size_t)q0 * S.ngs + j] : 0.;
        src[r] = 0.; bb[r] = 1.; cp[r] = 0.;
    }
    for (int j = lane; j < ng; j += 64) s_dge[j] = (j < ng - 1) ? D.e[j + 1] - D.e[j] : 0.;
    for (int k = lane; k < nf; k += 64) s_tf[k] = S.tf[(size_t)q0 * nf + k];
    __syncthreads();
    bool init = FL[0] != 0;
    double dr_comov = SC[0], vol = SC[1];
    int st = 0;
    {
        const double B = V(vB, it), gm = V(vgm, it), gM = V(vgM, it), gc = V(vgc, it), acc = V(vacc, it);
        if (!init) {
            dr_comov = V(vth, it);
            const double N_tot = V(vM, it) / kMp;
            vol = V(vM, it) / V(vrho, it);
            for (int r = 0; r < m; r++) {
                int j = j0 + r;
                if (j < ng) {
                    power_law(D.e, j, p, gm, gc, gM, N_tot, f[r]);
                    if (j < ng - 1) f[r] *= acc;
                }
            }
            init = true;
        } else {
            const double tc = V(PV_tcomov, it - 1), tc1 = V(PV_tcomov, it);
            const double mm = V(vM, it - 1), m1 = V(vM, it), rp = V(vrho, it - 1), rp1 = V(vrho, it);
            if ((B <= 0) || (gm < 1.) || (gM <= 1.) || (gm >= gM) || (acc > 1.)) st |= PBAGPU_S_BAD_MPHYS;
            dr_comov = V(vth, it - 1) * V(vGsh, it - 1);
            const double N_tot = m1 / kMp;
            const double dt = tc1 - tc;
            const double volume = mm / rp, volume_p1 = m1 / rp1;
            const double dlnVdt = D.adi_loss ? 1. / dt * (1. - volume / volume_p1) : 0.;
            if (!isfinite(dlnVdt) || dlnVdt < 0) st |= PBAGPU_S_NONFINITE;
            const double dlog_gam = (D.e[1] - D.e[0]) / D.e[0];
            const double dts = kSigmaT * gM * B * B / (6. * kPi * kMe * kC);
            const double dta = (gM * gM - 1.) / (3. * gM * gM) * dlnVdt;
            double delta_t = dlog_gam / (dts + dta);
            long n_sub;
            if (delta_t >= dt / 1.) { delta_t = dt / 1.; n_sub = 1; }
            else if (delta_t < dt / (double)D.max_substeps) { delta_t = dt / (double)D.max_substeps; n_sub = D.max_substeps; }
            else n_sub = (long)(dt / delta_t);
            vol = volume;
            if (gm <= 2) {
                for (int r = 0; r < m; r++) {
                    int j = j0 + r;
                    double t = 0.;
                    if (j < ng) power_law(D.e, j, p, gm, gc, gM, N_tot, t);
                    f[r] = t;
                }
            } else {
                const double Nsrc = (m1 - mm) / kMp / dt;
                const double index = -p;
                const double norm = (1 + index) / (pow(gM, index + 1) - pow(gm, index + 1));
                // photon weights for the SSC cooling integral
                if (D.ssc) {
                    for (int k = lane; k < nf; k += 64)
                        s_w[k] = (k < nf - 1) ? s_tf[k] / D.nu[k] * (D.nu[k + 1] - D.nu[k]) : 0.;
                    __syncthreads();
                }
                const double cf = kSigmaT / (6. * kPi * kMe * kC);
                const double csc = 3. / 4. * kH * kSigmaT / kMe / kC;
                // heating terms of the lane's bins and of the bin below the lane's first bin
                double heat[MPL + 1];
                for (int r = -1; r < m; r++) {
                    int j = j0 + r;
                    double hv = 0.;
                    if (j >= 0 && j < ng - 1) {
                        double ge = D.half_e[j];
                        double gssc = 0.;
                        if (D.ssc) {
                            double res = 0.;
                            const double *KIr = D.KI + (size_t)j * nf;
                            for (int k = 0; k < nf - 1; k++) res += s_w[k] * KIr[k];
                            gssc = csc * res / ge / ge;
                        }
                        const double gsyn = cf * B * B * ge * ge, gadi = dlnVdt * (ge * ge - 1.) / (3. * ge);
                        hv = gsyn + gadi + gssc;
                        if (S.gdot && r >= 0) {
                            double *G = S.gdot + (size_t)q0 * 3 * S.ngs;
                            G[j] = gsyn; G[S.ngs + j] = gadi; G[2 * S.ngs + j] = gssc;
                        }
                    }
                    heat[r + 1] = hv;
                }
                for (int r = 0; r < m; r++) {
                    int j = j0 + r;
                    if (j >= ng) { bb[r] = 1.; cp[r] = 0.; src[r] = 0.; continue; }
                    double ej = D.e[j];
                    src[r] = ((ej > gm) && (ej <= gM)) ? Nsrc * norm * pow(ej, index) : 0.;
                    double odgb = 1.0 / D.dgbar[j];
                    double braw, craw;
                    if (j == 0) { braw = -0.; craw = odgb * heat[r + 1]; }
                    else if (j == ng - 1) { braw = -0.; craw = 0.; }
                    else { braw = -odgb * heat[r]; craw = odgb * heat[r + 1]; }
                    double b_ = (1. - braw * delta_t) + 0. * delta_t;
                    double c_ = craw * -delta_t;
                    bb[r] = b_;
                    cp[r] = c_ / b_;
                }
                for (long s = 0; s < n_sub; s++) {
                    // lane map: f_first = A + Bm * f_next
                    double dp[MPL];
                    double A = 0., Bm = 1.;
                    for (int r = MPL - 1; r >= 0; r--) {
                        if (r >= m) continue;
                        int j = j0 + r;
                        if (j >= ng) continue;
                        dp[r] = (f[r] + src[r] * delta_t) / bb[r];
                        A = dp[r] - cp[r] * A;
                        Bm = -cp[r] * Bm;
                    }
                    // inclusive suffix scan of affine maps over the wavefront
                    for (int d = 1; d < 64; d <<= 1) {
                        double A2 = __shfl_down(A, d, 64), B2 = __shfl_down(Bm, d, 64);
                        if (lane + d < 64) { A = A + Bm * A2; Bm = Bm * B2; }
                    }
                    double Fnext = __shfl_down(A, 1, 64);
                    if (lane == 63) Fnext = 0.;
                    for (int r = MPL - 1; r >= 0; r--) {
                        if (r >= m) continue;
                        int j = j0 + r;
                        if (j >= ng) continue;
                        Fnext = dp[r] - cp[r] * Fnext;
                        f[r] = Fnext;
                    }
                }
            }
        }
        // --- emission: synchrotron part and what the Compton contraction needs ---
        double ne_part = 0.;
        for (int r = 0; r < m; r++) {
            int j = j0 + r;
            if (j < ng) {
                s_f[j] = f[r];
                if (j < ng - 1) ne_part += f[r] * s_dge[j];
                S.f[(size_t)q0 * S.ngs + j] = f[r];
                S.ft[(size_t)j * n_pop + q0] = f[r];
            }
        }
        if (D.ssc)
            for (int k = lane; k < nf; k += 64)
                S.wt[(size_t)k * n_pop + q0] = (k < nf - 1) ? s_tf[k] / D.nu[k] * (D.nu[k + 1] - D.nu[k]) : 0.;
        __syncthreads();
        const double n_ele = wave_sum(ne_part);
        if (n_ele <= 0) st |= PBAGPU_S_NONFINITE;
        const double fac = (dr_comov / kC) / vol;
        if (D.ssa) {
            // computeSSA + dfdx (numeric_model.h:190-238, base_state.h:482-536): derivative of
            // f / gamma^2, central in the interior (divided by the lower spacing), one-sided
            // second-order stencils at the ends (not divided by a spacing); bins with f = 0 keep
            // the value of the last step that set them
            auto tmp = [&](int j) { return s_f[j] / D.e[j] / D.e[j]; };
            double *dfde = S.dfde + (size_t)q0 * S.ngs;
            for (int j = lane; j < ng; j += 64) {
                const double tj = tmp(j);
                if (tj > 0) {
                    if (j == 0) dfde[j] = -1.5 * tj + 2. * tmp(1) - 0.5 * tmp(2);
                    else if (j == ng - 1) dfde[j] = 1.5 * tj - 2. * tmp(j - 1) + 0.5 * tmp(j - 2);
                    else dfde[j] = 0.5 * (tmp(j + 1) - tmp(j - 1)) / (D.e[j] - D.e[j - 1]);
                }
                if (j < ng - 1) s_wa[j] = D.e[j] * D.e[j] * (D.e[j + 1] - D.e[j]) * dfde[j];
            }
            __syncthreads();
        }
        const double acc_ssa = D.ssa ? V(vacc, it) : 0., nprime_ssa = D.ssa ? V(vrho, it) / kMp : 0.;
        for (int i = lane; i < nf; i += 64) {
            const double nui = D.nu[i];
            double synj = 0., asum = 0.;
            const bool emit = i < nf - 1;   // computeSynSpectrum leaves the last frequency at 0
            if (emit || D.ssa) {
                for (int j = 0; j < ng - 1; j++) {
                    const double k = D.syn == 0   ? cyc_synch(B, nui, D.e[j], pref)
                                     : D.syn == 1 ? dermer_synch(B, nui, D.e[j])
                                     : D.syn == 2 ? bessel_synch(B, nui, D.e[j], D.bes)
                                                  : gsl_synch(B, nui, D.e[j]);
                    if (emit) synj += k * s_f[j] * s_dge[j];
                    if (D.ssa) asum += s_wa[j] * k;
                }
                if (emit && (!isfinite(synj) || synj < 0.)) st |= PBAGPU_S_NONFINITE;
            }
            S.jsyn[(size_t)q0 * nf + i] = synj;
            if (D.ssa) {
                asum *= -1. / (8. * kPi * kMe * nui * nui);
                asum = fabs(asum);
                if (!isfinite(asum)) asum = 1e99;
                // total_rad.a = syn.a / n_ele * accel_frac * n_prime (microphysics.h:576-579)
                D.aall[(size_t)it * nf + i] = asum / n_ele * acc_ssa * nprime_ssa;
            }
        }
        if (D.fall)
            for (int r = 0; r < m; r++) {
                int j = j0 + r;
                if (j < ng) D.fall[(size_t)it * S.ngs + j] = f[r];
            }
        // cooling rates of the last Chang-Cooper step (Electrons::save_to_all, base_state.h:200-211)
        if (D.sp_gds && S.gdot) {
            const double *G = S.gdot + (size_t)q0 * 3 * S.ngs;
            for (int j = lane; j < ng; j += 64) {
                D.sp_gds[(size_t)it * S.ngs + j] = G[j];
                D.sp_gda[(size_t)it * S.ngs + j] = G[S.ngs + j];
                D.sp_gdc[(size_t)it * S.ngs + j] = G[2 * S.ngs + j];
            }
        }
        for (int d = 32; d >= 1; d >>= 1) st |= __shfl_xor(st, d, 64);
        if (lane == 0) {
            FL[0] = 1; FL[1] = 1; FL[2] |= st;
            SC[0] = dr_comov; SC[1] = vol; SC[2] = n_ele; SC[3] = fac;
        }
    }
}

// Step `it`, part 2: SSC emissivity of the populations of one Compton table
// (computeSSCSpectrum, src/microphysics/numeric_model.h:247-285):
//   j_ssc[i] = sum_j f_j h nu_i (3/4 sigma_T c / gamma_j^2) (sum_k w_k K[i][j][k]) (gamma_{j+1} - gamma_j)
// Lane = population, so the table entry K[i][j][k] is the same for all 64 lanes (one scalar
// load feeds 64 populations).  grid = (population blocks of 64, frequency i, j-block); each
// workgroup sums its block of Lorentz-factor bins with four independent k-chains and writes a
// partial sum, which ele_emit_kernel adds up in j-block order.
constexpr int SSC_JB = 8;
// Population q0 is checked against the plain contraction at step it (surrogate mode).
__device__ __forceinline__ bool sur_sampled(int q0, int it, int every) {
    return every > 0 && ((131LL * q0 + 37LL * it) % every) == 0;
}

// sel_fb: null (every population of the list) or, in surrogate mode, the per-population fallback
// flags: then only populations sampled for the check at step sel_it or fallen back are computed.
__global__ void __launch_bounds__(64) ssc_step_kernel(const int *__restrict__ plist, int npk, int n_pop,
                                                      const double *__restrict__ K, const double *__restrict__ e,
                                                      const double *__restrict__ nu, int ng, int nf, PopState S,
                                                      const int *__restrict__ sel_fb, int sel_it, int sel_every) {
    extern __shared__ double s_wl[];   // [nf][64] photon weights of the 64 populations
    const int lane = threadIdx.x;
    const int idx = blockIdx.x * 64 + lane;
    const int q0 = idx < npk ? plist[idx] : -1;
    bool act = q0 >= 0 && S.flag[3 * q0 + 1] != 0;
    if (sel_fb && act) act = sel_fb[q0] != 0 || sur_sampled(q0, sel_it, sel_every);
    if (__ballot(act) == 0ull) return;   // the whole (single-wavefront) workgroup leaves together
    for (int k = 0; k < nf; k++) s_wl[k * 64 + lane] = act ? S.wt[(size_t)k * n_pop + q0] : 0.;
    __syncthreads();
    const double c0 = 3. / 4. * kSigmaT * kC;
    const int i = blockIdx.y;
    const int jn = ng - 1;
    const int js = (jn + SSC_JB - 1) / SSC_JB;
    const int j0 = blockIdx.z * js, j1 = min(j0 + js, jn);
    const double nui = nu[i];
    double part = 0.;
    int j = j0;
    for (; j + 4 <= j1; j += 4) {
        const double *K0 = K + ((size_t)i * ng + j) * nf;
        const double *K1 = K0 + nf, *K2 = K1 + nf, *K3 = K2 + nf;
        double s0 = 0., s1 = 0., s2 = 0., s3 = 0.;
        for (int k = 0; k < nf - 1; k++) {
            const double w = s_wl[k * 64 + lane];
            s0 += w * K0[k]; s1 += w * K1[k]; s2 += w * K2[k]; s3 += w * K3[k];
        }
        const double sv[4] = {s0, s1, s2, s3};
        for (int u = 0; u < 4; u++) {
            const int jj = j + u;
            const double ej = e[jj];
            const double fj = act ? S.ft[(size_t)jj * n_pop + q0] : 0.;
            part += fj * kH * nui * (c0 / (ej * ej) * sv[u]) * (e[jj + 1] - e[jj]);
        }
    }
    for (; j < j1; j++) {
        const double *Kr = K + ((size_t)i * ng + j) * nf;
        double sv = 0.;
        for (int k = 0; k < nf - 1; k++) sv += s_wl[k * 64 + lane] * Kr[k];
        const double ej = e[j];
        const double fj = act ? S.ft[(size_t)j * n_pop + q0] : 0.;
        part += fj * kH * nui * (c0 / (ej * ej) * sv) * (e[j + 1] - e[j]);
    }
    if (act) S.jssc[((size_t)q0 * nf + i) * SSC_JB + blockIdx.z] = part;
}

// ---------------------------------------------------------------------------------------------
// Mixed-precision SSC contraction (precision = 1): the same sum as ssc_step_kernel, on MFMA in
// BF16 with FP32 accumulation, taken in two levels:
//   T[i][j][p] = sum_k K[i][j][k] wn_k^p                 (a GEMM over k, on v_mfma_f32_16x16x32_bf16)
//   C[i][p]    = sum_chunks amax_c^p sum_{j in chunk c} an_j^p T[i][j][p]   (FP32 within a chunk,
//                                                        FP64 across the chunks)
// with a_j = f_j (gamma_{j+1}-gamma_j) (3/4 sigma_T c)/gamma_j^2 normalised per chunk of 16
// Lorentz-factor bins (an_j = a_j / amax_c, amax_c = max of a_j over the chunk) and w_k normalised
// by its maximum (wn_k = w_k / wmax), so every operand lies in [0, 1] whatever the decades the
// spectra span.  The GEMM operands K and wn are split into BF16 pairs x = hi + lo (the table once,
// the photon weights once per step) and the product is hi*hi + hi*lo + lo*hi (the lo*lo term is
// below FP32 resolution), ~16 significant bits per operand; an_j enters in FP32.
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;

__device__ __forceinline__ unsigned short bf16_rn(float x) {
    unsigned u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);   // round to nearest even (operands are finite)
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f(unsigned short h) { return __uint_as_float(((unsigned)h) << 16); }

// K (FP64 [nf][ng][nf]) -> BF16 pair [nf][ng][kp], k padded with zeros to kp (a multiple of 32)
__global__ void ssc_table_bf16_kernel(const double *__restrict__ K, int nf, int ng, int kp,
                                      unsigned short *__restrict__ Khi, unsigned short *__restrict__ Klo) {
    size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t n = (size_t)nf * ng * kp;
    if (idx >= n) return;
    int k = (int)(idx % kp);
    size_t ij = idx / kp;
    float v = k < nf - 1 ? (float)K[ij * nf + k] : 0.f;   // the last frequency carries no photons
    unsigned short h = bf16_rn(v);
    Khi[idx] = h;
    Klo[idx] = bf16_rn(v - bf16_to_f(h));
}

constexpr int SSC_CH = 16;                 // Lorentz-factor bins per normalisation chunk (one MFMA row block)
constexpr int SSC_MAXKS = MAXF / 32;       // k-steps of 32 at the largest frequency grid

// Per population of the list: normalised operands and their scales for this step.
//   an  [npk][nch * 16] FP32 (0 past the last contributing bin), amax [npk][nch] FP64,
//   wnh, wnl [npk][kp] BF16 pair of the normalised photon weights, wmax [npk] FP64
__global__ void __launch_bounds__(64) ssc_prep_kernel(const int *__restrict__ plist, int npk, int n_pop,
                                                      const double *__restrict__ e, int ng, int nf, int kp, int nch,
                                                      PopState S, float *__restrict__ an, double *__restrict__ amax,
                                                      unsigned short *__restrict__ wnh, unsigned short *__restrict__ wnl,
                                                      double *__restrict__ wmax) {
    const int q = blockIdx.x, lane = threadIdx.x;
    const int q0 = plist[q];
    const bool act = S.flag[3 * q0 + 1] != 0;
    const double c0 = 3. / 4. * kSigmaT * kC;
    const int jn = ng - 1;
    for (int c = 0; c < nch; c++) {
        const int j = c * SSC_CH + (lane & (SSC_CH - 1));
        double aj = 0.;
        if (act && j < jn && lane < SSC_CH) {
            const double ej = e[j];
            aj = S.ft[(size_t)j * n_pop + q0] * (e[j + 1] - e[j]) * (c0 / (ej * ej));
        }
        double m = fabs(aj);
        for (int d = 8; d >= 1; d >>= 1) m = fmax(m, __shfl_xor(m, d, 64));   // over the chunk's 16 lanes
        if (lane < SSC_CH) an[(size_t)q * nch * SSC_CH + j] = m > 0. ? (float)(aj / m) : 0.f;
        if (lane == 0) amax[(size_t)q * nch + c] = m;
    }
    double w = 0.;
    for (int k = lane; k < nf - 1; k += 64) w = fmax(w, fabs(act ? S.wt[(size_t)k * n_pop + q0] : 0.));
    for (int d = 32; d >= 1; d >>= 1) w = fmax(w, __shfl_xor(w, d, 64));
    for (int k = lane; k < kp; k += 64) {
        const double wk = (act && k < nf - 1) ? S.wt[(size_t)k * n_pop + q0] : 0.;
        const float v = w > 0. ? (float)(wk / w) : 0.f;
        const unsigned short h = bf16_rn(v);
        wnh[(size_t)q * kp + k] = h;
        wnl[(size_t)q * kp + k] = bf16_rn(v - bf16_to_f(h));
    }
    if (lane == 0) wmax[q] = w;
}

__device__ __forceinline__ bf16x8 load_bf16x8(const unsigned short *p) {
    return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(p));
}

// grid = (nf, ceil(npk / 64)), 4 wavefronts; wavefront w takes the 16 populations
// blockIdx.y * 64 + 16 w + (lane mod 16) at frequency i = blockIdx.x, and runs over every chunk of
// 16 Lorentz-factor bins.  v_mfma_f32_16x16x32_bf16: lane l holds row l mod 16 of A (bin j) and
// column l mod 16 of B (population), eight consecutive k from 8 (l / 16); its four results are rows
// 4 (l / 16) + r of column l mod 16.  The photon weights of the lane's population stay in
// registers for all chunks; the table rows of a chunk are 16-byte loads shared through L1 by the
// four wavefronts.
__global__ void __launch_bounds__(256) ssc_mfma_kernel(const int *__restrict__ plist, int npk, const double *__restrict__ nu,
                                                       int ng, int nf, int kp, int nch,
                                                       const unsigned short *__restrict__ Khi,
                                                       const unsigned short *__restrict__ Klo,
                                                       const float *__restrict__ an, const double *__restrict__ amax,
                                                       const unsigned short *__restrict__ wnh,
                                                       const unsigned short *__restrict__ wnl,
                                                       const double *__restrict__ wmax, PopState S) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = blockIdx.x;
    const int col = lane & 15, g = lane >> 4;
    const int pc = blockIdx.y * 64 + wave * 16 + col;
    const bool pv = pc < npk;
    const int nks = kp / 32;
    bf16x8 bh[SSC_MAXKS], bl[SSC_MAXKS];
    const bf16x8 zero8 = __builtin_bit_cast(bf16x8, make_uint4(0u, 0u, 0u, 0u));
#pragma unroll
    for (int ks = 0; ks < SSC_MAXKS; ks++) {
        bh[ks] = zero8; bl[ks] = zero8;
        if (ks < nks && pv) {
            const size_t o = (size_t)pc * kp + ks * 32 + 8 * g;
            bh[ks] = load_bf16x8(wnh + o);
            bl[ks] = load_bf16x8(wnl + o);
        }
    }
    const int jn = ng - 1;
    double total = 0.;
    for (int c = 0; c < nch; c++) {
        const int j = c * SSC_CH + col;          // this lane's A row
        const bool jv = j < jn;
        const size_t rowo = ((size_t)i * ng + (jv ? j : 0)) * kp + 8 * g;
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < SSC_MAXKS; ks++) {   // compile-time bound keeps bh, bl in registers
            if (ks >= nks) break;
            const bf16x8 ah = jv ? load_bf16x8(Khi + rowo + ks * 32) : zero8;
            const bf16x8 al = jv ? load_bf16x8(Klo + rowo + ks * 32) : zero8;
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[ks], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[ks], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[ks], acc, 0, 0, 0);
        }
        // acc[r] = T[i][c * 16 + 4 g + r][pc]: weight by an and sum over the chunk's 16 bins
        float part = 0.f;
        if (pv) {
            const float4 a4 = *reinterpret_cast<const float4 *>(an + (size_t)pc * nch * SSC_CH + c * SSC_CH + 4 * g);
            part = a4.x * acc[0] + a4.y * acc[1] + a4.z * acc[2] + a4.w * acc[3];
        }
        part += __shfl_xor(part, 16, 64);
        part += __shfl_xor(part, 32, 64);
        if (pv) total += amax[(size_t)pc * nch + c] * (double)part;
    }
    if (!pv || g != 0) return;
    const int q0 = plist[pc];
    if (S.flag[3 * q0 + 1] == 0) return;
    double *J = S.jssc + ((size_t)q0 * nf + i) * SSC_JB;
    J[0] = kH * nu[i] * (wmax[pc] * total);
    for (int jb = 1; jb < SSC_JB; jb++) J[jb] = 0.;
}

// ------------------------------------------------------------ learned SSC emissivity
// Input features of one cell (one wavefront): f_j = fj(j), w_k = wk(k) for k < nf - 1 (the weights
// the contraction takes; the last one is unused and given as 0), magnetic field B.
template <class FA, class WA>
__device__ void sur_features_row(FA fj, WA wk, int ng, int nf, double B, float *__restrict__ x,
                                 double *__restrict__ scale) {
    const int lane = threadIdx.x & 63;
    double fm = 0., wm = 0.;
    for (int j = lane; j < ng; j += 64) fm = fmax(fm, fj(j));
    for (int k = lane; k < nf - 1; k += 64) wm = fmax(wm, wk(k));
    for (int d = 32; d >= 1; d >>= 1) { fm = fmax(fm, __shfl_xor(fm, d, 64)); wm = fmax(wm, __shfl_xor(wm, d, 64)); }
    for (int j = lane; j < ng; j += 64) {
        const double v = fm > 0. ? fj(j) / fm : 0.;
        x[j] = (float)((log10(fmax(v, 1e-30)) + 30.) / 30.);
    }
    for (int k = lane; k < nf; k += 64) {
        const double v = (k < nf - 1 && wm > 0.) ? wk(k) / wm : 0.;
        x[ng + k] = (float)((log10(fmax(v, 1e-30)) + 30.) / 30.);
    }
    if (lane == 0) {
        x[ng + nf] = B > 0. ? (float)(log10(B) / 10.) : 0.f;
        *scale = fm * wm;
    }
}

__global__ void __launch_bounds__(64) ssc_features_kernel(const double *__restrict__ f, const double *__restrict__ w,
                                                          const double *__restrict__ B, int ng, int nf,
                                                          float *__restrict__ x, double *__restrict__ scale) {
    const size_t c = blockIdx.x;
    const int dx = ng + nf + 1;
    sur_features_row([&](int j) { return f[c * ng + j]; }, [&](int k) { return w[c * nf + k]; }, ng, nf, B[c],
                     x + c * dx, scale + c);
}

// features of the populations of one Compton table at step it, from the electron stage's state
__global__ void __launch_bounds__(64) sur_feat_step_kernel(const PopDesc *__restrict__ pops,
                                                           const int *__restrict__ plist, int npk, int n_pop, int it,
                                                           int nt, int ng, int nf, PopState S, float *__restrict__ x,
                                                           double *__restrict__ scale) {
    const int q = blockIdx.x;
    const int q0 = plist[q];
    const PopDesc &D = pops[q0];
    const double B = D.T[(size_t)(D.shock ? PV_B3 : PV_B) * nt + it];
    sur_features_row([&](int j) { return S.ft[(size_t)j * n_pop + q0]; },
                     [&](int k) { return S.wt[(size_t)k * n_pop + q0]; }, ng, nf, B,
                     x + (size_t)q * (ng + nf + 1), scale + q);
}

// Surrogate statistics on the device: inferred, checked, failed, fallback, max error (bits)
struct SurStats { unsigned long long n[5]; };

// SSC emissivity of the populations of one Compton table at step it: the network's value, or the
// plain contraction's (already in S.jssc) for populations that failed a check; the sampled
// populations are compared with the plain contraction first.
__global__ void __launch_bounds__(64) sur_apply_kernel(const int *__restrict__ plist, int npk, int n_pop, int it,
                                                       int nf, const float *__restrict__ y,
                                                       const float *__restrict__ y_off, float y_scale,
                                                       const double *__restrict__ scale, PopState S,
                                                       int *__restrict__ fb, int every, double tol,
                                                       SurStats *__restrict__ st) {
    const int q = blockIdx.x, lane = threadIdx.x;
    const int q0 = plist[q];
    if (!S.flag[3 * q0 + 1]) return;
    double *J = S.jssc + (size_t)q0 * nf * SSC_JB;
    if (fb[q0]) {   // computed by the plain contraction this step
        if (lane == 0) atomicAdd(&st->n[3], 1ull);
        return;
    }
    const bool chk = sur_sampled(q0, it, every);
    const double sc = scale[q];
    double err = 0., peak = 0.;
    if (chk) {
        for (int i = lane; i < nf; i += 64) {
            double ex = 0.;
            for (int jb = 0; jb < SSC_JB; jb++) ex += J[(size_t)i * SSC_JB + jb];
            peak = fmax(peak, fabs(ex));
        }
        for (int d = 32; d >= 1; d >>= 1) peak = fmax(peak, __shfl_xor(peak, d, 64));
        for (int i = lane; i < nf; i += 64) {
            double ex = 0.;
            for (int jb = 0; jb < SSC_JB; jb++) ex += J[(size_t)i * SSC_JB + jb];
            const double sv = sc > 0. ? sc * exp10((double)y_off[i] + (double)y_scale * (double)y[(size_t)q * nf + i]) : 0.;
            if (fabs(ex) > 1e-6 * peak) err = fmax(err, fabs(sv - ex) / fabs(ex));
            // an all-zero plain emissivity is met only by an all-zero prediction
            if (peak == 0. && sv != 0.) err = INFINITY;
        }
        for (int d = 32; d >= 1; d >>= 1) err = fmax(err, __shfl_xor(err, d, 64));
        if (lane == 0) {
            atomicAdd(&st->n[1], 1ull);
            atomicMax(&st->n[4], (unsigned long long)__double_as_longlong(err));
        }
        if (err > tol) {   // keep the plain value and use the plain contraction from now on
            if (lane == 0) { fb[q0] = 1; atomicAdd(&st->n[2], 1ull); }
            return;
        }
    }
    if (lane == 0) atomicAdd(&st->n[0], 1ull);
    for (int i = lane; i < nf; i += 64) {
        const double sv = sc > 0. ? sc * exp10((double)y_off[i] + (double)y_scale * (double)y[(size_t)q * nf + i]) : 0.;
        J[(size_t)i * SSC_JB] = sv;
        for (int jb = 1; jb < SSC_JB; jb++) J[(size_t)i * SSC_JB + jb] = 0.;
    }
}

__global__ void transpose_kernel(const double *__restrict__ a, int rows, int cols, double *__restrict__ at) {
    const size_t n = (size_t)rows * cols;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const size_t r = i / cols, c = i % cols;
        at[c * rows + r] = a[i];
    }
}

__global__ void sum_jb_kernel(const double *__restrict__ jssc, size_t n, double *__restrict__ out) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        double v = 0.;
        for (int jb = 0; jb < SSC_JB; jb++) v += jssc[i * SSC_JB + jb];
        out[i] = v;
    }
}

__global__ void fill_int_kernel(int *__restrict__ a, int n, int v, int stride_val) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] = stride_val ? i : v;
}

__global__ void ele_status_kernel(const int *__restrict__ flag, int n_pop, int *__restrict__ status) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n_pop) status[q] = flag[3 * q + 2];
}

// computeIntensity with the iAPPROX rule (src/microphysics/microphysics.h:11-58): Dermer (2009)'s
// u(tau) for a depth above 1e-3, the emissivity itself below
__device__ __forceinline__ double intensity_approx(double em, double dtau) {
    if (dtau == 0) return em;
    const double e = exp(-dtau);
    const double u = 1. / 2. + e / dtau - (1 - e) / (dtau * dtau);
    return dtau > 1.e-3 ? em * (3. * u / dtau) : em;
}

// Step `it`, part 3: total emissivity per electron, photon density for the next step
// (computeRadiationNumericMixed, src/microphysics/microphysics.h:492-599).
__global__ void __launch_bounds__(64) ele_emit_kernel(const PopDesc *__restrict__ pops, int it, int nt, PopState S) {
    const PopDesc D = pops[blockIdx.x];
    const int q0 = blockIdx.x, lane = threadIdx.x, nf = D.nf;
    const int *FL = S.flag + 3 * q0;
    if (!FL[1]) return;
    const double *SC = S.scal + 4 * q0;
    const double n_ele = SC[2], fac = SC[3];
    const int vacc = D.shock ? PV_accel_frac_rs : PV_accel_frac;
    const double acc = D.T[(size_t)vacc * nt + it];
    int st = 0;
    for (int i = lane; i < nf; i += 64) {
        const double nui = D.nu[i];
        const double synj = S.jsyn[(size_t)q0 * nf + i];
        double sscj = 0.;
        if (D.ssc)
            for (int jb = 0; jb < SSC_JB; jb++) sscj += S.jssc[((size_t)q0 * nf + i) * SSC_JB + jb];
        const double tf = synj * fac / (kH * nui) + (D.ssc ? sscj * fac / (kH * nui) : 0.);
        S.tf[(size_t)q0 * nf + i] = tf;
        double jt = (synj + sscj) * acc / n_ele;
        if (!isfinite(jt) || jt < 0.) st |= PBAGPU_S_NONFINITE;
        D.jall[(size_t)it * nf + i] = jt;
        if (D.sp_synj) {
            // storeDenseOutput (microphysics.h:602-633) and total_rad's intensity (:580-583):
            // the depth is the normalised absorption times the comoving shell thickness
            const size_t o = (size_t)it * nf + i;
            const double dtau = D.aall ? D.aall[o] * SC[0] : 0.;
            D.sp_synj[o] = synj * acc;
            D.sp_syni[o] = intensity_approx(synj * acc, dtau);
            D.sp_sscj[o] = sscj * acc;
            D.sp_totf[o] = tf;
            D.sp_toti[o] = intensity_approx((synj + sscj) * acc, dtau);
        }
    }
    for (int d = 32; d >= 1; d >>= 1) st |= __shfl_xor(st, d, 64);
    if (lane == 0 && st) S.flag[3 * q0 + 2] |= st;
}

struct GridSet {
    std::vector<double> e, half_e, dg, dgbar, nu;
};

static void build_grids(const pba_shock_params &s, GridSet &G) {
    int n = s.ngam;
    G.e.assign(n, 0.); G.half_e.assign(n - 1, 0.); G.dg.assign(n + 1, 0.); G.dgbar.assign(n, 0.);
    double step = std::exp(std::log(s.gam2) / (double)n);
    for (int i = 0; i < n; i++) {
        G.e[i] = std::pow(step, (double)i);
        if (i < n - 1) G.half_e[i] = .5 * G.e[i] * (step + 1.0);
    }
    for (int i = 0; i < n - 1; i++) G.dg[i + 1] = G.e[i + 1] - G.e[i];
    G.dg[0] = G.e[0] * (1 - 1.0 / step);
    G.dg[n] = G.e[n - 1] * (step - 1);
    for (int i = 0; i < n; i++) G.dgbar[i] = (G.dg[i] + G.dg[i + 1]) / 2.;
    int nf = s.nfreq;
    G.nu.assign(nf, 0.);
    double cur = std::pow(10.0, std::log10(s.freq1));
    double ratio = std::pow(10.0, (std::log10(s.freq2) - std::log10(s.freq1)) / nf);
    for (int i = 0; i < nf; i++) { G.nu[i] = cur; cur *= ratio; }
}

}  // namespace ele
}  // namespace pbagpu

// precision 0: FP64; 1: SSC contraction in BF16 hi/lo on MFMA; 2: SSC emissivity from the surrogate
static int ele_evolve_impl(const pba_run_params *params, int n_runs, const pba_population *pops, int n_pop,
                           const double *d_table, int bw_total, int nt, double *d_jall, double *d_aall,
                           double *d_fall, int *d_status, int precision, const pba_ssc_surrogate *sur,
                           pba_surrogate_stats *stats, const pba_spectra *spec, void *stream) {
    using namespace pbagpu::ele;
    if (!params || !pops || !d_table || !d_jall || !d_status || n_pop < 1 || nt < 1 || n_runs < 1) return PBAGPU_E_ARG;
    if (pops[0].run < 0 || pops[0].run >= n_runs) return PBAGPU_E_ARG;
    if (precision < 0 || precision > 2 || (precision == 2) != (sur != nullptr)) return PBAGPU_E_UNSUPPORTED;
    hipStream_t s = (hipStream_t)stream;
    // spectra: the three cooling-rate rows go together (one kept state per population), and so do
    // the five emission rows
    const bool want_gdot = spec && spec->gam_dot_syn && spec->gam_dot_adi && spec->gam_dot_ssc;
    const bool want_emit = spec && spec->syn_j && spec->syn_i && spec->ssc_j && spec->total_f && spec->total_i;
    if (spec && ((!want_gdot && (spec->gam_dot_syn || spec->gam_dot_adi || spec->gam_dot_ssc)) ||
                 (!want_emit && (spec->syn_j || spec->syn_i || spec->ssc_j || spec->total_f || spec->total_i)))) {
        pbagpu::set_error_msg("ele_evolve_spectra: give all three gam_dot arrays or none, and all five spectra or none");
        return PBAGPU_E_ARG;
    }
    const pba_shock_params &s0 = pops[0].shock ? params[pops[0].run].rs : params[pops[0].run].fs;
    const int nf = s0.nfreq;
    if (nf < 2 || nf > MAXF) return PBAGPU_E_ARG;
    // every population's own electron grid size (its override or its shock's); NG = the largest
    auto pop_ng = [&](int q) {
        const pba_population &P = pops[q];
        const pba_shock_params &sp = P.shock ? params[P.run].rs : params[P.run].fs;
        return P.ngam > 0 ? P.ngam : sp.ngam;
    };
    const double h_by_mec2 = pbagpu::kH / pbagpu::kMe / pbagpu::kC / pbagpu::kC;
    // one grid set (and Compton table if any population of that grid needs it) per distinct
    // grid (gam2, freq1, freq2, ngam); runs of a batch usually share one
    struct Key { double gam2, freq1, freq2; int ng; };
    std::vector<Key> keys;
    std::vector<int> key_of(n_pop);
    std::vector<char> need_ssc;
    int NG = 0, ng_min = MAXG + 1;
    for (int q = 0; q < n_pop; q++) {
        const pba_population &P = pops[q];
        if (P.run < 0 || P.run >= n_runs || P.bw < 0 || P.bw >= bw_total || (P.shock != 0 && P.shock != 1))
            return PBAGPU_E_ARG;
        const pba_shock_params &sp = P.shock ? params[P.run].rs : params[P.run].fs;
        const int ngq = pop_ng(q);
        if (ngq < 2 || ngq > MAXG || P.ngam < 0 || sp.nfreq != nf) {
            pbagpu::set_error_msg("ele_evolve: a population's ngam must be 2..512 and nfreq equal in one call");
            return PBAGPU_E_ARG;
        }
        NG = std::max(NG, ngq);
        ng_min = std::min(ng_min, ngq);
        if (sp.syn < 0 || sp.syn > 3) return PBAGPU_E_UNSUPPORTED;
        if (sp.ssa && !d_aall) {
            pbagpu::set_error_msg("ele_evolve: a shock with ssa = 1 needs the absorption table d_aall");
            return PBAGPU_E_ARG;
        }
        size_t pos = 0;
        while (pos < keys.size() && !(keys[pos].gam2 == sp.gam2 && keys[pos].freq1 == sp.freq1 &&
                                      keys[pos].freq2 == sp.freq2 && keys[pos].ng == ngq))
            pos++;
        if (pos == keys.size()) { keys.push_back({sp.gam2, sp.freq1, sp.freq2, ngq}); need_ssc.push_back(0); }
        if (sp.ssc) need_ssc[pos] = 1;
        key_of[q] = (int)pos;
    }
    // grids of key kk at dgrids + goff[kk]: e [ng], half_e [ng-1], delta_grid [ng+1],
    // delta_grid_bar [ng], nu [nf]
    std::vector<size_t> goff(keys.size() + 1, 0);
    for (size_t kk = 0; kk < keys.size(); kk++) goff[kk + 1] = goff[kk] + 4 * (size_t)keys[kk].ng + nf;
    auto nu_of = [&](const double *base, int ngk) { return base + 4 * (size_t)ngk; };
    std::vector<double> hgrids(goff[keys.size()]);
    for (size_t kk = 0; kk < keys.size(); kk++) {
        const int ngk = keys[kk].ng;
        pba_shock_params sp = s0;
        sp.gam2 = keys[kk].gam2; sp.freq1 = keys[kk].freq1; sp.freq2 = keys[kk].freq2; sp.ngam = ngk;
        GridSet G;
        build_grids(sp, G);
        double *o = &hgrids[goff[kk]];
        std::copy(G.e.begin(), G.e.end(), o); o += ngk;
        std::copy(G.half_e.begin(), G.half_e.end(), o); o += ngk - 1;
        std::copy(G.dg.begin(), G.dg.end(), o); o += ngk + 1;
        std::copy(G.dgbar.begin(), G.dgbar.end(), o); o += ngk;
        std::copy(G.nu.begin(), G.nu.end(), o);
    }
    // Compton tables: slot t holds K [nf][ng][nf] then its integral [ng][nf], at dK + koff[t]
    size_t n_ssc = 0;
    std::vector<size_t> ssc_slot(keys.size(), 0), slot_key;
    for (size_t kk = 0; kk < keys.size(); kk++)
        if (need_ssc[kk]) { ssc_slot[kk] = n_ssc++; slot_key.push_back(kk); }
    std::vector<size_t> koff(n_ssc + 1, 0);
    for (size_t t = 0; t < n_ssc; t++) {
        const size_t ngk = keys[slot_key[t]].ng;
        koff[t + 1] = koff[t] + (size_t)nf * ngk * nf + ngk * nf;
    }
    auto *dgrids = (double *)pbagpu::workspace(pbagpu::WS_ELE_GRIDS, sizeof(double) * hgrids.size(), s);
    auto *dK = n_ssc ? (double *)pbagpu::workspace(pbagpu::WS_ELE_KTAB, sizeof(double) * koff[n_ssc], s) : nullptr;
    auto *dpops = (PopDesc *)pbagpu::workspace(pbagpu::WS_ELE_POPS, sizeof(PopDesc) * n_pop, s);
    if (!dgrids || (n_ssc && !dK) || !dpops) return PBAGPU_E_HIP;
    int rc = PBAGPU_OK;
    PBA_TRY(rc, hipMemcpyAsync(dgrids, hgrids.data(), sizeof(double) * hgrids.size(), hipMemcpyHostToDevice, s));
    pbagpu::KernelTimer ttab(PBAGPU_STAGE_SSC_TABLE), tele(PBAGPU_STAGE_ELE);
    if (n_ssc) ttab.start(s);
    for (size_t t = 0; t < n_ssc && rc == PBAGPU_OK; t++) {
        const size_t kk = slot_key[t];
        const int ngk = keys[kk].ng;
        const size_t ktab = (size_t)nf * ngk * nf;
        const double *base = dgrids + goff[kk];
        double *K = dK + koff[t];
        double *KI = K + ktab;
        const double *e = base, *nu = nu_of(base, ngk);
        hipLaunchKernelGGL(ssc_table_kernel, dim3((unsigned)((ktab + 255) / 256)), dim3(256), 0, s, nu, e, nf, ngk,
                           h_by_mec2, K);
        PBA_TRY(rc, hipGetLastError());
        hipLaunchKernelGGL(ssc_integral_kernel, dim3((unsigned)((ngk * nf + 255) / 256)), dim3(256), 0, s, nu, K, nf,
                           ngk, KI);
        PBA_TRY(rc, hipGetLastError());
    }
    // BF16 pair of every Compton table for the mixed-precision contraction: slot t's hi part at
    // dKb + kboff[t], its lo part nf * ng * kp entries after it
    const int kp = (nf + 31) / 32 * 32;
    std::vector<size_t> kboff(n_ssc + 1, 0);
    for (size_t t = 0; t < n_ssc; t++) kboff[t + 1] = kboff[t] + 2 * (size_t)nf * keys[slot_key[t]].ng * kp;
    unsigned short *dKb = nullptr;
    if (precision == 1 && n_ssc) {
        dKb = (unsigned short *)pbagpu::workspace(pbagpu::WS_ELE_KBF, sizeof(unsigned short) * kboff[n_ssc], s);
        if (!dKb) return PBAGPU_E_HIP;
        for (size_t t = 0; t < n_ssc && rc == PBAGPU_OK; t++) {
            const int ngk = keys[slot_key[t]].ng;
            const size_t kbf = (size_t)nf * ngk * kp;
            hipLaunchKernelGGL(ssc_table_bf16_kernel, dim3((unsigned)((kbf + 255) / 256)), dim3(256), 0, s,
                               (const double *)(dK + koff[t]), nf, ngk, kp, dKb + kboff[t], dKb + kboff[t] + kbf);
            PBA_TRY(rc, hipGetLastError());
        }
    }
    if (n_ssc) ttab.stop(s);
    std::vector<PopDesc> hp(n_pop);
    for (int q = 0; q < n_pop; q++) {
        const pba_population &P = pops[q];
        const pba_shock_params &sp = P.shock ? params[P.run].rs : params[P.run].fs;
        const int kk = key_of[q];
        const int ng = keys[kk].ng;
        double *base = dgrids + goff[kk];
        PopDesc &D = hp[q];
        D.T = d_table + (size_t)P.bw * PBA_NV * nt;
        D.e = base; D.half_e = base + ng; D.dg = base + ng + (ng - 1); D.dgbar = D.dg + (ng + 1); D.nu = D.dgbar + ng;
        D.K = (sp.ssc && need_ssc[kk]) ? dK + koff[ssc_slot[kk]] : nullptr;
        D.KI = D.K ? D.K + (size_t)nf * ng * nf : nullptr;
        D.jall = d_jall + (size_t)q * nt * nf;
        D.aall = sp.ssa ? d_aall + (size_t)q * nt * nf : nullptr;
        D.ssa = sp.ssa ? 1 : 0;
        D.syn = sp.syn;
        D.bes = pbagpu::bessel_third_constants();
        D.fall = d_fall ? d_fall + (size_t)q * nt * NG : nullptr;
        auto rows = [&](double *base, int n) { return base ? base + (size_t)q * nt * n : nullptr; };
        D.sp_synj = want_emit ? rows(spec->syn_j, nf) : nullptr;
        D.sp_syni = want_emit ? rows(spec->syn_i, nf) : nullptr;
        D.sp_sscj = want_emit ? rows(spec->ssc_j, nf) : nullptr;
        D.sp_totf = want_emit ? rows(spec->total_f, nf) : nullptr;
        D.sp_toti = want_emit ? rows(spec->total_i, nf) : nullptr;
        D.sp_gds = want_gdot ? rows(spec->gam_dot_syn, NG) : nullptr;
        D.sp_gda = want_gdot ? rows(spec->gam_dot_adi, NG) : nullptr;
        D.sp_gdc = want_gdot ? rows(spec->gam_dot_ssc, NG) : nullptr;
        D.p = sp.p; D.eps_e = sp.eps_e; D.shock = P.shock; D.ng = ng; D.nf = nf;
        D.max_substeps = sp.max_substeps; D.adi_loss = sp.adi_loss_ele; D.ssc = D.K != nullptr;
    }
    PBA_TRY(rc, hipMemcpyAsync(dpops, hp.data(), sizeof(PopDesc) * n_pop, hipMemcpyHostToDevice, s));
    PBA_TRY(rc, hipMemsetAsync(d_jall, 0, sizeof(double) * (size_t)n_pop * nt * nf, s));
    if (d_aall) PBA_TRY(rc, hipMemsetAsync(d_aall, 0, sizeof(double) * (size_t)n_pop * nt * nf, s));
    if (d_fall) PBA_TRY(rc, hipMemsetAsync(d_fall, 0, sizeof(double) * (size_t)n_pop * nt * NG, s));
    if (spec) {
        for (double *a : {spec->syn_j, spec->syn_i, spec->ssc_j, spec->total_f, spec->total_i})
            if (a) PBA_TRY(rc, hipMemsetAsync(a, 0, sizeof(double) * (size_t)n_pop * nt * nf, s));
        for (double *a : {spec->gam_dot_syn, spec->gam_dot_adi, spec->gam_dot_ssc})
            if (a) PBA_TRY(rc, hipMemsetAsync(a, 0, sizeof(double) * (size_t)n_pop * nt * NG, s));
    }
    // per-population state between the per-step launches
    const size_t np = (size_t)n_pop;
    const size_t n_dbl = np * NG * 3 + np * nf * (3 + SSC_JB) + np * 4 + (want_gdot ? np * 3 * NG : 0);
    double *wsd = (double *)pbagpu::workspace(pbagpu::WS_ELE_WORK, sizeof(double) * n_dbl + sizeof(int) * np * 3, s);
    // populations of each Compton table, for the batched contraction
    std::vector<int> hlist;
    std::vector<size_t> list_off(n_ssc + 1, 0);
    for (size_t t = 0; t < n_ssc; t++) {
        list_off[t] = hlist.size();
        for (int q = 0; q < n_pop; q++)
            if (hp[q].ssc && ssc_slot[key_of[q]] == t && need_ssc[key_of[q]]) hlist.push_back(q);
    }
    list_off[n_ssc] = hlist.size();
    int *dlist = hlist.empty() ? nullptr
                               : (int *)pbagpu::workspace(pbagpu::WS_ELE_LIST, sizeof(int) * hlist.size(), s);
    if (!wsd || (!hlist.empty() && !dlist)) return PBAGPU_E_HIP;
    // operands of the mixed-precision contraction (per population of the SSC lists)
    float *d_an = nullptr;
    unsigned short *d_wnh = nullptr, *d_wnl = nullptr;
    double *d_amax = nullptr, *d_wmax = nullptr;
    const int NCH = (NG - 1 + SSC_CH - 1) / SSC_CH;   // normalisation chunks of the largest grid
    if (precision == 1 && !hlist.empty()) {
        const size_t nl = hlist.size();
        char *mw = (char *)pbagpu::workspace(pbagpu::WS_ELE_MIXED,
                                             nl * (sizeof(float) * NCH * SSC_CH + sizeof(double) * (NCH + 1) +
                                                   2 * sizeof(unsigned short) * kp) + 256, s);
        if (!mw) return PBAGPU_E_HIP;
        // 16-byte vector loads read an, wnh and wnl: they come first, each a multiple of 64 bytes long
        d_an = (float *)mw;
        d_wnh = (unsigned short *)(d_an + nl * NCH * SSC_CH);
        d_wnl = d_wnh + nl * kp;
        d_amax = (double *)(d_wnl + nl * kp);
        d_wmax = d_amax + nl * NCH;
    }
    // surrogate: features, outputs and scales of the listed populations, fallback flags, counters
    const int dx = NG + nf + 1;
    float *d_sx = nullptr, *d_sy = nullptr;
    double *d_sscale = nullptr;
    int *d_fb = nullptr;
    SurStats *d_st = nullptr;
    if (precision == 2) {
        if (d_aall) {
            pbagpu::set_error_msg("ele_evolve_surrogate: synchrotron self-absorption is not implemented with the surrogate");
            return PBAGPU_E_UNSUPPORTED;
        }
        if (ng_min != NG) {
            pbagpu::set_error_msg("ele_evolve_surrogate: the populations of one call need one electron-grid size");
            return PBAGPU_E_ARG;
        }
        const pba_mlp &net = sur->net;
        if (pbagpu::mlp::check_network(&net) != PBAGPU_OK || net.dims[0] != dx || net.dims[net.n_layers] != nf ||
            !sur->y_offset || sur->check_every < 1 || !(sur->check_tol > 0.f)) {
            pbagpu::set_error_msg("ele_evolve_surrogate: network dims must be [ngam + nfreq + 1 ... nfreq]; y_offset, "
                                  "check_every >= 1 and check_tol > 0 are needed");
            return PBAGPU_E_ARG;
        }
        const size_t nl = std::max<size_t>(hlist.size(), 1);
        char *sw = (char *)pbagpu::workspace(pbagpu::WS_SUR, nl * (sizeof(float) * (dx + nf) + sizeof(double)) +
                                                                 sizeof(int) * np + sizeof(SurStats) + 1024, s);
        if (!sw) return PBAGPU_E_HIP;
        d_st = (SurStats *)sw;
        d_sscale = (double *)(sw + 256);
        d_sx = (float *)(d_sscale + nl);
        d_sy = d_sx + nl * dx;
        d_fb = (int *)(d_sy + nl * nf);
        PBA_TRY(rc, hipMemsetAsync(d_st, 0, sizeof(SurStats), s));
        PBA_TRY(rc, hipMemsetAsync(d_fb, 0, sizeof(int) * np, s));
    }
    PopState S;
    S.ngs = NG;
    S.f = wsd; S.ft = S.f + np * NG; S.tf = S.ft + np * NG; S.wt = S.tf + np * nf;
    S.jsyn = S.wt + np * nf; S.jssc = S.jsyn + np * nf; S.scal = S.jssc + np * nf * SSC_JB;
    S.dfde = S.scal + np * 4;
    S.gdot = want_gdot ? S.dfde + np * NG : nullptr;
    S.flag = (int *)(S.dfde + np * NG + (want_gdot ? np * 3 * NG : 0));
    PBA_TRY(rc, hipMemsetAsync(wsd, 0, sizeof(double) * n_dbl + sizeof(int) * np * 3, s));
    if (dlist) PBA_TRY(rc, hipMemcpyAsync(dlist, hlist.data(), sizeof(int) * hlist.size(), hipMemcpyHostToDevice, s));
    const size_t ssc_lds = sizeof(double) * (size_t)nf * 64;
    if (ssc_lds > 64 * 1024)
        (void)hipFuncSetAttribute((const void *)ssc_step_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)ssc_lds);
    // The per-step launches (3 or more per step) can be captured into one graph and launched once:
    // the library option PBAGPU_OPT_ELE_GRAPH, for the FP64 and BF16 contractions (the surrogate's
    // network grows its own workspace during the first steps, which a capture cannot record), on a
    // created stream (the null stream cannot be captured: its calls stay eager).
    const bool use_graph = pbagpu::option(PBAGPU_OPT_ELE_GRAPH) != 0 && precision != 2 && s != nullptr &&
                           s != hipStreamPerThread;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    if (rc == PBAGPU_OK) {
        tele.start(s);
        if (use_graph) PBA_TRY(rc, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        for (int it = 0; it < nt && rc == PBAGPU_OK; it++) {
            hipLaunchKernelGGL(ele_step_kernel, dim3(n_pop), dim3(64), 0, s, dpops, n_pop, it, nt, S);
            for (size_t t = 0; t < n_ssc; t++) {
                const int npk = (int)(list_off[t + 1] - list_off[t]);
                if (npk == 0) continue;
                size_t kk = 0;
                while (!(need_ssc[kk] && ssc_slot[kk] == t)) kk++;
                const int ng = keys[kk].ng;
                const double *base = dgrids + goff[kk];
                const double *e = base, *nu = nu_of(base, ng);
                const double *K = dK + koff[t];
                if (precision == 1) {
                    const size_t lo = list_off[t];
                    const int nch = (ng - 1 + SSC_CH - 1) / SSC_CH;   // chunks of this table's grid (<= NCH)
                    hipLaunchKernelGGL(ssc_prep_kernel, dim3((unsigned)npk), dim3(64), 0, s, dlist + lo, npk, n_pop, e, ng,
                                       nf, kp, nch, S, d_an + lo * NCH * SSC_CH, d_amax + lo * NCH, d_wnh + lo * kp,
                                       d_wnl + lo * kp, d_wmax + lo);
                    hipLaunchKernelGGL(ssc_mfma_kernel, dim3((unsigned)nf, (unsigned)((npk + 63) / 64)), dim3(256), 0,
                                       s, dlist + lo, npk, nu, ng, nf, kp, nch, dKb + kboff[t],
                                       dKb + kboff[t] + (size_t)nf * ng * kp, d_an + lo * NCH * SSC_CH,
                                       d_amax + lo * NCH, d_wnh + lo * kp, d_wnl + lo * kp, d_wmax + lo, S);
                } else if (precision == 2) {
                    const size_t lo = list_off[t];
                    hipLaunchKernelGGL(sur_feat_step_kernel, dim3((unsigned)npk), dim3(64), 0, s, dpops, dlist + lo, npk,
                                       n_pop, it, nt, ng, nf, S, d_sx, d_sscale);
                    int rm = pbagpu::mlp::forward_into(sur->net, npk, d_sx, d_sy, pbagpu::WS_MLP_ELE, s);
                    if (rm != PBAGPU_OK) { rc = rm; break; }
                    // the plain contraction for the sampled and the fallen-back populations
                    hipLaunchKernelGGL(ssc_step_kernel, dim3((unsigned)((npk + 63) / 64), (unsigned)nf, (unsigned)SSC_JB),
                                       dim3(64), ssc_lds, s, dlist + lo, npk, n_pop, K, e, nu, ng, nf, S,
                                       (const int *)d_fb, it, sur->check_every);
                    hipLaunchKernelGGL(sur_apply_kernel, dim3((unsigned)npk), dim3(64), 0, s, dlist + lo, npk, n_pop, it,
                                       nf, d_sy, sur->y_offset, sur->y_scale, d_sscale, S, d_fb, sur->check_every,
                                       (double)sur->check_tol, d_st);
                } else {
                    hipLaunchKernelGGL(ssc_step_kernel, dim3((unsigned)((npk + 63) / 64), (unsigned)nf, (unsigned)SSC_JB),
                                       dim3(64), ssc_lds, s, dlist + list_off[t], npk, n_pop, K, e, nu, ng, nf, S,
                                       (const int *)nullptr, 0, 0);
                }
            }
            hipLaunchKernelGGL(ele_emit_kernel, dim3(n_pop), dim3(64), 0, s, dpops, it, nt, S);
            PBA_TRY(rc, hipGetLastError());
        }
        if (use_graph) {
            // end the capture whatever happened inside it, then launch the step sequence once
            hipError_t e = hipStreamEndCapture(s, &graph);
            if (rc == PBAGPU_OK && e != hipSuccess) { pbagpu::set_error("hipStreamEndCapture", e); rc = PBAGPU_E_HIP; }
            if (rc == PBAGPU_OK) PBA_TRY(rc, hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0));
            if (rc == PBAGPU_OK) PBA_TRY(rc, hipGraphLaunch(graph_exec, s));
        }
        tele.stop(s);
        // status of every population
        if (rc == PBAGPU_OK)
            hipLaunchKernelGGL(ele_status_kernel, dim3((unsigned)((n_pop + 63) / 64)), dim3(64), 0, s, S.flag, n_pop,
                               d_status);
        PBA_TRY(rc, hipGetLastError());
    }
    SurStats hst{};
    if (d_st && rc == PBAGPU_OK) PBA_TRY(rc, hipMemcpyAsync(&hst, d_st, sizeof(SurStats), hipMemcpyDeviceToHost, s));
    PBA_TRY(rc, hipStreamSynchronize(s));
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    ttab.collect();
    tele.collect();
    if (stats) {
        stats->inferred = (long long)hst.n[0]; stats->checked = (long long)hst.n[1];
        stats->failed = (long long)hst.n[2]; stats->fallback = (long long)hst.n[3];
        stats->max_err = __builtin_bit_cast(double, hst.n[4]);
    }
    return rc;
}

extern "C" int pbagpu_ele_evolve(const pba_run_params *params, int n_runs, const pba_population *pops, int n_pop,
                                 const double *d_table, int bw_total, int nt, double *d_jall, double *d_aall,
                                 double *d_fall, int *d_status, int precision, void *stream) {
    if (precision != 0 && precision != 1) return PBAGPU_E_UNSUPPORTED;
    return ele_evolve_impl(params, n_runs, pops, n_pop, d_table, bw_total, nt, d_jall, d_aall, d_fall, d_status,
                           precision, nullptr, nullptr, nullptr, stream);
}

extern "C" int pbagpu_ele_evolve_spectra(const pba_run_params *params, int n_runs, const pba_population *pops,
                                         int n_pop, const double *d_table, int bw_total, int nt, double *d_jall,
                                         double *d_aall, double *d_fall, int *d_status, int precision,
                                         const pba_spectra *spec, void *stream) {
    if (precision != 0 && precision != 1) return PBAGPU_E_UNSUPPORTED;
    if (!spec) return PBAGPU_E_ARG;
    return ele_evolve_impl(params, n_runs, pops, n_pop, d_table, bw_total, nt, d_jall, d_aall, d_fall, d_status,
                           precision, nullptr, nullptr, spec, stream);
}

extern "C" int pbagpu_ele_evolve_surrogate(const pba_run_params *params, int n_runs, const pba_population *pops,
                                           int n_pop, const double *d_table, int bw_total, int nt, double *d_jall,
                                           double *d_fall, int *d_status, const pba_ssc_surrogate *sur,
                                           pba_surrogate_stats *stats, void *stream) {
    if (!sur) return PBAGPU_E_ARG;
    return ele_evolve_impl(params, n_runs, pops, n_pop, d_table, bw_total, nt, d_jall, nullptr, d_fall, d_status, 2,
                           sur, stats, nullptr, stream);
}

// grids and Compton table of one shock, for the batched entry points below
static int shock_tables(const pba_shock_params &sh, hipStream_t s, const double **e, const double **nu,
                        const double **K) {
    using namespace pbagpu::ele;
    const int ng = sh.ngam, nf = sh.nfreq;
    GridSet G;
    build_grids(sh, G);
    const size_t ktab = (size_t)nf * ng * nf;
    double *d = (double *)pbagpu::workspace(pbagpu::WS_SSCX, sizeof(double) * (ng + nf + ktab + (size_t)ng * nf), s);
    if (!d) return PBAGPU_E_HIP;
    int rc = PBAGPU_OK;
    PBA_TRY(rc, hipMemcpyAsync(d, G.e.data(), sizeof(double) * ng, hipMemcpyHostToDevice, s));
    PBA_TRY(rc, hipMemcpyAsync(d + ng, G.nu.data(), sizeof(double) * nf, hipMemcpyHostToDevice, s));
    double *dK = d + ng + nf;
    const double h_by_mec2 = pbagpu::kH / pbagpu::kMe / pbagpu::kC / pbagpu::kC;
    hipLaunchKernelGGL(ssc_table_kernel, dim3((unsigned)((ktab + 255) / 256)), dim3(256), 0, s, d + ng, d, nf, ng,
                       h_by_mec2, dK);
    PBA_TRY(rc, hipGetLastError());
    *e = d; *nu = d + ng; *K = dK;
    return rc;
}

extern "C" int pbagpu_ssc_features(const pba_shock_params *shock, int n, const double *d_f, const double *d_w,
                                   const double *d_B, float *d_x, double *d_scale, void *stream) {
    using namespace pbagpu::ele;
    if (!shock || n < 0 || (n > 0 && (!d_f || !d_w || !d_B || !d_x || !d_scale))) return PBAGPU_E_ARG;
    if (shock->ngam < 2 || shock->nfreq < 2) return PBAGPU_E_ARG;
    hipStream_t s = (hipStream_t)stream;
    int rc = PBAGPU_OK;
    if (n > 0) {
        hipLaunchKernelGGL(ssc_features_kernel, dim3((unsigned)n), dim3(64), 0, s, d_f, d_w, d_B, shock->ngam,
                           shock->nfreq, d_x, d_scale);
        PBA_TRY(rc, hipGetLastError());
    }
    PBA_TRY(rc, hipStreamSynchronize(s));
    return rc;
}

extern "C" int pbagpu_ssc_emissivity(const pba_shock_params *shock, int n, const double *d_f, const double *d_w,
                                     double *d_j, void *stream) {
    using namespace pbagpu::ele;
    if (!shock || n < 0 || (n > 0 && (!d_f || !d_w || !d_j))) return PBAGPU_E_ARG;
    const int ng = shock->ngam, nf = shock->nfreq;
    if (ng < 2 || ng > MAXG || nf < 2 || nf > MAXF) return PBAGPU_E_ARG;
    if (n == 0) return PBAGPU_OK;
    hipStream_t s = (hipStream_t)stream;
    const double *e, *nu, *K;
    int rc = shock_tables(*shock, s, &e, &nu, &K);
    if (rc != PBAGPU_OK) return rc;
    // cells go through the contraction in chunks, laid out as the electron stage's state
    const int CH = std::min(n, 16384);
    const size_t per = (size_t)CH * (ng + nf + (size_t)nf * SSC_JB);
    char *w = (char *)pbagpu::workspace(pbagpu::WS_SUR, sizeof(double) * per + sizeof(int) * 4 * (size_t)CH + 512, s);
    if (!w) return PBAGPU_E_HIP;
    PopState S{};
    S.ft = (double *)w;
    S.wt = S.ft + (size_t)CH * ng;
    S.jssc = S.wt + (size_t)CH * nf;
    S.flag = (int *)(S.jssc + (size_t)CH * nf * SSC_JB);
    int *plist = S.flag + 3 * (size_t)CH;
    hipLaunchKernelGGL(fill_int_kernel, dim3((unsigned)((3 * CH + 255) / 256)), dim3(256), 0, s, S.flag, 3 * CH, 1, 0);
    hipLaunchKernelGGL(fill_int_kernel, dim3((unsigned)((CH + 255) / 256)), dim3(256), 0, s, plist, CH, 0, 1);
    const size_t ssc_lds = sizeof(double) * (size_t)nf * 64;
    if (ssc_lds > 64 * 1024)
        (void)hipFuncSetAttribute((const void *)ssc_step_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)ssc_lds);
    for (int c0 = 0; c0 < n && rc == PBAGPU_OK; c0 += CH) {
        const int m = std::min(CH, n - c0);
        hipLaunchKernelGGL(transpose_kernel, dim3(1024), dim3(256), 0, s, d_f + (size_t)c0 * ng, m, ng, S.ft);
        hipLaunchKernelGGL(transpose_kernel, dim3(1024), dim3(256), 0, s, d_w + (size_t)c0 * nf, m, nf, S.wt);
        hipLaunchKernelGGL(ssc_step_kernel, dim3((unsigned)((m + 63) / 64), (unsigned)nf, (unsigned)SSC_JB), dim3(64),
                           ssc_lds, s, plist, m, m, K, e, nu, ng, nf, S, (const int *)nullptr, 0, 0);
        hipLaunchKernelGGL(sum_jb_kernel, dim3(1024), dim3(256), 0, s, S.jssc, (size_t)m * nf, d_j + (size_t)c0 * nf);
        PBA_TRY(rc, hipGetLastError());
    }
    PBA_TRY(rc, hipStreamSynchronize(s));
    return rc;
}

// Host evaluation of the Bessel kernel's K_{1/3}, K_{4/3} routine (the code the electron stage
// runs on the device), for checks against libm / scipy.  Touches no GPU.
extern "C" int pbagpu_bessel_k_thirds(int n, const double *x, double *k13, double *k43) {
    if (n < 0 || (n > 0 && (!x || !k13 || !k43))) return PBAGPU_E_ARG;
    const pbagpu::BesselThird b = pbagpu::bessel_third_constants();
    for (int i = 0; i < n; i++) pbagpu::bessel_k_thirds(x[i], b, k13[i], k43[i]);
    return PBAGPU_OK;
}

extern "C" int pbagpu_synchrotron_f(int n, const double *x, double *f) {
    if (n < 0 || (n > 0 && (!x || !f))) return PBAGPU_E_ARG;
    for (int i = 0; i < n; i++) f[i] = pbagpu::ele::synchrotron_F(x[i]);
    return PBAGPU_OK;
}
