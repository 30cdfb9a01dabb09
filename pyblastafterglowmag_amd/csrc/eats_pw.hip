// Stage 3, piece-wise method: the light curve of a layer as a sum over its phi cells.
//
// One lane per (layer, t_obs, nu_obs) pair; one workgroup per (layer, block of PW_WG
// consecutive pairs).  The workgroup stages the layer's table columns and the comoving
// frequency grid in LDS once; every lane then walks the layer's 2i+1 cells in the order of
// EATS::evalSkyMapPW (src/eats.h:1064-1213), primary jet then counter-jet
// (src/eats.h:859-883):
//   * the arrival time of every step, ttobs[k] = tt[k] + R[k]/c (1 - mu), is not stored:
//     it is evaluated only at the steps the bisection of findClosestIndex probes
//     (src/utilitites/interpolators.h:37-68), so one cell costs O(log nt), not O(nt);
//   * the previous cell's index is the initial guess of the next search, as there;
//   * the cell flux follows fluxDensPieceWiseWithComov (src/blastwave/blastwave_radiation.h:14-72)
//     and ElectronAndRadiation::fluxDens / computeFluxDensity
//     (src/microphysics/microphysics.h:799-867, 1185-1217), with the cell angle and the
//     jet's total cell count in the shell-thickness and luminosity terms and the
//     forward-shock emissivity for both shocks.
// With an image buffer the same kernel also writes every cell's raw sky-map entry (radius,
// angles, mu, sky-plane position in mas and intensity), as evalSkyMapPW fills its `out`
// arrays (src/eats.h:1163-1184); cells that do not contribute hold zeros, as there.
#include "pba_device.h"
#include "timing.h"
#include "runtime.h"
#include <cfloat>
#include <vector>

namespace pbagpu {
namespace eatspw {

constexpr int PW_WG = 256;

// Table columns the cell flux reads, staged in LDS per workgroup.
enum Col { C_tt, C_R, C_Gamma, C_beta, C_M2, C_rho2, C_thrs, C_M3, C_rho3, NCOL };
__device__ __forceinline__ constexpr int col_var(int c) {
    return c == C_tt ? PV_tt : c == C_R ? PV_R : c == C_Gamma ? PV_Gamma : c == C_beta ? PV_beta
         : c == C_M2 ? PV_M2 : c == C_rho2 ? PV_rho2 : c == C_thrs ? PV_thickness_rs : c == C_M3 ? PV_M3 : PV_rho3;
}

struct PwArgs {
    const double *T;    // table [PBA_NV][nt]
    const double *jfs;  // forward-shock emissivity [nt][nf]
    const double *afs;  // forward-shock absorption [nt][nf] (use_ssa) or null
    const double *nu;   // forward-shock comoving frequency grid [nf]
    double ctheta0, ncells;
    double z, theta_obs, d_l;
    int cil, rs, i_end_r, counter_jet;
    int img_off;     // first cell of this layer in the image (cells of earlier layers before it)
    int remove_mu;   // skymap_remove_mu
};

// Raw sky-map image of one (t_obs, nu_obs): PBA_IMG_NV arrays of 2 ncells entries, primary
// jet cells first, then the counter-jet cells.
struct Img {
    double *base;    // null: light curve only
    int ncells;
};

extern __shared__ double lds_pw[];

struct Lane {
    int nt, nf, N;    // N = i_end_r
    int status;
    __device__ double V(int c, int k) const { return lds_pw[c * nt + k]; }
    __device__ double nu(int i) const { return lds_pw[NCOL * nt + i]; }
    // arrival time of step k for a cell with 1 - mu = omu; the steps past i_end_r hold
    // DBL_MAX in the reference's vector.  Product and sum are rounded separately, as there
    // (hipcc contracts across statements by default, and __dadd_rn / __dmul_rn are plain
    // operators in its headers: only the pragma keeps the multiply-add unfused).
    __device__ double ttobs(int k, double omu) const {
#pragma clang fp contract(off)
        if (k >= N) return DBL_MAX;
        const double d = V(C_R, k) / kC * omu;
        return V(C_tt, k) + d;
    }
    // findClosestIndex over the nt-long arrival-time vector
    __device__ int closest(double t, double omu, int guess) const {
        int low = 0, high = nt - 1, best_i = guess;
        double best = fabs(ttobs(guess, omu) - t);
        while (low <= high) {
            int mid = low + (high - low) / 2;
            double v = ttobs(mid, omu);
            double d = fabs(v - t);
            if (d < best) { best = d; best_i = mid; }
            if (v == t) return mid;
            if (v < t) low = mid + 1; else high = mid - 1;
        }
        return best_i;
    }
    // fluxDens + computeFluxDensity (useNe, sameAsBW, smooth optical depth)
    __device__ double shock_flux(const double *jall, const double *aall, double freq, int ia, int ib, double Gamma,
                                 double mu, double r, double ne, double n_prime, double theta, double ncells) {
        if (freq >= nu(nf - 1) || freq <= nu(0)) { status |= PBAGPU_S_NU_OUT_OF_GRID; return 0.; }
        auto grid = [&](int i) { return nu(i); };
        const int ia_nu = find_index(freq, grid, nf), ib_nu = ia_nu + 1;
        const double x1 = nu(ia_nu), x2 = nu(ib_nu), y1 = V(C_R, ia), y2 = V(C_R, ib);
        const double em_prime = table_bilinear(jall, nf, ia, ib, ia_nu, ib_nu, x1, x2, y1, y2, freq, r);
        const double beta = beta_from_gamma(Gamma);
        const double delta_D = Gamma * (1.0 - beta * mu);
        const double omc = (1. - cos(theta) / ncells);
        const double dtau = (mu == beta) ? 1.0e100
                          : aall ? absorption_depth(aall, nf, ia, ib, ia_nu, ib_nu, x1, x2, y1, y2, freq, r, Gamma,
                                                    beta, delta_D, mu, ne, n_prime, omc)
                                 : 0.;
        bool bad = false;
        const double lum = shock_luminosity(em_prime, dtau, Gamma, beta, delta_D, mu, ne, omc, bad);
        if (bad) status |= PBAGPU_S_NONFINITE;
        return lum;
    }
};

// imageXXs / imageYYs and their counter-jet forms (src/image.h:379-392)
__device__ __forceinline__ void image_xy(bool cj, double theta, double phi, double sa, double ca, double &x,
                                         double &y) {
    if (!cj) {
        x = -1.0 * ca * (sin(theta) * sin(phi)) + sa * cos(theta);
        y = +1. * sin(theta) * cos(phi);
    } else {
        x = -1.0 * ca * sin(kPi - theta) * sin(phi) + sa * cos(kPi - theta);
        y = +1. * sin(kPi - theta) * cos(phi);
    }
}

constexpr double kRad2mas = 206264806.247;

// One jet of EATS::evalSkyMapPW: the sum of the cell fluxes, mJy.
__device__ double sky_map_pw(Lane &L, const PwArgs &A, double t_obs, double nu_obs, double sa, double ca, double st,
                             double ct, bool cj, const Img &im, unsigned long long &cells) {
    const int N = L.N;
    double *img = im.base ? im.base + (cj ? im.ncells : 0) + A.img_off : nullptr;
    const size_t vs = 2 * (size_t)im.ncells;   // stride between the image arrays
    if (L.V(C_R, 0) == 0. && L.V(C_Gamma, 0) == 0.) return 0.;
    if (L.V(C_tt, 0) == 0. && N < L.nt && L.V(C_tt, N) == 0.) L.status |= PBAGPU_S_NONFINITE;
    double tot = 0.;
    int ia_prev = 0;
    for (int i = 0; i < A.cil; i++) {
        const double cphi = (double)i * 2.0 * M_PI / (double)A.cil;
        // obsAngle / obsAngleCJ (src/image.h:361-378); st, ct are sin and cos of the cell angle,
        // or of pi minus it for the counter-jet
        const double mu = sa * (st * sin(cphi)) + ca * ct;
        const double omu = 1.0 - mu;
        int guess = i > 0 ? ia_prev : L.nt / 2;
        ia_prev = 0;   // a cell that ends before t_obs leaves its index at 0
        if (t_obs < L.ttobs(0, omu)) { L.status |= PBAGPU_S_NONFINITE; continue; }
        if (t_obs > L.ttobs(N - 1, omu)) continue;
        const int ia = L.closest(t_obs, omu, guess);
        ia_prev = ia;
        if (ia >= N - 1) continue;
        const int ib = ia + 1;
        const double ta = L.ttobs(ia, omu), tb = L.ttobs(ib, omu);
        // one exponent for every log-log interpolation of this cell (interpSegLog, interpolators.h:114-133)
        const double w = log(t_obs / ta) / log(tb / ta);
        auto lg = [&](int c) {
            double ya = L.V(c, ia);
            return ya * pow(L.V(c, ib) / ya, w);
        };
        const double r = lg(C_R);
        if (r <= 0.0 || !isfinite(r)) continue;
        const double Gamma = lg(C_Gamma), beta = lg(C_beta);
        const double nprime = lg(C_rho2) / kMp, nprotons = lg(C_M2) / kMp;
        const double nuprime = (1.0 + A.z) * nu_obs * (Gamma * (1.0 - beta * mu));
        double f = L.shock_flux(A.jfs, A.afs, nuprime, ia, ib, Gamma, mu, r, nprotons, nprime, A.ctheta0, A.ncells);
        if (A.rs && L.V(C_thrs, ia) > 0 && L.V(C_thrs, ib) > 0) {
            const double nprime_rs = lg(C_rho3) / kMp, ne_rs = lg(C_M3) / kMp;
            f += L.shock_flux(A.jfs, A.afs, nuprime, ia, ib, Gamma, mu, r, ne_rs, nprime_rs, A.ctheta0, A.ncells);
        }
        f *= (1.0 + A.z) / (2.0 * A.d_l * A.d_l);
        if (!isfinite(f) || f < 0) L.status |= PBAGPU_S_NONFINITE;
        if (img) {
            // fluxDensPieceWiseWithComov's ctheta: linear in time between the bracketing steps
            const double *cth = A.T + (size_t)PV_ctheta * L.nt;
            const double ctheta = cth[ia] + (cth[ib] - cth[ia]) * (t_obs - ta) / (tb - ta);
            double x, y;
            image_xy(cj, ctheta, cphi, sa, ca, x, y);
            double inten = f / (r * r) * kCgs2mJy;
            if (!A.remove_mu) inten /= fabs(mu);
            img[PBA_IMG_r * vs + i] = r;
            img[PBA_IMG_ctheta * vs + i] = A.ctheta0;
            img[PBA_IMG_cphi * vs + i] = cphi;
            img[PBA_IMG_mu * vs + i] = mu;
            img[PBA_IMG_xrs * vs + i] = r * x * (kRad2mas / A.d_l);
            img[PBA_IMG_yrs * vs + i] = r * y * (kRad2mas / A.d_l);
            img[PBA_IMG_intensity * vs + i] = inten * (A.d_l * A.d_l / kRad2mas / kRad2mas);
        }
        tot += f;
        cells++;
    }
    return tot * kCgs2mJy;
}

// grid: (ceil(n_pairs / PW_WG), n_layers)
__global__ void __launch_bounds__(PW_WG) eats_pw_kernel(const PwArgs *__restrict__ layers, int n_pairs, int nt, int nf,
                                                        const double *__restrict__ tobs,
                                                        const double *__restrict__ nuobs, double *__restrict__ out,
                                                        int *__restrict__ status,
                                                        unsigned long long *__restrict__ n_cells, double *image,
                                                        int ncells) {
    const int l = blockIdx.y;
    const PwArgs A = layers[l];
    for (int idx = threadIdx.x; idx < NCOL * nt; idx += PW_WG) {
        const int c = idx / nt, k = idx - c * nt;
        lds_pw[idx] = A.T[(size_t)col_var(c) * nt + k];
    }
    for (int i = threadIdx.x; i < nf; i += PW_WG) lds_pw[NCOL * nt + i] = A.nu[i];
    __syncthreads();

    const int k = blockIdx.x * PW_WG + threadIdx.x;
    if (k >= n_pairs) return;
    Lane L;
    L.nt = nt; L.nf = nf; L.N = A.i_end_r; L.status = 0;
    double fl = 0.;
    unsigned long long cells = 0;
    if (image) {   // this layer's cells of both jets start at zero, as the reference's image arrays
        double *b = image + (size_t)k * PBA_IMG_NV * 2 * ncells + A.img_off;
        for (int v = 0; v < PBA_IMG_NV; v++)
            for (int i = 0; i < A.cil; i++) {
                b[(size_t)v * 2 * ncells + i] = 0.;
                b[(size_t)v * 2 * ncells + ncells + i] = 0.;
            }
    }
    if (A.i_end_r > 0) {   // evalEjectaLightCurves skips blastwaves that never evolved
        const double t = tobs[k], nu = nuobs[k];
        // evalSkyMapPW(t_obs, nu_obs, offset): t_obs inside the evolved burster times
        if (L.V(C_R, 0) > 0) {
            const double tb0 = A.T[(size_t)PV_tburst * nt], tb1 = A.T[(size_t)PV_tburst * nt + A.i_end_r - 1];
            if (t < tb0 || t > tb1) L.status |= PBAGPU_S_NONFINITE;
        }
        const double th = A.ctheta0, th_cj = kPi - A.ctheta0;
        const double sa = sin(A.theta_obs), ca = cos(A.theta_obs);
        Img im;
        im.base = image ? image + (size_t)k * PBA_IMG_NV * 2 * ncells : nullptr;
        im.ncells = ncells;
        const double pj = sky_map_pw(L, A, t, nu, sa, ca, sin(th), cos(th), false, im, cells);
        double cjf = 0.;
        if (A.counter_jet) cjf = sky_map_pw(L, A, t, nu, sa, ca, sin(th_cj), cos(th_cj), true, im, cells);
        fl = pj + cjf;
    }
    out[(size_t)l * n_pairs + k] = fl;
    if (L.status) atomicOr(&status[l], L.status);
    if (n_cells) {
        // one atomic per wavefront
        unsigned long long s = cells;
        for (int d = 32; d > 0; d >>= 1) s += __shfl_down(s, d, 64);
        if ((threadIdx.x & 63) == 0) atomicAdd(n_cells, s);
    }
}

}  // namespace eatspw
}  // namespace pbagpu

namespace pbagpu {
namespace eatspw {

// Shared host part of the light-curve and sky-map entries; image == nullptr for light curves.
static int launch(const pba_run_params *params, int n_runs, const pba_eats_pw_layer *layers, int n_layers,
                  const double *d_table, int nt, const double *d_jall, const double *d_aall, int nfreq, int n_pairs, const double *d_tobs,
                  const double *d_nuobs, double *d_out, int *d_status, double *d_image, int remove_mu,
                  hipStream_t s) {
    if (!params || !layers || !d_table || !d_jall || !d_tobs || !d_nuobs || !d_out || !d_status || n_runs < 1 ||
        n_layers < 1 || n_pairs < 1 || nt < 2 || nfreq < 2 || n_layers > 65535)
        return PBAGPU_E_ARG;
    const size_t lds = sizeof(double) * ((size_t)NCOL * nt + (size_t)nfreq);
    if (lds > 160 * 1024) {
        set_error_msg("eats_pw: time grid too long for the LDS-staged table (nt <= 2200)");
        return PBAGPU_E_ARG;
    }
    long ncells = 0;
    for (int l = 0; l < n_layers; l++) {
        const pba_eats_pw_layer &Y = layers[l];
        if (Y.run < 0 || Y.run >= n_runs || Y.i_end_r > nt || Y.i_end_r < 0 || Y.pop_fs < 0 || Y.bw < 0 ||
            Y.cil < 1 || !(Y.ncells >= 1.0)) {
            set_error_msg("eats_pw: bad layer descriptor");
            return PBAGPU_E_ARG;
        }
        ncells += Y.cil;
    }
    if (d_image && ((size_t)n_pairs * PBA_IMG_NV * 2 * ncells > ((size_t)1 << 40) || ncells > (1L << 30))) {
        set_error_msg("eats_pw: sky-map image too large");
        return PBAGPU_E_ARG;
    }
    for (int r = 0; r < n_runs; r++)
        if (params[r].fs.nfreq != nfreq) return PBAGPU_E_ARG;
    // forward-shock frequency grid of every run (MakeLogspaceVec, utils.h:91-105)
    std::vector<double> hnu((size_t)n_runs * nfreq);
    for (int r = 0; r < n_runs; r++) {
        const pba_shock_params &sp = params[r].fs;
        double cur = std::pow(10.0, std::log10(sp.freq1));
        const double ratio = std::pow(10.0, (std::log10(sp.freq2) - std::log10(sp.freq1)) / sp.nfreq);
        for (int i = 0; i < nfreq; i++) { hnu[(size_t)r * nfreq + i] = cur; cur *= ratio; }
    }
    auto *dnu = (double *)workspace(WS_PW_NU, sizeof(double) * hnu.size(), s);
    auto *dL = (PwArgs *)workspace(WS_PW_LAYERS, sizeof(PwArgs) * n_layers, s);
    if (!dnu || !dL) return PBAGPU_E_HIP;
    std::vector<PwArgs> hl(n_layers);
    int off = 0;
    for (int l = 0; l < n_layers; l++) {
        const pba_eats_pw_layer &Y = layers[l];
        const pba_run_params &P = params[Y.run];
        PwArgs &A = hl[l];
        A.T = d_table + (size_t)Y.bw * PBA_NV * nt;
        A.jfs = d_jall + (size_t)Y.pop_fs * nt * nfreq;
        A.afs = (d_aall && P.fs.ssa) ? d_aall + (size_t)Y.pop_fs * nt * nfreq : nullptr;
        A.nu = dnu + (size_t)Y.run * nfreq;
        A.ctheta0 = Y.ctheta0; A.ncells = Y.ncells;
        A.z = P.z; A.theta_obs = P.theta_obs; A.d_l = P.d_l;
        A.cil = Y.cil; A.rs = Y.rs ? 1 : 0; A.i_end_r = Y.i_end_r; A.counter_jet = P.counter_jet;
        A.img_off = off; A.remove_mu = remove_mu ? 1 : 0;
        off += Y.cil;
    }
    int rc = PBAGPU_OK;
    PBA_TRY(rc, hipMemcpyAsync(dnu, hnu.data(), sizeof(double) * hnu.size(), hipMemcpyHostToDevice, s));
    PBA_TRY(rc, hipMemcpyAsync(dL, hl.data(), sizeof(PwArgs) * n_layers, hipMemcpyHostToDevice, s));
    PBA_TRY(rc, hipMemsetAsync(d_status, 0, sizeof(int) * n_layers, s));
    KernelTimer timer(PBAGPU_STAGE_EATS_PW);
    unsigned long long *d_cells = nullptr;
    if (rc == PBAGPU_OK && timing_enabled()) {
        d_cells = (unsigned long long *)workspace(WS_PW_SAMPLES, sizeof(unsigned long long), s);
        if (d_cells) PBA_TRY(rc, hipMemsetAsync(d_cells, 0, sizeof(unsigned long long), s));
    }
    if (rc == PBAGPU_OK) {
        (void)hipFuncSetAttribute((const void *)eats_pw_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        dim3 grid((unsigned)((n_pairs + PW_WG - 1) / PW_WG), (unsigned)n_layers);
        timer.start(s);
        hipLaunchKernelGGL(eats_pw_kernel, grid, dim3(PW_WG), lds, s, dL, n_pairs, nt, nfreq, d_tobs, d_nuobs, d_out,
                           d_status, d_cells, d_image, (int)ncells);
        PBA_TRY(rc, hipGetLastError());
        timer.stop(s);
        if (d_cells && rc == PBAGPU_OK) {
            unsigned long long hc = 0;
            PBA_TRY(rc, hipMemcpyAsync(&hc, d_cells, sizeof(hc), hipMemcpyDeviceToHost, s));
            PBA_TRY(rc, hipStreamSynchronize(s));
            timing_add_units(PBAGPU_STAGE_EATS_PW, (double)hc);
        }
    }
    PBA_TRY(rc, hipStreamSynchronize(s));
    timer.collect();
    return rc;
}

}  // namespace eatspw
}  // namespace pbagpu

extern "C" int pbagpu_eats_lightcurve_pw(const pba_run_params *params, int n_runs, const pba_eats_pw_layer *layers,
                                         int n_layers, const double *d_table, int nt, const double *d_jall,
                                         const double *d_aall, int nfreq, int n_pairs, const double *d_tobs,
                                         const double *d_nuobs, double *d_out, int *d_status, void *stream) {
    return pbagpu::eatspw::launch(params, n_runs, layers, n_layers, d_table, nt, d_jall, d_aall, nfreq, n_pairs, d_tobs,
                                  d_nuobs, d_out, d_status, nullptr, 0, (hipStream_t)stream);
}

extern "C" int pbagpu_eats_skymap_pw(const pba_run_params *params, const pba_eats_pw_layer *layers, int n_layers,
                                     const double *d_table, int nt, const double *d_jall, const double *d_aall,
                                     int nfreq, int n_images,
                                     const double *d_tobs, const double *d_nuobs, int remove_mu, double *d_image,
                                     double *d_flux, int *d_status, void *stream) {
    if (!d_image || !layers || n_layers < 1) return PBAGPU_E_ARG;
    for (int l = 0; l < n_layers; l++)
        if (layers[l].run != layers[0].run) {
            pbagpu::set_error_msg("eats_skymap_pw: the layers of one image belong to one run");
            return PBAGPU_E_ARG;
        }
    if (layers[0].run != 0) {
        pbagpu::set_error_msg("eats_skymap_pw: layers refer to params[0] (run = 0)");
        return PBAGPU_E_ARG;
    }
    return pbagpu::eatspw::launch(params, 1, layers, n_layers, d_table, nt, d_jall, d_aall, nfreq, n_images, d_tobs,
                                  d_nuobs, d_flux, d_status, d_image, remove_mu, (hipStream_t)stream);
}
