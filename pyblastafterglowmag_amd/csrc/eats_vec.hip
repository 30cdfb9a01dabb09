// Stage 3, frequency-vectorised: the observer light curves of every layer at every (t_obs, nu_obs)
// of a tensor grid, with one wavefront per (layer, t_obs) integrating all nu_obs at once.
//
// The integrand is the reference's (EATS::integrand + fluxDensAdaptiveWithComov +
// ElectronAndRadiation::fluxDens / computeFluxDensity, src/eats.h:175-458,
// src/blastwave/blastwave_radiation.h:75-141, src/microphysics/microphysics.h:799-867,
// 1185-1217) over the same domain (phi in [0, pi] times 2; 1 - cos(theta) between the layer edges
// of find_jet_edge, src/eats.h:103-173, 460-684; primary jet and counter-jet).  What differs is the
// quadrature: the reference integrates every (t_obs, nu_obs) pair with its own nested CADRE tree,
// so the surface geometry of a sample (mu, the arrival-time search, the state interpolation at the
// emission time) is recomputed for each of the 96 frequencies.  Here the nested quadrature is an
// adaptive Gauss-Kronrod 7-15 rule in phi and in 1 - cos(theta) shared by all frequencies: an
// interval is accepted when |K15 - G7| <= rtol * |I| * (its share of the domain) for every
// frequency, otherwise bisected (depth first, at most 12 levels).  The 15 nodes of an interval are
// evaluated by 15 lanes at once (geometry) and broadcast with readlane; every lane then evaluates
// its frequencies (lane, lane + 64) at all 15 nodes.  Arithmetic is FP64; the difference from the
// reference's light curve is the quadrature tolerance, stated by the caller.
#include "pba_device.h"
#include "timing.h"
#include "runtime.h"
#include <vector>

namespace pbagpu {
namespace eatsv {

constexpr int WAVE = 64;
constexpr int WPB = 4;           // wavefronts per workgroup: 4 t_obs of one layer
constexpr int NU_PER_LANE = 2;   // frequencies per lane: lane and lane + 64
constexpr int MAX_NU = WAVE * NU_PER_LANE;
constexpr int MAX_DEPTH = 12;    // bisection levels per integral
constexpr int NG = 15;           // Gauss-Kronrod nodes per interval

// Gauss-Kronrod 7-15 rule on [-1, 1]: Kronrod abscissae x[0..7] (x[1], x[3], x[5], x[7] = 0 are the
// 7-point Gauss abscissae), Kronrod weights wk[0..7], Gauss weights wg for x[1], x[3], x[5], x[7].
__constant__ double gk_x[8] = {0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
                               0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
                               0.586087235467691130294144845693013, 0.405845151377397166906606412076961,
                               0.207784955007898467600689403773245, 0.0};
__constant__ double gk_wk[8] = {0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
                                0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
                                0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
                                0.204432940075298892414161999234649, 0.209482141084727828012999174891714};
__constant__ double gk_wg[4] = {0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
                                0.381830050505118944950369775488975, 0.417959183673469387755102040816327};

// node k = 0 .. 14 of an interval: abscissa index m = min(k, 14 - k), sign -1 for k < 7
__device__ __forceinline__ double gk_node(int k) { const int m = k < 7 ? k : 14 - k; return k < 7 ? -gk_x[m] : gk_x[m]; }
__device__ __forceinline__ double gk_kw(int k) { return gk_wk[k < 7 ? k : 14 - k]; }
__device__ __forceinline__ double gk_gw(int k) { const int m = k < 7 ? k : 14 - k; return (m & 1) ? gk_wg[m >> 1] : 0.0; }

// value of `v` in lane k (k uniform over the wavefront)
__device__ __forceinline__ double bcast(double v, int k) {
    const unsigned long long u = __double_as_longlong(v);
    const unsigned lo = __builtin_amdgcn_readlane((unsigned)u, k), hi = __builtin_amdgcn_readlane((unsigned)(u >> 32), k);
    return __longlong_as_double(((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ int bcast_i(int v, int k) { return __builtin_amdgcn_readlane(v, k); }

struct VLayer {
    const double *T;     // dynamics table [PBA_NV][nt]
    const double *jfs;   // emissivity [nt][nf]
    const double *jrs;   // or null
    const double *nufs, *nurs;
    double theta_c_l, theta_c_h, z, theta_obs, d_l;
    int counter_jet, i_end_r, nt, nf;
};

// staged per workgroup: t_burst, R, theta columns of the layer, then its two frequency grids
extern __shared__ double vlds[];

struct Ctx {
    const double *T;
    int nt, nf, i_end_r;
    double z, tz, sa, ca, mu0, muN;
    double sphi;   // sin(phi) of the current outer node
    double lnu0[2], ilr[2];   // ln nu_0 and (nf - 1) / ln(nu_{nf-1} / nu_0) of the two comoving grids
    bool cj, has_rs;
    __device__ double tb(int i) const { return vlds[i]; }
    __device__ double R(int i) const { return vlds[nt + i]; }
    __device__ double th(int i) const { return vlds[2 * nt + i]; }
    __device__ double nufs(int i) const { return vlds[3 * nt + i]; }
    __device__ double nurs(int i) const { return vlds[3 * nt + nf + i]; }
    // 1 / (nu_{i+1} - nu_i) of the two grids
    __device__ double idfs(int i) const { return vlds[3 * nt + 2 * nf + i]; }
    __device__ double idrs(int i) const { return vlds[3 * nt + 3 * nf + i]; }
    __device__ double col(int v, int i) const { return T[(size_t)v * nt + i]; }
    __device__ double mu_at(int i) const { return (tb(i) - tz) / R(i) * kC; }
    __device__ int mu_index(double mu) const {
        auto arr = [&](int i) { return mu_at(i); };
        return find_index(mu, arr, i_end_r);
    }
    // ImageExtend obsAngle / obsAngleCJ (src/image.h:361-378)
    __device__ double obs_angle(double theta, double sphi_) const {
        if (!cj) return sa * (sin(theta) * sphi_) + ca * cos(theta);
        return sa * (sin(kPi - theta) * sphi_) + ca * cos(kPi - theta);
    }
};

// One node of the inner integral: the surface geometry at 1 - cos(theta) = x (EATS::integrand and
// fluxDensAdaptiveWithComov up to the emissivity lookup).  A node with valid = 0 contributes 0.
// Everything here depends on the node only, so the per-frequency work of a node is a grid lookup,
// four table loads and a bilinear blend.
struct Node {
    int valid, ia;
    double Dnu;        // Gamma (1 - beta mu) with the interpolated beta: nu' = (1 + z) nu Dnu
    double lDnu;       // ln Dnu: ln nu' = ln((1 + z) nu) + ln Dnu estimates the frequency-grid index
    double iD2;        // 1 / D^2, D the Doppler factor with beta from the interpolated Gamma
    double wy2, wy1;   // radial weights (r2 - r) / (r2 - r1) and (r - r1) / (r2 - r1) of the bracketing steps
    double fac_fs;     // ne / (2 pi (1 - cos theta_bw)) / Gamma / (1 - mu beta), times the absorption
    double fac_rs;     // the same with the reverse shock's ne, or 0
};

__device__ Node node_geom(const Ctx &C, double x) {
    Node n;
    n.valid = 0; n.ia = 0; n.Dnu = 1.; n.lDnu = 0.; n.iD2 = 1.; n.wy2 = 0.; n.wy1 = 0.; n.fac_fs = 0.; n.fac_rs = 0.;
    const int N = C.i_end_r;
    if (C.R(0) == 0.0 && C.R(N - 1) == 0.0) return n;
    const double cos_theta = 1.0 - x;
    const double sh = sqrt(0.5 * (1.0 - cos_theta));
    const double st = 2.0 * sh * sqrt(1.0 - sh * sh), ct = 1.0 - 2.0 * sh * sh;
    const double mu = C.cj ? C.sa * (st * C.sphi) - C.ca * ct : C.sa * (st * C.sphi) + C.ca * ct;
    if (mu < C.mu0 || mu > C.muN) return n;
    if (C.mu0 >= mu) return n;
    const int ia = C.mu_index(mu), ib = ia + 1;
    const double mua = C.mu_at(ia), mub = C.mu_at(ib);
    const double tba = C.tb(ia), tbb = C.tb(ib);
    const double t_e = interp_lin(mu, mua, mub, tba, tbb);
    if (t_e < 0.0 || !isfinite(t_e)) return n;
    const double Ra = C.R(ia), Rb = C.R(ib);
    const double w = log(t_e / tba) / log(tbb / tba);
    const double r = Ra * pow(Rb / Ra, w);
    if (r <= 0.0 || !isfinite(r)) return n;
    auto li = [&](int v) { const double ya = C.col(v, ia), yb = C.col(v, ib); return ya * pow(yb / ya, w); };
    const double Gamma = li(PV_Gamma), beta = li(PV_beta);
    const double ne = li(PV_M2) / kMp;
    const double theta = li(PV_theta);
    const double omc = 1. - cos(theta);
    // the frequency shift takes the interpolated beta (fluxDensAdaptiveWithComov); the luminosity
    // (computeFluxDensity) recomputes beta from the interpolated Gamma
    const double bg = beta_from_gamma(Gamma);
    // no absorption table: optical depth 0 unless mu equals the shock velocity
    const double att = (mu == bg) ? -expm1(-1.0e100) / 1.0e100 : 1.0;
    const double geo = 1.0 / (2. * kPi * omc) / Gamma / (1.0 - mu * bg) * att;
    n.valid = 1; n.ia = ia; n.Dnu = Gamma * (1.0 - beta * mu);
    n.lDnu = log(n.Dnu);
    const double D = Gamma * (1.0 - bg * mu);
    n.iD2 = 1.0 / (D * D);
    const double idy = 1.0 / (Rb - Ra);
    n.wy2 = (Rb - r) * idy;
    n.wy1 = (r - Ra) * idy;
    n.fac_fs = ne * geo;
    if (C.has_rs && C.col(PV_thickness_rs, ia) > 0 && C.col(PV_thickness_rs, ib) > 0)
        n.fac_rs = li(PV_M3) / kMp * geo;
    return n;
}

// luminosity of one shock at comoving frequency freq (computeFluxDensity with useNe / smooth tau);
// frequencies off the comoving grid set `st` as where the reference stops
// lfreq = ln((1 + z) nu_obs) + ln Dnu, an estimate of ln freq for the grid index
template <bool RS>
__device__ __forceinline__ double shock_lum(const Ctx &C, const double *jall, const Node &n, double freq, double lfreq,
                                            double fac, int &st) {
    const int nf = C.nf;
    auto nu = [&](int i) { return RS ? C.nurs(i) : C.nufs(i); };
    if (freq >= nu(nf - 1) || freq <= nu(0)) { st |= PBAGPU_S_NU_OUT_OF_GRID; return 0.; }
    // find_index on the log-spaced grid: the index from the logarithm, then moved to the bracketing
    // grid points (the grid is built by repeated multiplication, so it is checked against its values)
    int ia_nu = (int)((lfreq - C.lnu0[RS]) * C.ilr[RS]);
    ia_nu = ia_nu < 0 ? 0 : (ia_nu > nf - 2 ? nf - 2 : ia_nu);
    while (ia_nu > 0 && nu(ia_nu) > freq) ia_nu--;
    while (ia_nu < nf - 2 && nu(ia_nu + 1) <= freq) ia_nu++;
    const int ib_nu = ia_nu + 1;
    // bilinear interpolation of the [step][frequency] table inside the cell (ia, ia + 1) x (ia_nu, ib_nu)
    // (Interp2d::InterpolateBilinear, interpolators.h:724-808), with the node's radial weights
    const double *r1 = jall + (size_t)n.ia * nf, *r2 = r1 + nf;
    const double q11 = r1[ia_nu], q21 = r1[ib_nu], q12 = r2[ia_nu], q22 = r2[ib_nu];
    const double idx = RS ? C.idrs(ia_nu) : C.idfs(ia_nu);
    const double em = ((nu(ib_nu) - freq) * idx) * (q11 * n.wy2 + q12 * n.wy1) +
                      ((freq - nu(ia_nu)) * idx) * (q21 * n.wy2 + q22 * n.wy1);
    const double em_lab = em * n.iD2;
    if (em_lab < 0 || !isfinite(em_lab)) st |= PBAGPU_S_NONFINITE;
    const double lum = em_lab * fac;   // fac holds 1 / (1 - mu beta) with beta from Gamma
    if (!(lum >= 0) || !isfinite(lum)) st |= PBAGPU_S_NONFINITE;
    return lum;
}

struct Wave {
    Ctx C;
    const double *jfs, *jrs;
    int lane;
    double nuz[NU_PER_LANE];   // (1 + z) nu_obs of this lane's frequencies (0: no frequency)
    double lnuz[NU_PER_LANE];  // their logarithms
    double rtol;
    int st;
    unsigned long long samples;
};

// one Gauss-Kronrod interval of the inner integral over x = 1 - cos(theta) in [a, b]
__device__ void gk_inner(Wave &W, double a, double b, double K[NU_PER_LANE], double G[NU_PER_LANE]) {
    const double c = 0.5 * (a + b), h = 0.5 * (b - a);
    Node mine;
    if (W.lane < NG) mine = node_geom(W.C, c + h * gk_node(W.lane));
    W.samples += NG;
    for (int j = 0; j < NU_PER_LANE; j++) { K[j] = 0.; G[j] = 0.; }
    for (int k = 0; k < NG; k++) {
        if (!bcast_i(mine.valid, k)) continue;
        Node n;
        n.valid = 1; n.ia = bcast_i(mine.ia, k); n.Dnu = bcast(mine.Dnu, k); n.lDnu = bcast(mine.lDnu, k);
        n.iD2 = bcast(mine.iD2, k);
        n.wy2 = bcast(mine.wy2, k); n.wy1 = bcast(mine.wy1, k); n.fac_fs = bcast(mine.fac_fs, k);
        n.fac_rs = bcast(mine.fac_rs, k);
        const double wk = gk_kw(k), wg = gk_gw(k);
        for (int j = 0; j < NU_PER_LANE; j++) {
            if (W.nuz[j] == 0.) continue;
            const double freq = W.nuz[j] * n.Dnu, lfreq = W.lnuz[j] + n.lDnu;
            double f = shock_lum<false>(W.C, W.jfs, n, freq, lfreq, n.fac_fs, W.st);
            if (n.fac_rs != 0.) f += shock_lum<true>(W.C, W.jrs, n, freq, lfreq, n.fac_rs, W.st);
            K[j] += wk * f;
            G[j] += wg * f;
        }
    }
    for (int j = 0; j < NU_PER_LANE; j++) { K[j] *= h; G[j] *= h; }
}

// every frequency of the wavefront accepts the interval estimate
__device__ __forceinline__ bool accept_all(const Wave &W, const double K[NU_PER_LANE], const double G[NU_PER_LANE],
                                           const double scale[NU_PER_LANE], double share) {
    bool ok = true;
    for (int j = 0; j < NU_PER_LANE; j++) {
        const double e = fabs(K[j] - G[j]);
        if (!(e <= W.rtol * scale[j] * share || e <= 1e-300)) ok = false;
    }
    return __ballot(!ok) == 0ull;
}

// Depth-first adaptive integral of the inner rule over [x0, x1]; the interval stack is one entry
// per lane (entry i in lane i), so it lives in registers.
__device__ void inner(Wave &W, double x0, double x1, double T[NU_PER_LANE]) {
    for (int j = 0; j < NU_PER_LANE; j++) T[j] = 0.;
    if (!(x1 > x0)) return;
    const double L = x1 - x0;
    double K[NU_PER_LANE], G[NU_PER_LANE], scale[NU_PER_LANE];
    gk_inner(W, x0, x1, K, G);
    for (int j = 0; j < NU_PER_LANE; j++) scale[j] = fabs(K[j]);
    if (accept_all(W, K, G, scale, 1.0)) { for (int j = 0; j < NU_PER_LANE; j++) T[j] = K[j]; return; }
    double sa = 0., sb = 0.;
    int sd = 0, sp = 0;
    const double m0 = 0.5 * (x0 + x1);
    if (W.lane == 0) { sa = m0; sb = x1; sd = 1; }
    if (W.lane == 1) { sa = x0; sb = m0; sd = 1; }
    sp = 2;
    while (sp > 0) {
        sp--;
        const double a = bcast(sa, sp), b = bcast(sb, sp);
        const int d = bcast_i(sd, sp);
        gk_inner(W, a, b, K, G);
        for (int j = 0; j < NU_PER_LANE; j++) scale[j] = fmax(scale[j], fabs(T[j] + K[j]));
        if (d >= MAX_DEPTH || accept_all(W, K, G, scale, (b - a) / L)) {
            for (int j = 0; j < NU_PER_LANE; j++) T[j] += K[j];
        } else {
            const double m = 0.5 * (a + b);
            if (W.lane == sp) { sa = m; sb = b; sd = d + 1; }
            if (W.lane == sp + 1) { sa = a; sb = m; sd = d + 1; }
            sp += 2;
        }
    }
}

// layer edge at sin(phi) = sphi (find_jet_edge, src/eats.h:103-173): bisection in theta on where
// the emission-time surface leaves the layer's angular extent
__device__ double jet_edge(const Ctx &C, double sphi, double th_con_hi) {
    double mu = C.obs_angle(th_con_hi, sphi);
    int ia = C.mu_index(mu);
    const double a0 = C.th(ia), a1 = C.th(ia + 1);
    if (a0 <= th_con_hi && th_con_hi <= a1) return th_con_hi;
    double ta, tb;
    if (th_con_hi < a0) { ta = th_con_hi; tb = 0.5 * M_PI; }
    else { ta = 0.0; tb = th_con_hi; }
    for (int i = 0; tb - ta > 1.0e-5 && i < 100; i++) {
        const double tc = 0.5 * (ta + tb);
        mu = C.obs_angle(tc, sphi);
        ia = C.mu_index(mu);
        if (tc < C.th(ia)) ta = tc; else tb = tc;
    }
    return ta;
}

// one Gauss-Kronrod interval of the outer integral over phi in [a, b]
__device__ void gk_outer(Wave &W, double th0c, double th1c, double a, double b, double K[NU_PER_LANE],
                         double G[NU_PER_LANE]) {
    const double c = 0.5 * (a + b), h = 0.5 * (b - a);
    double my_sphi = 0., my_x0 = 0., my_x1 = 0.;
    if (W.lane < NG) {
        const double phi = c + h * gk_node(W.lane);
        my_sphi = sin(phi);
        const double th_1 = jet_edge(W.C, my_sphi, th1c);
        const double th_0 = (th0c / th1c) * th_1;
        const double t0 = th_0 > 0.5 * M_PI ? 0.5 * M_PI : th_0;
        const double t1 = th_1 > 0.5 * M_PI ? 0.5 * M_PI : th_1;
        const double s0 = sin(0.5 * t0), s1 = sin(0.5 * t1);
        my_x0 = 2.0 * s0 * s0;
        my_x1 = 2.0 * s1 * s1;
    }
    for (int j = 0; j < NU_PER_LANE; j++) { K[j] = 0.; G[j] = 0.; }
    for (int k = 0; k < NG; k++) {
        W.C.sphi = bcast(my_sphi, k);
        double J[NU_PER_LANE];
        inner(W, bcast(my_x0, k), bcast(my_x1, k), J);
        const double wk = gk_kw(k), wg = gk_gw(k);
        for (int j = 0; j < NU_PER_LANE; j++) {
            const double v = (J[j] != J[j] || J[j] < 0.0) ? 0. : J[j];   // phi_integrand's guard
            K[j] += wk * v;
            G[j] += wg * v;
        }
    }
    for (int j = 0; j < NU_PER_LANE; j++) { K[j] *= h; G[j] *= h; }
}

__device__ void outer(Wave &W, double th0c, double th1c, double T[NU_PER_LANE]) {
    const double A = 0.0, B = M_PI, L = B - A;
    double K[NU_PER_LANE], G[NU_PER_LANE], scale[NU_PER_LANE];
    for (int j = 0; j < NU_PER_LANE; j++) T[j] = 0.;
    gk_outer(W, th0c, th1c, A, B, K, G);
    for (int j = 0; j < NU_PER_LANE; j++) scale[j] = fabs(K[j]);
    if (accept_all(W, K, G, scale, 1.0)) { for (int j = 0; j < NU_PER_LANE; j++) T[j] = K[j]; return; }
    double sa = 0., sb = 0.;
    int sd = 0, sp = 0;
    const double m0 = 0.5 * (A + B);
    if (W.lane == 0) { sa = m0; sb = B; sd = 1; }
    if (W.lane == 1) { sa = A; sb = m0; sd = 1; }
    sp = 2;
    while (sp > 0) {
        sp--;
        const double a = bcast(sa, sp), b = bcast(sb, sp);
        const int d = bcast_i(sd, sp);
        gk_outer(W, th0c, th1c, a, b, K, G);
        for (int j = 0; j < NU_PER_LANE; j++) scale[j] = fmax(scale[j], fabs(T[j] + K[j]));
        if (d >= MAX_DEPTH || accept_all(W, K, G, scale, (b - a) / L)) {
            for (int j = 0; j < NU_PER_LANE; j++) T[j] += K[j];
        } else {
            const double m = 0.5 * (a + b);
            if (W.lane == sp) { sa = m; sb = b; sd = d + 1; }
            if (W.lane == sp + 1) { sa = a; sb = m; sd = d + 1; }
            sp += 2;
        }
    }
}

// grid: (ceil(n_t / WPB), n_layers); wavefront w of a workgroup integrates t_obs index
// blockIdx.x * WPB + w of layer blockIdx.y at every frequency
__global__ void __launch_bounds__(WAVE * WPB, 2) eats_vec_kernel(const VLayer *__restrict__ layers, int n_t,
                                                              const double *__restrict__ tobs, int n_nu,
                                                              const double *__restrict__ nuobs, double rtol,
                                                              double *__restrict__ out, int *__restrict__ status,
                                                              unsigned long long *__restrict__ n_samples) {
    const VLayer &L = layers[blockIdx.y];
    const int nt = L.nt, nf = L.nf;
    for (int i = threadIdx.x; i < nt; i += blockDim.x) {
        vlds[i] = L.T[(size_t)PV_tburst * nt + i];
        vlds[nt + i] = L.T[(size_t)PV_R * nt + i];
        vlds[2 * nt + i] = L.T[(size_t)PV_theta * nt + i];
    }
    for (int i = threadIdx.x; i < nf; i += blockDim.x) {
        vlds[3 * nt + i] = L.nufs[i];
        vlds[3 * nt + nf + i] = L.nurs[i];
        vlds[3 * nt + 2 * nf + i] = i < nf - 1 ? 1.0 / (L.nufs[i + 1] - L.nufs[i]) : 0.;
        vlds[3 * nt + 3 * nf + i] = i < nf - 1 ? 1.0 / (L.nurs[i + 1] - L.nurs[i]) : 0.;
    }
    __syncthreads();
    const int it = blockIdx.x * WPB + threadIdx.x / WAVE;
    if (it >= n_t) return;   // after the one barrier of the kernel
    Wave W;
    W.lane = threadIdx.x % WAVE;
    W.rtol = rtol;
    W.st = 0;
    W.samples = 0;
    W.jfs = L.jfs; W.jrs = L.jrs;
    Ctx &C = W.C;
    C.T = L.T; C.nt = nt; C.nf = nf; C.i_end_r = L.i_end_r; C.z = L.z;
    C.has_rs = L.jrs != nullptr;
    for (int g = 0; g < 2; g++) {
        const double a = g ? C.nurs(0) : C.nufs(0), b = g ? C.nurs(nf - 1) : C.nufs(nf - 1);
        C.lnu0[g] = log(a);
        C.ilr[g] = (nf - 1) / (log(b) - log(a));
    }
    C.sa = sin(L.theta_obs);
    C.ca = cos(L.theta_obs);
    C.sphi = 0.;
    C.cj = false;
    const double t_obs = tobs[it];
    C.tz = t_obs / (1.0 + L.z);
    for (int j = 0; j < NU_PER_LANE; j++) {
        const int q = W.lane + WAVE * j;
        W.nuz[j] = q < n_nu ? (1.0 + L.z) * nuobs[q] : 0.;
        W.lnuz[j] = q < n_nu ? log(W.nuz[j]) : 0.;
    }
    double flux[NU_PER_LANE] = {0., 0.};
    if (L.i_end_r == 1) W.st |= PBAGPU_S_NONFINITE;   // one usable step: no surface
    if (L.i_end_r > 1) {
        if (t_obs < C.tb(0) || t_obs > C.tb(L.i_end_r - 1)) W.st |= PBAGPU_S_NONFINITE;
        C.mu0 = C.mu_at(0);
        C.muN = C.mu_at(L.i_end_r - 1);
        const double Fcoeff = kCgs2mJy / (4 * M_PI * L.d_l * L.d_l);
        for (int cj = 0; cj < (L.counter_jet ? 2 : 1); cj++) {
            C.cj = cj == 1;
            double O[NU_PER_LANE];
            outer(W, L.theta_c_l, L.theta_c_h, O);
            for (int j = 0; j < NU_PER_LANE; j++) flux[j] += 2 * Fcoeff * O[j];
        }
    }
    for (int j = 0; j < NU_PER_LANE; j++) {
        const int q = W.lane + WAVE * j;
        if (q < n_nu) out[((size_t)blockIdx.y * n_t + it) * n_nu + q] = 0. + flux[j];
    }
    // status of the layer and the sample count, one atomic per wavefront
    unsigned long long m = __ballot(W.st != 0);
    int st = W.st;
    for (int off = 32; off > 0; off >>= 1) st |= __shfl_xor(st, off, WAVE);
    if (W.lane == 0) {
        if (m) atomicOr(&status[blockIdx.y], st);
        if (n_samples) atomicAdd(n_samples, W.samples);
    }
}

}  // namespace eatsv
}  // namespace pbagpu

extern "C" int pbagpu_eats_lightcurve_vec(const pba_run_params *params, int n_runs, const pba_eats_layer *layers,
                                          int n_layers, const double *d_table, int nt, const double *d_jall,
                                          int nfreq, int n_t, const double *d_tobs, int n_nu, const double *d_nuobs,
                                          double rtol, double *d_out, int *d_status, void *stream) {
    using namespace pbagpu::eatsv;
    if (!params || !layers || !d_table || !d_jall || !d_out || !d_status || !d_tobs || !d_nuobs || n_runs < 1 ||
        n_layers < 1 || n_t < 1 || n_nu < 1 || nt < 2 || nfreq < 2 || !(rtol > 0.0))
        return PBAGPU_E_ARG;
    if (n_nu > MAX_NU) {
        pbagpu::set_error_msg("eats_lightcurve_vec: at most 128 observer frequencies per call");
        return PBAGPU_E_ARG;
    }
    const size_t lds = sizeof(double) * (3 * (size_t)nt + 4 * (size_t)nfreq);
    if (lds > 160 * 1024) {
        pbagpu::set_error_msg("eats_lightcurve_vec: time grid too long for the LDS-staged columns");
        return PBAGPU_E_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    std::vector<double> hnu((size_t)n_runs * 2 * nfreq);
    for (int r = 0; r < n_runs; r++)
        for (int sh = 0; sh < 2; sh++) {
            const pba_shock_params &sp = sh ? params[r].rs : params[r].fs;
            if (sp.nfreq != nfreq) return PBAGPU_E_ARG;
            if (sp.ssa) {
                pbagpu::set_error_msg("eats_lightcurve_vec: synchrotron self-absorption is not implemented here");
                return PBAGPU_E_UNSUPPORTED;
            }
            double cur = std::pow(10.0, std::log10(sp.freq1));
            double ratio = std::pow(10.0, (std::log10(sp.freq2) - std::log10(sp.freq1)) / sp.nfreq);
            for (int i = 0; i < nfreq; i++) { hnu[((size_t)r * 2 + sh) * nfreq + i] = cur; cur *= ratio; }
        }
    std::vector<VLayer> hl(n_layers);
    for (int l = 0; l < n_layers; l++) {
        const pba_eats_layer &Y = layers[l];
        if (Y.run < 0 || Y.run >= n_runs || Y.i_end_r > nt || Y.i_end_r < 0 || Y.pop_fs < 0 || Y.pop_rs < -1 ||
            Y.bw < 0) {
            pbagpu::set_error_msg("eats_lightcurve_vec: bad layer descriptor (run, bw, pop_fs >= 0, pop_rs >= -1, "
                                  "0 <= i_end_r <= nt)");
            return PBAGPU_E_ARG;
        }
    }
    auto *dnu = (double *)pbagpu::workspace(pbagpu::WS_VEC_NU, sizeof(double) * hnu.size(), s);
    auto *dL = (VLayer *)pbagpu::workspace(pbagpu::WS_VEC_LAYERS, sizeof(VLayer) * n_layers, s);
    if (!dnu || !dL) return PBAGPU_E_HIP;
    for (int l = 0; l < n_layers; l++) {
        const pba_eats_layer &Y = layers[l];
        const pba_run_params &P = params[Y.run];
        VLayer &A = hl[l];
        A.T = d_table + (size_t)Y.bw * PBA_NV * nt;
        A.jfs = d_jall + (size_t)Y.pop_fs * nt * nfreq;
        A.jrs = Y.pop_rs >= 0 ? d_jall + (size_t)Y.pop_rs * nt * nfreq : nullptr;
        A.nufs = dnu + ((size_t)Y.run * 2) * nfreq;
        A.nurs = dnu + ((size_t)Y.run * 2 + 1) * nfreq;
        A.theta_c_l = Y.theta_c_l; A.theta_c_h = Y.theta_c_h;
        A.z = P.z; A.theta_obs = P.theta_obs; A.d_l = P.d_l;
        A.counter_jet = P.counter_jet; A.i_end_r = Y.i_end_r; A.nt = nt; A.nf = nfreq;
    }
    int rc = PBAGPU_OK;
    PBA_TRY(rc, hipMemcpyAsync(dnu, hnu.data(), sizeof(double) * hnu.size(), hipMemcpyHostToDevice, s));
    PBA_TRY(rc, hipMemcpyAsync(dL, hl.data(), sizeof(VLayer) * n_layers, hipMemcpyHostToDevice, s));
    PBA_TRY(rc, hipMemsetAsync(d_status, 0, sizeof(int) * n_layers, s));
    pbagpu::KernelTimer timer(PBAGPU_STAGE_EATS_VEC);
    unsigned long long *d_samples = nullptr;
    if (rc == PBAGPU_OK && pbagpu::timing_enabled()) {
        d_samples = (unsigned long long *)pbagpu::workspace(pbagpu::WS_VEC_SAMPLES, sizeof(unsigned long long), s);
        if (d_samples) PBA_TRY(rc, hipMemsetAsync(d_samples, 0, sizeof(unsigned long long), s));
    }
    if (rc == PBAGPU_OK) {
        (void)hipFuncSetAttribute((const void *)eats_vec_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        timer.start(s);
        hipLaunchKernelGGL(eats_vec_kernel, dim3((unsigned)((n_t + WPB - 1) / WPB), (unsigned)n_layers),
                           dim3(WAVE * WPB), lds, s, dL, n_t, d_tobs, n_nu, d_nuobs, rtol, d_out, d_status, d_samples);
        PBA_TRY(rc, hipGetLastError());
        timer.stop(s);
    }
    if (d_samples && rc == PBAGPU_OK) {
        unsigned long long hs = 0;
        PBA_TRY(rc, hipMemcpyAsync(&hs, d_samples, sizeof(hs), hipMemcpyDeviceToHost, s));
        PBA_TRY(rc, hipStreamSynchronize(s));
        pbagpu::timing_add_units(PBAGPU_STAGE_EATS_VEC, (double)hs);
    }
    PBA_TRY(rc, hipStreamSynchronize(s));
    timer.collect();
    return rc;
}
