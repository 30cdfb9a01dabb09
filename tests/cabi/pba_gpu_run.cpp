// A C++ caller of the C ABI (include/pbagpu.h) with no Python in the process: the sequence a
// maintainer would put into the reference program's Ejecta::processEvolved (the run_on_gpu
// sketch of the integration notes), for one run on the adaptive or the piece-wise EATS grid.
//
//   pba_gpu_run <input.bin> <output.bin>
//
// input.bin (native byte order): int32 method (0 adaptive, 1 piece-wise), int32 n_bw, int32 nt,
// int32 n_t_obs, int32 n_nu_obs; pba_run_params; pba_layer_init[n_bw]; double t_grid[nt];
// double lc_times[n_t_obs]; double lc_freqs[n_nu_obs]; piece-wise only: int32 cells_in_layer[n_bw]
// (EjectaID2::CellsInLayer, 2 i + 1).
// output.bin: int32 n_bw, int32 n_pairs; double out[n_bw][n_pairs] (the reference's VecVector out
// of EjectaBase::evalEjectaLightCurves, pairs t-major as cast_times_freqs).
// Exit status 0 on success; on an error it prints pbagpu_last_error() and exits with 1, the
// reference's convention.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "pbagpu.h"

static void check(int rc, const char *what) {
    if (rc != PBAGPU_OK) {
        std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, pbagpu_last_error());
        std::exit(1);
    }
}

static void hip(hipError_t e, const char *what) {
    if (e != hipSuccess) {
        std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
        std::exit(1);
    }
}

template <class T>
static void rd(std::FILE *f, T *p, size_t n) {
    if (std::fread(p, sizeof(T), n, f) != n) { std::fprintf(stderr, "short input file\n"); std::exit(1); }
}

int main(int argc, char **argv) {
    if (argc != 3) { std::fprintf(stderr, "usage: %s input.bin output.bin\n", argv[0]); return 2; }
    std::FILE *f = std::fopen(argv[1], "rb");
    if (!f) { std::perror(argv[1]); return 1; }
    int hdr[5];
    rd(f, hdr, 5);
    const int method = hdr[0], nbw = hdr[1], nt = hdr[2], n_t = hdr[3], n_nu = hdr[4];
    pba_run_params P;
    rd(f, &P, 1);
    std::vector<pba_layer_init> init(nbw);
    rd(f, init.data(), nbw);
    std::vector<double> t_grid(nt), lc_times(n_t), lc_freqs(n_nu);
    rd(f, t_grid.data(), nt);
    rd(f, lc_times.data(), n_t);
    rd(f, lc_freqs.data(), n_nu);
    std::vector<int> cil(nbw, 1);
    if (method == 1) rd(f, cil.data(), nbw);
    std::fclose(f);

    double *d_t, *d_table;
    hip(hipMalloc(&d_t, sizeof(double) * nt), "hipMalloc");
    hip(hipMalloc(&d_table, sizeof(double) * nbw * PBA_NV * nt), "hipMalloc");
    hip(hipMemcpy(d_t, t_grid.data(), sizeof(double) * nt, hipMemcpyHostToDevice), "hipMemcpy");
    std::vector<pba_bw_result> res(nbw);
    check(pbagpu_dyn_evolve(&P, 1, nbw, init.data(), nt, d_t, d_table, res.data(), nullptr), "dynamics");
    for (int i = 0; i < nbw; i++)
        if (res[i].status) { std::fprintf(stderr, "blastwave %d status %d\n", i, res[i].status); return 1; }

    // populations: the forward shock of every blastwave, the reverse shock where it radiates
    std::vector<pba_population> pops;
    std::vector<int> pop_fs(nbw), pop_rs(nbw, -1);
    for (int i = 0; i < nbw; i++) {
        pop_fs[i] = (int)pops.size();
        pops.push_back({i, 0, 0, 0});
        if (res[i].fsrs && P.do_rs_radiation) { pop_rs[i] = (int)pops.size(); pops.push_back({i, 0, 1, 0}); }
    }
    const int npop = (int)pops.size(), nf = P.fs.nfreq;
    double *d_jall, *d_aall = nullptr;
    int *d_est, *d_lst;
    hip(hipMalloc(&d_jall, sizeof(double) * npop * nt * nf), "hipMalloc");
    if (P.fs.ssa || P.rs.ssa) hip(hipMalloc(&d_aall, sizeof(double) * npop * nt * nf), "hipMalloc");
    hip(hipMalloc(&d_est, sizeof(int) * npop), "hipMalloc");
    check(pbagpu_ele_evolve(&P, 1, pops.data(), npop, d_table, nbw, nt, d_jall, d_aall, nullptr, d_est, 0, nullptr),
          "electrons");
    std::vector<int> est(npop);
    hip(hipMemcpy(est.data(), d_est, sizeof(int) * npop, hipMemcpyDeviceToHost), "hipMemcpy");
    for (int q = 0; q < npop; q++)
        if (est[q]) { std::fprintf(stderr, "population %d status %d\n", q, est[q]); return 1; }

    // observer pairs, t-major (cast_times_freqs, src/utilitites/utils.h:1419-1448)
    std::vector<double> t_obs, nu_obs;
    for (double t : lc_times)
        for (double nu : lc_freqs) { t_obs.push_back(t); nu_obs.push_back(nu); }
    const int np = (int)t_obs.size();
    double *d_to, *d_no, *d_out;
    hip(hipMalloc(&d_to, sizeof(double) * np), "hipMalloc");
    hip(hipMalloc(&d_no, sizeof(double) * np), "hipMalloc");
    hip(hipMalloc(&d_out, sizeof(double) * nbw * np), "hipMalloc");
    hip(hipMalloc(&d_lst, sizeof(int) * nbw), "hipMalloc");
    hip(hipMemcpy(d_to, t_obs.data(), sizeof(double) * np, hipMemcpyHostToDevice), "hipMemcpy");
    hip(hipMemcpy(d_no, nu_obs.data(), sizeof(double) * np, hipMemcpyHostToDevice), "hipMemcpy");
    if (method == 1) {
        std::vector<pba_eats_pw_layer> pw(nbw);
        for (int i = 0; i < nbw; i++)
            pw[i] = {init[i].ctheta0, init[i].ncells, cil[i], i, 0, pop_fs[i], pop_rs[i] >= 0 ? 1 : 0, res[i].i_end_r};
        check(pbagpu_eats_lightcurve_pw(&P, 1, pw.data(), nbw, d_table, nt, d_jall, d_aall, nf, np, d_to, d_no,
                                        d_out, d_lst, nullptr), "light curve (piece-wise)");
    } else {
        std::vector<pba_eats_layer> lays(nbw);
        for (int i = 0; i < nbw; i++)
            lays[i] = {init[i].theta_c_l, init[i].theta_c_h, i, 0, pop_fs[i], pop_rs[i], res[i].i_end_r};
        check(pbagpu_eats_lightcurve(&P, 1, lays.data(), nbw, d_table, nt, d_jall, d_aall, nf, np, d_to, d_no, d_out,
                                     d_lst, nullptr), "light curve");
    }
    std::vector<int> lst(nbw);
    hip(hipMemcpy(lst.data(), d_lst, sizeof(int) * nbw, hipMemcpyDeviceToHost), "hipMemcpy");
    for (int i = 0; i < nbw; i++)
        if (lst[i]) { std::fprintf(stderr, "layer %d observer status %d\n", i, lst[i]); return 1; }
    std::vector<double> out((size_t)nbw * np);
    hip(hipMemcpy(out.data(), d_out, sizeof(double) * out.size(), hipMemcpyDeviceToHost), "hipMemcpy");
    std::FILE *g = std::fopen(argv[2], "wb");
    if (!g) { std::perror(argv[2]); return 1; }
    const int oh[2] = {nbw, np};
    std::fwrite(oh, sizeof(int), 2, g);
    std::fwrite(out.data(), sizeof(double), out.size(), g);
    std::fclose(g);
    for (void *p : {(void *)d_t, (void *)d_table, (void *)d_jall, (void *)d_aall, (void *)d_est, (void *)d_lst,
                    (void *)d_to, (void *)d_no, (void *)d_out})
        if (p) hip(hipFree(p), "hipFree");
    std::printf("pba_gpu_run: %d blastwaves, %d populations, %d pairs\n", nbw, npop, np);
    return 0;
}
