// Sky-map binning: intensity-weighted 2-D histogram, flux centroid and data extent of raw
// sky maps (ProcessRawSkymap._process_file, package/src/PyBlastAfterglowMag/skymap_process.py:465-485).
//
// Four launches on the caller's stream:
//   sky_zero_kernel    clears the histogram (and the optional point counts) with vector stores;
//   sky_partial_kernel one workgroup per work item (<= 4096 points of one entry): min / max of
//                      x and y and the sums of w, w x, w y;
//   sky_extent_kernel  one workgroup per image: combines its items in order into the per-entry
//                      extent and the centroid, then writes the edges and the status;
//   sky_bin_kernel     a grid over every image's points: bin search in LDS-staged edges and an
//                      atomic add of the intensity into the image's histogram.
// The work is HBM / atomic bound and tiny next to the observer stage (one 256 x 256 image of a
// 64-layer piece-wise jet is 8192 points); it is kept on the device so that sky maps made by
// pbagpu_eats_skymap_pw never have to leave HBM to be binned.
#include <hip/hip_runtime.h>
#include <float.h>
#include <algorithm>
#include <math.h>
#include <climits>
#include <cstdio>
#include <mutex>
#include <vector>
#include "runtime.h"
#include "timing.h"

namespace pbagpu {
namespace skybin {

constexpr int BLOCK = 256;
constexpr int MAX_EDGES = 4096;   // nx + ny edges staged in LDS by the binning kernel (32 KiB)

// Reductions: over the 64 lanes of a wave by butterfly shuffles, then over the 4 waves of the
// workgroup in LDS, in a fixed order (the result does not depend on scheduling).
constexpr int WAVES = BLOCK / 64;
struct Add { __device__ double operator()(double a, double b) const { return a + b; } };
// NaN never wins a comparison, so non-finite positions are screened separately.
struct Min { __device__ double operator()(double a, double b) const { return b < a ? b : a; } };
struct Max { __device__ double operator()(double a, double b) const { return b > a ? b : a; } };

template <typename Op>
__device__ double wave_reduce(double v, Op op) {
    for (int off = 32; off > 0; off >>= 1) v = op(v, __shfl_xor(v, off, 64));
    return v;
}

// np.linspace(lo * scale, hi * scale, num) with num >= 2 (numpy/_core/function_base.py): step = (stop - start)
// / (num - 1); e_i = i * step + start as two roundings (no fused multiply-add: the pragma keeps
// hipcc from contracting them), e_last = stop; a zero step (denormal range) takes
// e_i = (i / div) * delta + start instead.
__device__ void linspace(double lo, double hi, double scale, int num, double *out) {
#pragma clang fp contract(off)
    const double start = lo * scale, stop = hi * scale;
    const double div = (double)(num - 1);
    const double delta = stop - start;
    const double step = delta / div;
    for (int i = threadIdx.x; i < num; i += BLOCK) {
        double y;
        if (step == 0.0) y = ((double)i / div) * delta;
        else y = (double)i * step;
        y = y + start;
        out[i] = (i == num - 1) ? stop : y;
    }
}

__global__ void sky_zero_kernel(double *hist, unsigned *count, size_t n) {
    for (size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (size_t)gridDim.x * BLOCK) {
        hist[i] = 0.0;
        if (count) count[i] = 0u;
    }
}

// Partial statistics of one work item (at most ITEM points of one entry), one workgroup each.
constexpr int ITEM = 4096;
enum { P_X0 = 0, P_X1, P_Y0, P_Y1, P_W, P_WX, P_WY, P_BAD, P_N };

__global__ __launch_bounds__(BLOCK) void sky_partial_kernel(const long *item_range, const double *x, const double *y,
                                                            const double *w, double *part) {
    __shared__ double red[P_N][WAVES];
    const long b = item_range[2 * blockIdx.x], e = item_range[2 * blockIdx.x + 1];
    double v[P_N] = {DBL_MAX, -DBL_MAX, DBL_MAX, -DBL_MAX, 0., 0., 0., 0.};
    for (long i = b + threadIdx.x; i < e; i += BLOCK) {
        const double xi = x[i], yi = y[i], wi = w[i];
        if (!isfinite(xi) || !isfinite(yi) || !isfinite(wi)) v[P_BAD] = 1.;
        v[P_X0] = fmin(v[P_X0], xi); v[P_X1] = fmax(v[P_X1], xi);
        v[P_Y0] = fmin(v[P_Y0], yi); v[P_Y1] = fmax(v[P_Y1], yi);
        v[P_W] += wi; v[P_WX] += wi * xi; v[P_WY] += wi * yi;
    }
    v[P_X0] = wave_reduce(v[P_X0], Min()); v[P_X1] = wave_reduce(v[P_X1], Max());
    v[P_Y0] = wave_reduce(v[P_Y0], Min()); v[P_Y1] = wave_reduce(v[P_Y1], Max());
    v[P_W] = wave_reduce(v[P_W], Add()); v[P_WX] = wave_reduce(v[P_WX], Add());
    v[P_WY] = wave_reduce(v[P_WY], Add()); v[P_BAD] = wave_reduce(v[P_BAD], Max());
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0)
        for (int k = 0; k < P_N; k++) red[k][wv] = v[k];
    __syncthreads();
    if (threadIdx.x < P_N) {
        const int k = threadIdx.x;
        double r = red[k][0];
        for (int q = 1; q < WAVES; q++) {
            const double o = red[k][q];
            r = (k == P_X0 || k == P_Y0) ? Min()(r, o) : (k == P_X1 || k == P_Y1 || k == P_BAD) ? Max()(r, o) : r + o;
        }
        part[(size_t)blockIdx.x * P_N + k] = r;
    }
}

// One workgroup per image: combines its entries' partials in entry and item order.
//  * _get_data_extend (skymap_process.py:107-189): an entry counts when it has at least two
//    points and its intensities do not sum to zero; the image extent is the min / max of x and
//    y over the points of those entries; an entry with min == max in x or y is an error there.
//  * _compute_position_of_the_flux_centroid (:98-105): sum(w x) / sum(w) over every point.
//  * edges = np.linspace(xmin * extend_grid, xmax * extend_grid, nx) (:480-481) unless given.
__global__ __launch_bounds__(BLOCK) void sky_extent_kernel(const int *img_seg, const long *seg_off, const int *item_off,
                                                           const double *part, int nx, int ny, double extend_grid,
                                                           int edges_given, double *edges_x, double *edges_y,
                                                           double *mom, int *status) {
    __shared__ double s_r[8];
    __shared__ int s_st, s_dec;
    const int img = blockIdx.x;
    const int t = threadIdx.x;
    if (t == 0) {
        double sxw = 0., syw = 0., sw = 0.;
        double xmin = DBL_MAX, xmax = -DBL_MAX, ymin = DBL_MAX, ymax = -DBL_MAX;
        int kept = 0, flat = 0, bad = 0;
        for (int k = img_seg[img]; k < img_seg[img + 1]; k++) {
            double ex0 = DBL_MAX, ex1 = -DBL_MAX, ey0 = DBL_MAX, ey1 = -DBL_MAX, es = 0.;
            for (int it = item_off[k]; it < item_off[k + 1]; it++) {
                const double *P = part + (size_t)it * P_N;
                ex0 = fmin(ex0, P[P_X0]); ex1 = fmax(ex1, P[P_X1]);
                ey0 = fmin(ey0, P[P_Y0]); ey1 = fmax(ey1, P[P_Y1]);
                es += P[P_W];
                sw += P[P_W]; sxw += P[P_WX]; syw += P[P_WY];
                if (P[P_BAD] != 0.) bad = 1;
            }
            if (seg_off[k + 1] - seg_off[k] >= 2 && es != 0.0) {
                kept++;
                if (ex0 == ex1 || ey0 == ey1) flat = 1;
                xmin = fmin(xmin, ex0); xmax = fmax(xmax, ex1);
                ymin = fmin(ymin, ey0); ymax = fmax(ymax, ey1);
            }
        }
        int st = PBA_SKY_OK;
        if (bad) st = PBA_SKY_NONFINITE;
        else if (sw == 0.0) st = PBA_SKY_ZERO_WEIGHT;   // the centroid is computed first there
        // the extent's checks run only where the extent makes the edges
        else if (!edges_given && kept == 0) st = PBA_SKY_NO_DATA;
        else if (!edges_given && flat) st = PBA_SKY_FLAT;
        double *m = mom + (size_t)img * PBA_SKY_NM;
        m[PBA_SKY_XC] = sw != 0.0 ? sxw / sw : 0.0;
        m[PBA_SKY_YC] = sw != 0.0 ? syw / sw : 0.0;
        m[PBA_SKY_XMIN] = kept ? xmin : 0.0; m[PBA_SKY_XMAX] = kept ? xmax : 0.0;
        m[PBA_SKY_YMIN] = kept ? ymin : 0.0; m[PBA_SKY_YMAX] = kept ? ymax : 0.0;
        m[PBA_SKY_SUMW] = sw;
        m[PBA_SKY_NPTS] = (double)(seg_off[img_seg[img + 1]] - seg_off[img_seg[img]]);
        s_r[0] = xmin; s_r[1] = xmax; s_r[2] = ymin; s_r[3] = ymax;
        s_st = st; s_dec = 0;
    }
    __syncthreads();
    int st = s_st;   // uniform over the workgroup from here on
    double *ex = edges_x + (size_t)img * nx, *ey = edges_y + (size_t)img * ny;
    if (st == PBA_SKY_OK && !edges_given) {
        linspace(s_r[0], s_r[1], extend_grid, nx, ex);
        linspace(s_r[2], s_r[3], extend_grid, ny, ey);
    }
    if (st == PBA_SKY_OK && edges_given) {
        int dec = 0;
        for (int i = t; i + 1 < nx; i += BLOCK) dec |= ex[i] > ex[i + 1] || !isfinite(ex[i]) || !isfinite(ex[i + 1]);
        for (int i = t; i + 1 < ny; i += BLOCK) dec |= ey[i] > ey[i + 1] || !isfinite(ey[i]) || !isfinite(ey[i + 1]);
        if (dec) atomicOr(&s_dec, 1);
        __syncthreads();
        if (s_dec) st = PBA_SKY_BAD_EDGES;
    }
    if (t == 0) status[img] = st;
}

// Number of edges <= v (np.searchsorted(edges, v, side='right')); NaN gives 0.
__device__ inline int upper_count(const double *e, int n, double v) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (e[mid] <= v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// np.histogramdd((x, y), bins=(edges_x, edges_y), weights=w) (numpy/lib/_histograms_impl.py):
// bin index = searchsorted(edges, v, 'right'), minus one for a value equal to the last edge;
// indices 0 and len(edges) are outliers.  grid = (chunks, n_images).
__global__ __launch_bounds__(BLOCK) void sky_bin_kernel(const int *img_seg, const long *seg_off, const double *x,
                                                        const double *y, const double *w, int nx, int ny,
                                                        const double *edges_x, const double *edges_y,
                                                        const int *status, double *hist, unsigned *count) {
    extern __shared__ double s_e[];
    const int img = blockIdx.y;
    if (status[img] != PBA_SKY_OK) return;   // uniform over the workgroup
    double *sx = s_e, *sy = s_e + nx;
    for (int i = threadIdx.x; i < nx; i += BLOCK) sx[i] = edges_x[(size_t)img * nx + i];
    for (int i = threadIdx.x; i < ny; i += BLOCK) sy[i] = edges_y[(size_t)img * ny + i];
    __syncthreads();
    const long p0 = seg_off[img_seg[img]], p1 = seg_off[img_seg[img + 1]];
    const double xl = sx[nx - 1], yl = sy[ny - 1];
    const size_t nb = (size_t)(nx - 1) * (ny - 1);
    double *H = hist + (size_t)img * nb;
    unsigned *Cn = count ? count + (size_t)img * nb : nullptr;
    for (long i = p0 + (long)blockIdx.x * BLOCK + threadIdx.x; i < p1; i += (long)gridDim.x * BLOCK) {
        const double xi = x[i], yi = y[i];
        int ix = upper_count(sx, nx, xi), iy = upper_count(sy, ny, yi);
        if (xi == xl) ix--;
        if (yi == yl) iy--;
        if (ix < 1 || ix > nx - 1 || iy < 1 || iy > ny - 1) continue;
        const size_t b = (size_t)(ix - 1) * (ny - 1) + (iy - 1);
        atomicAdd(H + b, w[i]);
        if (Cn) atomicAdd(Cn + b, 1u);
    }
}

// Item table of the allocation-free entry, built on the device: entry k is split into
// max(1, ceil(len_k / ITEM)) work items; item_off[k] is the first item of entry k
// (item_off[n_segs] = the item count) and items[2 i], items[2 i + 1] is item i's point range,
// the same table the host builds for pbagpu_skymap_hist.  One workgroup: every thread takes a
// contiguous run of entries, and an exclusive prefix over the threads' item counts places them.
__host__ __device__ __forceinline__ long items_of(long b, long e) { return e > b ? (e - b + ITEM - 1) / ITEM : 1; }

__global__ __launch_bounds__(BLOCK) void sky_items_kernel(int n_segs, const long *seg_off, int *item_off,
                                                          long *items) {
    __shared__ int s_sum[BLOCK];
    const int t = threadIdx.x;
    const int per = (n_segs + BLOCK - 1) / BLOCK;
    const int k0 = min(t * per, n_segs), k1 = min(k0 + per, n_segs);
    int local = 0;
    for (int k = k0; k < k1; k++) local += (int)items_of(seg_off[k], seg_off[k + 1]);
    s_sum[t] = local;
    __syncthreads();
    if (t == 0) {
        int acc = 0;
        for (int q = 0; q < BLOCK; q++) { const int v = s_sum[q]; s_sum[q] = acc; acc += v; }
        item_off[n_segs] = acc;
    }
    __syncthreads();
    int it = s_sum[t];
    for (int k = k0; k < k1; k++) {
        item_off[k] = it;
        long b = seg_off[k];
        do {
            const long e = min(b + (long)ITEM, seg_off[k + 1]);
            items[2 * (size_t)it] = b; items[2 * (size_t)it + 1] = e;
            it++;
            b = e;
        } while (b < seg_off[k + 1]);
    }
}

// Scratch of the allocation-free entry: one copy per device, held until the process ends.
// Laid out per call as img_seg | item_off (ints), seg_off | items (longs), partials (doubles),
// each part 256-byte aligned.
constexpr size_t SCRATCH_BYTES = (size_t)4 << 20;
__device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[MAX_DEVICES] = {};

// Device address of the current device's scratch (hipGetSymbolAddress once per device, with
// that device current).  The array's name is a host-side stand-in and is never passed on.
void *scratch_address() {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) { set_error("hipGetDevice", e); return nullptr; }
    if (dev < 0 || dev >= MAX_DEVICES) { set_error_msg("skymap_hist_noalloc: device index above 63"); return nullptr; }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) { set_error("hipGetSymbolAddress(scratch)", e); return nullptr; }
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Argument checks shared by both entries.  On success sets the entry count, the point count and
// the largest image's point count.
int check_args(int n_images, const int *img_seg, const long *seg_off, const double *d_x, const double *d_y,
               const double *d_w, int nx, int ny, double extend_grid, double *d_edges_x, double *d_edges_y,
               double *d_hist, double *d_moments, int *d_status, const char *who, int &n_segs, long &n_pts,
               long &max_pts) {
    char msg[192];
    if (n_images < 1 || n_images > 65535 || !img_seg || !seg_off || !d_edges_x || !d_edges_y || !d_hist ||
        !d_moments || !d_status || nx < 2 || ny < 2 || nx + ny > MAX_EDGES) {
        snprintf(msg, sizeof(msg), "%s: bad arguments (1 <= n_images <= 65535, nx, ny >= 2, nx + ny <= 4096)", who);
        set_error_msg(msg);
        return PBAGPU_E_ARG;
    }
    if (img_seg[0] != 0) { snprintf(msg, sizeof(msg), "%s: img_seg[0] must be 0", who); set_error_msg(msg); return PBAGPU_E_ARG; }
    for (int i = 0; i < n_images; i++)
        if (img_seg[i + 1] < img_seg[i]) {
            snprintf(msg, sizeof(msg), "%s: img_seg decreases", who); set_error_msg(msg); return PBAGPU_E_ARG;
        }
    n_segs = img_seg[n_images];
    if (seg_off[0] != 0) { snprintf(msg, sizeof(msg), "%s: seg_off[0] must be 0", who); set_error_msg(msg); return PBAGPU_E_ARG; }
    for (int k = 0; k < n_segs; k++)
        if (seg_off[k + 1] < seg_off[k]) {
            snprintf(msg, sizeof(msg), "%s: seg_off decreases", who); set_error_msg(msg); return PBAGPU_E_ARG;
        }
    n_pts = seg_off[n_segs];
    if (n_pts > 0 && (!d_x || !d_y || !d_w)) return PBAGPU_E_ARG;
    if (!(extend_grid == extend_grid)) {
        snprintf(msg, sizeof(msg), "%s: extend_grid is NaN", who); set_error_msg(msg); return PBAGPU_E_ARG;
    }
    max_pts = 0;
    for (int i = 0; i < n_images; i++) {
        const long np_i = seg_off[img_seg[i + 1]] - seg_off[img_seg[i]];
        if (np_i > max_pts) max_pts = np_i;
    }
    return PBAGPU_OK;
}

// The four binning launches on stream s, given the device copies of img_seg / item_off
// (dseg, ditem_off), seg_off / items (doff, ditems) and the partials buffer.
void launch_binning(hipStream_t s, int n_images, int n_items, long max_pts, const int *dseg, const int *ditem_off,
                    const long *doff, const long *ditems, double *dpart, const double *d_x, const double *d_y,
                    const double *d_w, int nx, int ny, double extend_grid, int edges_given, double *d_edges_x,
                    double *d_edges_y, double *d_hist, unsigned *d_count, double *d_moments, int *d_status) {
    const size_t nb = (size_t)n_images * (nx - 1) * (ny - 1);
    const unsigned zgrid = (unsigned)std::min<size_t>((nb + BLOCK - 1) / BLOCK, 4096);
    hipLaunchKernelGGL(sky_zero_kernel, dim3(zgrid), dim3(BLOCK), 0, s, d_hist, d_count, nb);
    if (n_items > 0)
        hipLaunchKernelGGL(sky_partial_kernel, dim3(n_items), dim3(BLOCK), 0, s, ditems, d_x, d_y, d_w, dpart);
    hipLaunchKernelGGL(sky_extent_kernel, dim3(n_images), dim3(BLOCK), 0, s, dseg, doff, ditem_off, dpart, nx, ny,
                       extend_grid, edges_given, d_edges_x, d_edges_y, d_moments, d_status);
    if (max_pts > 0) {
        const unsigned chunks = (unsigned)std::min<long>((max_pts + BLOCK - 1) / BLOCK, 1024);
        hipLaunchKernelGGL(sky_bin_kernel, dim3(chunks, n_images), dim3(BLOCK), sizeof(double) * (nx + ny), s, dseg,
                           doff, d_x, d_y, d_w, nx, ny, d_edges_x, d_edges_y, d_status, d_hist, d_count);
    }
}

}  // namespace skybin
}  // namespace pbagpu

using namespace pbagpu;

extern "C" int pbagpu_skymap_hist(int n_images, const int *img_seg, const long *seg_off, const double *d_x,
                                  const double *d_y, const double *d_w, int nx, int ny, double extend_grid,
                                  int edges_given, double *d_edges_x, double *d_edges_y, double *d_hist,
                                  unsigned *d_count, double *d_moments, int *d_status, void *stream) {
    using namespace skybin;
    hipStream_t s = (hipStream_t)stream;
    int n_segs = 0;
    long n_pts = 0, max_pts = 0;
    int rc = check_args(n_images, img_seg, seg_off, d_x, d_y, d_w, nx, ny, extend_grid, d_edges_x, d_edges_y, d_hist,
                        d_moments, d_status, "skymap_hist", n_segs, n_pts, max_pts);
    if (rc) return rc;
    // work items of at most ITEM points; every entry has at least one (empty entries: an empty one)
    std::vector<int> item_off(n_segs + 1, 0);
    std::vector<long> items;
    for (int k = 0; k < n_segs; k++) {
        item_off[k] = (int)(items.size() / 2);
        long b = seg_off[k];
        do {
            const long e = std::min<long>(b + ITEM, seg_off[k + 1]);
            items.push_back(b); items.push_back(e);
            b = e;
        } while (b < seg_off[k + 1]);
    }
    item_off[n_segs] = (int)(items.size() / 2);
    const int n_items = item_off[n_segs];
    auto *dseg = (int *)workspace(WS_SKY_SEG, sizeof(int) * ((size_t)n_images + 1 + n_segs + 1), s);
    auto *doff = (long *)workspace(WS_SKY_OFF, sizeof(long) * ((size_t)n_segs + 1 + items.size()), s);
    auto *dpart = (double *)workspace(WS_SKY_PART, sizeof(double) * (size_t)n_items * P_N, s);
    if (!dseg || !doff || !dpart) return PBAGPU_E_HIP;
    int *ditem_off = dseg + n_images + 1;
    long *ditems = doff + n_segs + 1;
    PBA_TRY(rc, hipMemcpyAsync(dseg, img_seg, sizeof(int) * (n_images + 1), hipMemcpyHostToDevice, s));
    PBA_TRY(rc, hipMemcpyAsync(ditem_off, item_off.data(), sizeof(int) * (n_segs + 1), hipMemcpyHostToDevice, s));
    PBA_TRY(rc, hipMemcpyAsync(doff, seg_off, sizeof(long) * (n_segs + 1), hipMemcpyHostToDevice, s));
    if (!items.empty())
        PBA_TRY(rc, hipMemcpyAsync(ditems, items.data(), sizeof(long) * items.size(), hipMemcpyHostToDevice, s));
    if (rc) return rc;
    KernelTimer timer(PBAGPU_STAGE_SKY_HIST);
    timer.start(s);
    launch_binning(s, n_images, n_items, max_pts, dseg, ditem_off, doff, ditems, dpart, d_x, d_y, d_w, nx, ny,
                   extend_grid, edges_given, d_edges_x, d_edges_y, d_hist, d_count, d_moments, d_status);
    timer.stop(s);
    PBA_TRY(rc, hipGetLastError());
    PBA_TRY(rc, hipStreamSynchronize(s));
    timer.collect();
    return rc;
}

extern "C" int pbagpu_skymap_hist_noalloc(int n_images, const int *img_seg, const long *seg_off, const double *d_x,
                                          const double *d_y, const double *d_w, int nx, int ny, double extend_grid,
                                          int edges_given, double *d_edges_x, double *d_edges_y, double *d_hist,
                                          unsigned *d_count, double *d_moments, int *d_status, void *stream) {
    using namespace skybin;
    hipStream_t s = (hipStream_t)stream;
    int n_segs = 0;
    long n_pts = 0, max_pts = 0;
    int rc = check_args(n_images, img_seg, seg_off, d_x, d_y, d_w, nx, ny, extend_grid, d_edges_x, d_edges_y, d_hist,
                        d_moments, d_status, "skymap_hist_noalloc", n_segs, n_pts, max_pts);
    if (rc) return rc;
    long n_items = 0;
    for (int k = 0; k < n_segs; k++) n_items += items_of(seg_off[k], seg_off[k + 1]);
    // scratch layout of this call; sizes beyond the scratch are rejected before any device work
    const size_t b_seg = align256(sizeof(int) * ((size_t)n_images + 1 + (size_t)n_segs + 1));
    const size_t b_off = align256(sizeof(long) * ((size_t)n_segs + 1 + 2 * (size_t)n_items));
    const size_t b_part = align256(sizeof(double) * (size_t)n_items * P_N);
    if (n_items > INT_MAX || b_seg + b_off + b_part > SCRATCH_BYTES) {
        char msg[224];
        snprintf(msg, sizeof(msg),
                 "skymap_hist_noalloc: %d images, %d entries and %ld work items need %zu bytes of scratch; "
                 "the limit is %zu bytes", n_images, n_segs, n_items, b_seg + b_off + b_part, SCRATCH_BYTES);
        set_error_msg(msg);
        return PBAGPU_E_ARG;
    }
    auto *base = (unsigned char *)scratch_address();
    if (!base) return PBAGPU_E_HIP;
    int *dseg = (int *)base;
    int *ditem_off = dseg + n_images + 1;
    long *doff = (long *)(base + b_seg);
    long *ditems = doff + n_segs + 1;
    double *dpart = (double *)(base + b_seg + b_off);
    PBA_TRY(rc, hipMemcpyAsync(dseg, img_seg, sizeof(int) * (n_images + 1), hipMemcpyHostToDevice, s));
    PBA_TRY(rc, hipMemcpyAsync(doff, seg_off, sizeof(long) * (n_segs + 1), hipMemcpyHostToDevice, s));
    if (rc) return rc;
    hipLaunchKernelGGL(sky_items_kernel, dim3(1), dim3(BLOCK), 0, s, n_segs, (const long *)doff, ditem_off, ditems);
    launch_binning(s, n_images, (int)n_items, max_pts, dseg, ditem_off, doff, ditems, dpart, d_x, d_y, d_w, nx, ny,
                   extend_grid, edges_given, d_edges_x, d_edges_y, d_hist, d_count, d_moments, d_status);
    PBA_TRY(rc, hipGetLastError());
    PBA_TRY(rc, hipStreamSynchronize(s));
    return rc;
}
