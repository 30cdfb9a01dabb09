// Multilayer perceptron for the learned SSC emissivity: BF16 operands on MFMA with FP32
// accumulation, SiLU hidden layers, linear output; forward pass, mean-square-error backward pass
// and Adam update, all on the GPU.  The network has no counterpart in the reference; it is checked
// against the plain SSC contraction (computeSSCSpectrum, src/microphysics/numeric_model.h:247-285).
//
// Every product is one kernel, gemm_bf16: C[M][N] (FP32) = sum_k A(m, k) B(k, n) with A and B read
// through strides, so that the forward product X W, the input gradient dZ W^T and the weight
// gradient H^T dZ are the same kernel.  A workgroup of four wavefronts computes a 32 x 32 tile; each
// wavefront one 16 x 16 block with v_mfma_f32_16x16x32_bf16, taking K in steps of 32.  In that
// instruction lane l holds row (l mod 16) of A and column (l mod 16) of B, eight consecutive k of
// the step from 8 (l / 16); the four results of lane l are rows 4 (l / 16) + j of column (l mod 16).
#include "pba_device.h"
#include "timing.h"
#include "runtime.h"
#include "mlp_internal.h"
#include <vector>

namespace pbagpu {
namespace mlp {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

__device__ __forceinline__ __bf16 to_bf16(float x) {
    // round to nearest even (NaN kept a NaN)
    unsigned u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return __builtin_bit_cast(__bf16, (unsigned short)((u >> 16) | 0x40));
    u += 0x7fffu + ((u >> 16) & 1u);
    return __builtin_bit_cast(__bf16, (unsigned short)(u >> 16));
}
__device__ __forceinline__ float from_bf16(__bf16 h) {
    return __uint_as_float(((unsigned)__builtin_bit_cast(unsigned short, h)) << 16);
}

__global__ void __launch_bounds__(256) gemm_bf16(const __bf16 *__restrict__ A, long sam, long sak,
                                                 const __bf16 *__restrict__ B, long sbk, long sbn,
                                                 float *__restrict__ C, int M, int N, int K) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m0 = blockIdx.y * 32 + (wave >> 1) * 16, n0 = blockIdx.x * 32 + (wave & 1) * 16;
    const int r = lane & 15, g = lane >> 4;
    const int m = m0 + r, n = n0 + r;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    const __bf16 zero = __builtin_bit_cast(__bf16, (unsigned short)0);
    for (int k0 = 0; k0 < K; k0 += 32) {
        bf16x8 a, b;
        for (int j = 0; j < 8; j++) {
            const int k = k0 + 8 * g + j;
            a[j] = (m < M && k < K) ? A[(long)m * sam + (long)k * sak] : zero;
            b[j] = (n < N && k < K) ? B[(long)k * sbk + (long)n * sbn] : zero;
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
    for (int j = 0; j < 4; j++) {
        const int mm = m0 + 4 * g + j;
        if (mm < M && n < N) C[(long)mm * N + n] = acc[j];
    }
}

__global__ void to_bf16_kernel(const float *__restrict__ x, __bf16 *__restrict__ y, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) y[i] = to_bf16(x[i]);
}

__device__ __forceinline__ float silu(float z) { return z / (1.f + __expf(-z)); }
__device__ __forceinline__ float silu_grad(float z) {
    const float s = 1.f / (1.f + __expf(-z));
    return s * (1.f + z * (1.f - s));
}

// Z += b (row broadcast); hidden layers also write H = SiLU(Z) in BF16 for the next product
__global__ void bias_act_kernel(float *__restrict__ Z, const float *__restrict__ bias, int rows, int cols, int act,
                                __bf16 *__restrict__ H, float *__restrict__ out) {
    const long n = (long)rows * cols;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float z = Z[i] + bias[i % cols];
        Z[i] = z;
        if (act) H[i] = to_bf16(silu(z));
        if (out) out[i] = z;
    }
}

// output gradient of the mean square error over rows x cols: dZ = 2 (Y - T) / (rows cols); loss sum
__global__ void mse_grad_kernel(const float *__restrict__ Y, const float *__restrict__ T, int rows, int cols,
                                __bf16 *__restrict__ dZ, float *__restrict__ dZf, float *__restrict__ loss) {
    const long n = (long)rows * cols;
    const float sc = 2.f / (float)n;
    float part = 0.f;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float d = Y[i] - T[i];
        part += d * d;
        dZf[i] = sc * d;
        dZ[i] = to_bf16(sc * d);
    }
    for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
    if ((threadIdx.x & 63) == 0 && loss) atomicAdd(loss, part / (float)n);
}

// dZ_prev = dH * SiLU'(Z_prev), dH in FP32 from the product with W^T
__global__ void act_back_kernel(const float *__restrict__ dH, const float *__restrict__ Zp, long n,
                                __bf16 *__restrict__ dZ, float *__restrict__ dZf) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float v = dH[i] * silu_grad(Zp[i]);
        dZf[i] = v;
        dZ[i] = to_bf16(v);
    }
}

// bias gradient: column sums of dZ (FP32), one thread per column
__global__ void colsum_kernel(const float *__restrict__ dZ, int rows, int cols, float *__restrict__ db) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    float s = 0.f;
    for (int r = 0; r < rows; r++) s += dZ[(long)r * cols + c];
    db[c] = s;
}

__global__ void adam_kernel(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                            float *__restrict__ v, long n, float lr, float b1, float b2, float eps, float c1, float c2) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

inline unsigned grid1d(long n) { return (unsigned)std::min<long>((n + 255) / 256, 4096); }

void gemm(hipStream_t s, const __bf16 *A, long sam, long sak, const __bf16 *B, long sbk, long sbn, float *C, int M,
          int N, int K) {
    hipLaunchKernelGGL(gemm_bf16, dim3((unsigned)((N + 31) / 32), (unsigned)((M + 31) / 32)), dim3(256), 0, s, A, sam,
                       sak, B, sbk, sbn, C, M, N, K);
}

// Device workspace of one call: BF16 weights, BF16 activations, FP32 pre-activations and gradients.
struct Work {
    __bf16 *wb[PBA_MLP_MAXL];     // BF16 copies of the weights
    __bf16 *hb[PBA_MLP_MAXL + 1]; // BF16 layer inputs (hb[0] = x)
    float *z[PBA_MLP_MAXL];       // FP32 pre-activations (z[L-1] = output)
    float *dh;                    // FP32 product dZ W^T, widest layer
    float *dzf;                   // FP32 dZ of the current layer
    __bf16 *dzb;                  // BF16 dZ of the current layer
    __bf16 *dzb2;                 // BF16 dZ of the previous layer
};

static size_t carve(const pba_mlp &net, int batch, Work *w, char *base) {
    size_t off = 0;
    auto take = [&](size_t bytes) { char *p = base ? base + off : nullptr; off += (bytes + 255) / 256 * 256; return p; };
    int wmax = 0;
    for (int l = 0; l <= net.n_layers; l++) wmax = std::max(wmax, net.dims[l]);
    for (int l = 0; l < net.n_layers; l++) {
        char *p = take(sizeof(__bf16) * (size_t)net.dims[l] * net.dims[l + 1]);
        if (w) w->wb[l] = (__bf16 *)p;
    }
    for (int l = 0; l <= net.n_layers; l++) {
        char *p = take(sizeof(__bf16) * (size_t)batch * net.dims[l]);
        if (w) w->hb[l] = (__bf16 *)p;
    }
    for (int l = 0; l < net.n_layers; l++) {
        char *p = take(sizeof(float) * (size_t)batch * net.dims[l + 1]);
        if (w) w->z[l] = (float *)p;
    }
    char *p1 = take(sizeof(float) * (size_t)batch * wmax), *p2 = take(sizeof(float) * (size_t)batch * wmax);
    char *p3 = take(sizeof(__bf16) * (size_t)batch * wmax), *p4 = take(sizeof(__bf16) * (size_t)batch * wmax);
    if (w) { w->dh = (float *)p1; w->dzf = (float *)p2; w->dzb = (__bf16 *)p3; w->dzb2 = (__bf16 *)p4; }
    return off;
}

static int check_net(const pba_mlp *net) {
    if (!net || net->n_layers < 1 || net->n_layers > PBA_MLP_MAXL) return PBAGPU_E_ARG;
    for (int l = 0; l <= net->n_layers; l++)
        if (net->dims[l] < 1 || net->dims[l] > 65536) return PBAGPU_E_ARG;
    for (int l = 0; l < net->n_layers; l++)
        if (!net->w[l] || !net->b[l]) return PBAGPU_E_ARG;
    return PBAGPU_OK;
}

// Forward pass into the workspace; the output (pre-activation of the last layer) is w.z[L-1].
int forward(const pba_mlp &net, int batch, const float *x, Work &w, hipStream_t s) {
    const int L = net.n_layers;
    for (int l = 0; l < L; l++) {
        const long nw = (long)net.dims[l] * net.dims[l + 1];
        hipLaunchKernelGGL(to_bf16_kernel, dim3(grid1d(nw)), dim3(256), 0, s, net.w[l], w.wb[l], nw);
    }
    const long nx = (long)batch * net.dims[0];
    hipLaunchKernelGGL(to_bf16_kernel, dim3(grid1d(nx)), dim3(256), 0, s, x, w.hb[0], nx);
    for (int l = 0; l < L; l++) {
        const int din = net.dims[l], dout = net.dims[l + 1];
        gemm(s, w.hb[l], din, 1, w.wb[l], dout, 1, w.z[l], batch, dout, din);
        const long n = (long)batch * dout;
        hipLaunchKernelGGL(bias_act_kernel, dim3(grid1d(n)), dim3(256), 0, s, w.z[l], net.b[l], batch, dout,
                           l < L - 1 ? 1 : 0, w.hb[l + 1], (float *)nullptr);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) { set_error("mlp forward", e); return PBAGPU_E_HIP; }
    return PBAGPU_OK;
}

int forward_into(const pba_mlp &net, int batch, const float *x, float *y, int ws_slot, hipStream_t s) {
    Work w;
    const size_t bytes = carve(net, batch, nullptr, nullptr);
    char *base = (char *)workspace(ws_slot, bytes, s);
    if (!base) return PBAGPU_E_HIP;
    carve(net, batch, &w, base);
    int rc = forward(net, batch, x, w, s);
    const int dout = net.dims[net.n_layers];
    if (rc == PBAGPU_OK)
        PBA_TRY(rc, hipMemcpyAsync(y, w.z[net.n_layers - 1], sizeof(float) * (size_t)batch * dout,
                                   hipMemcpyDeviceToDevice, s));
    return rc;
}

int check_network(const pba_mlp *net) { return check_net(net); }

}  // namespace mlp
}  // namespace pbagpu

using namespace pbagpu::mlp;

extern "C" int pbagpu_mlp_forward(const pba_mlp *net, int batch, const float *d_x, float *d_y, void *stream) {
    if (check_net(net) != PBAGPU_OK || batch < 1 || !d_x || !d_y) return PBAGPU_E_ARG;
    hipStream_t s = (hipStream_t)stream;
    int rc = forward_into(*net, batch, d_x, d_y, pbagpu::WS_MLP, s);
    PBA_TRY(rc, hipStreamSynchronize(s));
    return rc;
}

extern "C" int pbagpu_mlp_train_step(const pba_mlp *net, int batch, const float *d_x, const float *d_t, float lr,
                                     int step, float *d_loss, void *stream) {
    if (check_net(net) != PBAGPU_OK || batch < 1 || !d_x || !d_t || step < 1 || !(lr >= 0.f)) return PBAGPU_E_ARG;
    const int L = net->n_layers;
    for (int l = 0; l < L; l++)
        if (!net->gw[l] || !net->gb[l] || !net->mw[l] || !net->vw[l] || !net->mb[l] || !net->vb[l]) return PBAGPU_E_ARG;
    hipStream_t s = (hipStream_t)stream;
    Work w;
    const size_t bytes = carve(*net, batch, nullptr, nullptr);
    char *base = (char *)pbagpu::workspace(pbagpu::WS_MLP, bytes, s);
    if (!base) return PBAGPU_E_HIP;
    carve(*net, batch, &w, base);
    int rc = PBAGPU_OK;
    if (d_loss) PBA_TRY(rc, hipMemsetAsync(d_loss, 0, sizeof(float), s));
    if (rc == PBAGPU_OK) rc = forward(*net, batch, d_x, w, s);
    if (rc != PBAGPU_OK) return rc;
    const int dout = net->dims[L];
    {
        const long n = (long)batch * dout;
        hipLaunchKernelGGL(mse_grad_kernel, dim3(grid1d(n)), dim3(256), 0, s, w.z[L - 1], d_t, batch, dout, w.dzb,
                           w.dzf, d_loss);
    }
    for (int l = L - 1; l >= 0; l--) {
        const int din = net->dims[l], dcur = net->dims[l + 1];
        // dW_l = H_l^T dZ_l: A(m = i, k = r) = H[r][i], B(k = r, n = o) = dZ[r][o]
        gemm(s, w.hb[l], 1, din, w.dzb, dcur, 1, net->gw[l], din, dcur, batch);
        hipLaunchKernelGGL(colsum_kernel, dim3((unsigned)((dcur + 255) / 256)), dim3(256), 0, s, w.dzf, batch, dcur,
                           net->gb[l]);
        if (l > 0) {
            // dH = dZ_l W_l^T: A(m = r, k = o) = dZ[r][o], B(k = o, n = i) = W[i][o]
            gemm(s, w.dzb, dcur, 1, w.wb[l], 1, dcur, w.dh, batch, din, dcur);
            const long n = (long)batch * din;
            hipLaunchKernelGGL(act_back_kernel, dim3(grid1d(n)), dim3(256), 0, s, w.dh, w.z[l - 1], n, w.dzb2, w.dzf);
            std::swap(w.dzb, w.dzb2);
        }
    }
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    for (int l = 0; l < L; l++) {
        const long nw = (long)net->dims[l] * net->dims[l + 1];
        hipLaunchKernelGGL(adam_kernel, dim3(grid1d(nw)), dim3(256), 0, s, net->w[l], net->gw[l], net->mw[l],
                           net->vw[l], nw, lr, b1, b2, eps, c1, c2);
        hipLaunchKernelGGL(adam_kernel, dim3(grid1d(net->dims[l + 1])), dim3(256), 0, s, net->b[l], net->gb[l],
                           net->mb[l], net->vb[l], (long)net->dims[l + 1], lr, b1, b2, eps, c1, c2);
    }
    PBA_TRY(rc, hipGetLastError());
    PBA_TRY(rc, hipStreamSynchronize(s));
    return rc;
}

// Plain product on the MFMA path (the unit the network is built from), for tests: C = A B with
// A [M][K] and B [K][N] row-major FP32 on the device, rounded to BF16, FP32 accumulation.
extern "C" int pbagpu_gemm_bf16(int M, int N, int K, const float *d_a, int trans_a, const float *d_b, int trans_b,
                                float *d_c, void *stream) {
    if (M < 1 || N < 1 || K < 1 || !d_a || !d_b || !d_c) return PBAGPU_E_ARG;
    hipStream_t s = (hipStream_t)stream;
    const size_t na = (size_t)M * K, nb = (size_t)K * N;
    char *base = (char *)pbagpu::workspace(pbagpu::WS_MLP, sizeof(__bf16) * (na + nb) + 512, s);
    if (!base) return PBAGPU_E_HIP;
    __bf16 *a = (__bf16 *)base, *b = (__bf16 *)(base + (sizeof(__bf16) * na + 255) / 256 * 256);
    hipLaunchKernelGGL(to_bf16_kernel, dim3(grid1d((long)na)), dim3(256), 0, s, d_a, a, (long)na);
    hipLaunchKernelGGL(to_bf16_kernel, dim3(grid1d((long)nb)), dim3(256), 0, s, d_b, b, (long)nb);
    // trans_a: the array holds A^T as [K][M]; trans_b: B^T as [N][K]
    gemm(s, a, trans_a ? 1 : K, trans_a ? M : 1, b, trans_b ? 1 : N, trans_b ? K : 1, d_c, M, N, K);
    int rc = PBAGPU_OK;
    PBA_TRY(rc, hipGetLastError());
    PBA_TRY(rc, hipStreamSynchronize(s));
    return rc;
}
