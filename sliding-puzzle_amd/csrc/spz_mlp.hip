// spz_mlp.hip -- bf16 MFMA kernels of the subtree-size predictor (see
// spz_mlp.h): forward, backward, Adam, and frontier prediction.
#include "spz_mlp.h"
#include "spz_board.h"

#include <vector>

namespace spz {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__host__ __device__ inline uint16_t f2bf(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u); // round to nearest even
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t b) {
    uint32_t u = (uint32_t)b << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

// C(m,n) [+]= sum_k A(m,k) B(k,n), bf16 operands, fp32 accumulate, using
// v_mfma_f32_16x16x16_bf16.  A(m,k) = A[m*sam + k*sak], B(k,n) =
// B[k*sbk + n*sbn], C row-major with leading dimension ldc.  One wave per
// 16x16 output tile, four waves per block along M.  Operand layout of the
// instruction: lane l holds A(l%16, 4(l/16)+j) and B(4(l/16)+j, l%16), and
// receives D(4(l/16)+j, l%16), j = 0..3.
__global__ void __launch_bounds__(256) k_gemm_bf16(const uint16_t *__restrict__ A, long sam, long sak,
                                                   const uint16_t *__restrict__ B, long sbk, long sbn, float *C,
                                                   int ldc, int M, int N, int K, int accumulate) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m0 = (blockIdx.y * 4 + wave) * 16, n0 = blockIdx.x * 16;
    if (m0 >= M) return; // wave-uniform
    const int r = lane & 15, q = lane >> 4;
    const bool mrow = (m0 + r) < M, ncol = (n0 + r) < N;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 16) {
        bf16x4 a, b;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int k = k0 + q * 4 + j;
            const bool kin = k < K;
            a[j] = (short)((mrow && kin) ? A[(long)(m0 + r) * sam + (long)k * sak] : 0);
            b[j] = (short)((ncol && kin) ? B[(long)k * sbk + (long)(n0 + r) * sbn] : 0);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int row = m0 + q * 4 + j, col = n0 + r;
        if (row < M && col < N) {
            float v = acc[j];
            if (accumulate) v += C[(long)row * ldc + col];
            C[(long)row * ldc + col] = v;
        }
    }
}

static hipError_t gemm(const uint16_t *A, long sam, long sak, const uint16_t *B, long sbk, long sbn, float *C,
                       int ldc, int M, int N, int K, int acc, hipStream_t s) {
    dim3 grid((unsigned)((N + 15) / 16), (unsigned)((M + 63) / 64));
    hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(256), 0, s, A, sam, sak, B, sbk, sbn, C, ldc, M, N, K, acc);
    return hipGetLastError();
}

// features of Items at thresholds (one wave per row, 10 columns per lane);
// with perm, row r is built from items[perm[r]] (and its target copied)
template <int W>
__global__ void k_features(const Item *items, const int *thr, int n, uint16_t *x, const int *perm, const float *tgt,
                           float *tout) {
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    const int src = perm ? perm[row] : row;
    const Item it = items[src];
    if (perm && lane == 0) tout[row] = tgt[src];
    Board<W> bd;
    bd.unpack(it.b);
    const int h = meta_h(it.meta);
    const int L = thr[src];
    constexpr int N2 = W * W;
    for (int k = lane; k < kMlpIn; k += 64) {
        float v = 0.f;
        if (k < N2 * N2) {
            int t = k / N2, c = k - t * N2;
            v = (bd.tile(c) == t) ? 1.f : 0.f;
        } else if (k == 625) {
            v = h * (1.f / 32.f);
        } else if (k == 626) {
            v = (L - h) * (1.f / 32.f);
        }
        x[(long)row * kMlpIn + k] = f2bf(v);
    }
}

// z += bias; h = bf16(relu(z))
__global__ void k_bias_relu(float *z, const float *bias, uint16_t *h, int n) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)n * kMlpH) return;
    float v = z[i] + bias[i % kMlpH];
    z[i] = v;
    h[i] = f2bf(v > 0.f ? v : 0.f);
}

// y = h2 . w3 + b3 (one wave per row; bf16 weights, fp32 sum)
__global__ void k_out(const uint16_t *h2, const uint16_t *wbf, const float *master, float *y, int n) {
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    float acc = 0.f;
    for (int j = lane; j < kMlpH; j += 64) acc += bf2f(h2[(long)row * kMlpH + j]) * bf2f(wbf[kMlpOffW3 + j]);
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane == 0) y[row] = acc + master[kMlpOffB3];
}

// dy = 2 (y - t) / n; loss += (y - t)^2 / n
__global__ void k_loss(const float *y, const float *t, int n, float *dy, float *loss) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    float e = 0.f;
    if (i < n) {
        float d = y[i] - t[i];
        dy[i] = 2.f * d / n;
        e = d * d / n;
    }
    for (int off = 32; off > 0; off >>= 1) e += __shfl_xor(e, off);
    if ((threadIdx.x & 63) == 0) atomicAdd(loss, e);
}

// output layer: dz2[i][j] = dy[i] w3[j] [z2 > 0] (fp32 and bf16), one
// thread per element
__global__ void k_out_dz(const float *z2, const float *dy, const uint16_t *wbf, float *dz, uint16_t *dzb, int n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)n * kMlpH) return;
    const int row = (int)(i / kMlpH), j = (int)(i - (long)row * kMlpH);
    const float v = z2[i] > 0.f ? dy[row] * bf2f(wbf[kMlpOffW3 + j]) : 0.f;
    dz[i] = v;
    dzb[i] = f2bf(v);
}

// Column sums over n rows of 256 columns in two passes with a fixed order
// (so training is deterministic): block b sums rows [b*kSumRows, ...) into
// part[b][j], then k_colsum_fin adds the parts.  Weighted form: sum_i
// bf16 a[i][j] * w[i] (the output weights' gradient, h2^T dy).
constexpr int kSumRows = 64;
__global__ void __launch_bounds__(256) k_colsum_part(const float *a, int n, float *part) {
    const int j = threadIdx.x, b = blockIdx.x;
    const int r1 = min(n, (b + 1) * kSumRows);
    float s = 0.f;
    for (int i = b * kSumRows; i < r1; ++i) s += a[(long)i * kMlpH + j];
    part[(long)b * kMlpH + j] = s;
}
__global__ void __launch_bounds__(256) k_colsum_part_w(const uint16_t *a, const float *w, int n, float *part) {
    const int j = threadIdx.x, b = blockIdx.x;
    const int r1 = min(n, (b + 1) * kSumRows);
    float s = 0.f;
    for (int i = b * kSumRows; i < r1; ++i) s += bf2f(a[(long)i * kMlpH + j]) * w[i];
    part[(long)b * kMlpH + j] = s;
}
__global__ void __launch_bounds__(256) k_colsum_fin(const float *part, int nparts, float *out) {
    const int j = threadIdx.x;
    float s = 0.f;
    for (int b = 0; b < nparts; ++b) s += part[(long)b * kMlpH + j];
    out[j] = s;
}
// sum of n values, one block (the output bias gradient)
__global__ void __launch_bounds__(256) k_sum(const float *v, int n, float *out) {
    __shared__ float sh[4];
    float s = 0.f;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += v[i];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[0] = sh[0] + sh[1] + sh[2] + sh[3];
}

static hipError_t colsum(const float *a, int n, float *part, float *out, hipStream_t s) {
    const int nb = (n + kSumRows - 1) / kSumRows;
    hipLaunchKernelGGL(k_colsum_part, dim3((unsigned)nb), dim3(kMlpH), 0, s, a, n, part);
    hipLaunchKernelGGL(k_colsum_fin, dim3(1), dim3(kMlpH), 0, s, part, nb, out);
    return hipGetLastError();
}

// dz1 = dh1 [z1 > 0] in place (dh1 given in dz), bf16 copy
__global__ void k_relu_grad(float *dz, const float *z, uint16_t *dzb, int n) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)n * kMlpH) return;
    float v = z[i] > 0.f ? dz[i] : 0.f;
    dz[i] = v;
    dzb[i] = f2bf(v);
}

__global__ void k_adam(float *w, const float *g, float *m, float *v, uint16_t *wbf, long n, float lr, float b1t,
                       float b2t) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    float gi = g[i];
    float mi = b1 * m[i] + (1.f - b1) * gi;
    float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    float mh = mi / (1.f - b1t), vh = vi / (1.f - b2t);
    float wi = w[i] - lr * mh / (sqrtf(vh) + eps);
    w[i] = wi;
    wbf[i] = f2bf(wi);
}

__global__ void k_to_bf16(const float *w, uint16_t *wbf, long n) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) wbf[i] = f2bf(w[i]);
}

// ---------------------------------------------------------------------------
hipError_t mlp_alloc(MlpState &m) {
    hipError_t e;
    if ((e = hipMalloc(&m.master, kMlpParams * 4)) != hipSuccess) return e;
    if ((e = hipMalloc(&m.wbf, kMlpParams * 2)) != hipSuccess) return e;
    if ((e = hipMalloc(&m.grad, kMlpParams * 4)) != hipSuccess) return e;
    if ((e = hipMalloc(&m.adam_m, kMlpParams * 4)) != hipSuccess) return e;
    if ((e = hipMalloc(&m.adam_v, kMlpParams * 4)) != hipSuccess) return e;
    if ((e = hipMalloc(&m.loss, 4)) != hipSuccess) return e;
    return hipSuccess;
}

static void fr(void *p) {
    if (p) (void)hipFree(p);
}

static void free_scratch(MlpState &m) {
    fr(m.x);
    fr(m.z1);
    fr(m.z2);
    fr(m.h1);
    fr(m.h2);
    fr(m.y);
    fr(m.dz);
    fr(m.dzb);
    fr(m.part);
    m.part = nullptr;
    m.x = nullptr;
    m.z1 = m.z2 = nullptr;
    m.h1 = m.h2 = nullptr;
    m.y = nullptr;
    m.dz = nullptr;
    m.dzb = nullptr;
    m.cap = 0;
}

void mlp_free(MlpState &m) {
    free_scratch(m);
    fr(m.master);
    fr(m.wbf);
    fr(m.grad);
    fr(m.adam_m);
    fr(m.adam_v);
    fr(m.loss);
    fr(m.w2frag);
    m = MlpState();
}

hipError_t mlp_reserve(MlpState &m, size_t rows) {
    if (rows <= m.cap) return hipSuccess;
    free_scratch(m);
    hipError_t e;
    if ((e = hipMalloc(&m.x, rows * kMlpIn * 2)) != hipSuccess) return e;
    if ((e = hipMalloc(&m.z1, rows * kMlpH * 4)) != hipSuccess) return e;
    if ((e = hipMalloc(&m.z2, rows * kMlpH * 4)) != hipSuccess) return e;
    if ((e = hipMalloc(&m.h1, rows * kMlpH * 2)) != hipSuccess) return e;
    if ((e = hipMalloc(&m.h2, rows * kMlpH * 2)) != hipSuccess) return e;
    if ((e = hipMalloc(&m.y, rows * 4)) != hipSuccess) return e;
    if ((e = hipMalloc(&m.dz, rows * kMlpH * 4)) != hipSuccess) return e;
    if ((e = hipMalloc(&m.dzb, rows * kMlpH * 2)) != hipSuccess) return e;
    if ((e = hipMalloc(&m.part, ((rows + kSumRows - 1) / kSumRows) * kMlpH * 4)) != hipSuccess) return e;
    m.cap = rows;
    return hipSuccess;
}

static hipError_t refresh_bf16(MlpState &m, hipStream_t s) {
    hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((kMlpParams + 255) / 256)), dim3(256), 0, s, m.master, m.wbf,
                       (long)kMlpParams);
    return hipGetLastError();
}

hipError_t mlp_set_params(MlpState &m, const float *host, hipStream_t s) {
    hipError_t e = hipMemcpyAsync(m.master, host, kMlpParams * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemsetAsync(m.adam_m, 0, kMlpParams * 4, s);
    if (e == hipSuccess) e = hipMemsetAsync(m.adam_v, 0, kMlpParams * 4, s);
    if (e == hipSuccess) e = refresh_bf16(m, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    m.adam_t = 0;
    m.ready = (e == hipSuccess);
    return e;
}

hipError_t mlp_get_params(MlpState &m, float *host, hipStream_t s) {
    hipError_t e = hipMemcpyAsync(host, m.master, kMlpParams * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e;
}

// He-uniform initialisation from a seeded splitmix64 stream
hipError_t mlp_init_random(MlpState &m, uint64_t seed, hipStream_t s) {
    std::vector<float> p(kMlpParams, 0.f);
    uint64_t x = seed;
    auto next = [&]() {
        uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        z ^= z >> 31;
        return (double)(z >> 11) / 9007199254740992.0; // [0,1)
    };
    const double a1 = sqrt(6.0 / 64.0), a2 = sqrt(6.0 / kMlpH), a3 = sqrt(6.0 / kMlpH);
    for (size_t i = 0; i < (size_t)kMlpIn * kMlpH; ++i) p[kMlpOffW1 + i] = (float)((2 * next() - 1) * a1 * 0.25);
    for (size_t i = 0; i < (size_t)kMlpH * kMlpH; ++i) p[kMlpOffW2 + i] = (float)((2 * next() - 1) * a2);
    for (int i = 0; i < kMlpH; ++i) p[kMlpOffW3 + i] = (float)((2 * next() - 1) * a3);
    for (int i = 0; i < kMlpH; ++i) p[kMlpOffB1 + i] = 0.05f;
    for (int i = 0; i < kMlpH; ++i) p[kMlpOffB2 + i] = 0.05f;
    p[kMlpOffB3] = 10.f; // log2 sizes are O(10)
    return mlp_set_params(m, p.data(), s);
}

hipError_t mlp_features_items(MlpState &m, int width, const Item *items, const int *thr, int n, hipStream_t s) {
    dim3 grid((unsigned)((n + 3) / 4));
    if (width == 4)
        hipLaunchKernelGGL((k_features<4>), grid, dim3(256), 0, s, items, thr, n, m.x, nullptr, nullptr, nullptr);
    else
        hipLaunchKernelGGL((k_features<5>), grid, dim3(256), 0, s, items, thr, n, m.x, nullptr, nullptr, nullptr);
    return hipGetLastError();
}

hipError_t mlp_features_gather(MlpState &m, int width, const Item *items, const int *thr, const float *tgt,
                               const int *perm, int n, float *tout, hipStream_t s) {
    dim3 grid((unsigned)((n + 3) / 4));
    if (width == 4)
        hipLaunchKernelGGL((k_features<4>), grid, dim3(256), 0, s, items, thr, n, m.x, perm, tgt, tout);
    else
        hipLaunchKernelGGL((k_features<5>), grid, dim3(256), 0, s, items, thr, n, m.x, perm, tgt, tout);
    return hipGetLastError();
}

hipError_t mlp_forward(MlpState &m, int n, hipStream_t s) {
    hipError_t e;
    // z1 = x W1
    if ((e = gemm(m.x, kMlpIn, 1, m.wbf + kMlpOffW1, kMlpH, 1, m.z1, kMlpH, n, kMlpH, kMlpIn, 0, s)) != hipSuccess)
        return e;
    unsigned g = (unsigned)(((long)n * kMlpH + 255) / 256);
    hipLaunchKernelGGL(k_bias_relu, dim3(g), dim3(256), 0, s, m.z1, m.master + kMlpOffB1, m.h1, n);
    // z2 = h1 W2
    if ((e = gemm(m.h1, kMlpH, 1, m.wbf + kMlpOffW2, kMlpH, 1, m.z2, kMlpH, n, kMlpH, kMlpH, 0, s)) != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_bias_relu, dim3(g), dim3(256), 0, s, m.z2, m.master + kMlpOffB2, m.h2, n);
    hipLaunchKernelGGL(k_out, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, m.h2, m.wbf, m.master, m.y, n);
    return hipGetLastError();
}

// Requires mlp_forward on the same rows.  Weights and activations are bf16,
// biases fp32 (their gradients are fp32 sums of the fp32 dz).  The batch's
// mean squared error is added to *loss_acc (loss_acc = m.loss after zeroing
// it when null), so an epoch's losses can add up on the device.
hipError_t mlp_backward(MlpState &m, const float *target, int n, hipStream_t s, float *loss_acc) {
    hipError_t e;
    float *dy = m.y + 0; // overwritten by dy after the loss (forward outputs consumed)
    if (!loss_acc) {
        if ((e = hipMemsetAsync(m.loss, 0, 4, s)) != hipSuccess) return e;
        loss_acc = m.loss;
    }
    // dy overwrites y in place: each thread reads y[i] before writing dy[i]
    hipLaunchKernelGGL(k_loss, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, m.y, target, n, dy, loss_acc);
    const int nb = (n + kSumRows - 1) / kSumRows;
    // gw3 = h2^T dy, gb3 = sum dy
    hipLaunchKernelGGL(k_colsum_part_w, dim3((unsigned)nb), dim3(kMlpH), 0, s, m.h2, dy, n, m.part);
    hipLaunchKernelGGL(k_colsum_fin, dim3(1), dim3(kMlpH), 0, s, m.part, nb, m.grad + kMlpOffW3);
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, s, dy, n, m.grad + kMlpOffB3);
    unsigned g = (unsigned)(((long)n * kMlpH + 255) / 256);
    hipLaunchKernelGGL(k_out_dz, dim3(g), dim3(256), 0, s, m.z2, dy, m.wbf, m.dz, m.dzb, n);
    // gW2 = h1^T dz2 ; gb2
    if ((e = gemm(m.h1, 1, kMlpH, m.dzb, kMlpH, 1, m.grad + kMlpOffW2, kMlpH, kMlpH, kMlpH, n, 0, s)) != hipSuccess)
        return e;
    if ((e = colsum(m.dz, n, m.part, m.grad + kMlpOffB2, s)) != hipSuccess) return e;
    // dh1 = dz2 W2^T  (into z2's buffer is not allowed: z1 needed; use dz)
    float *dh1 = m.dz; // dz2 (fp32) no longer needed after gW2/gb2; dzb still holds dz2 in bf16
    if ((e = gemm(m.dzb, kMlpH, 1, m.wbf + kMlpOffW2, 1, kMlpH, dh1, kMlpH, n, kMlpH, kMlpH, 0, s)) != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_relu_grad, dim3(g), dim3(256), 0, s, dh1, m.z1, m.dzb, n);
    // gW1 = x^T dz1 ; gb1
    if ((e = gemm(m.x, 1, kMlpIn, m.dzb, kMlpH, 1, m.grad + kMlpOffW1, kMlpH, kMlpIn, kMlpH, n, 0, s)) != hipSuccess)
        return e;
    if ((e = colsum(dh1, n, m.part, m.grad + kMlpOffB1, s)) != hipSuccess) return e;
    return hipGetLastError();
}

hipError_t mlp_adam_step(MlpState &m, float lr, hipStream_t s) {
    m.adam_t += 1;
    float b1t = powf(0.9f, (float)m.adam_t), b2t = powf(0.999f, (float)m.adam_t);
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((kMlpParams + 255) / 256)), dim3(256), 0, s, m.master, m.grad,
                       m.adam_m, m.adam_v, m.wbf, (long)kMlpParams, lr, b1t, b2t);
    return hipGetLastError();
}

// W2 in the B-fragment order of v_mfma_f32_16x16x16_bf16: fragment
// (kb, nb) of lane l holds W2[kb*16 + 4(l/16) + j][nb*16 + l%16], j = 0..3,
// so each lane loads its 4 values as one 8-byte word.
__global__ void k_w2frag(const uint16_t *wbf, uint16_t *frag) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x; // (kb*16 + nb)*64 + lane
    if (t >= 16 * 16 * 64) return;
    const int lane = t & 63, f = t >> 6, kb = f >> 4, nb = f & 15;
    const int q = lane >> 4, r = lane & 15;
#pragma unroll
    for (int j = 0; j < 4; ++j)
        frag[(size_t)t * 4 + j] = wbf[kMlpOffW2 + (size_t)(kb * 16 + q * 4 + j) * kMlpH + nb * 16 + r];
}

// Fused prediction: one wave per node, its kPredThr = 16 thresholds
// L0 + 2m (m = 0..15) as the 16 rows of one MFMA tile.  Layer 1 of a row is
// the node's one-hot and h/32 columns of W1 (shared by the 16 rows) plus
// (L - h)/32 times row 626 of W1; layers 2 and 3 follow in registers.
// Waves map to node slots (g, i), g < n_groups, i < stride; a slot is live
// if i < count[g] (or count is null).  L0 = h of roots[g] if given, else
// thr0[g*stride + i].  Rounding follows the forward pass: bf16 inputs and
// weights, fp32 sums, bf16 activations after ReLU.
template <int W>
__global__ void __launch_bounds__(256) k_predict16(const Item *items, int stride, const unsigned *count, int n_groups,
                                                   const Item *roots, const int *thr0, const uint16_t *wbf,
                                                   const float *master, const uint16_t *w2frag, float *out) {
    __shared__ float base_s[4][kMlpH];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long slot = (long)blockIdx.x * 4 + wave;
    const int g = (int)(slot / stride), i = (int)(slot - (long)g * stride);
    bool live = g < n_groups;
    if (live && count) live = (unsigned)i < (count[g] < (unsigned)stride ? count[g] : (unsigned)stride);
    if (!live) return; // wave-uniform
    const Item it = items[slot];
    Board<W> bd;
    bd.unpack(it.b);
    const int h = meta_h(it.meta);
    const int L0 = roots ? meta_h(roots[g].meta) : thr0[slot];
    constexpr int N2 = W * W;
    const float xh = bf2f(f2bf(h * (1.f / 32.f)));
    // layer 1, shared part: sum of the one-hot rows, the h row and b1
    float *base = base_s[wave];
#pragma unroll
    for (int cc = 0; cc < 4; ++cc) {
        const int col = cc * 64 + lane;
        float acc = 0.f;
#pragma unroll
        for (int c = 0; c < N2; ++c) acc += bf2f(wbf[kMlpOffW1 + (size_t)(bd.tile(c) * N2 + c) * kMlpH + col]);
        acc += xh * bf2f(wbf[kMlpOffW1 + (size_t)625 * kMlpH + col]);
        base[col] = acc + master[kMlpOffB1 + col];
    }
    // base is this wave's own: LDS accesses of a wave complete in order, so
    // only the compiler must keep the reads after the writes
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // layer 1 rows in A-fragment order: lane l holds row m = l%16 (threshold
    // L0 + 2m), columns kb*16 + 4(l/16) + j
    const int r = lane & 15, q = lane >> 4;
    const float xs = bf2f(f2bf((L0 + 2 * r - h) * (1.f / 32.f)));
    bf16x4 a[16];
#pragma unroll
    for (int kb = 0; kb < 16; ++kb) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int col = kb * 16 + q * 4 + j;
            float v = base[col] + xs * bf2f(wbf[kMlpOffW1 + (size_t)626 * kMlpH + col]);
            a[kb][j] = (short)f2bf(v > 0.f ? v : 0.f);
        }
    }
    // layers 2 and 3: per output block nb, 16 MFMAs over K, then bias, ReLU,
    // bf16, and the dot product with w3 (lane holds rows 4q+j, column nb*16+r)
    float part[4] = {0.f, 0.f, 0.f, 0.f};
    const uint2 *fr = (const uint2 *)w2frag;
#pragma unroll 1
    for (int nb = 0; nb < 16; ++nb) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < 16; ++kb) {
            const uint2 w = fr[(kb * 16 + nb) * 64 + lane];
            bf16x4 b;
            b[0] = (short)(w.x & 0xffffu);
            b[1] = (short)(w.x >> 16);
            b[2] = (short)(w.y & 0xffffu);
            b[3] = (short)(w.y >> 16);
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[kb], b, acc, 0, 0, 0);
        }
        const int col = nb * 16 + r;
        const float b2 = master[kMlpOffB2 + col];
        const float w3 = bf2f(wbf[kMlpOffW3 + col]);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float z = acc[j] + b2;
            part[j] += bf2f(f2bf(z > 0.f ? z : 0.f)) * w3;
        }
    }
    // sum over the 16 lanes holding the same rows
#pragma unroll
    for (int j = 0; j < 4; ++j)
        for (int off = 8; off > 0; off >>= 1) part[j] += __shfl_xor(part[j], off, 16);
    if (r == 0) {
        const float b3 = master[kMlpOffB3];
#pragma unroll
        for (int j = 0; j < 4; ++j) out[slot * kPredThr + q * 4 + j] = part[j] + b3;
    }
}

static hipError_t predict_launch(MlpState &m, int width, const Item *items, int stride, const unsigned *count,
                                 int n_groups, const Item *roots, const int *thr0, float *out, hipStream_t s) {
    hipError_t e = hipSuccess;
    if (!m.w2frag && (e = hipMalloc(&m.w2frag, (size_t)16 * 16 * 64 * 4 * 2)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_w2frag, dim3(64), dim3(256), 0, s, m.wbf, m.w2frag);
    const long waves = (long)n_groups * stride;
    dim3 grid((unsigned)((waves + 3) / 4));
    if (width == 4)
        hipLaunchKernelGGL((k_predict16<4>), grid, dim3(256), 0, s, items, stride, count, n_groups, roots, thr0, m.wbf,
                           m.master, m.w2frag, out);
    else
        hipLaunchKernelGGL((k_predict16<5>), grid, dim3(256), 0, s, items, stride, count, n_groups, roots, thr0, m.wbf,
                           m.master, m.w2frag, out);
    return hipGetLastError();
}

hipError_t mlp_predict_frontier(MlpState &m, int width, const Item *frontier, const unsigned *nfront,
                                const Item *roots, int n_inst, int capl, float *pred, hipStream_t s) {
    return predict_launch(m, width, frontier, capl, nfront, n_inst, roots, nullptr, pred, s);
}

hipError_t mlp_predict_items(MlpState &m, int width, const Item *items, const int *thr0, int n, float *out,
                             hipStream_t s) {
    return predict_launch(m, width, items, n, nullptr, 1, nullptr, thr0, out, s);
}

} // namespace spz
