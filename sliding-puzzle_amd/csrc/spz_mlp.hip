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

// features of Items at thresholds (one wave per row, 10 columns per lane)
template <int W>
__global__ void k_features(const Item *items, const int *thr, int n, uint16_t *x) {
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    const Item it = items[row];
    Board<W> bd;
    bd.unpack(it.b);
    const int h = meta_h(it.meta);
    const int L = thr[row];
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

// output layer gradients: gw3[j] = sum_i h2[i][j] dy[i], gb3 = sum dy;
// dz2[i][j] = dy[i] w3[j] [z2 > 0]
__global__ void k_out_grad(const uint16_t *h2, const float *z2, const float *dy, const uint16_t *wbf, float *grad,
                           float *dz, uint16_t *dzb, int n) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x; // column
    if (j >= kMlpH) return;
    float w3 = bf2f(wbf[kMlpOffW3 + j]);
    float g = 0.f, gb = 0.f;
    for (int i = 0; i < n; ++i) {
        g += bf2f(h2[(long)i * kMlpH + j]) * dy[i];
        gb += dy[i];
        float v = z2[(long)i * kMlpH + j] > 0.f ? dy[i] * w3 : 0.f;
        dz[(long)i * kMlpH + j] = v;
        dzb[(long)i * kMlpH + j] = f2bf(v);
    }
    grad[kMlpOffW3 + j] = g;
    if (j == 0) grad[kMlpOffB3] = gb;
}

// column sums (bias gradients) of dz
__global__ void k_colsum(const float *dz, int n, float *out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= kMlpH) return;
    float s = 0.f;
    for (int i = 0; i < n; ++i) s += dz[(long)i * kMlpH + j];
    out[j] = s;
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
    if (p) hipFree(p);
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
        hipLaunchKernelGGL((k_features<4>), grid, dim3(256), 0, s, items, thr, n, m.x);
    else
        hipLaunchKernelGGL((k_features<5>), grid, dim3(256), 0, s, items, thr, n, m.x);
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
// biases fp32 (their gradients are fp32 sums of the fp32 dz).
hipError_t mlp_backward(MlpState &m, const float *target, int n, hipStream_t s) {
    hipError_t e;
    float *dy = m.y + 0; // overwritten by dy after the loss (forward outputs consumed)
    if ((e = hipMemsetAsync(m.loss, 0, 4, s)) != hipSuccess) return e;
    // dy overwrites y in place: each thread reads y[i] before writing dy[i]
    hipLaunchKernelGGL(k_loss, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, m.y, target, n, dy, m.loss);
    hipLaunchKernelGGL(k_out_grad, dim3(1), dim3(kMlpH), 0, s, m.h2, m.z2, dy, m.wbf, m.grad, m.dz, m.dzb, n);
    // gW2 = h1^T dz2 ; gb2
    if ((e = gemm(m.h1, 1, kMlpH, m.dzb, kMlpH, 1, m.grad + kMlpOffW2, kMlpH, kMlpH, kMlpH, n, 0, s)) != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_colsum, dim3(1), dim3(kMlpH), 0, s, m.dz, n, m.grad + kMlpOffB2);
    // dh1 = dz2 W2^T  (into z2's buffer is not allowed: z1 needed; use dz)
    float *dh1 = m.dz; // dz2 (fp32) no longer needed after gW2/gb2; dzb still holds dz2 in bf16
    if ((e = gemm(m.dzb, kMlpH, 1, m.wbf + kMlpOffW2, 1, kMlpH, dh1, kMlpH, n, kMlpH, kMlpH, 0, s)) != hipSuccess)
        return e;
    unsigned g = (unsigned)(((long)n * kMlpH + 255) / 256);
    hipLaunchKernelGGL(k_relu_grad, dim3(g), dim3(256), 0, s, dh1, m.z1, m.dzb, n);
    // gW1 = x^T dz1 ; gb1
    if ((e = gemm(m.x, 1, kMlpIn, m.dzb, kMlpH, 1, m.grad + kMlpOffW1, kMlpH, kMlpIn, kMlpH, n, 0, s)) != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_colsum, dim3(1), dim3(kMlpH), 0, s, dh1, n, m.grad + kMlpOffB1);
    return hipGetLastError();
}

hipError_t mlp_adam_step(MlpState &m, float lr, hipStream_t s) {
    m.adam_t += 1;
    float b1t = powf(0.9f, (float)m.adam_t), b2t = powf(0.999f, (float)m.adam_t);
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((kMlpParams + 255) / 256)), dim3(256), 0, s, m.master, m.grad,
                       m.adam_m, m.adam_v, m.wbf, (long)kMlpParams, lr, b1t, b2t);
    return hipGetLastError();
}

// rows of one instance's frontier x K thresholds -> features
template <int W>
__global__ void k_features_frontier(const Item *F, unsigned nf, int K, int L0, long r0, long nrows, uint16_t *x) {
    const long row = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= nrows) return;
    const long r = r0 + row;
    const unsigned i = (unsigned)(r / K);
    const int k = (int)(r % K);
    if (i >= nf) return;
    const Item it = F[i];
    Board<W> bd;
    bd.unpack(it.b);
    const int h = meta_h(it.meta), L = L0 + 2 * k;
    constexpr int N2 = W * W;
    for (int c = lane; c < kMlpIn; c += 64) {
        float v = 0.f;
        if (c < N2 * N2) {
            int t = c / N2, cc = c - t * N2;
            v = (bd.tile(cc) == t) ? 1.f : 0.f;
        } else if (c == 625) {
            v = h * (1.f / 32.f);
        } else if (c == 626) {
            v = (L - h) * (1.f / 32.f);
        }
        x[row * kMlpIn + c] = f2bf(v);
    }
}

__global__ void k_scatter_pred(const float *y, long r0, long nrows, float *pred_inst) {
    long row = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (row < nrows) pred_inst[r0 + row] = y[row];
}

hipError_t mlp_predict_frontier(MlpState &m, int width, const Item *frontier, const unsigned *nfront,
                                const Item *roots, int n_inst, int capl, int K, float *pred, hipStream_t s) {
    std::vector<unsigned> nf((size_t)n_inst);
    std::vector<Item> rt((size_t)n_inst);
    hipError_t e = hipMemcpyAsync(nf.data(), nfront, nf.size() * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(rt.data(), roots, rt.size() * sizeof(Item), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    const long chunk = 1 << 16;
    if ((e = mlp_reserve(m, (size_t)chunk)) != hipSuccess) return e;
    for (int inst = 0; inst < n_inst; ++inst) {
        const unsigned n = nf[inst] < (unsigned)capl ? nf[inst] : (unsigned)capl;
        const long rows = (long)n * K;
        const int L0 = meta_h(rt[inst].meta);
        const Item *F = frontier + (size_t)inst * capl;
        float *P = pred + (size_t)inst * capl * K;
        for (long r0 = 0; r0 < rows; r0 += chunk) {
            const long nr = rows - r0 < chunk ? rows - r0 : chunk;
            dim3 g((unsigned)((nr + 3) / 4));
            if (width == 4)
                hipLaunchKernelGGL((k_features_frontier<4>), g, dim3(256), 0, s, F, n, K, L0, r0, nr, m.x);
            else
                hipLaunchKernelGGL((k_features_frontier<5>), g, dim3(256), 0, s, F, n, K, L0, r0, nr, m.x);
            if ((e = mlp_forward(m, (int)nr, s)) != hipSuccess) return e;
            hipLaunchKernelGGL(k_scatter_pred, dim3((unsigned)((nr + 255) / 256)), dim3(256), 0, s, m.y, r0, nr, P);
            if ((e = hipGetLastError()) != hipSuccess) return e;
        }
    }
    return hipSuccess;
}

} // namespace spz
