// spz_mlp.h -- subtree-size predictor for the learned load balancer.
//
// No reference counterpart: the reference balances by the previous
// iteration's per-root load (src/15puzzle_block_parallel.cu:490-563,
// src_horie/idas_dynamic.cu:1527-1614).  Here a small MLP predicts
// log2(1 + subtree size) of a frontier node at a threshold, and the
// restart of an iteration queues the frontier longest-predicted-first.
//
// Network: x[640] -> 256 -> 256 -> 1, ReLU, bf16 weights and activations,
// fp32 biases, fp32 accumulation (MFMA 16x16x16 bf16).  Features of a board b at
// threshold L with heuristic h:
//   x[t*N2 + c] = 1 if tile t is at cell c   (N2*N2 <= 625 one-hot)
//   x[625] = h / 32,  x[626] = (L - h) / 32,  x[627..639] = 0.
// Flat fp32 parameter layout: W1[640][256], b1[256], W2[256][256], b2[256],
// w3[256], b3[1].
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace spz {

constexpr int kMlpIn = 640, kMlpH = 256;
constexpr size_t kMlpOffW1 = 0;
constexpr size_t kMlpOffB1 = kMlpOffW1 + (size_t)kMlpIn * kMlpH;
constexpr size_t kMlpOffW2 = kMlpOffB1 + kMlpH;
constexpr size_t kMlpOffB2 = kMlpOffW2 + (size_t)kMlpH * kMlpH;
constexpr size_t kMlpOffW3 = kMlpOffB2 + kMlpH;
constexpr size_t kMlpOffB3 = kMlpOffW3 + kMlpH;
constexpr size_t kMlpParams = kMlpOffB3 + 1;

struct Item;

// Device state of the network and its training scratch (grown per batch).
struct MlpState {
    float *master = nullptr;    // fp32 parameters [kMlpParams]
    uint16_t *wbf = nullptr;    // bf16 copy of the parameters
    float *grad = nullptr;      // fp32 gradient [kMlpParams]
    float *adam_m = nullptr, *adam_v = nullptr;
    int adam_t = 0;
    bool ready = false;         // parameters set
    // batch scratch
    size_t cap = 0;
    uint16_t *x = nullptr;      // [cap][640] bf16 features
    float *z1 = nullptr, *z2 = nullptr;  // pre-activations [cap][256]
    uint16_t *h1 = nullptr, *h2 = nullptr; // bf16 activations
    float *y = nullptr;         // [cap]
    float *dz = nullptr;        // [cap][256] scratch gradient
    uint16_t *dzb = nullptr;    // bf16 copy
    float *part = nullptr;      // [cap/64][256] partial column sums
    float *loss = nullptr;      // [1]
    uint16_t *w2frag = nullptr; // W2 in MFMA B-fragment order (prediction)
};

hipError_t mlp_alloc(MlpState &m);
void mlp_free(MlpState &m);
hipError_t mlp_reserve(MlpState &m, size_t rows);
hipError_t mlp_set_params(MlpState &m, const float *host, hipStream_t s);
hipError_t mlp_get_params(MlpState &m, float *host, hipStream_t s);
hipError_t mlp_init_random(MlpState &m, uint64_t seed, hipStream_t s);

// features of n boards (packed Items with h in meta) at thresholds thr[n]
hipError_t mlp_features_items(MlpState &m, int width, const Item *items, const int *thr, int n, hipStream_t s);
// forward over rows [0,n) of the feature buffer; y in m.y
hipError_t mlp_forward(MlpState &m, int n, hipStream_t s);
// loss = mean (y - target)^2 (into m.loss, or added to *loss_acc) and the
// full gradient into m.grad
hipError_t mlp_backward(MlpState &m, const float *target, int n, hipStream_t s, float *loss_acc = nullptr);
// features of rows items[perm[i]] at thr[perm[i]], i < n, and their targets
// tgt[perm[i]] into tout[i] (a training batch gathered on the device)
hipError_t mlp_features_gather(MlpState &m, int width, const Item *items, const int *thr, const float *tgt,
                               const int *perm, int n, float *tout, hipStream_t s);
hipError_t mlp_adam_step(MlpState &m, float lr, hipStream_t s);

// Predictions at kPredThr consecutive thresholds L0, L0+2, ... per node, by
// one fused kernel (one wave per node; the thresholds of a node are the 16
// rows of one MFMA tile and share layer 1 up to a rank-1 term).
constexpr int kPredThr = 16;
// every frontier node of every instance:
// pred[(inst*capl + i)*kPredThr + k] = y(frontier node i, L = h0[inst] + 2k)
hipError_t mlp_predict_frontier(MlpState &m, int width, const Item *frontier, const unsigned *nfront,
                                const Item *roots, int n_inst, int capl, float *pred, hipStream_t s);
// n nodes with h in meta: out[i*kPredThr + k] = y(node i, L = thr0[i] + 2k)
hipError_t mlp_predict_items(MlpState &m, int width, const Item *items, const int *thr0, int n, float *out,
                             hipStream_t s);

} // namespace spz
