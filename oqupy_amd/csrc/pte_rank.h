// Learned truncation-rank form: the rank predictor (a small MLP), its
// features, and the per-bond history the engine keeps on the device.
#pragma once

#include <hip/hip_runtime.h>

#include "pte_kernels.h"

namespace pte {

// MLP: kRankFeatures -> 128 -> 128 -> 128 -> 1, tanh hidden activations.
constexpr int kRankFeatures = 24;
constexpr int kRankWidth = 128;
constexpr int kRankHist = 16;        // log singular values kept per bond
constexpr int kRankKeysPerKind = 4096;
constexpr int kRankKeys = 2 * kRankKeysPerKind;
// record: features, target log2(need) / 8, need, kept rank, min(m, n)
constexpr int kRankRecord = kRankFeatures + 4;

// flat parameter layout (float): W1 [128][24], b1, W2 [128][128], b2,
// W3 [128][128], b3, W4 [128], b4
constexpr int kRankOffW1 = 0;
constexpr int kRankOffB1 = kRankOffW1 + kRankWidth * kRankFeatures;
constexpr int kRankOffW2 = kRankOffB1 + kRankWidth;
constexpr int kRankOffB2 = kRankOffW2 + kRankWidth * kRankWidth;
constexpr int kRankOffW3 = kRankOffB2 + kRankWidth;
constexpr int kRankOffB3 = kRankOffW3 + kRankWidth * kRankWidth;
constexpr int kRankOffW4 = kRankOffB3 + kRankWidth;
constexpr int kRankOffB4 = kRankOffW4 + kRankWidth;
constexpr int kRankParams = kRankOffB4 + 1;

// scalar inputs of one truncated SVD's features
struct RankQuery {
  int key;             // bond key: kind * kRankKeysPerKind + distance from the newest site
  float step_frac;     // step / num_steps
  float rel_frac;      // distance from the newest site / dkmax
  float kind;          // 0 zip-up, 1 sweep
  float log_n;         // log2(min(m, n)) / 8
  float alpha, temperature, log_eps;   // bath and tolerance (log10(epsrel) / 10)
  int m, n;
};

struct RankPredictArgs {
  const float* params;     // null: heuristic (previous rank + oversample)
  const float* hist;       // [kRankKeys][kRankHist]
  const int* hist_rank;    // [kRankKeys] pivoted-QR steps the bond needed last (0: none)
  RankQuery q;
  int oversample;
  float margin;
  int nmax;                // min(m, n): no cap above it
  int* cap;                // out: step cap of the pivoted QR
};

struct RankRecordArgs {
  float* hist;
  int* hist_rank;
  RankQuery q;
  const double* sorted;    // singular values, descending
  const SvdInfo* info;     // kept rank and s0
  const double* need;      // steps the pivoted QR needed to reach the threshold
  float* rec;              // trace records (null: none)
  int* rec_count;
  int rec_cap;
};

__global__ void rank_predict_kernel(RankPredictArgs a);
__global__ void rank_record_kernel(RankRecordArgs a);

// training and batched inference of the MLP (samples idx[0..n) of X / y;
// H and D hold the three hidden layers' activations and back-propagated
// pre-activation gradients of each sample, [n][3][128])
__global__ void mlp_forward_kernel(const float* P, const float* X, const int* idx, int n,
                                   float* H, float* yhat);
__global__ void mlp_backward_kernel(const float* P, const float* H, const float* yhat,
                                    const float* y, const int* idx, int n, float* D,
                                    float* d4);
__global__ void mlp_grad_kernel(const float* X, const int* idx, const float* H,
                                const float* D, const float* d4, int n, float* G);
__global__ void mlp_loss_kernel(const float* yhat, const float* y, const int* idx, int n,
                                float* out);
__global__ void mlp_adam_kernel(float* P, const float* G, float* M, float* V, int count,
                                float lr, float beta1, float beta2, float eps, int t);

}  // namespace pte
