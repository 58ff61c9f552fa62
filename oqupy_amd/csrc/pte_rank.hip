// Learned truncation-rank form: device kernels of the rank predictor.
//
// A small MLP (24 features -> 3 x 128 tanh -> 1) predicts the change, from the
// bond's previous truncation, of log2 of the number of pivoted-QR steps a
// truncated SVD needs before its residual falls below
// the acceptance threshold (a fraction of epsrel times the largest column
// norm), from the bond's previous truncation (its last 16 log singular values
// and step count), the position of the bond, the step, the bath and the
// tolerance. The engine caps its pivoted QR at the prediction plus an
// oversampling margin and resumes the exact factorisation when the residual
// of the capped one is too large. The network is trained
// here too: forward, back-propagation, gradient accumulation and Adam, all
// in FP32.
#include <hip/hip_runtime.h>

#include "pte_device.h"
#include "pte_rank.h"

namespace pte {

__device__ __forceinline__ float rank_target(int k) { return log2f((float)k) / 8.0f; }

// feature vector of one query (all threads of the block call it; x in LDS)
__device__ void rank_features(const RankQuery& q, const float* hist, const int* hist_rank,
                              float* x) {
  const int t = threadIdx.x;
  const int kp = hist_rank[q.key];
  if (t < kRankHist) x[3 + t] = kp > 0 ? hist[(long)q.key * kRankHist + t] : 0.0f;
  if (t == 0) {
    x[0] = q.step_frac;
    x[1] = q.rel_frac;
    x[2] = q.kind;
    x[19] = kp > 0 ? rank_target(kp) : 0.0f;
    x[20] = q.alpha;
    x[21] = q.temperature;
    x[22] = q.log_eps;
    x[23] = q.log_n;
  }
  __syncthreads();
}

// one sample through the network with a block of kRankWidth threads:
// h[0..2][128] receive the hidden activations; returns the output
__device__ float mlp_forward_one(const float* P, const float* x, float* h, float* red) {
  const int o = threadIdx.x;
  {
    float acc = P[kRankOffB1 + o];
    const float* w = P + kRankOffW1 + o * kRankFeatures;
    for (int i = 0; i < kRankFeatures; ++i) acc = fmaf(w[i], x[i], acc);
    h[o] = tanhf(acc);
  }
  __syncthreads();
  {
    float acc = P[kRankOffB2 + o];
    const float* w = P + kRankOffW2 + o * kRankWidth;
    for (int i = 0; i < kRankWidth; ++i) acc = fmaf(w[i], h[i], acc);
    h[kRankWidth + o] = tanhf(acc);
  }
  __syncthreads();
  {
    float acc = P[kRankOffB3 + o];
    const float* w = P + kRankOffW3 + o * kRankWidth;
    for (int i = 0; i < kRankWidth; ++i) acc = fmaf(w[i], h[kRankWidth + i], acc);
    h[2 * kRankWidth + o] = tanhf(acc);
  }
  __syncthreads();
  red[o] = P[kRankOffW4 + o] * h[2 * kRankWidth + o];
  __syncthreads();
  for (int s = kRankWidth / 2; s > 0; s >>= 1) {
    if (o < s) red[o] += red[o + s];
    __syncthreads();
  }
  const float y = red[0] + P[kRankOffB4];
  __syncthreads();
  return y;
}

// step cap of the pivoted QR for one SVD (one block of kRankWidth threads)
__global__ void __launch_bounds__(kRankWidth)
rank_predict_kernel(RankPredictArgs a) {
  __shared__ float x[kRankFeatures];
  __shared__ float h[3 * kRankWidth];
  __shared__ float red[kRankWidth];
  rank_features(a.q, a.hist, a.hist_rank, x);
  const int kp = a.hist_rank[a.q.key];
  int cap = a.nmax;
  if (kp > 0) {
    if (a.params) {
      // the network predicts the change of log2(steps) / 8 from the bond's
      // previous truncation (x[19])
      const float y = mlp_forward_one(a.params, x, h, red) + x[19];
      const float k = exp2f(fminf(8.0f * y, 20.0f));
      cap = (int)ceilf(k * (1.0f + a.margin)) + a.oversample;
    } else {
      cap = kp + a.oversample;
    }
  }
  if (threadIdx.x == 0) *a.cap = max(1, min(cap, a.nmax));
}

// after a truncation: optional trace record (features before the update,
// target log2(need) / 8, need, kept rank, min(m, n)), then the bond's history
__global__ void __launch_bounds__(kRankWidth)
rank_record_kernel(RankRecordArgs a) {
  __shared__ float x[kRankFeatures];
  rank_features(a.q, a.hist, a.hist_rank, x);
  const int k = a.info->rank;
  const int need = max(1, (int)(*a.need));
  const double s0 = a.info->s0;
  const int t = threadIdx.x;
  if (a.rec && t < kRankRecord) {
    const int c = *a.rec_count;
    if (c < a.rec_cap) {
      float v;
      if (t < kRankFeatures) v = x[t];
      else if (t == kRankFeatures) v = rank_target(need);
      else if (t == kRankFeatures + 1) v = (float)need;
      else if (t == kRankFeatures + 2) v = (float)k;
      else v = (float)min(a.q.m, a.q.n);
      a.rec[(long)c * kRankRecord + t] = v;
    }
  }
  __syncthreads();
  if (t == 0 && a.rec && *a.rec_count < a.rec_cap) *a.rec_count += 1;
  if (t < kRankHist) {
    const int i = max(0, k - kRankHist + t);
    const double s = a.sorted[i];
    const float v = (s > 0.0 && s0 > 0.0) ? (float)(log10(s / s0) / 8.0) : -2.0f;
    a.hist[(long)a.q.key * kRankHist + t] = v;
  }
  if (t == 0) a.hist_rank[a.q.key] = need;
}

// ---- training ----

__global__ void __launch_bounds__(kRankWidth)
mlp_forward_kernel(const float* P, const float* X, const int* idx, int n, float* H,
                   float* yhat) {
  __shared__ float x[kRankFeatures];
  __shared__ float h[3 * kRankWidth];
  __shared__ float red[kRankWidth];
  const int b = blockIdx.x;
  if (b >= n) return;
  const int s = idx ? idx[b] : b;
  if (threadIdx.x < kRankFeatures) x[threadIdx.x] = X[(long)s * kRankFeatures + threadIdx.x];
  __syncthreads();
  const float y = mlp_forward_one(P, x, h, red);
  for (int i = threadIdx.x; i < 3 * kRankWidth; i += kRankWidth)
    H[(long)b * 3 * kRankWidth + i] = h[i];
  if (threadIdx.x == 0) yhat[b] = y;
}

// loss = mean (yhat - y)^2 over the n samples; D = d loss / d pre-activation
__global__ void __launch_bounds__(kRankWidth)
mlp_backward_kernel(const float* P, const float* H, const float* yhat, const float* y,
                    const int* idx, int n, float* D, float* d4) {
  __shared__ float g[kRankWidth];
  const int b = blockIdx.x;
  if (b >= n) return;
  const int s = idx ? idx[b] : b;
  const int o = threadIdx.x;
  const float* h = H + (long)b * 3 * kRankWidth;
  float* d = D + (long)b * 3 * kRankWidth;
  const float dy = 2.0f * (yhat[b] - y[s]) / (float)n;
  if (o == 0) d4[b] = dy;
  // layer 3
  float h3 = h[2 * kRankWidth + o];
  float z = P[kRankOffW4 + o] * dy * (1.0f - h3 * h3);
  d[2 * kRankWidth + o] = z;
  g[o] = z;
  __syncthreads();
  // layer 2: dh2[i] = sum_o W3[o][i] dz3[o]
  float acc = 0.0f;
  for (int q = 0; q < kRankWidth; ++q) acc = fmaf(P[kRankOffW3 + q * kRankWidth + o], g[q], acc);
  const float h2 = h[kRankWidth + o];
  z = acc * (1.0f - h2 * h2);
  d[kRankWidth + o] = z;
  __syncthreads();
  g[o] = z;
  __syncthreads();
  acc = 0.0f;
  for (int q = 0; q < kRankWidth; ++q) acc = fmaf(P[kRankOffW2 + q * kRankWidth + o], g[q], acc);
  const float h1 = h[o];
  d[o] = acc * (1.0f - h1 * h1);
}

// G[param] = sum over the samples (one thread per parameter)
__global__ void __launch_bounds__(256)
mlp_grad_kernel(const float* X, const int* idx, const float* H, const float* D,
                const float* d4, int n, float* G) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= kRankParams) return;
  float acc = 0.0f;
  if (p < kRankOffB1) {
    const int o = p / kRankFeatures, i = p % kRankFeatures;
    for (int b = 0; b < n; ++b) {
      const int s = idx ? idx[b] : b;
      acc = fmaf(D[(long)b * 3 * kRankWidth + o], X[(long)s * kRankFeatures + i], acc);
    }
  } else if (p < kRankOffW2) {
    const int o = p - kRankOffB1;
    for (int b = 0; b < n; ++b) acc += D[(long)b * 3 * kRankWidth + o];
  } else if (p < kRankOffB2) {
    const int o = (p - kRankOffW2) / kRankWidth, i = (p - kRankOffW2) % kRankWidth;
    for (int b = 0; b < n; ++b)
      acc = fmaf(D[(long)b * 3 * kRankWidth + kRankWidth + o], H[(long)b * 3 * kRankWidth + i],
                 acc);
  } else if (p < kRankOffW3) {
    const int o = p - kRankOffB2;
    for (int b = 0; b < n; ++b) acc += D[(long)b * 3 * kRankWidth + kRankWidth + o];
  } else if (p < kRankOffB3) {
    const int o = (p - kRankOffW3) / kRankWidth, i = (p - kRankOffW3) % kRankWidth;
    for (int b = 0; b < n; ++b)
      acc = fmaf(D[(long)b * 3 * kRankWidth + 2 * kRankWidth + o],
                 H[(long)b * 3 * kRankWidth + kRankWidth + i], acc);
  } else if (p < kRankOffW4) {
    const int o = p - kRankOffB3;
    for (int b = 0; b < n; ++b) acc += D[(long)b * 3 * kRankWidth + 2 * kRankWidth + o];
  } else if (p < kRankOffB4) {
    const int i = p - kRankOffW4;
    for (int b = 0; b < n; ++b) acc = fmaf(d4[b], H[(long)b * 3 * kRankWidth + 2 * kRankWidth + i], acc);
  } else {
    for (int b = 0; b < n; ++b) acc += d4[b];
  }
  G[p] = acc;
}

// out[0] = mean (yhat - y)^2 (one block of 256 threads)
__global__ void __launch_bounds__(256)
mlp_loss_kernel(const float* yhat, const float* y, const int* idx, int n, float* out) {
  __shared__ float red[256];
  float acc = 0.0f;
  for (int b = threadIdx.x; b < n; b += 256) {
    const float e = yhat[b] - y[idx ? idx[b] : b];
    acc = fmaf(e, e, acc);
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0] / (float)max(n, 1);
}

// Adam (Kingma & Ba 2015) with bias correction, step t >= 1
__global__ void __launch_bounds__(256)
mlp_adam_kernel(float* P, const float* G, float* M, float* V, int count, float lr,
                float beta1, float beta2, float eps, int t) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= count) return;
  const float g = G[p];
  const float m = beta1 * M[p] + (1.0f - beta1) * g;
  const float v = beta2 * V[p] + (1.0f - beta2) * g * g;
  M[p] = m;
  V[p] = v;
  const float mh = m / (1.0f - powf(beta1, (float)t));
  const float vh = v / (1.0f - powf(beta2, (float)t));
  P[p] -= lr * mh / (sqrtf(vh) + eps);
}

}  // namespace pte
