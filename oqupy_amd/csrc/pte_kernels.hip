// HIP kernels of the process-tensor engine for gfx950 (MI355X).
//
//  * zgemm_kernel      complex128 GEMM with arbitrary strides and optional
//                      conjugation (site x carry contractions, S V^H absorption)
//  * mpo_apply_kernel  applies one influence node W[w,p,s,v] to the GEMM result
//                      G[(l,p),(x,v)] and writes the SVD input matrix or a site
//  * jacobi_kernel     blocked one-sided Jacobi SVD (Hestenes) of a column-major
//                      matrix, one workgroup per block pair and round, grid
//                      barriers between rounds (cooperative launch)
//  * column_norm_kernel / rank_sort_kernel / truncate_kernel
//                      singular values, descending order, epsrel cut
//  * emit kernels      new sites, carries and S V^H rows from the Jacobi output
//
// Every kernel takes its sizes as arguments and bounds-checks its indices
// against them; the host checks that the sizes match the buffers.
#include "pte_device.h"
#include "pte_kernels.h"

namespace pte {

// Optional phase timers of the cooperative kernels (workgroup 0, thread 0,
// s_memrealtime ticks), compiled in with -DPTE_PHASE_TIMING only.
#ifdef PTE_PHASE_TIMING
__device__ unsigned long long g_phase[16];
#define PHASE_INIT() unsigned long long ph_t_ = wall_clock64()
#define PHASE(k)                                                       \
  do {                                                                 \
    if (blockIdx.x == 0 && threadIdx.x == 0) {                         \
      const unsigned long long t_ = wall_clock64();                    \
      g_phase[k] += t_ - ph_t_;                                        \
      ph_t_ = t_;                                                      \
    }                                                                  \
  } while (0)
// Segment cycle counts of the one-workgroup SVD (thread 0, s_memtime), in
// the same build: SEG_INIT() starts, SEG(k) adds the cycles since the last
// mark to slot k, SEG_FLUSH(k0, n) adds slots k0.. to g_phase.
#define SEG_INIT()                                 \
  unsigned long long sg_[6] = {0, 0, 0, 0, 0, 0}; \
  unsigned long long sg_t_ = clock64()
#define SEG(k)                                \
  do {                                        \
    const unsigned long long t_ = clock64(); \
    sg_[k] += t_ - sg_t_;                     \
    sg_t_ = t_;                               \
  } while (0)
#define SEG_FLUSH(k0, cnt)                                                 \
  do {                                                                     \
    if (threadIdx.x == 0)                                                  \
      for (int k_ = 0; k_ < (cnt); ++k_) atomicAdd(&g_phase[(k0) + k_], sg_[k_]); \
  } while (0)
#else
#define PHASE_INIT() (void)0
#define PHASE(k) (void)0
#define SEG_INIT() (void)0
#define SEG(k) (void)0
#define SEG_FLUSH(k0, cnt) (void)0
#endif

// Optional per-round timestamps of the Jacobi (thread 0 of every workgroup,
// s_memrealtime ticks, for the first kRtRounds rounds and kRtBlocks
// workgroups), compiled in with -DPTE_ROUND_TRACE only.
#ifdef PTE_ROUND_TRACE
constexpr int kRtRounds = 1024, kRtBlocks = 512, kRtMarks = 8;
__device__ unsigned long long g_rtrace[(long)kRtRounds * kRtBlocks * kRtMarks];
#define RTRACE(k)                                                                     \
  do {                                                                                \
    if (threadIdx.x == 0 && rt_round_ < kRtRounds && bid < kRtBlocks)                 \
      g_rtrace[((long)rt_round_ * kRtBlocks + bid) * kRtMarks + (k)] = wall_clock64(); \
  } while (0)
#else
#define RTRACE(k) (void)0
#endif

// ---------------------------------------------------------------------------
// complex GEMM:  C(i,j) = alpha * sum_k opA(A)(i,k) * opB(B)(k,j) [+ C(i,j)]
// 64x64 output tile per 256-thread workgroup, 4x4 outputs per thread, K in
// slices of 16 staged through LDS.
// ---------------------------------------------------------------------------
constexpr int GB_M = 64, GB_N = 64, GB_K = 16;

__global__ void __launch_bounds__(256)
zgemm_kernel(GemmArgs a) {
  __shared__ double2 As[GB_K][GB_M + 1];
  __shared__ double2 Bs[GB_K][GB_N + 1];
  const int tid = threadIdx.x;
  const int tx = tid & 15, ty = tid >> 4;
  const int i0 = blockIdx.y * GB_M, j0 = blockIdx.x * GB_N;
  double2 acc[4][4];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[u][v] = make_double2(0.0, 0.0);

  const bool a_kfast = (a.sak == 1);
  const bool b_jfast = (a.sbj == 1);
  for (int k0 = 0; k0 < a.K; k0 += GB_K) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = tid + 256 * q;
      int row, kk;
      if (a_kfast) { row = e >> 4; kk = e & 15; } else { row = e & 63; kk = e >> 6; }
      const int gi = i0 + row, gk = k0 + kk;
      double2 val = make_double2(0.0, 0.0);
      if (gi < a.M && gk < a.K) {
        val = a.A[(long)gi * a.sai + (long)gk * a.sak];
        if (a.conjA) val.y = -val.y;
      }
      As[kk][row] = val;
      int col;
      if (b_jfast) { kk = e >> 6; col = e & 63; } else { kk = e & 15; col = e >> 4; }
      const int gj = j0 + col, gk2 = k0 + kk;
      double2 vb = make_double2(0.0, 0.0);
      if (gj < a.N && gk2 < a.K) {
        vb = a.B[(long)gk2 * a.sbk + (long)gj * a.sbj];
        if (a.conjB) vb.y = -vb.y;
      }
      Bs[kk][col] = vb;
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < GB_K; ++kk) {
      double2 ra[4], rb[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) ra[u] = As[kk][ty + 16 * u];
#pragma unroll
      for (int v = 0; v < 4; ++v) rb[v] = Bs[kk][tx + 16 * v];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) cfma(acc[u][v], ra[u], rb[v]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int gi = i0 + ty + 16 * u;
    if (gi >= a.M) continue;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int gj = j0 + tx + 16 * v;
      if (gj >= a.N) continue;
      double2 r = cmul(a.alpha, acc[u][v]);
      double2* c = a.C + (long)gi * a.sci + (long)gj * a.scj;
      if (!a.beta_zero) r = cadd(r, *c);
      *c = r;
    }
  }
}

// ---------------------------------------------------------------------------
// complex128 GEMM on the FP64 matrix cores (the plain form's contraction):
// C(i,j) = alpha * sum_k opA(A)(i,k) opB(B)(k,j) [+ C(i,j)], with a 3M split
// (three real products per complex product, Gauss):
//   P1 = Ar Br,  P2 = Ai Bi,  P3 = (Ar + Ai)(Br + Bi),
//   Re = P1 - P2,  Im = P3 - P1 - P2,
// accumulated in FP64 by v_mfma_f64_16x16x4_f64. 64x64 output tile per
// 256-thread workgroup, each wavefront a 32x32 quarter as 2x2 tiles of
// 16x16; K in slices of 16 staged through LDS as separate real and
// imaginary planes. Operand layout of the FP64 16x16x4 instruction (measured
// on gfx950): lane l supplies A[l%16][l/16] and B[l/16][l%16] and holds
// D[4*r + l/16][l%16] in its item r = 0..3 (rows interleaved by 4, unlike
// the FP32 16x16x4 instruction's D[4*(l/16)+r][l%16]).
// ---------------------------------------------------------------------------
typedef double doublex4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256)
zgemm_mfma_kernel(GemmArgs a) {
  __shared__ double Ar[GB_K][GB_M + 1], Ai[GB_K][GB_M + 1];
  __shared__ double Br[GB_K][GB_N + 1], Bi[GB_K][GB_N + 1];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int i0 = blockIdx.y * GB_M, j0 = blockIdx.x * GB_N;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;
  doublex4 p1[2][2], p2[2][2], p3[2][2];
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int v = 0; v < 2; ++v) {
      p1[u][v] = doublex4{0.0, 0.0, 0.0, 0.0};
      p2[u][v] = doublex4{0.0, 0.0, 0.0, 0.0};
      p3[u][v] = doublex4{0.0, 0.0, 0.0, 0.0};
    }
  const bool a_kfast = (a.sak == 1);
  const bool b_jfast = (a.sbj == 1);
  for (int k0 = 0; k0 < a.K; k0 += GB_K) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = tid + 256 * q;
      int row, kk;
      if (a_kfast) { row = e >> 4; kk = e & 15; } else { row = e & 63; kk = e >> 6; }
      const int gi = i0 + row, gk = k0 + kk;
      double2 va = make_double2(0.0, 0.0);
      if (gi < a.M && gk < a.K) {
        va = a.A[(long)gi * a.sai + (long)gk * a.sak];
        if (a.conjA) va.y = -va.y;
      }
      Ar[kk][row] = va.x;
      Ai[kk][row] = va.y;
      int col;
      if (b_jfast) { kk = e >> 6; col = e & 63; } else { kk = e & 15; col = e >> 4; }
      const int gj = j0 + col, gk2 = k0 + kk;
      double2 vb = make_double2(0.0, 0.0);
      if (gj < a.N && gk2 < a.K) {
        vb = a.B[(long)gk2 * a.sbk + (long)gj * a.sbj];
        if (a.conjB) vb.y = -vb.y;
      }
      Br[kk][col] = vb.x;
      Bi[kk][col] = vb.y;
    }
    __syncthreads();
#pragma unroll
    for (int k4 = 0; k4 < GB_K; k4 += 4) {
      double ar[2], ai[2], br[2], bi[2];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        ar[u] = Ar[k4 + lk][wr + 16 * u + lr];
        ai[u] = Ai[k4 + lk][wr + 16 * u + lr];
      }
#pragma unroll
      for (int v = 0; v < 2; ++v) {
        br[v] = Br[k4 + lk][wc + 16 * v + lr];
        bi[v] = Bi[k4 + lk][wc + 16 * v + lr];
      }
#pragma unroll
      for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int v = 0; v < 2; ++v) {
          p1[u][v] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[u], br[v], p1[u][v], 0, 0, 0);
          p2[u][v] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[u], bi[v], p2[u][v], 0, 0, 0);
          p3[u][v] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[u] + ai[u], br[v] + bi[v],
                                                          p3[u][v], 0, 0, 0);
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int v = 0; v < 2; ++v)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gi = i0 + wr + 16 * u + 4 * r + lk;
        const int gj = j0 + wc + 16 * v + lr;
        if (gi < a.M && gj < a.N) {
          const double re = p1[u][v][r] - p2[u][v][r];
          const double im = p3[u][v][r] - p1[u][v][r] - p2[u][v][r];
          double2 c = cmul(a.alpha, make_double2(re, im));
          double2* dst = a.C + (long)gi * a.sci + (long)gj * a.scj;
          if (!a.beta_zero) c = cadd(c, *dst);
          *dst = c;
        }
      }
}

// ---------------------------------------------------------------------------
// complex GEMM in complex64 with FP32 accumulation on the matrix cores
// (precision form): C(i,j) = sum_k A(i,k) B(k,j), inputs and output stored
// as complex128 and rounded to complex64 when staged. 64x64 output tile per
// 256-thread workgroup; each wavefront owns a 32x32 quarter as 2x2 tiles of
// v_mfma_f32_16x16x4_f32, four real products per complex product:
//   Re += Ar Br - Ai Bi,   Im += Ar Bi + Ai Br.
// Operand layout of the 16x16x4 instruction: lane l supplies A[l%16][l/16]
// and B[l/16][l%16]; it holds D[4*(l/16)+r][l%16], r = 0..3.
// ---------------------------------------------------------------------------
typedef float floatx4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256)
cgemm_mfma_kernel(GemmArgs a) {
  __shared__ float2 As[GB_K][GB_M + 1];
  __shared__ float2 Bs[GB_K][GB_N + 1];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int i0 = blockIdx.y * GB_M, j0 = blockIdx.x * GB_N;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;
  floatx4 accr[2][2], acci[2][2];
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int v = 0; v < 2; ++v) {
      accr[u][v] = floatx4{0.f, 0.f, 0.f, 0.f};
      acci[u][v] = floatx4{0.f, 0.f, 0.f, 0.f};
    }
  const bool a_kfast = (a.sak == 1);
  const bool b_jfast = (a.sbj == 1);
  for (int k0 = 0; k0 < a.K; k0 += GB_K) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = tid + 256 * q;
      int row, kk;
      if (a_kfast) { row = e >> 4; kk = e & 15; } else { row = e & 63; kk = e >> 6; }
      const int gi = i0 + row, gk = k0 + kk;
      float2 va = make_float2(0.f, 0.f);
      if (gi < a.M && gk < a.K) {
        const double2 v = a.A[(long)gi * a.sai + (long)gk * a.sak];
        va = make_float2((float)v.x, a.conjA ? -(float)v.y : (float)v.y);
      }
      As[kk][row] = va;
      int col;
      if (b_jfast) { kk = e >> 6; col = e & 63; } else { kk = e & 15; col = e >> 4; }
      const int gj = j0 + col, gk2 = k0 + kk;
      float2 vb = make_float2(0.f, 0.f);
      if (gj < a.N && gk2 < a.K) {
        const double2 v = a.B[(long)gk2 * a.sbk + (long)gj * a.sbj];
        vb = make_float2((float)v.x, a.conjB ? -(float)v.y : (float)v.y);
      }
      Bs[kk][col] = vb;
    }
    __syncthreads();
#pragma unroll
    for (int k4 = 0; k4 < GB_K; k4 += 4) {
      float2 av[2], bv[2];
#pragma unroll
      for (int u = 0; u < 2; ++u) av[u] = As[k4 + lk][wr + 16 * u + lr];
#pragma unroll
      for (int v = 0; v < 2; ++v) bv[v] = Bs[k4 + lk][wc + 16 * v + lr];
#pragma unroll
      for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int v = 0; v < 2; ++v) {
          accr[u][v] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[u].x, bv[v].x, accr[u][v], 0, 0, 0);
          accr[u][v] = __builtin_amdgcn_mfma_f32_16x16x4f32(-av[u].y, bv[v].y, accr[u][v], 0, 0, 0);
          acci[u][v] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[u].x, bv[v].y, acci[u][v], 0, 0, 0);
          acci[u][v] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[u].y, bv[v].x, acci[u][v], 0, 0, 0);
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int v = 0; v < 2; ++v)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gi = i0 + wr + 16 * u + 4 * lk + r;
        const int gj = j0 + wc + 16 * v + lr;
        if (gi < a.M && gj < a.N) {
          double2 c = make_double2((double)accr[u][v][r], (double)acci[u][v][r]);
          if (!a.beta_zero) c = cadd(c, a.C[(long)gi * a.sci + (long)gj * a.scj]);
          a.C[(long)gi * a.sci + (long)gj * a.scj] = c;
        }
      }
}

// ---------------------------------------------------------------------------
// T[x,s,l,w] = sum_{p,v} W[w,p,s,v] * G[(l,p),(x,v)]
//   mode MPO_OUT_SVD_H : X[(l*Wl+w) + (x*S+s)*ldx] = conj(T)   (X = M^H)
//   mode MPO_OUT_SITE  : out[(l*S+s)*Dx + x] = T   (requires Wl == 1)
//   mode MPO_OUT_M     : out[(x*S+s) + (l*Wl+w)*ldx] = T     (M column-major)
// node_kind DELTA: W[w,p,s,v] = delta(w,v) delta(p,s) I[w,p] (I is Wl x P).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
mpo_apply_kernel(MpoApplyArgs a) {
  const long total = (long)a.Dx * a.S * a.Dl * a.Wl;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x) {
    // e enumerates (x, s, l, w) with w fastest
    long t = e;
    const int w = t % a.Wl; t /= a.Wl;
    const int l = t % a.Dl; t /= a.Dl;
    const int s = t % a.S;  t /= a.S;
    const int x = (int)t;
    const long gcols = (long)a.Dx * a.V;
    double2 acc = make_double2(0.0, 0.0);
    if (a.kind == NODE_DELTA) {
      // p = s, v = w
      const double2 g = a.G[((long)l * a.P + s) * gcols + (long)x * a.V + w];
      acc = cmul(a.I[(long)w * a.P + s], g);
    } else {
      for (int p = 0; p < a.P; ++p) {
        const double2* grow = a.G + ((long)l * a.P + p) * gcols + (long)x * a.V;
        const double2* wrow = a.W + (((long)w * a.P + p) * a.S + s) * a.V;
        for (int v = 0; v < a.V; ++v) cfma(acc, wrow[v], grow[v]);
      }
    }
    if (a.mode == MPO_OUT_SVD_H) {
      a.out[((long)l * a.Wl + w) + ((long)x * a.S + s) * a.ldx] = cconj(acc);
    } else if (a.mode == MPO_OUT_M) {
      a.out[((long)x * a.S + s) + ((long)l * a.Wl + w) * a.ldx] = acc;
    } else {
      a.out[((long)l * a.S + s) * a.Dx + x] = acc;
    }
  }
}

// ---------------------------------------------------------------------------
// TEMPO zip (left to right):
// T[x,s,r,e] = sum_{p,w} W[w,p,s,e] * G[(x,w),(p,r)]
//   mode MPO_OUT_SVD_H : X[(r*E+e) + (x*S+s)*ldx] = conj(T)
//   mode MPO_OUT_M     : out[(x*S+s) + (r*E+e)*ldx] = T   (M column-major)
//   mode MPO_OUT_SITE  : out[(x*S+s)*E + e] = T (requires Dr == 1)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
mpo_apply_right_kernel(MpoApplyArgs a) {
  // reuse of fields: Dl <- Dr (right bond of the site), Wl <- E (east leg),
  // V <- W (west leg of the node)
  const int Dr = a.Dl, E = a.Wl, Wn = a.V;
  const long total = (long)a.Dx * a.S * Dr * E;
  for (long e_ = blockIdx.x * (long)blockDim.x + threadIdx.x; e_ < total;
       e_ += (long)gridDim.x * blockDim.x) {
    long t = e_;
    const int e = t % E; t /= E;
    const int r = t % Dr; t /= Dr;
    const int s = t % a.S; t /= a.S;
    const int x = (int)t;
    const long gcols = (long)a.P * Dr;
    double2 acc = make_double2(0.0, 0.0);
    if (a.kind == NODE_DELTA) {
      // W[w,p,s,e] = delta(w,e) delta(p,s) I[w,p]  ->  w = e, p = s
      const double2 g = a.G[((long)x * Wn + e) * gcols + (long)s * Dr + r];
      acc = cmul(a.I[(long)e * a.P + s], g);
    } else {
      for (int w = 0; w < Wn; ++w) {
        const double2* grow = a.G + ((long)x * Wn + w) * gcols;
        for (int p = 0; p < a.P; ++p) {
          const double2 wv = a.W[(((long)w * a.P + p) * a.S + s) * E + e];
          cfma(acc, wv, grow[(long)p * Dr + r]);
        }
      }
    }
    if (a.mode == MPO_OUT_SVD_H) {
      a.out[((long)r * E + e) + ((long)x * a.S + s) * a.ldx] = cconj(acc);
    } else if (a.mode == MPO_OUT_M) {
      a.out[((long)x * a.S + s) + ((long)r * E + e) * a.ldx] = acc;
    } else {
      a.out[((long)x * a.S + s) * E + e] = acc;
    }
  }
}

// ---------------------------------------------------------------------------
// Blocked one-sided Jacobi SVD.
// X (m x n, column-major, leading dimension ldx) is orthogonalised column-wise:
// X <- X Q with Q unitary accumulated in V (n x n, column-major). Columns are
// grouped into blocks of JB; in every round each block pair of the
// round-robin tournament over nb blocks is handled by `slices` workgroups,
// each owning a slice of the rows of X and of V. The pair's 2JB x 2JB Gram
// matrix is summed from the slices' partial Grams, diagonalised (cyclic
// two-sided Jacobi in LDS, at most inner_sweeps sweeps, identically in every
// workgroup of the pair) and the resulting unitary applied to the slice's rows
// of X and V. A sweep is nb-1 rounds; the iteration stops after a sweep in
// which no pair needed a rotation. Rounds are not separated by grid barriers:
// every block carries a version (rounds completed on it), a pair starts a
// round once both of its blocks have reached it and, when all of its slices
// are done, releases them for the next round; the slices of a pair exchange
// partial Grams through a pair-local barrier (double-buffered by round
// parity); a grid barrier closes each sweep. A pair needs one when some
// |g_ij| > tol*sqrt(g_ii g_jj) and |g_ij| > abs_tol*sqrt(max(g_ii, g_jj)).
// ---------------------------------------------------------------------------
constexpr int JB = 16;          // block width
constexpr int JP = 2 * JB;      // columns per block pair
constexpr int JR = 32;          // rows per streamed chunk
constexpr int JT = 256;         // threads per workgroup
#ifndef PTE_JACOBI_SUM_BATCH
#define PTE_JACOBI_SUM_BATCH 4
#endif
// slices of partial Grams whose loads are in flight together
constexpr int kJacobiSumBatch = PTE_JACOBI_SUM_BATCH;

__device__ __forceinline__ void tournament_pair(int nplayers, int round, int slot,
                                                int& a, int& b) {
  if (slot == 0) {
    a = nplayers - 1;
    b = round % (nplayers - 1);
  } else {
    a = (round + slot) % (nplayers - 1);
    b = (round - slot + nplayers - 1) % (nplayers - 1);
  }
}

__device__ __forceinline__ void ordered_pair(int sub, int slot, int& p, int& q) {
  tournament_pair(JP, sub, slot, p, q);
  if (p > q) { const int t = p; p = q; q = t; }
}

// Does the Hermitian 2x2 block [[app, apq], [conj(apq), aqq]] need a rotation:
// |apq| > tol * sqrt(|app aqq|) and |apq| > abs_tol * sqrt(max(app, aqq)),
// compared in squares.
__device__ __forceinline__ bool needs_rotation(double app, double aqq, double mag2,
                                               double tol2, double abs_tol2) {
  return mag2 > tol2 * fabs(app * aqq) && mag2 > abs_tol2 * fmax(fmax(app, aqq), 0.0) &&
         mag2 > 0.0;
}

// 1/sqrt(x) and 1/x for normal positive x: the hardware estimates refined
// by Newton steps (quadratic convergence) to full double precision.
__device__ __forceinline__ double rsqrt_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
  y = y * fma(-0.5 * x * y, y, 1.5);
  y = y * fma(-0.5 * x * y, y, 1.5);
  return y;
}
__device__ __forceinline__ double rcp_nr(double x) {
  double y = __builtin_amdgcn_rcp(x);
  y = y * fma(-x, y, 2.0);
  y = y * fma(-x, y, 2.0);
  return y;
}

// Rotation of that block: J = [[cs, sn], [-sn conj(ph), cs conj(ph)]],
// ph = apq / |apq|, chosen so that J^H [[app, apq], [conj(apq), aqq]] J is
// diagonal (Rutishauser's formulas). r = {cs, sn, Re ph, Im ph}.
__device__ __forceinline__ bool jacobi_rotation(double app, double aqq, double2 apq,
                                                double tol2, double abs_tol2, double* r) {
  const double mag2 = cabs2(apq);
  r[0] = 1.0; r[1] = 0.0; r[2] = 1.0; r[3] = 0.0;
  if (!needs_rotation(app, aqq, mag2, tol2, abs_tol2)) return false;
  const double inv = rsqrt_nr(mag2);
  const double zeta = (aqq - app) * 0.5 * inv;
  double t;
  if (zeta == 0.0) {
    t = 1.0;
  } else if (fabs(zeta) > 1e150) {
    t = copysign(0.5, zeta) * rcp_nr(fabs(zeta));
  } else {
    const double q = fma(zeta, zeta, 1.0);
    const double root = q * rsqrt_nr(q);             // sqrt(1 + zeta^2)
    t = copysign(1.0, zeta) * rcp_nr(fabs(zeta) + root);
  }
  const double cs = rsqrt_nr(fma(t, t, 1.0));
  r[0] = cs; r[1] = cs * t; r[2] = apq.x * inv; r[3] = apq.y * inv;
  return true;
}

// The same rotation with two dependent square-root chains instead of four:
// with d = aqq - app and h = sqrt(d^2 + 4|apq|^2), cos(theta) =
// sqrt((h + |d|) / (2h)) and sin(theta) = sign(d) |apq| / (h cos(theta))
// (theta = pi/4 for d = 0); no sum in it cancels. The phase apq / |apq| is an
// independent chain. The 2x2 block is first scaled by a power of two that
// brings max(app, aqq) into [1, 2) (the rotation is scale-invariant and
// |apq|^2 <= app aqq keeps every square finite), and everything is evaluated
// without branches (a pair that needs no rotation runs the formula on
// harmless operands and keeps the identity). r is the same
// {cs, sn, Re ph, Im ph} as jacobi_rotation's up to rounding.
__device__ __forceinline__ bool jacobi_rotation_fast(bool valid, double app, double aqq,
                                                     double2 apq, double tol2, double abs_tol2,
                                                     double* r) {
  const double mag2 = cabs2(apq);
  // needs_rotation, without short-circuit branches
  const bool need = valid & (mag2 > tol2 * fabs(app * aqq)) &
                    (mag2 > abs_tol2 * fmax(fmax(app, aqq), 0.0)) & (mag2 > 0.0);
  const double sc = need ? ldexp(1.0, -ilogb(fmax(fmax(app, aqq), 1e-300))) : 0.0;
  const double2 c = make_double2(apq.x * sc, apq.y * sc);
  const double m2 = need ? cabs2(c) : 1.0;
  const double d = (aqq - app) * sc;
  const double h2 = fma(d, d, 4.0 * m2);
  // 1 / |apq| and 1 / h: hardware estimates and two Newton steps each
  double inv = __builtin_amdgcn_rsq(m2), rh = __builtin_amdgcn_rsq(h2);
#pragma unroll
  for (int it = 0; it < 2; ++it) {
    inv = inv * fma(-0.5 * m2 * inv, inv, 1.5);
    rh = rh * fma(-0.5 * h2 * rh, rh, 1.5);
  }
  const double cos2 = fma(0.5 * fabs(d), rh, 0.5);   // cos^2(theta)
  double rc = __builtin_amdgcn_rsq(cos2);             // 1 / cos(theta)
#pragma unroll
  for (int it = 0; it < 2; ++it) rc = rc * fma(-0.5 * cos2 * rc, rc, 1.5);
  const double sn = copysign(m2 * inv * rh * rc, d < 0.0 ? -1.0 : 1.0);
  r[0] = need ? cos2 * rc : 1.0;
  r[1] = need ? sn : 0.0;
  r[2] = need ? c.x * inv : 1.0;
  r[3] = need ? c.y * inv : 0.0;
  return need;
}

// Valid columns of the block pair (bi, bj) as a 32-bit mask (the last block
// may be partial).
__device__ __forceinline__ unsigned pair_valid_mask(int bi, int bj, int n) {
  const unsigned mi = bi * JB + JB <= n ? 0xffffu : (bi * JB < n ? (1u << (n - bi * JB)) - 1u : 0u);
  const unsigned mj = bj * JB + JB <= n ? 0xffffu : (bj * JB < n ? (1u << (n - bj * JB)) - 1u : 0u);
  return mi | (mj << JB);
}

// Cyclic two-sided Jacobi sweeps on the pair's 2JB x 2JB Gram Gs[0],
// accumulating the rotations in Ws (identity on entry). Gs is
// double-buffered: a sub-step reads Gs[cur] and writes every entry of
// Gs[cur ^ 1], so one barrier per sub-step suffices. Each thread owns the
// 2x2 block (P1, P2) of the sub-step's pairing (ptab); each lane computes the
// rotation of pair P2 = lane & 15 and takes the one of pair P1 from lane P1
// of its wavefront; it also rotates rows P1 and P1+16 of Ws for pair P2.
__device__ __forceinline__ void pair_inner_sweeps(double2 (*Gs)[JP][JP + 2],
                                                  double2 (*Ws)[JP + 2],
                                                  const int (*ptab)[JB], unsigned vmask,
                                                  double tol2, double abs_tol2,
                                                  int inner_sweeps) {
  const int tid = threadIdx.x;
  const int P1 = tid >> 4, P2 = tid & 15;
  int cur = 0;
  for (int isw = 0; isw < inner_sweeps; ++isw) {
    int any = 0;
    for (int sub = 0; sub < JP - 1; ++sub) {
      const int e1 = ptab[sub][P1], e2 = ptab[sub][P2];
      const int p1 = e1 & 255, q1 = e1 >> 8, p2 = e2 & 255, q2 = e2 >> 8;
      // every LDS operand of the sub-step is read at once
      const double2 g22 = Gs[cur][p2][q2];
      const double d2p = Gs[cur][p2][p2].x, d2q = Gs[cur][q2][q2].x;
      double2 b00 = Gs[cur][p1][p2], b01 = Gs[cur][p1][q2];
      double2 b10 = Gs[cur][q1][p2], b11 = Gs[cur][q1][q2];
      double2 w0p = Ws[P1][p2], w0q = Ws[P1][q2];
      double2 w1p = Ws[P1 + 16][p2], w1q = Ws[P1 + 16][q2];
      // keep every read above issued before the rotation's arithmetic
      __builtin_amdgcn_sched_barrier(0);
      double r1[4], r2[4];
      const bool on2 = jacobi_rotation_fast(((vmask >> p2) & (vmask >> q2) & 1u) != 0u, d2p,
                                            d2q, g22, tol2, abs_tol2, r2);
#pragma unroll
      for (int h = 0; h < 4; ++h) r1[h] = __shfl(r2[h], P1, 64);
      any |= on2;
      // applied without branches: a pair that needs no rotation has the
      // identity (cs 1, sn 0, ph 1), which leaves its entries exactly as
      // they are
      {
        const double c = r2[0], sn = r2[1];
        const double2 cph = make_double2(r2[2], -r2[3]);
        const double2 x01 = cmul(b01, cph), x11 = cmul(b11, cph);
        const double2 n00 = make_double2(c * b00.x - sn * x01.x, c * b00.y - sn * x01.y);
        const double2 n01 = make_double2(sn * b00.x + c * x01.x, sn * b00.y + c * x01.y);
        const double2 n10 = make_double2(c * b10.x - sn * x11.x, c * b10.y - sn * x11.y);
        const double2 n11 = make_double2(sn * b10.x + c * x11.x, sn * b10.y + c * x11.y);
        b00 = n00; b01 = n01; b10 = n10; b11 = n11;
        const double2 y0 = cmul(w0q, cph), y1 = cmul(w1q, cph);
        Ws[P1][p2] = make_double2(c * w0p.x - sn * y0.x, c * w0p.y - sn * y0.y);
        Ws[P1][q2] = make_double2(sn * w0p.x + c * y0.x, sn * w0p.y + c * y0.y);
        Ws[P1 + 16][p2] = make_double2(c * w1p.x - sn * y1.x, c * w1p.y - sn * y1.y);
        Ws[P1 + 16][q2] = make_double2(sn * w1p.x + c * y1.x, sn * w1p.y + c * y1.y);
      }
      {
        const double c = r1[0], sn = r1[1];
        const double2 ph = make_double2(r1[2], r1[3]);
        const double2 y10 = cmul(ph, b10), y11 = cmul(ph, b11);
        const double2 n00 = make_double2(c * b00.x - sn * y10.x, c * b00.y - sn * y10.y);
        const double2 n01 = make_double2(c * b01.x - sn * y11.x, c * b01.y - sn * y11.y);
        const double2 n10 = make_double2(sn * b00.x + c * y10.x, sn * b00.y + c * y10.y);
        const double2 n11 = make_double2(sn * b01.x + c * y11.x, sn * b01.y + c * y11.y);
        b00 = n00; b01 = n01; b10 = n10; b11 = n11;
      }
      if (P1 == P2 && on2) {
        b01 = make_double2(0.0, 0.0);
        b10 = make_double2(0.0, 0.0);
        b00.y = 0.0;
        b11.y = 0.0;
      }
      const int nxt = cur ^ 1;
      Gs[nxt][p1][p2] = b00; Gs[nxt][p1][q2] = b01;
      Gs[nxt][q1][p2] = b10; Gs[nxt][q1][q2] = b11;
      __syncthreads();
      cur = nxt;
    }
    if (!__syncthreads_or(any)) break;
  }
}

// One 32-row chunk of the pair's 32 columns (rows r0.., below hi), four
// entries per thread (entry e = tid + 256 q: row e % 32, column e / 32), read
// with agent-scope loads; PF chunks are issued before any is consumed so
// that their latencies overlap.
constexpr int PF = 4;

__device__ __forceinline__ void load_chunk(double2 (&out)[4], const double2* base, long ld,
                                           int r0, int hi, const int* cols) {
  // 16-byte agent-coherent loads within the first 2^31 bytes of the
  // matrix (the buffer offset range), the pair of 8-byte loads beyond
  const __amdgpu_buffer_rsrc_t rs = agent_rsrc(base, 0x7fffffffu);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int e = threadIdx.x + JT * q;
    const int rr = e % JR, cc = e / JR;
    const int gr = r0 + rr, gc = cols[cc];
    const long off = gr + (long)gc * ld;
    out[q] = (gr < hi && gc >= 0)
                 ? (off < kBufMaxEntries ? ld_agent16(rs, (unsigned)off * 16u) : ld_agent(base + off))
                 : make_double2(0.0, 0.0);
  }
}

__device__ __forceinline__ void put_chunk(double2 (*Xs)[JP + 1], const double2 (&in)[4]) {
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int e = threadIdx.x + JT * q;
    Xs[e % JR][e / JR] = in[q];
  }
}

__device__ __forceinline__ void jacobi_body(const JacobiArgs& a, const int bid,
                                            const unsigned nblk) {
  __shared__ double2 Xs[JR][JP + 1];
  // rows padded by two entries: the sub-steps' scattered 16-byte accesses
  // conflict less in the LDS banks than with one
  __shared__ double2 Gs[2][JP][JP + 2];
  __shared__ double2 Ws[JP][JP + 2];
  // the pairs (p | q << 8, p < q) of every sub-step of the inner sweep
  __shared__ int ptab[JP - 1][JB];
  __shared__ int cols[JP];
  __shared__ int need;
  __shared__ bool skip_pair;
  __shared__ double red[JT];

  const int tid = threadIdx.x;
  const int R = a.slices;
  const int pair = bid / R, slice = bid % R;
  const int x_lo = min(a.m, slice * a.xrows_per), x_hi = min(a.m, x_lo + a.xrows_per);
  const int v_lo = min(a.n, slice * a.vrows_per), v_hi = min(a.n, v_lo + a.vrows_per);
  unsigned epoch = 0;
  for (int e = tid; e < (JP - 1) * JB; e += JT) {
    int p, q;
    ordered_pair(e / JB, e % JB, p, q);
    ptab[e / JB][e % JB] = p | (q << 8);
  }

  // V = identity, Frobenius norm of X
  for (long e = bid * (long)JT + tid; e < (long)a.n * a.n;
       e += (long)nblk * JT) {
    const long i = e % a.n, j = e / a.n;
    st_agent(a.V + (i + j * a.ldv), make_double2(i == j ? 1.0 : 0.0, 0.0));
  }
  double part = 0.0;
  for (long e = bid * (long)JT + tid; e < (long)a.m * a.n;
       e += (long)nblk * JT) {
    const long i = e % a.m, j = e / a.m;
    part += cabs2(a.X[i + j * a.ldx]);
  }
  red[tid] = part;
  __syncthreads();
  for (int s = JT / 2; s > 0; s >>= 1) {
    if (tid < s) red[tid] += red[tid + s];
    __syncthreads();
  }
  if (tid == 0) atomicAdd(a.fro2, red[0]);
  grid_barrier_all(a.bar, nblk, bid, epoch);
  const double fro = sqrt(__hip_atomic_load(a.fro2, __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_AGENT));
  const double abs_tol = a.abs_tol_scale * fro;
  const double tol2 = a.tol * a.tol, abs_tol2 = abs_tol * abs_tol;
  const int ti = tid >> 4, tj = tid & 15;        // 2x2 Gram blocking
  const int npairs = a.nb / 2;
  unsigned* pair_bar = a.bar + kBarrierWords;
  unsigned* pair_done = pair_bar + npairs;
  unsigned* blockver = pair_done + npairs;
  unsigned gen = 0, pair_epoch = 0;

  int sweep = 0;
  bool converged = false;
  PHASE_INIT();
  for (; sweep < a.max_sweeps; ++sweep) {
    for (int round = 0; round < a.nb - 1; ++round) {
      PHASE(7);
#ifdef PTE_ROUND_TRACE
      const int rt_round_ = sweep * (a.nb - 1) + round;
#endif
      RTRACE(0);
      int bi, bj;
      tournament_pair(a.nb, round, pair, bi, bj);
      if (tid < JP) {
        const int c = (tid < JB) ? bi * JB + tid : bj * JB + (tid - JB);
        cols[tid] = (c < a.n) ? c : -1;
      }
      // both blocks released by the pairs that held them in the previous round
      const int plo = min(bi, bj), phi = max(bi, bj);
      unsigned* const pair_ok = a.skip + a.nb + ((long)phi * (phi - 1)) / 2 + plo;
      if (tid == 0) {
        while (ld_agent(blockver + bi) < gen || ld_agent(blockver + bj) < gen)
          __builtin_amdgcn_s_sleep(1);
        // a pair found converged whose blocks have not been rotated since has
        // the same Gram matrix: it needs no rotation and no pass over X
        const unsigned ok = ld_agent(pair_ok);
        skip_pair = ok != 0u && ld_agent(a.skip + bi) < ok && ld_agent(a.skip + bj) < ok;
      }
      __syncthreads();
      RTRACE(1);
      if (skip_pair) {
        // the pair barrier is kept in a skipped round: without it a slice
        // could finish the next round before a slower slice of its pair
        // finished this one, and the arrival count on pair_done, which
        // releases the blocks, would then never reach this round's total
        if (R > 1) grid_barrier(pair_bar + pair, R, pair_epoch);
      } else {
      // ---- partial Gram of the pair's columns over this slice's rows ----
      double2 g[2][2];
#pragma unroll
      for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int v = 0; v < 2; ++v) g[u][v] = make_double2(0.0, 0.0);
      for (int gbase = x_lo; gbase < x_hi; gbase += PF * JR) {
        double2 pre[PF][4];
#pragma unroll
        for (int c = 0; c < PF; ++c) load_chunk(pre[c], a.X, a.ldx, gbase + c * JR, x_hi, cols);
#pragma unroll
        for (int c = 0; c < PF; ++c) {
          if (gbase + c * JR >= x_hi) break;
          put_chunk(Xs, pre[c]);
          __syncthreads();
#pragma unroll 4
          for (int rr = 0; rr < JR; ++rr) {
            const double2 a0 = Xs[rr][ti], a1 = Xs[rr][ti + 16];
            const double2 b0 = Xs[rr][tj], b1 = Xs[rr][tj + 16];
            cfmac(g[0][0], a0, b0);
            cfmac(g[0][1], a0, b1);
            cfmac(g[1][0], a1, b0);
            cfmac(g[1][1], a1, b1);
          }
          __syncthreads();
        }
      }
      if (R == 1) {
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
          for (int v = 0; v < 2; ++v) Gs[0][ti + 16 * u][tj + 16 * v] = g[u][v];
      } else {
        // double-buffered by round parity (every round of a pair, skipped or
        // not, passes its pair barrier)
        double2* gbuf = a.gpart + (long)(pair_epoch & 1) * npairs * R * (JP * JP);
        double2* mine = gbuf + ((long)pair * R + slice) * (JP * JP);
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
          for (int v = 0; v < 2; ++v) st_agent(mine + ((ti + 16 * u) * JP + tj + 16 * v), g[u][v]);
        PHASE(0);
        RTRACE(2);
        grid_barrier(pair_bar + pair, R, pair_epoch);
        PHASE(1);
        RTRACE(3);
        // identical summation order (slice 0, 1, ...) in every workgroup of
        // the pair; loads of eight slices are in flight together
        constexpr int SB = kJacobiSumBatch;
        double2 acc[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[q] = make_double2(0.0, 0.0);
        const double2* gp = gbuf + (long)pair * R * (JP * JP);
        for (int sl0 = 0; sl0 < R; sl0 += SB) {
          double2 v[SB][4];
#pragma unroll
          for (int t = 0; t < SB; ++t)
#pragma unroll
            for (int q = 0; q < 4; ++q)
              v[t][q] = (sl0 + t < R) ? ld_agent(gp + (long)(sl0 + t) * (JP * JP) + tid + JT * q)
                                      : make_double2(0.0, 0.0);
#pragma unroll
          for (int t = 0; t < SB; ++t)
#pragma unroll
            for (int q = 0; q < 4; ++q)
              if (sl0 + t < R) acc[q] = cadd(acc[q], v[t][q]);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int e = tid + JT * q;
          Gs[0][e / JP][e % JP] = acc[q];
        }
      }
      for (int e = tid; e < JP * JP; e += JT)
        Ws[e / JP][e % JP] = make_double2((e / JP) == (e % JP) ? 1.0 : 0.0, 0.0);
      if (tid == 0) need = 0;
      __syncthreads();
      // ---- does the pair need a rotation? ----
      {
        int mine = 0;
        for (int e = tid; e < JP * JP; e += JT) {
          const int gi = e / JP, gj = e % JP;
          if (gi < gj && cols[gi] >= 0 && cols[gj] >= 0) {
            if (needs_rotation(Gs[0][gi][gi].x, Gs[0][gj][gj].x, cabs2(Gs[0][gi][gj]), tol2,
                               abs_tol2))
              mine = 1;
          }
        }
        if (mine) need = 1;
      }
      __syncthreads();
      PHASE(2);
      RTRACE(4);
      // record the outcome for later rounds: a rotated pair's blocks change,
      // an unrotated pair is converged as long as they do not
      if (slice == 0 && tid == 0) {
        if (need) {
          st_agent(a.skip + bi, gen + 1);
          st_agent(a.skip + bj, gen + 1);
          st_agent(pair_ok, 0u);
        } else {
          st_agent(pair_ok, gen + 1);
        }
      }
      if (need) {
        const unsigned vmask = pair_valid_mask(bi, bj, a.n);
        pair_inner_sweeps(Gs, Ws, ptab, vmask, tol2, abs_tol2, a.inner_sweeps);
        PHASE(3);
        RTRACE(5);
        // ---- X_pair <- X_pair Ws, V_pair <- V_pair Ws over this slice ----
        const int rr = tid >> 3, c0 = (tid & 7) * 4;
        const int nx = (x_hi - x_lo + JR - 1) / JR, nv = (v_hi - v_lo + JR - 1) / JR;
        for (int t0 = 0; t0 < nx + nv; t0 += PF) {
          double2 pre[PF][4];
#pragma unroll
          for (int c = 0; c < PF; ++c) {
            const int t = t0 + c;
            if (t < nx) load_chunk(pre[c], a.X, a.ldx, x_lo + t * JR, x_hi, cols);
            else if (t < nx + nv) load_chunk(pre[c], a.V, a.ldv, v_lo + (t - nx) * JR, v_hi, cols);
          }
#pragma unroll
          for (int c = 0; c < PF; ++c) {
            const int t = t0 + c;
            if (t >= nx + nv) break;
            const bool isx = t < nx;
            double2* base = isx ? a.X : a.V;
            const long ld = isx ? a.ldx : a.ldv;
            const int r0 = isx ? x_lo + t * JR : v_lo + (t - nx) * JR;
            const int hi = isx ? x_hi : v_hi;
            put_chunk(Xs, pre[c]);
            __syncthreads();
            double2 o[4];
#pragma unroll
            for (int cc = 0; cc < 4; ++cc) o[cc] = make_double2(0.0, 0.0);
#pragma unroll 8
            for (int i = 0; i < JP; ++i) {
              const double2 xv = Xs[rr][i];
#pragma unroll
              for (int cc = 0; cc < 4; ++cc) cfma(o[cc], xv, Ws[i][c0 + cc]);
            }
            __syncthreads();
            // through LDS, so that each store covers 32 consecutive rows of
            // one column (512 contiguous bytes) instead of scattered entries
#pragma unroll
            for (int cc = 0; cc < 4; ++cc) Xs[rr][c0 + cc] = o[cc];
            __syncthreads();
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const int e = tid + JT * q;
              const int row = e % JR, col = e / JR;
              const int gr = r0 + row, gc = cols[col];
              const long off = gr + (long)gc * ld;
              if (gr < hi && gc >= 0) {
                if (off < kBufMaxEntries)
                  st_agent16(agent_rsrc(base, 0x7fffffffu), (unsigned)off * 16u, Xs[row][col]);
                else
                  st_agent(base + off, Xs[row][col]);
              }
            }
            __syncthreads();
          }
        }
        if (tid == 0) st_agent(a.flags + sweep, 1);
        PHASE(4);
        RTRACE(6);
      }
      }   // skip_pair
      // the last slice of the pair to finish releases both blocks
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
      if (tid == 0) {
        const unsigned old = __hip_atomic_fetch_add(pair_done + pair, 1u, __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT);
        if (old == (gen + 1) * (unsigned)R - 1) {
          st_agent(blockver + bi, gen + 1);
          st_agent(blockver + bj, gen + 1);
        }
      }
      ++gen;
      PHASE(5);
      RTRACE(7);
    }
    grid_barrier_all(a.bar, nblk, bid, epoch);
    if (__hip_atomic_load(&a.flags[sweep], __ATOMIC_RELAXED,
                          __HIP_MEMORY_SCOPE_AGENT) == 0) {
      converged = true;
      break;
    }
  }
  if (bid == 0 && tid == 0) *a.info = converged ? sweep + 1 : -1;
}

__global__ void __launch_bounds__(JT)
jacobi_kernel(JacobiArgs a) {
  jacobi_body(a, blockIdx.x, gridDim.x);
}

// Several independent SVDs in one cooperative launch: workgroup b works on
// problem block_prob[b] as its (b - starts[p])-th workgroup of
// starts[p+1] - starts[p]; every problem has its own barrier counter.
__global__ void __launch_bounds__(JT)
jacobi_batch_kernel(const JacobiArgs* probs, const int* starts, const int* block_prob) {
  const int p = block_prob[blockIdx.x];
  const int s0 = starts[p];
  const JacobiArgs a = probs[p];
  jacobi_body(a, blockIdx.x - s0, starts[p + 1] - s0);
}

// The same Jacobi with the slice kept in LDS for the round (one workgroup
// per compute unit): each round loads the slice's rows of X and of V in the
// pair's 32 columns into an LDS tile with one batch of coalesced loads,
// computes the partial Gram from it and, after the pair's inner sweeps, the
// rotation X <- X W, V <- V W in place on the FP64 matrix cores
// (v_mfma_f64_16x16x4_f64, 3M split), and writes the tile back with coalesced
// stores. The round protocol (block versions, converged-pair skipping, pair
// barrier, release count) is the one of jacobi_body. The tile has
// a.xrows_per + a.vrows_per rows of JP + 1 entries (dynamic LDS).
constexpr int kTileLd = kJacobiTileLd;
static_assert(kTileLd == JP + 1, "tile row stride");

__device__ __forceinline__ void jacobi_tile_body(const JacobiArgs& a, const int bid,
                                                 const unsigned nblk) {
  extern __shared__ double2 tile[];
  __shared__ double2 Gs[2][JP][JP + 2];
  __shared__ double2 Ws[JP][JP + 2];
  __shared__ int ptab[JP - 1][JB];
  __shared__ int cols[JP];
  __shared__ int need;
  __shared__ bool skip_pair;
  __shared__ double red[JT];

  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int lr = lane & 15, lk = lane >> 4;
  const int R = a.slices;
  const int pair = bid / R, slice = bid % R;
  const int xr = a.xrows_per, vr = a.vrows_per, tr = xr + vr;
  const int x_lo = min(a.m, slice * xr), x_hi = min(a.m, x_lo + xr);
  const int v_lo = min(a.n, slice * vr), v_hi = min(a.n, v_lo + vr);
  // 16-byte agent-coherent accesses to X, V and the partial Grams (every
  // offset of the tile path fits 31 bits: at most 16 slices of 192 rows and
  // 8192 columns)
  const __amdgpu_buffer_rsrc_t xrs =
      agent_rsrc(a.X, (unsigned)min((long)a.ldx * a.n * 16, 0x7fffffffL));
  const __amdgpu_buffer_rsrc_t vrs =
      agent_rsrc(a.V, (unsigned)min((long)a.ldv * a.n * 16, 0x7fffffffL));
  const __amdgpu_buffer_rsrc_t grs =
      agent_rsrc(a.gpart, (unsigned)min(2L * (a.nb / 2) * a.slices * (JP * JP) * 16, 0x7fffffffL));
  unsigned epoch = 0;
  for (int e = tid; e < (JP - 1) * JB; e += JT) {
    int p, q;
    ordered_pair(e / JB, e % JB, p, q);
    ptab[e / JB][e % JB] = p | (q << 8);
  }

  // V = identity, Frobenius norm of X
  for (long e = bid * (long)JT + tid; e < (long)a.n * a.n; e += (long)nblk * JT) {
    const long i = e % a.n, j = e / a.n;
    st_agent(a.V + (i + j * a.ldv), make_double2(i == j ? 1.0 : 0.0, 0.0));
  }
  double part = 0.0;
  for (long e = bid * (long)JT + tid; e < (long)a.m * a.n; e += (long)nblk * JT) {
    const long i = e % a.m, j = e / a.m;
    part += cabs2(a.X[i + j * a.ldx]);
  }
  red[tid] = part;
  __syncthreads();
  for (int s = JT / 2; s > 0; s >>= 1) {
    if (tid < s) red[tid] += red[tid + s];
    __syncthreads();
  }
  if (tid == 0) atomicAdd(a.fro2, red[0]);
  grid_barrier_all(a.bar, nblk, bid, epoch);
  const double fro = sqrt(__hip_atomic_load(a.fro2, __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_AGENT));
  const double abs_tol = a.abs_tol_scale * fro;
  const double tol2 = a.tol * a.tol, abs_tol2 = abs_tol * abs_tol;
  const int npairs = a.nb / 2;
  unsigned* pair_bar = a.bar + kBarrierWords;
  unsigned* pair_done = pair_bar + npairs;
  unsigned* blockver = pair_done + npairs;
  unsigned gen = 0, pair_epoch = 0;
  const int ntile = tr * JP;          // entries of the tile

  int sweep = 0;
  bool converged = false;
  for (; sweep < a.max_sweeps; ++sweep) {
    for (int round = 0; round < a.nb - 1; ++round) {
#ifdef PTE_ROUND_TRACE
      const int rt_round_ = sweep * (a.nb - 1) + round;
#endif
      RTRACE(0);
      int bi, bj;
      tournament_pair(a.nb, round, pair, bi, bj);
      if (tid < JP) {
        const int c = (tid < JB) ? bi * JB + tid : bj * JB + (tid - JB);
        cols[tid] = (c < a.n) ? c : -1;
      }
      // both blocks released by the pairs that held them in the previous round
      const int plo = min(bi, bj), phi = max(bi, bj);
      unsigned* const pair_ok = a.skip + a.nb + ((long)phi * (phi - 1)) / 2 + plo;
      if (tid == 0) {
        while (ld_agent(blockver + bi) < gen || ld_agent(blockver + bj) < gen)
          __builtin_amdgcn_s_sleep(1);
        const unsigned ok = ld_agent(pair_ok);
        skip_pair = ok != 0u && ld_agent(a.skip + bi) < ok && ld_agent(a.skip + bj) < ok;
      }
      __syncthreads();
      RTRACE(1);
      if (skip_pair) {
        // the pair barrier is kept in a skipped round (see jacobi_body)
        if (R > 1) grid_barrier(pair_bar + pair, R, pair_epoch);
      } else {
        // ---- the slice's rows of X and V into the tile: entry e is row
        // e % tr of column e / tr, so that a wavefront reads runs of
        // consecutive rows of one column; every load of the tile is in
        // flight at once (at most kTileMaxRows * JP / JT per thread) ----
        {
          constexpr int kLoads = kTileMaxRows * JP / JT;
          double2 v[kLoads];
#pragma unroll
          for (int q = 0; q < kLoads; ++q) {
            const int e = tid + JT * q;
            const int row = e % tr, col = e / tr;
            v[q] = make_double2(0.0, 0.0);
            if (e < ntile && cols[col] >= 0) {
              if (row < xr) {
                if (x_lo + row < x_hi)
                  v[q] = ld_agent16(xrs, (unsigned)(x_lo + row + (long)cols[col] * a.ldx) * 16u);
              } else if (v_lo + row - xr < v_hi) {
                v[q] = ld_agent16(vrs, (unsigned)(v_lo + row - xr + (long)cols[col] * a.ldv) * 16u);
              }
            }
          }
#pragma unroll
          for (int q = 0; q < kLoads; ++q) {
            const int e = tid + JT * q;
            if (e < ntile) tile[(e % tr) * kTileLd + e / tr] = v[q];
          }
        }
        __syncthreads();
        // partial Grams are double-buffered by the parity of the pair
        // barrier's epoch before this round's barrier
        const double2* const gbuf =
            a.gpart + (long)(pair_epoch & 1) * npairs * R * (JP * JP) + (long)pair * R * (JP * JP);
        {
          const int I = wave >> 1, J = wave & 1;
          doublex4 p1 = {0.0, 0.0, 0.0, 0.0}, p2 = p1, p3 = p1;
          for (int k4 = 0; k4 < xr; k4 += 4) {
            const double2 x = tile[(k4 + lk) * kTileLd + 16 * I + lr];
            const double2 y = tile[(k4 + lk) * kTileLd + 16 * J + lr];
            const double ar = x.x, ai = -x.y;
            p1 = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, y.x, p1, 0, 0, 0);
            p2 = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, y.y, p2, 0, 0, 0);
            p3 = __builtin_amdgcn_mfma_f64_16x16x4f64(ar + ai, y.x + y.y, p3, 0, 0, 0);
          }
          double2* mine = const_cast<double2*>(gbuf) + (long)slice * (JP * JP);
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int gi = 16 * I + 4 * r + lk, gj = 16 * J + lr;
            const double2 g = make_double2(p1[r] - p2[r], p3[r] - p1[r] - p2[r]);
            if (R == 1) Gs[0][gi][gj] = g;
            else st_agent16(grs, (unsigned)((mine - a.gpart) + gi * JP + gj) * 16u, g);
          }
        }
        if (R > 1) {
          RTRACE(2);
          grid_barrier(pair_bar + pair, R, pair_epoch);
          RTRACE(3);
          // identical summation order (slice 0, 1, ...) in every workgroup
          // of the pair; the loads of all slices (at most 16) in flight
          // together
          constexpr int SB = 16;
          double2 acc[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) acc[q] = make_double2(0.0, 0.0);
          const double2* gp = gbuf;
          for (int sl0 = 0; sl0 < R; sl0 += SB) {
            double2 v[SB][4];
#pragma unroll
            for (int t = 0; t < SB; ++t)
#pragma unroll
              for (int q = 0; q < 4; ++q)
                v[t][q] = (sl0 + t < R)
                              ? ld_agent16(grs, (unsigned)((gp - a.gpart) + (long)(sl0 + t) * (JP * JP) +
                                                           tid + JT * q) * 16u)
                              : make_double2(0.0, 0.0);
#pragma unroll
            for (int t = 0; t < SB; ++t)
#pragma unroll
              for (int q = 0; q < 4; ++q)
                if (sl0 + t < R) acc[q] = cadd(acc[q], v[t][q]);
          }
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int e = tid + JT * q;
            Gs[0][e / JP][e % JP] = acc[q];
          }
        }
        for (int e = tid; e < JP * JP; e += JT)
          Ws[e / JP][e % JP] = make_double2((e / JP) == (e % JP) ? 1.0 : 0.0, 0.0);
        if (tid == 0) need = 0;
        __syncthreads();
        // ---- does the pair need a rotation? ----
        {
          int mine = 0;
          for (int e = tid; e < JP * JP; e += JT) {
            const int gi = e / JP, gj = e % JP;
            if (gi < gj && cols[gi] >= 0 && cols[gj] >= 0) {
              if (needs_rotation(Gs[0][gi][gi].x, Gs[0][gj][gj].x, cabs2(Gs[0][gi][gj]), tol2,
                                 abs_tol2))
                mine = 1;
            }
          }
          if (mine) need = 1;
        }
        __syncthreads();
        RTRACE(4);
        // record the outcome for later rounds: a rotated pair's blocks change,
        // an unrotated pair is converged as long as they do not
        if (slice == 0 && tid == 0) {
          if (need) {
            st_agent(a.skip + bi, gen + 1);
            st_agent(a.skip + bj, gen + 1);
            st_agent(pair_ok, 0u);
          } else {
            st_agent(pair_ok, gen + 1);
          }
        }
        if (need) {
          pair_inner_sweeps(Gs, Ws, ptab, pair_valid_mask(bi, bj, a.n), tol2, abs_tol2,
                            a.inner_sweeps);
          RTRACE(5);
          // ---- tile <- tile W on the matrix cores: B = W (lane: k = 4 k4 +
          // lk, column lr of the 16-column half v), A = 16 tile rows ----
          double wr[8][2], wi[8][2];
#pragma unroll
          for (int k4 = 0; k4 < 8; ++k4)
#pragma unroll
            for (int v = 0; v < 2; ++v) {
              const double2 w = Ws[4 * k4 + lk][16 * v + lr];
              wr[k4][v] = w.x;
              wi[k4][v] = w.y;
            }
          for (int t = wave; t < tr / 16; t += 4) {
            doublex4 p1[2], p2[2], p3[2];
#pragma unroll
            for (int v = 0; v < 2; ++v) {
              p1[v] = doublex4{0.0, 0.0, 0.0, 0.0};
              p2[v] = p1[v];
              p3[v] = p1[v];
            }
#pragma unroll
            for (int k4 = 0; k4 < 8; ++k4) {
              const double2 x = tile[(16 * t + lr) * kTileLd + 4 * k4 + lk];
#pragma unroll
              for (int v = 0; v < 2; ++v) {
                p1[v] = __builtin_amdgcn_mfma_f64_16x16x4f64(x.x, wr[k4][v], p1[v], 0, 0, 0);
                p2[v] = __builtin_amdgcn_mfma_f64_16x16x4f64(x.y, wi[k4][v], p2[v], 0, 0, 0);
                p3[v] = __builtin_amdgcn_mfma_f64_16x16x4f64(x.x + x.y, wr[k4][v] + wi[k4][v],
                                                             p3[v], 0, 0, 0);
              }
            }
            // the wavefront's reads of these 16 rows are complete (in-order
            // LDS operations of one wavefront) before it overwrites them
#pragma unroll
            for (int v = 0; v < 2; ++v)
#pragma unroll
              for (int r = 0; r < 4; ++r)
                tile[(16 * t + 4 * r + lk) * kTileLd + 16 * v + lr] =
                    make_double2(p1[v][r] - p2[v][r], p3[v][r] - p1[v][r] - p2[v][r]);
          }
          __syncthreads();
          // ---- the tile back, entry e as in the load ----
          for (int e = tid; e < ntile; e += JT) {
            const int row = e % tr, col = e / tr;
            const int gc = cols[col];
            if (gc < 0) continue;
            const double2 val = tile[row * kTileLd + col];
            if (row < xr) {
              if (x_lo + row < x_hi)
                st_agent16(xrs, (unsigned)(x_lo + row + (long)gc * a.ldx) * 16u, val);
            } else if (v_lo + row - xr < v_hi) {
              st_agent16(vrs, (unsigned)(v_lo + row - xr + (long)gc * a.ldv) * 16u, val);
            }
          }
          if (tid == 0) st_agent(a.flags + sweep, 1);
          RTRACE(6);
        }
      }   // skip_pair
      // the last slice of the pair to finish releases both blocks
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
      if (tid == 0) {
        const unsigned old = __hip_atomic_fetch_add(pair_done + pair, 1u, __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT);
        if (old == (gen + 1) * (unsigned)R - 1) {
          st_agent(blockver + bi, gen + 1);
          st_agent(blockver + bj, gen + 1);
        }
      }
      ++gen;
      RTRACE(7);
    }
    grid_barrier_all(a.bar, nblk, bid, epoch);
    if (__hip_atomic_load(&a.flags[sweep], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
      converged = true;
      break;
    }
  }
  if (bid == 0 && tid == 0) *a.info = converged ? sweep + 1 : -1;
}

__global__ void __launch_bounds__(JT)
jacobi_tile_kernel(JacobiArgs a) {
  jacobi_tile_body(a, blockIdx.x, gridDim.x);
}

// Several independent SVDs in one cooperative launch of the LDS-tile
// Jacobi (problem map as in jacobi_batch_kernel; the dynamic LDS holds the
// largest problem's tile).
__global__ void __launch_bounds__(JT)
jacobi_tile_batch_kernel(const JacobiArgs* probs, const int* starts, const int* block_prob) {
  const int p = block_prob[blockIdx.x];
  const int s0 = starts[p];
  const JacobiArgs a = probs[p];
  jacobi_tile_body(a, blockIdx.x - s0, starts[p + 1] - s0);
}

// sigma_j = || X[:, j] ||  (one workgroup per column)
__device__ __forceinline__ void column_norm_body(const double2* X, int m, int n, long ldx,
                                                 double* sigma, int j) {
  __shared__ double red[256];
  if (j >= n) return;
  double s = 0.0;
  for (int i = threadIdx.x; i < m; i += 256) s += cabs2(X[i + (long)j * ldx]);
  red[threadIdx.x] = s;
  __syncthreads();
  for (int t = 128; t > 0; t >>= 1) {
    if (threadIdx.x < t) red[threadIdx.x] += red[threadIdx.x + t];
    __syncthreads();
  }
  if (threadIdx.x == 0) sigma[j] = sqrt(red[0]);
}

__global__ void __launch_bounds__(256)
column_norm_kernel(const double2* X, int m, int n, long ldx, double* sigma) {
  column_norm_body(X, m, n, ldx, sigma, blockIdx.x);
}

// descending order, ties by index: perm[rank] = j, sorted[rank] = sigma[j]
__device__ __forceinline__ void rank_sort_body(const double* sigma, int n, int* perm,
                                               double* sorted, int j) {
  if (j >= n) return;
  const double sj = sigma[j];
  int rank = 0;
  for (int i = 0; i < n; ++i) {
    const double si = sigma[i];
    rank += (si > sj) || (si == sj && i < j);
  }
  perm[rank] = j;
  sorted[rank] = sj;
}

__global__ void __launch_bounds__(256)
rank_sort_kernel(const double* sigma, int n, int* perm, double* sorted) {
  rank_sort_body(sigma, n, perm, sorted, blockIdx.x * 256 + threadIdx.x);
}

// extra2: squared norm of a part of the matrix left out of `sorted` (the
// residual of a pivoted QR stopped at a rank cap), counted in every tail
__device__ __forceinline__ void truncate_body(const double* sorted, int n, double epsrel,
                                              SvdInfo* info, int max_rank, SvdInfo* copy,
                                              const double* extra2 = nullptr) {
  __shared__ double part[1024];
  __shared__ int cnt[1024];
  const int tid = threadIdx.x;
  const double base2 = extra2 ? *extra2 : 0.0;
  const int per = (n + 1023) / 1024;
  const int lo = min(n, tid * per), hi = min(n, lo + per);
  double s = 0.0;
  for (int i = lo; i < hi; ++i) s += sorted[i] * sorted[i];
  part[tid] = s;
  __syncthreads();
  // suffix sums of the per-thread partials (serial over 1024 values by thread 0)
  if (tid == 0) {
    double acc = base2;
    for (int t = 1023; t >= 0; --t) {
      const double v = part[t];
      part[t] = acc;        // sum of everything after thread t's segment
      acc += v;
    }
  }
  __syncthreads();
  const double s0 = (n > 0) ? sorted[0] : 0.0;
  const double thr = epsrel * s0;
  double tail = part[tid];
  int c = 0;
  for (int i = hi - 1; i >= lo; --i) {
    tail += sorted[i] * sorted[i];
    if (sqrt(tail) > thr) ++c;
  }
  cnt[tid] = c;
  __syncthreads();
  for (int t = 512; t > 0; t >>= 1) {
    if (tid < t) cnt[tid] += cnt[tid + t];
    __syncthreads();
  }
  if (tid == 0) {
    int k = cnt[0];
    info->rank = k;
    info->s0 = s0;
    info->overflow = (k > max_rank) ? 1 : 0;
    // discarded 2-norm
    double d = base2;
    for (int i = n - 1; i >= k; --i) d += sorted[i] * sorted[i];
    info->discarded = sqrt(d);
    if (copy) *copy = *info;
  }
}

__global__ void __launch_bounds__(1024)
truncate_kernel(const double* sorted, int n, double epsrel, SvdInfo* info,
                int max_rank, const double* extra2) {
  truncate_body(sorted, n, epsrel, info, max_rank, nullptr, extra2);
}

// ---- the same post-processing for many SVDs at once (problem = blockIdx.y,
// or blockIdx.x for the truncation); see SvdPostArgs ----
__global__ void __launch_bounds__(256)
column_norm_batch_kernel(const SvdPostArgs* pa) {
  const SvdPostArgs a = pa[blockIdx.y];
  column_norm_body(a.X, a.m, a.n, a.m, a.sigma, blockIdx.x);
}

__global__ void __launch_bounds__(256)
rank_sort_batch_kernel(const SvdPostArgs* pa) {
  const SvdPostArgs a = pa[blockIdx.y];
  rank_sort_body(a.sigma, a.n, a.perm, a.sorted, blockIdx.x * 256 + threadIdx.x);
}

// also gathers every SvdInfo into `out` for a single copy to the host
__global__ void __launch_bounds__(1024)
truncate_batch_kernel(const SvdPostArgs* pa, SvdInfo* out) {
  const SvdPostArgs a = pa[blockIdx.x];
  truncate_body(a.sorted, a.n, a.epsrel, a.info, a.max_rank, out + blockIdx.x);
}

// zero a list of device words (barrier counters, sweep flags): item = blockIdx.x
__global__ void __launch_bounds__(256)
zero_items_kernel(const ZeroItem* items) {
  const ZeroItem it = items[blockIdx.x];
  for (int i = threadIdx.x; i < it.words; i += 256) it.p[i] = 0u;
}

// out[i] = *probs[i].rank (pivoted-QR ranks for a single copy to the host)
__global__ void __launch_bounds__(256)
gather_qr_rank_kernel(const QrcpArgs* probs, int n, int* out) {
  for (int i = threadIdx.x; i < n; i += 256) out[i] = *probs[i].rank;
}



// ---------------------------------------------------------------------------
// Emission of the truncated factors. After the Jacobi on X = M^H (m' x n'):
// M = V Sigma (Z/Sigma)^H with Z the orthogonalised X. Left singular vector j
// of M is V[:, perm[j]]; Sigma * right^H row j is conj(Z[:, perm[j]]).
// ---------------------------------------------------------------------------
// dst[j*dj + a*da + b*db] = src[(a*B + b) + perm[j]*lds]     (conj if right)
// for a < A, b < B, j < k, with each stride d = d0 + dk * k. Left factors
// read V (no conjugation), Sigma*right^H rows read Z (conjugated).
__global__ void __launch_bounds__(256)
emit_kernel(EmitArgs a) {
  const int k = a.info->rank;
  const long dj = a.dj0 + a.djk * (long)k;
  const long da = a.da0 + a.dak * (long)k;
  const long db = a.db0 + a.dbk * (long)k;
  const int rows = a.A * a.B;
  const long total = (long)rows * k;
  for (long e = blockIdx.x * 256L + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    const int s = e % rows;
    const int j = e / rows;
    const int c = a.rowmap ? a.rowmap[s] : s;
    const int aa = c / a.B, b = c % a.B;
    double2 v = a.src[s + (long)a.perm[j] * a.lds];
    if (a.conj) v.y = -v.y;
    a.dst[j * dj + aa * da + b * db] = v;
  }
}

// ---------------------------------------------------------------------------
// Householder QR with column pivoting (preconditioner of the Jacobi SVD).
// A (m x n, column-major, lda) -> A P = Q R with P the column permutation
// piv (piv[t] = physical column of step t), Q = H_0 ... H_{r-1},
// H_t = I - tau_t v_t v_t^H, v_t[t] = 1 and v_t[i] = A[i + piv[t]*lda] *
// vscale[t] for i > t (the pivot column keeps its unscaled entries),
// R[t][piv[t]] = beta[t] and R[t][c] = A[t + c*lda] for t above the diagonal.
// The reflector convention is LAPACK's zlarfg, the column-norm downdate with
// its cancellation guard is zlaqp2's. The factorisation stops at step r when
// the Frobenius norm of the unfactored block is at most drop * ||A||_F; the
// rows below r of R are dropped (a perturbation of that size).
// Column c belongs to workgroup c % gridDim.x; every workgroup forms the
// step's reflector from the pivot column itself, so one grid barrier per
// step suffices: candidates (largest residual norm per workgroup) are posted
// before it and reduced identically by every workgroup after it.
// ---------------------------------------------------------------------------
// (value, column) maximum with ties to the smaller column; -1 = none
__device__ __forceinline__ void better(double& v, int& c, double ov, int oc) {
  if (oc >= 0 && (c < 0 || ov > v || (ov == v && oc < c))) { v = ov; c = oc; }
}

// global winner and residual total among the posted candidates; every
// thread of the workgroup gets identical results
__device__ void qrcp_reduce(const QrcpCand* cand, int nblk, double (*slot)[4],
                            double (*slotv)[4], int (*slotc)[4], int& parity,
                            int* bestc, double* total) {
  double bv = -1.0, sv = 0.0;
  int bc = -1;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
    sv += ld_agent(&cand[b].vsum);
    better(bv, bc, ld_agent(&cand[b].vmax), ld_agent(&cand[b].col));
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double ov = __shfl_xor(bv, off, 64);
    const int oc = __shfl_xor(bc, off, 64);
    better(bv, bc, ov, oc);
  }
  if ((threadIdx.x & 63) == 0) {
    slotv[parity][threadIdx.x >> 6] = bv;
    slotc[parity][threadIdx.x >> 6] = bc;
  }
  // block_sum256 synchronises before reading, which also publishes the slots
  *total = block_sum256(sv, slot, parity);
  double v = -1.0;
  int c = -1;
  const int par = parity ^ 1;
#pragma unroll
  for (int w = 0; w < 4; ++w) better(v, c, slotv[par][w], slotc[par][w]);
  *bestc = c;
}

// wavefront-level column work of qrcp_kernel: the four wavefronts of a
// workgroup take its columns in turn
__device__ __forceinline__ void qrcp_post(QrcpCand* dst, double wmax, int wcol, double wsum,
                                          double* cw_max, int* cw_col, double* cw_sum) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) { cw_max[wave] = wmax; cw_col[wave] = wcol; cw_sum[wave] = wsum; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = -1.0, sm = 0.0;
    int c = -1;
    for (int w = 0; w < 4; ++w) {
      sm += cw_sum[w];
      better(v, c, cw_max[w], cw_col[w]);
    }
    st_agent(&dst->vmax, v);
    st_agent(&dst->vsum, sm);
    st_agent(&dst->col, c);
  }
}

// loads of the pivoted QR issued together (per thread), and the rows of a
// column (per lane, times 64) kept in registers between its dot product and
// its update
constexpr int kQrBatch = 8;
constexpr int kQrCache = 16;

__device__ __forceinline__ void qrcp_body(const QrcpArgs& a, const int blk, const int nblk) {
  extern __shared__ double2 vcol[];            // reflector of the step (m)
  __shared__ double slot[2][4], slotv[2][4];
  __shared__ int slotc[2][4];
  __shared__ double vn1[kQrcpMaxLocal], vn2[kQrcpMaxLocal];
  __shared__ int done[kQrcpMaxLocal];
  __shared__ double cw_max[4], cw_sum[4];
  __shared__ int cw_col[4];
  int parity = 0;
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int m = a.m, n = a.n;
  const long lda = a.lda;
  const int kmax = m < n ? m : n;
  const int nloc = (n - blk + nblk - 1) / nblk;   // own columns blk, blk+nblk, ...
  unsigned epoch = 0;

  // initial column norms (one wavefront per column); when resuming, the
  // norms of the unfactored rows start..m-1 of the columns not yet pivoted
  const int j0 = a.start;
  double wmax = -1.0, wsum = 0.0;
  int wcol = -1;
  // each wavefront keeps its first own column (q = wave) in registers for the
  // whole factorisation, rows lane + 64 u (u < kQrCache) at absolute row
  // indices (rows below the current step are left as they are); its
  // updates are still stored every step, for the pivot reads and the
  // factors. It stays resident until it is pivoted.
  double2 res[kQrCache];
  bool resident = false;
#pragma unroll
  for (int u = 0; u < kQrCache; ++u) res[u] = make_double2(0.0, 0.0);
  if (wave < nloc && !(j0 > 0 && ld_agent(a.done + blk + wave * nblk))) {
    const double2* col = a.A + (long)(blk + wave * nblk) * lda;
    const __amdgpu_buffer_rsrc_t crs = agent_rsrc(col, (unsigned)m * 16u);
#pragma unroll
    for (int u = 0; u < kQrCache; ++u) {
      const int i = lane + 64 * u;
      if (i >= j0 && i < m) res[u] = ld_agent16(crs, (unsigned)(i) * 16u);
    }
    resident = true;
  }
  for (int q = wave; q < nloc; q += 4) {
    const int c = blk + q * nblk;
    if (j0 > 0 && ld_agent(a.done + c)) {
      if (lane == 0) done[q] = 1;
      continue;
    }
    const double2* col = a.A + (long)c * lda;
    const __amdgpu_buffer_rsrc_t crs = agent_rsrc(col, (unsigned)m * 16u);
    double part = 0.0;
    for (int base = j0 + lane; base < m; base += 64 * kQrBatch) {
      double2 t[kQrBatch];
#pragma unroll
      for (int u = 0; u < kQrBatch; ++u) {
        const int i = base + 64 * u;
        t[u] = (i < m) ? ld_agent16(crs, (unsigned)(i) * 16u) : make_double2(0.0, 0.0);
      }
#pragma unroll
      for (int u = 0; u < kQrBatch; ++u) part += cabs2(t[u]);
    }
    const double nrm = sqrt(wave_sum(part));
    if (lane == 0) {
      vn1[q] = nrm; vn2[q] = nrm; done[q] = 0;
      if (j0 == 0) st_agent(a.done + c, 0);
    }
    wsum += nrm * nrm;
    better(wmax, wcol, nrm, c);
  }
  qrcp_post(a.cand + blk, wmax, wcol, wsum, cw_max, cw_col, cw_sum);
  grid_barrier_all(a.bar, nblk, blk, epoch);
  double total;
  int p;
  qrcp_reduce(a.cand, nblk, slot, slotv, slotc, parity, &p, &total);
  const double total0 = total;
  const double drop2 = j0 > 0 ? a.drop_abs2 : a.drop * a.drop * total;
  const double tol3z = 1.4901161193847656e-08;   // sqrt(eps), zlaqp2
  const int jmax = a.max_steps ? min(kmax, max(1, *a.max_steps)) : kmax;
  int j = j0;
  // mark: the first step count with residual^2 <= (mark_rel |beta_0|)^2;
  // every workgroup reduces the same candidates in the same order, so the
  // test (and a stop on it) is uniform over the grid
  int mark = -1;
  double mark2 = -1.0;
  if (a.mark_rel > 0.0 && j0 > 0) {
    const double b0 = a.mark_rel * fabs(a.beta[0]);
    mark2 = b0 * b0;
  }
  PHASE_INIT();
  while (j < jmax && p >= 0 && total > drop2) {
    // reflector j from the pivot column p (rows j..m-1), whole workgroup: the
    // column is read once into vcol (loads issued in batches, so that their
    // latencies overlap), then scaled in LDS
    const double2* pc = a.A + (long)p * lda;
    const __amdgpu_buffer_rsrc_t prs = agent_rsrc(pc, (unsigned)m * 16u);
    double part = 0.0;
    for (int base = j + tid; base < m; base += 256 * kQrBatch) {
      double2 t[kQrBatch];
#pragma unroll
      for (int q = 0; q < kQrBatch; ++q) {
        const int i = base + 256 * q;
        t[q] = (i < m) ? ld_agent16(prs, (unsigned)(i) * 16u) : make_double2(0.0, 0.0);
      }
#pragma unroll
      for (int q = 0; q < kQrBatch; ++q) {
        const int i = base + 256 * q;
        if (i < m) {
          vcol[i] = t[q];
          if (i > j) part += cabs2(t[q]);
        }
      }
    }
    const double xnorm2 = block_sum256(part, slot, parity);
    const double2 alpha = vcol[j];
    double2 tau, scale;
    double beta;
    if (xnorm2 == 0.0 && alpha.y == 0.0) {
      tau = make_double2(0.0, 0.0);
      scale = make_double2(0.0, 0.0);
      beta = alpha.x;
    } else {
      beta = -copysign(sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xnorm2), alpha.x);
      const double dr = alpha.x - beta, di = alpha.y;
      const double den = dr * dr + di * di;
      scale = make_double2(dr / den, -di / den);
      tau = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
    }
    __syncthreads();   // every thread has read alpha = vcol[j]
    for (int i = j + tid; i < m; i += 256)
      vcol[i] = (i == j) ? make_double2(1.0, 0.0) : cmul(vcol[i], scale);
    if (blk == 0 && tid == 0) {
      a.piv[j] = p;
      a.tau[j] = tau;
      a.vscale[j] = scale;
      a.beta[j] = beta;
    }
    if (p % nblk == blk && tid == 0) {
      done[p / nblk] = 1;
      st_agent(a.done + p, 1);
    }
    if (p % nblk == blk && p / nblk == wave) resident = false;
    __syncthreads();
    PHASE(8);
    // apply H_j^H to the own remaining columns, downdate their norms
    const double2 tc = make_double2(tau.x, -tau.y);
    const bool reflect = (tau.x != 0.0 || tau.y != 0.0);
    wmax = -1.0; wsum = 0.0; wcol = -1;
    for (int q = wave; q < nloc; q += 4) {
      if (done[q]) continue;
      const int c = blk + q * nblk;
      double2* col = a.A + (long)c * lda;
      const __amdgpu_buffer_rsrc_t crs = agent_rsrc(col, (unsigned)m * 16u);
      if (q == wave && resident) {
        // the resident column: rows below 64 kQrCache from registers, the
        // rest (if any) streamed. The register rows are processed without
        // branches: rows outside [j, m) take a zero reflector entry (their
        // update is then exact: x - 0 = x); only the stores are predicated,
        // after all the arithmetic.
        const int beyond = max(j, 64 * kQrCache);
        if (reflect) {
          double2 vr[kQrCache];
          double2 w = make_double2(0.0, 0.0);
#pragma unroll
          for (int u = 0; u < kQrCache; ++u) {
            const int i = lane + 64 * u;
            const double2 vi = vcol[min(i, m - 1)];
            vr[u] = (i >= j && i < m) ? vi : make_double2(0.0, 0.0);
            cfmac(w, vr[u], res[u]);
          }
          for (int base = beyond + lane; base < m; base += 64 * kQrBatch) {
            double2 t[kQrBatch];
#pragma unroll
            for (int u = 0; u < kQrBatch; ++u) {
              const int i = base + 64 * u;
              t[u] = (i < m) ? ld_agent16(crs, (unsigned)(i) * 16u) : make_double2(0.0, 0.0);
            }
#pragma unroll
            for (int u = 0; u < kQrBatch; ++u) {
              const int i = base + 64 * u;
              if (i < m) cfmac(w, vcol[i], t[u]);
            }
          }
          w.x = wave_sum(w.x);
          w.y = wave_sum(w.y);
          const double2 tw = cmul(tc, w);
#pragma unroll
          for (int u = 0; u < kQrCache; ++u) res[u] = csub(res[u], cmul(vr[u], tw));
#pragma unroll
          for (int u = 0; u < kQrCache; ++u) {
            const int i = lane + 64 * u;
            if (i >= j && i < m) st_agent16(crs, (unsigned)(i) * 16u, res[u]);
          }
          for (int base = beyond + lane; base < m; base += 64 * kQrBatch) {
            double2 t[kQrBatch];
#pragma unroll
            for (int u = 0; u < kQrBatch; ++u) {
              const int i = base + 64 * u;
              t[u] = (i < m) ? ld_agent16(crs, (unsigned)(i) * 16u) : make_double2(0.0, 0.0);
            }
#pragma unroll
            for (int u = 0; u < kQrBatch; ++u) {
              const int i = base + 64 * u;
              if (i < m) st_agent16(crs, (unsigned)(i) * 16u, csub(t[u], cmul(vcol[i], tw)));
            }
          }
        }
        // row j of the updated column: register j / 64 of lane j % 64 (or
        // memory beyond the registers)
        double2 rj2 = make_double2(0.0, 0.0);
        if (j < 64 * kQrCache) {
#pragma unroll
          for (int u = 0; u < kQrCache; ++u)
            if (u == (j >> 6)) rj2 = res[u];
          rj2.x = __shfl(rj2.x, j & 63, 64);
          rj2.y = __shfl(rj2.y, j & 63, 64);
        } else {
          __threadfence_block();
          rj2 = ld_agent16(crs, (unsigned)(j) * 16u);
        }
        double nv = vn1[q];
        if (nv != 0.0) {
          const double rj = sqrt(cabs2(rj2));
          double temp = 1.0 - (rj / nv) * (rj / nv);
          temp = temp > 0.0 ? temp : 0.0;
          const double ratio = nv / vn2[q];
          const double temp2 = temp * ratio * ratio;
          if (temp2 <= tol3z) {
            // recompute the norm of rows j+1.. from the registers and, beyond
            // them, from memory (this lane's own stores: complete them before
            // reading back)
            __threadfence_block();
            double pp = 0.0;
#pragma unroll
            for (int u = 0; u < kQrCache; ++u) {
              const int i = lane + 64 * u;
              if (i > j && i < m) pp += cabs2(res[u]);
            }
            for (int i = max(j + 1, 64 * kQrCache) + lane; i < m; i += 64)
              pp += cabs2(ld_agent16(crs, (unsigned)(i) * 16u));
            nv = (j + 1 < m) ? sqrt(wave_sum(pp)) : 0.0;
            if (lane == 0) { vn1[q] = nv; vn2[q] = nv; }
          } else {
            nv = nv * sqrt(temp);
            if (lane == 0) vn1[q] = nv;
          }
        }
        wsum += nv * nv;
        better(wmax, wcol, nv, c);
        continue;
      }
      // rows j .. j + 64 kQrCache - 1 stay in registers between the dot
      // product and the update; all loads are issued in batches
      double2 cache[kQrCache];
#pragma unroll
      for (int u = 0; u < kQrCache; ++u) {
        const int i = j + lane + 64 * u;
        cache[u] = (i < m) ? ld_agent16(crs, (unsigned)(i) * 16u) : make_double2(0.0, 0.0);
      }
      if (reflect) {
        double2 w = make_double2(0.0, 0.0);
#pragma unroll
        for (int u = 0; u < kQrCache; ++u) {
          const int i = j + lane + 64 * u;
          if (i < m) cfmac(w, vcol[i], cache[u]);
        }
        for (int base = j + lane + 64 * kQrCache; base < m; base += 64 * kQrBatch) {
          double2 t[kQrBatch];
#pragma unroll
          for (int u = 0; u < kQrBatch; ++u) {
            const int i = base + 64 * u;
            t[u] = (i < m) ? ld_agent16(crs, (unsigned)(i) * 16u) : make_double2(0.0, 0.0);
          }
#pragma unroll
          for (int u = 0; u < kQrBatch; ++u) {
            const int i = base + 64 * u;
            if (i < m) cfmac(w, vcol[i], t[u]);
          }
        }
        w.x = wave_sum(w.x);
        w.y = wave_sum(w.y);
        const double2 tw = cmul(tc, w);
#pragma unroll
        for (int u = 0; u < kQrCache; ++u) {
          const int i = j + lane + 64 * u;
          if (i < m) {
            cache[u] = csub(cache[u], cmul(vcol[i], tw));
            st_agent16(crs, (unsigned)(i) * 16u, cache[u]);
          }
        }
        for (int base = j + lane + 64 * kQrCache; base < m; base += 64 * kQrBatch) {
          double2 t[kQrBatch];
#pragma unroll
          for (int u = 0; u < kQrBatch; ++u) {
            const int i = base + 64 * u;
            t[u] = (i < m) ? ld_agent16(crs, (unsigned)(i) * 16u) : make_double2(0.0, 0.0);
          }
#pragma unroll
          for (int u = 0; u < kQrBatch; ++u) {
            const int i = base + 64 * u;
            if (i < m) st_agent16(crs, (unsigned)(i) * 16u, csub(t[u], cmul(vcol[i], tw)));
          }
        }
      }
      // row j of the updated column sits in lane 0's first cached entry
      double2 rj2 = cache[0];
      rj2.x = __shfl(rj2.x, 0, 64);
      rj2.y = __shfl(rj2.y, 0, 64);
      double nv = vn1[q];
      if (nv != 0.0) {
        const double rj = sqrt(cabs2(rj2));
        double temp = 1.0 - (rj / nv) * (rj / nv);
        temp = temp > 0.0 ? temp : 0.0;
        const double ratio = nv / vn2[q];
        const double temp2 = temp * ratio * ratio;
        if (temp2 <= tol3z) {
          // recompute the norm of rows j+1.. from the cached entries and,
          // beyond them, from memory (this lane's own stores: complete them
          // before reading back)
          __threadfence_block();
          double pp = 0.0;
#pragma unroll
          for (int u = 0; u < kQrCache; ++u) {
            const int i = j + lane + 64 * u;
            if (i > j && i < m) pp += cabs2(cache[u]);
          }
          for (int i = j + lane + 64 * kQrCache; i < m; i += 64) pp += cabs2(ld_agent16(crs, (unsigned)(i) * 16u));
          nv = (j + 1 < m) ? sqrt(wave_sum(pp)) : 0.0;
          if (lane == 0) { vn1[q] = nv; vn2[q] = nv; }
        } else {
          nv = nv * sqrt(temp);
          if (lane == 0) vn1[q] = nv;
        }
      }
      wsum += nv * nv;
      better(wmax, wcol, nv, c);
    }
    PHASE(9);
    qrcp_post(a.cand + ((j + 1) & 1) * nblk + blk, wmax, wcol, wsum, cw_max, cw_col,
              cw_sum);
    PHASE(10);
    grid_barrier_all(a.bar, nblk, blk, epoch);
    PHASE(11);
    qrcp_reduce(a.cand + ((j + 1) & 1) * nblk, nblk, slot, slotv, slotc, parity, &p,
                &total);
    PHASE(12);
    ++j;
    if (a.mark_rel > 0.0) {
      if (j == 1) {
        const double b0 = a.mark_rel * fabs(beta);
        mark2 = b0 * b0;
      }
      if (mark < 0 && total <= mark2) {
        mark = j;
        if (a.stop_at_mark) break;
      }
    }
  }
  // remaining columns in ascending order after the pivots
  if (blk == 0 && tid == 0) {
    int t = j;
    for (int c = 0; c < n; ++c)
      if (!ld_agent(a.done + c)) a.piv[t++] = c;
    *a.rank = j;
    if (a.out) {
      a.out[0] = (p >= 0 && j < kmax) ? total : 0.0;
      if (j0 == 0) a.out[1] = total0;
      a.out[2] = j > 0 ? fabs(a.beta[0]) : 0.0;
      a.out[3] = (double)(mark >= 0 ? mark : j);
    }
  }
}

__global__ void __launch_bounds__(256)
qrcp_kernel(QrcpArgs a) {
  qrcp_body(a, blockIdx.x, gridDim.x);
}

// Several pivoted QRs in one cooperative launch (see jacobi_batch_kernel).
__global__ void __launch_bounds__(256)
qrcp_batch_kernel(const QrcpArgs* probs, const int* starts, const int* block_prob) {
  const int p = block_prob[blockIdx.x];
  const int s0 = starts[p];
  const QrcpArgs a = probs[p];
  qrcp_body(a, blockIdx.x - s0, starts[p + 1] - s0);
}

// Y = R^H (n x r, column-major, ld = n) in the pivoted column order:
// Y[s + t*n] = conj(R[t][piv[s]])
__device__ __forceinline__ void build_rh_body(const double2* A, long lda, int n,
                                              const int* piv, const double* beta,
                                              const int* rank, double2* Y, int bid,
                                              int nblocks) {
  const int r = *rank;
  const long total = (long)n * r;
  for (long e = bid * 256L + threadIdx.x; e < total; e += (long)nblocks * 256) {
    const int s = e % n;
    const int t = e / n;
    double2 v = make_double2(0.0, 0.0);
    if (t < s) v = cconj(A[t + (long)piv[s] * lda]);
    else if (t == s) v = make_double2(beta[s], 0.0);
    Y[e] = v;
  }
}

__global__ void __launch_bounds__(256)
build_rh_kernel(const double2* A, long lda, int n, const int* piv, const double* beta,
                const int* rank, double2* Y) {
  build_rh_body(A, lda, n, piv, beta, rank, Y, blockIdx.x, gridDim.x);
}

__global__ void __launch_bounds__(256)
build_rh_batch_kernel(const RhArgs* pa) {
  const RhArgs a = pa[blockIdx.y];
  build_rh_body(a.A, a.lda, a.n, a.piv, a.beta, a.rank, a.Y, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------
// Small truncated SVD in one workgroup of NT threads (M with at most
// kSmallMaxCols columns, staged in LDS): the pivoted QR of qrcp_kernel (same
// reflectors, pivoting, norm downdating and eps-rank stop), then the
// one-sided Jacobi of Y = R^H (n x r) with scalar column pairs in the
// round-robin order (the rotation and stopping rule of jacobi_kernel), then
// singular values, order and cut. Barriers are workgroup barriers only;
// column pairs (and QR columns) are reduced by lane groups with wavefront
// shuffles. Y stays in LDS; the accumulated rotation W (r x r) is kept in LDS
// when Y and W fit in the dynamic LDS (a.lds_bytes), else in a.W.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double group_sum(double v, int G) { return group_sum1(v, G); }

__device__ __forceinline__ int pow2_ceil(int x) {
  int p = 1;
  while (p < x) p <<= 1;
  return p;
}

// Sum over an NT-thread workgroup with one __syncthreads (wavefront sums in
// slot[parity][wave], parity alternating between calls); every thread gets
// the same bits.
template <int NT>
__device__ __forceinline__ double block_sum_nt(double v, double (*slot)[NT / 64], int& parity) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) slot[parity][threadIdx.x >> 6] = v;
  __syncthreads();
  double r = 0.0;
#pragma unroll
  for (int w = 0; w < NT / 64; ++w) r += slot[parity][w];
  parity ^= 1;
  return r;
}

// Lanes per pair (a power of two, NT/64..64, all pairs in one pass): the
// count that minimises the issue slots of the most loaded SIMD, busy
// wavefronts per SIMD x instructions per wavefront (rows per lane of the two
// Y and two W columns, plus the reduction and the rotation); ties go to more
// lanes (shorter chains).
template <int NT>
__device__ __forceinline__ int small_pair_lanes(int P, int n, int r) {
  int best = NT / 64, best_cost = 1 << 30;
  for (int gj = NT / 64; gj <= 64; gj *= 2) {
    if (P * gj > NT || (gj > NT / 64 && gj > pow2_ceil(n))) break;
    const int waves = (P * gj + 63) / 64;
    const int per_simd = (waves + 3) / 4;
    const int instr = ((n + gj - 1) / gj) * 35 + ((r + gj - 1) / gj) * 15 + 150;
    const int cost = per_simd * instr;
    if (cost <= best_cost) { best_cost = cost; best = gj; }
  }
  return best;
}

// Cyclic one-sided Jacobi sweeps on the columns of Ys (n x r, ld n, LDS),
// rotations accumulated in W (r x r, ld r): in LDS after Y when WL, else Wg.
// Lane group q (GJ lanes) owns pair q of each round; a lane holds rows
// gj, gj + GJ, ... (at most RP) of the pair's two columns in registers. The
// row loops are branch-free: a block of GJ rows is skipped when it starts
// past the end (a test uniform over the workgroup), and inside the last block
// a lane past the end works on the last row instead, masked out of the inner
// products; it then stores the same bits as the lane that owns that row.
// A pair is skipped when neither column has been rotated since the round
// one sweep earlier (the order repeats every sweep, so that round checked
// this pair and found it converged): recomputing it would give the same bits.
// last_rot[c] (LDS) is the round of column c's last rotation; a wavefront
// whose pairs are all skipped only takes part in the barrier.
// Returns the sweep count; `converged` when the last sweep rotated nothing.
template <int NT, bool WL>
__device__ __forceinline__ int small_jacobi(double2* Ys, double2* Wg, int n, int r,
                                            double tol2, double abs_tol2, int max_sweeps,
                                            int* last_rot, bool& converged) {
  constexpr int RP = kSmallMaxCols * 64 / NT;   // GJ >= NT / 64 lanes per pair
  double2* const Wp = WL ? Ys + (long)n * r : Wg;
  const int tid = threadIdx.x;
  const int re = r + (r & 1);
  const int P = re / 2;
  const int GJ = small_pair_lanes<NT>(P, n, r);
  const int q = tid / GJ, gj = tid % GJ;
  const int L = re > 1 ? re - 1 : 1;
  // a wavefront whose lane groups all lie beyond the last pair only takes
  // part in the barriers
  const bool wave_busy = (tid & ~63) / GJ < P;
  const double2 zero = make_double2(0.0, 0.0);
  int sweeps = 0;
  converged = false;
  SEG_INIT();
  for (; sweeps < max_sweeps; ++sweeps) {
    int any = 0;
    for (int round = 0; round < L; ++round) {
      const int gr = sweeps * L + round;   // global round index
      if (!wave_busy) {
        __syncthreads();
        continue;
      }
      int ci = 0, cj = 0;
      bool valid = false;
      if (q < P && re > 1) {
        // round-robin pairing of tournament_pair without integer division
        const int Lr = re - 1;
        if (q == 0) {
          ci = Lr;
          cj = round;
        } else {
          ci = round + q;
          if (ci >= Lr) ci -= Lr;
          cj = round - q;
          if (cj < 0) cj += Lr;
        }
        if (ci > cj) { const int tt = ci; ci = cj; cj = tt; }
        valid = cj < r;
        if (valid && gr >= L && last_rot[ci] < gr - L && last_rot[cj] < gr - L)
          valid = false;   // converged at its last check, columns unchanged
      }
      // a wavefront whose pairs are all skipped only takes part in the barrier
      if (__ballot(valid) != 0) {
        if (!valid) { ci = 0; cj = 0; }   // harmless addresses; no rotation
        double2* y0 = Ys + (long)ci * n;
        double2* y1 = Ys + (long)cj * n;
        double2 yi[RP], yj[RP];
        double al = 0.0, be = 0.0;
        double2 ga = zero;
#pragma unroll
        for (int t = 0; t < RP; ++t) {
          if (t * GJ < n) {
            const int i0 = gj + t * GJ;
            const int i = min(i0, n - 1);
            yi[t] = y0[i];
            yj[t] = y1[i];
            const double2 u = i0 < n ? yi[t] : zero;
            const double2 v = i0 < n ? yj[t] : zero;
            al += cabs2(u);
            be += cabs2(v);
            cfmac(ga, u, v);
          }
        }
        SEG(0);
        group_sum4(al, be, ga.x, ga.y, GJ);
        SEG(1);
        double rot[4];
        const bool rotate = jacobi_rotation_fast(valid, al, be, ga, tol2, abs_tol2, rot);
        SEG(2);
        if (rotate) {
          any = 1;
          const double c = rot[0], sn = rot[1];
          const double2 cph = make_double2(rot[2], -rot[3]);
#pragma unroll
          for (int t = 0; t < RP; ++t) {
            if (t * GJ < n) {
              const int i = min(gj + t * GJ, n - 1);
              const double2 xp = yi[t];
              const double2 xq = cmul(yj[t], cph);
              y0[i] = make_double2(c * xp.x - sn * xq.x, c * xp.y - sn * xq.y);
              y1[i] = make_double2(sn * xp.x + c * xq.x, sn * xp.y + c * xq.y);
            }
          }
          SEG(3);
          double2* w0 = Wp + (long)ci * r;
          double2* w1 = Wp + (long)cj * r;
#pragma unroll
          for (int t = 0; t < RP; ++t) {
            if (t * GJ < r) {
              const int i = min(gj + t * GJ, r - 1);
              const double2 xp = w0[i];
              const double2 xq = cmul(w1[i], cph);
              w0[i] = make_double2(c * xp.x - sn * xq.x, c * xp.y - sn * xq.y);
              w1[i] = make_double2(sn * xp.x + c * xq.x, sn * xp.y + c * xq.y);
            }
          }
          if (gj == 0) { last_rot[ci] = gr; last_rot[cj] = gr; }
          SEG(4);
        }
      }
      __syncthreads();
      SEG(5);
    }
    if (!__syncthreads_or(any)) {
      converged = true;
      break;
    }
  }
  SEG_FLUSH(0, 6);
  return sweeps;
}

template <int NT>
__device__ void small_svd_body(const SmallSvdArgs& a) {
  static_assert(NT % 64 == 0 && (kSmallMaxCols * 64) % NT == 0, "workgroup size");
  extern __shared__ double2 lds_small[];
  __shared__ double slot[2][NT / 64];
  int parity = 0;
  const int m = a.m, n = a.n, tid = threadIdx.x, lane = tid & 63;
  // QR phase: M (m x n, ld m), then the per-column scalars (64 n bytes)
  double2* As = lds_small;
  double2* q_tau = As + (long)m * n;
  double2* q_vscale = q_tau + n;
  double* vn1 = reinterpret_cast<double*>(q_vscale + n);
  double* vn2 = vn1 + n;
  double* q_beta = vn2 + n;
  int* done = reinterpret_cast<int*>(q_beta + n);
  int* q_piv = done + n;
  const double eps = 2.220446049250313e-16;

  PHASE_INIT();
  for (long e = tid; e < (long)m * n; e += NT) As[e] = a.A[e];
  __syncthreads();
  PHASE(6);
  // ---- pivoted QR in LDS ----
  const int NP = pow2_ceil(n);
  // threads per column: at most 64, at most NT / NP, no more than the rows
  const int G = min(min(NP >= NT / 64 ? NT / NP : 64, pow2_ceil(m)), 64);
  const int myc = tid / G, g = tid % G;
  {
    double part = 0.0;
    if (myc < n)
      for (int i = g; i < m; i += G) part += cabs2(As[i + (long)myc * m]);
    const double nrm = sqrt(group_sum(part, G));
    if (myc < n && g == 0) { vn1[myc] = nrm; vn2[myc] = nrm; done[myc] = 0; }
  }
  __syncthreads();
  // Each step takes two barriers: after the first, every wavefront reads the
  // step's pivot and reflector from LDS and the group of each remaining
  // column applies the reflector to it and downdates its norm; after the
  // second, wavefront 0 alone chooses the next pivot (largest residual norm,
  // ties to the smaller column), decides whether to stop, computes the next
  // reflector and stores it into the pivot column already scaled (v_i =
  // M_ij * scale below the diagonal, so vscale is 1), so the scalar work is
  // not repeated by every wavefront.
  __shared__ int s_ctl[2];          // pivot column, 1 while the QR continues
  __shared__ double2 s_tau;
  const int wave = tid >> 6;
  int j = 0;
  // adaptive stop (stop_rel > 0): once the residual is at most
  // stop_rel * |beta_0|; the residual is then counted in the cut (folded2,
  // used by thread 0 only)
  double drop2 = 0.0, stop2 = -1.0, folded2 = 0.0;
  const double tol3z = 1.4901161193847656e-08;
  const int kmax = m < n ? m : n;
  // wavefront 0: pivot and reflector of step j, or the stop
  auto pivot_step = [&]() {
    double bv = -1.0, sv = 0.0;
    int bc = -1;
    for (int c = lane; c < n; c += 64)
      if (!done[c]) { sv += vn1[c] * vn1[c]; better(bv, bc, vn1[c], c); }
    better(bv, bc, dpp_d<kDppXor1>(bv), dpp_i<kDppXor1>(bc));
    better(bv, bc, dpp_d<kDppXor2>(bv), dpp_i<kDppXor2>(bc));
    better(bv, bc, dpp_d<kDppHalfMirror>(bv), dpp_i<kDppHalfMirror>(bc));
    better(bv, bc, dpp_d<kDppMirror>(bv), dpp_i<kDppMirror>(bc));
    better(bv, bc, __shfl_xor(bv, 16, 64), __shfl_xor(bc, 16, 64));
    better(bv, bc, __shfl_xor(bv, 32, 64), __shfl_xor(bc, 32, 64));
    const int p = bc;
    const double total = wave_sum(sv);
    if (j == 0) drop2 = eps * eps * total;
    bool go;
    if (a.stop_rel > 0.0 && j >= 1) {
      if (j == 1) stop2 = (a.stop_rel * q_beta[0]) * (a.stop_rel * q_beta[0]);
      if (total <= stop2) {
        folded2 = (p >= 0 && j < kmax) ? total : 0.0;
        go = false;
      } else {
        go = j < kmax && p >= 0 && total > drop2;
      }
    } else {
      go = j < kmax && p >= 0 && total > drop2;
    }
    if (go) {
      double2* pc = As + (long)p * m;
      double part = 0.0;
      for (int i = j + 1 + lane; i < m; i += 64) part += cabs2(pc[i]);
      const double xnorm2 = wave_sum(part);
      const double2 alpha = pc[j];
      double2 tau, scale;
      double beta;
      if (xnorm2 == 0.0 && alpha.y == 0.0) {
        tau = make_double2(0.0, 0.0);
        scale = make_double2(0.0, 0.0);
        beta = alpha.x;
      } else {
        beta = -copysign(sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xnorm2), alpha.x);
        const double dr = alpha.x - beta, di = alpha.y;
        const double den = dr * dr + di * di;
        scale = make_double2(dr / den, -di / den);
        tau = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
      }
      for (int i = j + 1 + lane; i < m; i += 64) pc[i] = cmul(pc[i], scale);
      if (lane == 0) {
        q_piv[j] = p;
        q_tau[j] = tau;
        q_vscale[j] = make_double2(1.0, 0.0);
        q_beta[j] = beta;
        s_tau = tau;
        s_ctl[0] = p;
      }
    }
    if (lane == 0) s_ctl[1] = go ? 1 : 0;
  };
  SEG_INIT();
  if (wave == 0) pivot_step();
  SEG(0);
  for (;;) {
    __syncthreads();
    SEG(1);
    if (!s_ctl[1]) break;
    const int p = s_ctl[0];
    const double2 tau = s_tau;
    const double2* pv = As + (long)p * m;   // reflector, v_j = 1 implicit
    const bool reflect = (tau.x != 0.0 || tau.y != 0.0);
    // the group of column myc applies H_j^H and downdates the column norm;
    // only that group touches the column, its norms and its done flag
    if (myc < n && !done[myc]) {
      if (myc == p) {
        if (g == 0) done[myc] = 1;
      } else {
        double2* col = As + (long)myc * m;
        double2 rj2 = col[j];
        if (reflect) {
          double2 w = make_double2(0.0, 0.0);
#pragma unroll 4
          for (int i = j + g; i < m; i += G) {
            const double2 vi = (i == j) ? make_double2(1.0, 0.0) : pv[i];
            cfmac(w, vi, col[i]);
          }
          {
            double z0 = 0.0, z1 = 0.0;
            group_sum4(w.x, w.y, z0, z1, G);
          }
          const double2 tw = cmul(make_double2(tau.x, -tau.y), w);
#pragma unroll 4
          for (int i = j + g; i < m; i += G) {
            const double2 vi = (i == j) ? make_double2(1.0, 0.0) : pv[i];
            col[i] = csub(col[i], cmul(vi, tw));
          }
          // row j of the column after the reflection (held by lane g = 0)
          rj2 = csub(rj2, tw);
        }
        double nv = vn1[myc];
        if (nv != 0.0) {
          const double rj = sqrt(cabs2(rj2));
          double temp = 1.0 - (rj / nv) * (rj / nv);
          temp = temp > 0.0 ? temp : 0.0;
          const double ratio = nv / vn2[myc];
          if (temp * ratio * ratio <= tol3z) {
            double pp = 0.0;
            for (int i = j + 1 + g; i < m; i += G) pp += cabs2(col[i]);
            nv = (j + 1 < m) ? sqrt(group_sum(pp, G)) : 0.0;
            if (g == 0) { vn1[myc] = nv; vn2[myc] = nv; }
          } else if (g == 0) {
            vn1[myc] = nv * sqrt(temp);
          }
        }
      }
    }
    SEG(2);
    __syncthreads();
    ++j;
    if (wave == 0) pivot_step();
    SEG(3);
  }
  SEG_FLUSH(8, 4);
  const int r = j;
  PHASE(13);
  if (tid == 0) {
    int t = r;
    for (int c = 0; c < n; ++c)
      if (!done[c]) q_piv[t++] = c;
    *a.qrank = r;
  }
  __syncthreads();
  for (int t = tid; t < n; t += NT) {
    a.piv[t] = q_piv[t];
    if (t < r) {
      a.tau[t] = q_tau[t];
      a.vscale[t] = q_vscale[t];
      a.beta[t] = q_beta[t];
    }
  }
  // reflectors and R back to global memory (apply_q_kernel reads them)
  for (long e = tid; e < (long)m * n; e += NT) a.A[e] = As[e];
  __syncthreads();
  if (r == 0) {
    if (tid == 0) {
      SvdInfo inf;
      inf.rank = 0; inf.overflow = 0; inf.sweeps = 1; inf.pad = 0;
      inf.s0 = 0.0; inf.discarded = 0.0;
      *a.info = inf;
      if (a.copy) *a.copy = inf;
    }
    return;
  }
  // ---- Y = R^H (n x r, ld n) in LDS; W = I (r x r) in LDS or a.W ----
  // Y lies inside the space of M (r <= m), below the scalars it reads; W
  // (when it fits) follows Y and may cover the scalars, so it is set after
  // a barrier.
  double2* Ys = lds_small;
  const bool wl = (long)(n + r) * r * (long)sizeof(double2) <= a.lds_bytes;
  double2* Wp = wl ? Ys + (long)n * r : a.W;
  for (long e = tid; e < (long)n * r; e += NT) {
    const int sidx = e % n, t = e / n;
    const int ps = q_piv[sidx];
    double2 v = make_double2(0.0, 0.0);
    if (t < sidx) v = cconj(a.A[t + (long)ps * m]);   // As is being overwritten
    else if (t == sidx) v = make_double2(q_beta[sidx], 0.0);
    Ys[e] = v;
  }
  __syncthreads();
  for (long e = tid; e < (long)r * r; e += NT)
    Wp[e] = make_double2((e % r) == (e / r) ? 1.0 : 0.0, 0.0);
  __syncthreads();
  double fpart = 0.0;
  for (long e = tid; e < (long)n * r; e += NT) fpart += cabs2(Ys[e]);
  const double fro = sqrt(block_sum_nt<NT>(fpart, slot, parity));
  const double tol = (n > 16 ? n : 16) * eps, abs_tol = eps * fro;
  // ---- one-sided Jacobi on the columns of Y ----
  __shared__ int last_rot[kSmallMaxCols];
  for (int c = tid; c < r; c += NT) last_rot[c] = -(1 << 30);
  __syncthreads();
  bool converged = false;
  const int sweeps = wl ? small_jacobi<NT, true>(Ys, a.W, n, r, tol * tol, abs_tol * abs_tol,
                                                 a.max_sweeps, last_rot, converged)
                        : small_jacobi<NT, false>(Ys, a.W, n, r, tol * tol, abs_tol * abs_tol,
                                                  a.max_sweeps, last_rot, converged);
  PHASE(14);
  // ---- singular values, order, cut; Z and W to global memory ----
  for (long e = tid; e < (long)n * r; e += NT) a.Y[e] = Ys[e];
  if (wl)
    for (long e = tid; e < (long)r * r; e += NT) a.W[e] = Wp[e];
  __syncthreads();
  // two r-vectors after Y (over W, which is in global memory now)
  double* s_sig = reinterpret_cast<double*>(Ys + (long)n * r);
  double* s_sorted = s_sig + r;
  for (int c = tid / 64; c < r; c += NT / 64) {
    double part2 = 0.0;
    for (int i = lane; i < n; i += 64) part2 += cabs2(Ys[i + (long)c * n]);
    const double sg = sqrt(wave_sum(part2));
    if (lane == 0) { s_sig[c] = sg; a.sigma[c] = sg; }
  }
  __syncthreads();
  if (tid < r) {
    const double sj = s_sig[tid];
    int rank = 0;
    for (int i = 0; i < r; ++i) {
      const double si = s_sig[i];
      rank += (si > sj) || (si == sj && i < tid);
    }
    a.perm[rank] = tid;
    a.sorted[rank] = sj;
    s_sorted[rank] = sj;
  }
  __syncthreads();
  if (tid == 0) {
    // smallest k with sqrt(sum_{i>=k} s_i^2) <= epsrel s_0 (suffix sums)
    const double s0 = s_sorted[0];
    const double thr = a.epsrel * s0;
    double tail = folded2;
    int k = 0;
    for (int i = r - 1; i >= 0; --i) {
      tail += s_sorted[i] * s_sorted[i];
      if (sqrt(tail) > thr) { k = i + 1; break; }
    }
    double d = folded2;
    for (int i = r - 1; i >= k; --i) d += s_sorted[i] * s_sorted[i];
    SvdInfo inf;
    inf.rank = k;
    inf.overflow = k > a.max_rank ? 1 : 0;
    inf.sweeps = converged ? sweeps + 1 : -1;
    inf.pad = r;
    inf.s0 = s0;
    inf.discarded = sqrt(d);
    *a.info = inf;
    if (a.copy) *a.copy = inf;
  }
  PHASE(15);
}

__global__ void __launch_bounds__(kSmallThreads)
small_svd_kernel(SmallSvdArgs a) {
  small_svd_body<kSmallThreads>(a);
}

__global__ void __launch_bounds__(kSmallBatchThreads)
small_svd_batch_kernel(const SmallSvdArgs* probs) {
  const SmallSvdArgs a = probs[blockIdx.x];
  small_svd_body<kSmallBatchThreads>(a);
}

// Left singular vectors Q W[:, perm[j]] (j < k) of the pivoted QR, the
// reflectors applied from the last to the first to a column held in LDS;
// row i = a*B + b goes to dst[j*dj + a*da + b*db].
// One column per workgroup, the reflector entries of each step loaded into
// registers in one batch (rows t + tid + 256 u, u < kApplyQRegs; rows beyond
// on demand) and those of the next step issued before this step's
// reduction, so that their latency overlaps the work.
constexpr int kApplyQRegs = 8;

__global__ void __launch_bounds__(256)
apply_q_kernel(ApplyQArgs a) {
  extern __shared__ double2 cvec[];
  __shared__ double slot[2][4], sloti[2][4];
  int parity = 0;
  const int tid = threadIdx.x;
  const int k = a.info->rank;
  const int r = *a.qrank;
  const int m = a.m;
  const long dj = a.dj0 + a.djk * (long)k;
  const long da = a.da0 + a.dak * (long)k;
  const long db = a.db0 + a.dbk * (long)k;
  for (int j = blockIdx.x; j < k; j += gridDim.x) {
    const double2* w = a.W + (long)a.perm[j] * a.ldw;
    for (int i = tid; i < m; i += blockDim.x)
      cvec[i] = (i < r) ? w[i] : make_double2(0.0, 0.0);
    __syncthreads();
    // reflector t's entries (unscaled) for rows t + tid + 256 u
    double2 vn[kApplyQRegs];
    auto fetch = [&](int t) {
      const double2* v = a.A + (long)a.piv[t] * a.lda;
#pragma unroll
      for (int u = 0; u < kApplyQRegs; ++u) {
        const int i = t + tid + 256 * u;
        vn[u] = (i < m) ? v[i] : make_double2(0.0, 0.0);
      }
    };
    if (r > 0) fetch(r - 1);
    for (int t = r - 1; t >= 0; --t) {
      double2 vc[kApplyQRegs];
#pragma unroll
      for (int u = 0; u < kApplyQRegs; ++u) vc[u] = vn[u];
      const double2 tau = a.tau[t];
      const double2 sc = a.vscale[t];
      const double2* v = a.A + (long)a.piv[t] * a.lda;
      if (t > 0) fetch(t - 1);
      if (tau.x == 0.0 && tau.y == 0.0) continue;
      double2 acc = make_double2(0.0, 0.0);
#pragma unroll
      for (int u = 0; u < kApplyQRegs; ++u) {
        const int i = t + tid + 256 * u;
        if (i < m) {
          vc[u] = (i == t) ? make_double2(1.0, 0.0) : cmul(vc[u], sc);
          cfmac(acc, vc[u], cvec[i]);
        }
      }
      for (int i = t + tid + 256 * kApplyQRegs; i < m; i += blockDim.x) {
        const double2 vi = cmul(v[i], sc);
        cfmac(acc, vi, cvec[i]);
      }
      const double2 tw = cmul(tau, block_sum256c(acc, slot, sloti, parity));
#pragma unroll
      for (int u = 0; u < kApplyQRegs; ++u) {
        const int i = t + tid + 256 * u;
        if (i < m) cvec[i] = csub(cvec[i], cmul(vc[u], tw));
      }
      for (int i = t + tid + 256 * kApplyQRegs; i < m; i += blockDim.x) {
        const double2 vi = cmul(v[i], sc);
        cvec[i] = csub(cvec[i], cmul(vi, tw));
      }
      __syncthreads();
    }
    for (int i = tid; i < m; i += blockDim.x) {
      const int aa = i / a.B, b = i % a.B;
      a.dst[(long)j * dj + (long)aa * da + (long)b * db] = cvec[i];
    }
    __syncthreads();
  }
}

// The same for Q of at most 128 rows: one wavefront per column, rows lane and
// lane + 64 in registers, wavefront reductions only (no barrier, no LDS);
// reflector t - 1 is loaded while step t is reduced.
__global__ void __launch_bounds__(256)
apply_q_wave_kernel(ApplyQArgs a) {
  const int lane = threadIdx.x & 63;
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int k = a.info->rank;
  if (j >= k) return;   // a whole wavefront; the kernel has no barrier
  const int r = *a.qrank;
  const int m = a.m;
  const double2 zero = make_double2(0.0, 0.0);
  const int i0 = lane, i1 = lane + 64;
  const double2* w = a.W + (long)a.perm[j] * a.ldw;
  double2 x0 = i0 < r ? w[i0] : zero;
  double2 x1 = i1 < r ? w[i1] : zero;
  double2 n0 = zero, n1 = zero;
  auto fetch = [&](int t) {
    const double2* v = a.A + (long)a.piv[t] * a.lda;
    n0 = i0 < m ? v[i0] : zero;
    n1 = i1 < m ? v[i1] : zero;
  };
  if (r > 0) fetch(r - 1);
  for (int t = r - 1; t >= 0; --t) {
    const double2 c0 = n0, c1 = n1;
    const double2 tau = a.tau[t];
    const double2 sc = a.vscale[t];
    if (t > 0) fetch(t - 1);
    if (tau.x == 0.0 && tau.y == 0.0) continue;
    const double2 one = make_double2(1.0, 0.0);
    const double2 v0 = i0 == t ? one : (i0 > t ? cmul(c0, sc) : zero);
    const double2 v1 = i1 == t ? one : (i1 > t && i1 < m ? cmul(c1, sc) : zero);
    double2 acc = zero;
    cfmac(acc, v0, x0);
    cfmac(acc, v1, x1);
    double z0 = 0.0, z1 = 0.0;
    group_sum4(acc.x, acc.y, z0, z1, 64);
    const double2 tw = cmul(tau, acc);
    x0 = csub(x0, cmul(v0, tw));
    x1 = csub(x1, cmul(v1, tw));
  }
  const long dj = a.dj0 + a.djk * (long)k;
  const long da = a.da0 + a.dak * (long)k;
  const long db = a.db0 + a.dbk * (long)k;
  if (i0 < m) a.dst[(long)j * dj + (long)(i0 / a.B) * da + (long)(i0 % a.B) * db] = x0;
  if (i1 < m) a.dst[(long)j * dj + (long)(i1 / a.B) * da + (long)(i1 % a.B) * db] = x1;
}

// X (rows x cols, column-major, ld = rows) = conj(src) with src row-major
// (cols x rows): X[r + c*rows] = conj(src[c*rows + r]) -- a plain conjugate
__global__ void __launch_bounds__(256)
conj_copy_kernel(const double2* src, long count, double2* dst) {
  for (long e = blockIdx.x * 256L + threadIdx.x; e < count; e += (long)gridDim.x * 256)
    dst[e] = cconj(src[e]);
}

// X[r + c*R] = conj(src[r*C + c])  (src row-major R x C)
__global__ void __launch_bounds__(256)
conj_transpose_kernel(const double2* src, int R, int C, double2* dst) {
  const long total = (long)R * C;
  for (long e = blockIdx.x * 256L + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    const int r = e % R;
    const long c = e / R;
    dst[e] = cconj(src[(long)r * C + c]);
  }
}

// X[r + c*R] = src[r*C + c]: row-major R x C to column-major
__global__ void __launch_bounds__(256)
transpose_kernel(const double2* src, int R, int C, double2* dst) {
  const long total = (long)R * C;
  for (long e = blockIdx.x * 256L + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    const int r = e % R;
    const long c = e / R;
    dst[e] = src[(long)r * C + c];
  }
}

// out[i] = sum_j in[j*stride_j ... ] : contract one leg of a tensor [a][p][b]
// with a vector: out[a][b] = sum_p T[a][p][b] * vec[p]
__global__ void __launch_bounds__(256)
leg_vector_kernel(const double2* T, int A, int P, int B, const double2* vec,
                  double2* out) {
  const long total = (long)A * B;
  for (long e = blockIdx.x * 256L + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    const int b = e % B;
    const long aa = e / B;
    double2 acc = make_double2(0.0, 0.0);
    for (int p = 0; p < P; ++p) cfma(acc, T[(aa * P + p) * B + b], vec[p]);
    out[e] = acc;
  }
}

__global__ void __launch_bounds__(256)
identity_carry_kernel(double2* dst, int R, int V) {
  // dst[(r*X + x)*V + v] = delta(x, r*V + v), X = R*V
  const int X = R * V;
  const long total = (long)R * X * V;
  for (long e = blockIdx.x * 256L + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    const int v = e % V;
    const long t = e / V;
    const int x = t % X;
    const int r = t / X;
    dst[e] = make_double2(x == r * V + v ? 1.0 : 0.0, 0.0);
  }
}

}  // namespace pte

#ifdef PTE_ROUND_TRACE
// per-round timestamp readout (n entries at most); clears the buffer
extern "C" int pte_debug_round_trace(unsigned long long* out, long n) {
  const long total = (long)pte::kRtRounds * pte::kRtBlocks * pte::kRtMarks;
  if (n > total) n = total;
  if (hipDeviceSynchronize() != hipSuccess) return -3;
  if (hipMemcpyFromSymbol(out, HIP_SYMBOL(pte::g_rtrace), sizeof(unsigned long long) * n) !=
      hipSuccess)
    return -3;
  void* dev = nullptr;
  if (hipGetSymbolAddress(&dev, HIP_SYMBOL(pte::g_rtrace)) != hipSuccess) return -3;
  if (hipMemset(dev, 0, sizeof(unsigned long long) * total) != hipSuccess) return -3;
  if (hipDeviceSynchronize() != hipSuccess) return -3;
  return 0;
}
#endif

#ifdef PTE_PHASE_TIMING
// phase timer readout (s_memrealtime ticks); resets the timers
extern "C" int pte_debug_phases(unsigned long long* out) {
  if (hipDeviceSynchronize() != hipSuccess) return -3;
  if (hipMemcpyFromSymbol(out, HIP_SYMBOL(pte::g_phase), sizeof(unsigned long long) * 16) !=
      hipSuccess)
    return -3;
  unsigned long long z[16] = {0};
  if (hipMemcpyToSymbol(HIP_SYMBOL(pte::g_phase), z, sizeof(z)) != hipSuccess) return -3;
  return 0;
}
#endif
