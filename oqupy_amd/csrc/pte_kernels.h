// Argument blocks and kernel declarations of the process-tensor engine.
#pragma once

#include <hip/hip_runtime.h>

namespace pte {

enum NodeKind { NODE_DENSE = 0, NODE_DELTA = 1 };
enum MpoOut { MPO_OUT_SVD_H = 0, MPO_OUT_SITE = 1, MPO_OUT_M = 2 };

struct GemmArgs {
  int M, N, K;
  const double2* A; long sai, sak;
  const double2* B; long sbk, sbj;
  double2* C; long sci, scj;
  int conjA, conjB;
  double2 alpha;
  int beta_zero;
};

struct MpoApplyArgs {
  const double2* G;   // GEMM result
  const double2* W;   // dense node [w][p][s][v]
  const double2* I;   // delta node weights [w][p]
  int kind, mode;
  int Dl, P, Dx, V, Wl, S;
  double2* out;
  long ldx;
};

// Summary of one truncated SVD, written by the device.
struct SvdInfo {
  int rank;        // kept singular values
  int overflow;    // rank above the bond capacity
  int sweeps;      // Jacobi sweeps (or -1 if not converged)
  int pad;
  double s0;       // largest singular value
  double discarded;  // 2-norm of the discarded singular values
};

struct JacobiArgs {
  double2* X; int m, n; long ldx;
  double2* V; long ldv;
  int nb;
  int slices;         // workgroups per block pair (row slices)
  int xrows_per;      // rows of X per slice (multiple of 32)
  int vrows_per;      // rows of V per slice (multiple of 32)
  double2* gpart;     // partial Grams [pair][slice][32*32] when slices > 1
  double tol, abs_tol_scale;
  int max_sweeps;
  int inner_sweeps;   // cyclic sweeps of the pair's 2JB x 2JB eigenproblem
  unsigned* bar;
  int* flags;
  double* fro2;
  int* info;
  // skipping of converged pairs (zeroed before the launch): skip[b], b < nb,
  // is 1 + the round in which block b was last rotated (0: never);
  // skip[nb + pair_index(bi, bj)] is 1 + the round in which that block pair
  // was last found to need no rotation (0: not since its last rotation)
  unsigned* skip;
};

struct EmitArgs {
  const double2* src; long lds;
  const int* perm;
  const SvdInfo* info;
  int A, B;
  long dj0, djk, da0, dak, db0, dbk;
  int conj;
  const int* rowmap;   // destination row of source row s (null: identity)
  double2* dst;
};

// Pivoted QR of a column-major matrix (see qrcp_kernel).
struct QrcpCand {
  double vmax;   // largest residual column norm of a workgroup
  double vsum;   // sum of squared residual norms of a workgroup
  int col;       // column of vmax (-1: none)
  int pad;
};

struct QrcpArgs {
  double2* A; int m, n; long lda;
  double drop;                // stop when the residual norm <= drop * ||A||_F
  double2* tau; double2* vscale; double* beta;
  int* piv;                   // [n] physical column of each step, then the rest
  int* done;                  // [n] scratch
  QrcpCand* cand;             // [2 * gridDim.x] scratch
  int* rank;                  // steps taken
  unsigned* bar;
  const int* max_steps;       // optional cap on the steps (device int)
  double* out;                // optional: [0] squared residual norm at the stop,
                              // [1] ||A||_F^2 (first run), [2] |beta_0|, [3] the
                              // first step count whose residual norm is at most
                              // mark_rel * |beta_0| (the steps taken if none)
  double mark_rel;            // 0: no mark
  int stop_at_mark;           // stop at the mark (adaptive cap)
  int start;                  // > 0: resume a factorisation stopped after `start`
                              // steps (done, piv, tau, vscale, beta kept)
  double drop_abs2;           // stop threshold on the squared residual when resuming
};

struct ApplyQArgs {
  const double2* A; int m; long lda;      // reflectors from qrcp_kernel
  const double2* tau; const double2* vscale; const int* piv; const int* qrank;
  const double2* W; long ldw;             // Jacobi unitary (qrank x qrank)
  const int* perm;
  const SvdInfo* info;
  int B;                                  // row i = a*B + b
  long dj0, djk, da0, dak, db0, dbk;      // strides d0 + dk * rank
  double2* dst;
};

constexpr int kQrcpMaxLocal = 64;         // columns per workgroup in qrcp_kernel

// One SVD's singular values, order and cut (batched post-processing).
struct SvdPostArgs {
  const double2* X; int m, n;               // orthogonalised X (ld = m)
  double* sigma; int* perm; double* sorted;
  SvdInfo* info;
  double epsrel; int max_rank; int pad;
};

// Y = R^H of one pivoted QR (batched).
struct RhArgs {
  const double2* A; long lda; int n; int pad;
  const int* piv; const double* beta; const int* rank;
  double2* Y;
};

// One small truncated SVD done by one workgroup (small_svd_kernel): the same
// outputs as the pivoted QR + Jacobi path (reflectors in A, tau, vscale,
// beta, piv, *qrank; Z = Y W in Y (n x r, ld n); W (r x r, ld r); sigma,
// perm, sorted; info with info->pad = r).
struct SmallSvdArgs {
  double2* A; int m, n;
  double2* tau; double2* vscale; double* beta; int* piv; int* qrank;
  double2* Y; double2* W;
  double* sigma; int* perm; double* sorted; SvdInfo* info;
  double epsrel; int max_rank; int max_sweeps;
  SvdInfo* copy;    // optional second copy of the summary (batched gather)
  double stop_rel;  // > 0: stop the QR at residual <= stop_rel * |beta_0|
  long lds_bytes;   // dynamic LDS of the launch (W stays in LDS when it fits)
};

constexpr int kSmallMaxCols = 96;           // columns of M in small_svd_kernel
constexpr int kSmallThreads = 1024;         // workgroup of small_svd_kernel
constexpr int kSmallBatchThreads = 512;     // workgroup of small_svd_batch_kernel

struct ZeroItem {
  unsigned* p; int words; int pad;
};

__global__ void qrcp_kernel(QrcpArgs a);
__global__ void qrcp_batch_kernel(const QrcpArgs* probs, const int* starts,
                                  const int* block_prob);
__global__ void build_rh_kernel(const double2* A, long lda, int n, const int* piv,
                                const double* beta, const int* rank, double2* Y);
__global__ void build_rh_batch_kernel(const RhArgs* pa);
__global__ void column_norm_batch_kernel(const SvdPostArgs* pa);
__global__ void rank_sort_batch_kernel(const SvdPostArgs* pa);
__global__ void truncate_batch_kernel(const SvdPostArgs* pa, SvdInfo* out);
__global__ void zero_items_kernel(const ZeroItem* items);
__global__ void small_svd_kernel(SmallSvdArgs a);
__global__ void small_svd_batch_kernel(const SmallSvdArgs* probs);
__global__ void gather_qr_rank_kernel(const QrcpArgs* probs, int n, int* out);
__global__ void apply_q_kernel(ApplyQArgs a);
__global__ void apply_q_wave_kernel(ApplyQArgs a);   // Q of at most kApplyQWaveRows rows
constexpr int kApplyQWaveRows = 128;
__global__ void zgemm_kernel(GemmArgs a);
__global__ void zgemm_mfma_kernel(GemmArgs a);
__global__ void cgemm_mfma_kernel(GemmArgs a);
__global__ void mpo_apply_kernel(MpoApplyArgs a);
__global__ void mpo_apply_right_kernel(MpoApplyArgs a);
__global__ void jacobi_kernel(JacobiArgs a);
// the same with the slice held in an LDS tile of xrows_per + vrows_per rows
// (dynamic LDS: rows * kJacobiTileLd complex entries; at most kTileMaxRows rows)
__global__ void jacobi_tile_kernel(JacobiArgs a);
__global__ void jacobi_tile_batch_kernel(const JacobiArgs* probs, const int* starts,
                                         const int* block_prob);
__global__ void jacobi_batch_kernel(const JacobiArgs* probs, const int* starts,
                                    const int* block_prob);
__global__ void column_norm_kernel(const double2* X, int m, int n, long ldx, double* sigma);
__global__ void rank_sort_kernel(const double* sigma, int n, int* perm, double* sorted);
__global__ void truncate_kernel(const double* sorted, int n, double epsrel, SvdInfo* info,
                                int max_rank, const double* extra2);
__global__ void emit_kernel(EmitArgs a);
__global__ void conj_copy_kernel(const double2* src, long count, double2* dst);
__global__ void conj_transpose_kernel(const double2* src, int R, int C, double2* dst);
__global__ void transpose_kernel(const double2* src, int R, int C, double2* dst);
__global__ void leg_vector_kernel(const double2* T, int A, int P, int B, const double2* vec,
                                  double2* out);
__global__ void identity_carry_kernel(double2* dst, int R, int V);

constexpr int kJacobiBlock = 16;
constexpr int kJacobiTileLd = 2 * kJacobiBlock + 1;
constexpr int kTileMaxRows = 192;

}  // namespace pte
