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
// T[x,s,l,w] = sum_{p,v} W[w,p,s,v] * G[(l,p),(x,v)]
//   mode MPO_OUT_SVD_H : X[(l*Wl+w) + (x*S+s)*ldx] = conj(T)   (X = M^H)
//   mode MPO_OUT_SITE  : out[(l*S+s)*Dx + x] = T   (requires Wl == 1)
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
    } else {
      a.out[((long)l * a.S + s) * a.Dx + x] = acc;
    }
  }
}

// ---------------------------------------------------------------------------
// TEMPO zip (left to right):
// T[x,s,r,e] = sum_{p,w} W[w,p,s,e] * G[(x,w),(p,r)]
//   mode MPO_OUT_SVD_H : X[(r*E+e) + (x*S+s)*ldx] = conj(T)
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
    } else {
      a.out[((long)x * a.S + s) * E + e] = acc;
    }
  }
}

// ---------------------------------------------------------------------------
// Blocked one-sided Jacobi SVD.
// X (m x n, column-major, leading dimension ldx) is orthogonalised column-wise:
// X <- X Q with Q unitary accumulated in V (n x n, column-major). Columns are
// grouped into blocks of JB; in every round each workgroup takes one block pair
// (round-robin tournament over nb blocks), forms the 2JB x 2JB Gram matrix,
// diagonalises it with cyclic two-sided Jacobi in LDS and applies the unitary
// to the pair's columns of X and V. A sweep is nb-1 rounds; the iteration stops
// after a sweep in which no pair needed a rotation. A pair needs one when some
// |g_ij| > tol*sqrt(g_ii g_jj) and |g_ij| > abs_tol*sqrt(max(g_ii, g_jj)).
// ---------------------------------------------------------------------------
constexpr int JB = 16;          // block width
constexpr int JP = 2 * JB;      // columns per block pair
constexpr int JR = 32;          // rows per streamed chunk
constexpr int JT = 256;         // threads per workgroup
constexpr int INNER_SWEEPS = 4;

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

__global__ void __launch_bounds__(JT)
jacobi_kernel(JacobiArgs a) {
  __shared__ double2 Xs[JR][JP + 1];
  __shared__ double2 Gs[JP][JP + 1];
  __shared__ double2 Ws[JP][JP + 1];
  __shared__ double rot[JB][4];     // cs, sn, phase_re, phase_im
  __shared__ int rotflag[JB];
  __shared__ int cols[JP];
  __shared__ int need;
  __shared__ double red[JT];

  const int tid = threadIdx.x;
  const unsigned nblk = gridDim.x;
  unsigned epoch = 0;

  // V = identity, Frobenius norm of X
  for (long e = blockIdx.x * (long)JT + tid; e < (long)a.n * a.n;
       e += (long)nblk * JT) {
    const long i = e % a.n, j = e / a.n;
    a.V[i + j * a.ldv] = make_double2(i == j ? 1.0 : 0.0, 0.0);
  }
  double part = 0.0;
  for (long e = blockIdx.x * (long)JT + tid; e < (long)a.m * a.n;
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
  grid_barrier(a.bar, nblk, epoch);
  const double fro = sqrt(__hip_atomic_load(a.fro2, __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_AGENT));
  const double abs_tol = a.abs_tol_scale * fro;
  const double tol = a.tol;

  int sweep = 0;
  bool converged = false;
  for (; sweep < a.max_sweeps; ++sweep) {
    for (int round = 0; round < a.nb - 1; ++round) {
      int bi, bj;
      tournament_pair(a.nb, round, blockIdx.x, bi, bj);
      if (tid < JP) {
        const int c = (tid < JB) ? bi * JB + tid : bj * JB + (tid - JB);
        cols[tid] = (c < a.n) ? c : -1;
      }
      __syncthreads();
      // ---- Gram matrix of the pair's columns ----
      double2 g[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) g[q] = make_double2(0.0, 0.0);
      for (int r0 = 0; r0 < a.m; r0 += JR) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int e = tid + JT * q;      // JR*JP = 1024 entries
          const int rr = e % JR, cc = e / JR;
          const int gr = r0 + rr, gc = cols[cc];
          Xs[rr][cc] = (gr < a.m && gc >= 0) ? a.X[gr + (long)gc * a.ldx]
                                             : make_double2(0.0, 0.0);
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int e = tid + JT * q;      // JP*JP = 1024 Gram entries
          const int gi = e / JP, gj = e % JP;
#pragma unroll 8
          for (int rr = 0; rr < JR; ++rr) cfmac(g[q], Xs[rr][gi], Xs[rr][gj]);
        }
        __syncthreads();
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int e = tid + JT * q;
        Gs[e / JP][e % JP] = g[q];
        Ws[e / JP][e % JP] = make_double2((e / JP) == (e % JP) ? 1.0 : 0.0, 0.0);
      }
      if (tid == 0) need = 0;
      __syncthreads();
      // ---- does the pair need a rotation? ----
      {
        int mine = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int e = tid + JT * q;
          const int gi = e / JP, gj = e % JP;
          if (gi < gj && cols[gi] >= 0 && cols[gj] >= 0) {
            const double dii = Gs[gi][gi].x, djj = Gs[gj][gj].x;
            const double mag = sqrt(cabs2(Gs[gi][gj]));
            if (mag > tol * sqrt(fabs(dii * djj)) &&
                mag > abs_tol * sqrt(fmax(fmax(dii, djj), 0.0)))
              mine = 1;
          }
        }
        if (mine) need = 1;
      }
      __syncthreads();
      if (need) {
        // ---- cyclic two-sided Jacobi on Gs, accumulating Ws ----
        for (int isw = 0; isw < INNER_SWEEPS; ++isw) {
          int any = 0;
          for (int sub = 0; sub < JP - 1; ++sub) {
            if (tid < JB) {
              int p, q;
              tournament_pair(JP, sub, tid, p, q);
              if (p > q) { const int t = p; p = q; q = t; }
              double cs = 1.0, sn = 0.0, phr = 1.0, phi = 0.0;
              int flag = 0;
              if (cols[p] >= 0 && cols[q] >= 0) {
                const double app = Gs[p][p].x, aqq = Gs[q][q].x;
                const double2 apq = Gs[p][q];
                const double mag = sqrt(cabs2(apq));
                if (mag > tol * sqrt(fabs(app * aqq)) &&
                    mag > abs_tol * sqrt(fmax(fmax(app, aqq), 0.0)) && mag > 0.0) {
                  phr = apq.x / mag;
                  phi = apq.y / mag;
                  const double zeta = (aqq - app) / (2.0 * mag);
                  const double t = (zeta == 0.0)
                      ? 1.0
                      : copysign(1.0, zeta) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                  cs = 1.0 / sqrt(1.0 + t * t);
                  sn = cs * t;
                  flag = 1;
                }
              }
              rot[tid][0] = cs; rot[tid][1] = sn; rot[tid][2] = phr; rot[tid][3] = phi;
              rotflag[tid] = flag;
            }
            __syncthreads();
            // J = [[cs, sn], [-sn*conj(ph), cs*conj(ph)]]; columns: (Gp,Gq) <- (Gp,Gq) J
            for (int task = tid; task < JB * JP * 2; task += JT) {
              const int which = task / (JB * JP);        // 0: G, 1: W
              const int rem = task % (JB * JP);
              const int pr = rem / JP, row = rem % JP;
              if (!rotflag[pr]) continue;
              int p, q;
              tournament_pair(JP, sub, pr, p, q);
              if (p > q) { const int t = p; p = q; q = t; }
              const double cs = rot[pr][0], sn = rot[pr][1];
              const double2 cph = make_double2(rot[pr][2], -rot[pr][3]);  // conj(ph)
              double2 (*M)[JP + 1] = which ? Ws : Gs;
              const double2 xp = M[row][p], xq = M[row][q];
              const double2 xqc = cmul(xq, cph);
              M[row][p] = make_double2(cs * xp.x - sn * xqc.x, cs * xp.y - sn * xqc.y);
              M[row][q] = make_double2(sn * xp.x + cs * xqc.x, sn * xp.y + cs * xqc.y);
            }
            __syncthreads();
            // rows: (Gp;Gq) <- J^H (Gp;Gq)
            for (int task = tid; task < JB * JP; task += JT) {
              const int pr = task / JP, col = task % JP;
              if (!rotflag[pr]) continue;
              int p, q;
              tournament_pair(JP, sub, pr, p, q);
              if (p > q) { const int t = p; p = q; q = t; }
              const double cs = rot[pr][0], sn = rot[pr][1];
              const double2 ph = make_double2(rot[pr][2], rot[pr][3]);
              const double2 yp = Gs[p][col], yq = Gs[q][col];
              const double2 yqp = cmul(ph, yq);   // conj(conj(ph)) * yq
              Gs[p][col] = make_double2(cs * yp.x - sn * yqp.x, cs * yp.y - sn * yqp.y);
              Gs[q][col] = make_double2(sn * yp.x + cs * yqp.x, sn * yp.y + cs * yqp.y);
            }
            __syncthreads();
            if (tid < JB && rotflag[tid]) {
              int p, q;
              tournament_pair(JP, sub, tid, p, q);
              Gs[p][q] = make_double2(0.0, 0.0);
              Gs[q][p] = make_double2(0.0, 0.0);
              Gs[p][p].y = 0.0;
              Gs[q][q].y = 0.0;
              any = 1;
            }
            __syncthreads();
          }
          if (tid < JB) rotflag[tid] = any;   // reuse as a sweep flag
          __syncthreads();
          int anyrot = 0;
          for (int i = 0; i < JB; ++i) anyrot |= rotflag[i];
          __syncthreads();
          if (!anyrot) break;
        }
        // ---- X_pair <- X_pair Ws, V_pair <- V_pair Ws ----
        for (int mat = 0; mat < 2; ++mat) {
          double2* base = mat ? a.V : a.X;
          const long ld = mat ? a.ldv : a.ldx;
          const int rows = mat ? a.n : a.m;
          for (int r0 = 0; r0 < rows; r0 += JR) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const int e = tid + JT * q;
              const int rr = e % JR, cc = e / JR;
              const int gr = r0 + rr, gc = cols[cc];
              Xs[rr][cc] = (gr < rows && gc >= 0) ? base[gr + (long)gc * ld]
                                                  : make_double2(0.0, 0.0);
            }
            __syncthreads();
            double2 outv[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const int e = tid + JT * q;
              const int rr = e % JR, cc = e / JR;
              double2 acc = make_double2(0.0, 0.0);
#pragma unroll 8
              for (int i = 0; i < JP; ++i) cfma(acc, Xs[rr][i], Ws[i][cc]);
              outv[q] = acc;
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const int e = tid + JT * q;
              const int rr = e % JR, cc = e / JR;
              const int gr = r0 + rr, gc = cols[cc];
              if (gr < rows && gc >= 0) base[gr + (long)gc * ld] = outv[q];
            }
          }
        }
        if (tid == 0) a.flags[sweep] = 1;
      }
      grid_barrier(a.bar, nblk, epoch);
    }
    if (__hip_atomic_load(&a.flags[sweep], __ATOMIC_RELAXED,
                          __HIP_MEMORY_SCOPE_AGENT) == 0) {
      converged = true;
      break;
    }
  }
  if (blockIdx.x == 0 && tid == 0) *a.info = converged ? sweep + 1 : -1;
}

// sigma_j = || X[:, j] ||  (one workgroup per column)
__global__ void __launch_bounds__(256)
column_norm_kernel(const double2* X, int m, int n, long ldx, double* sigma) {
  __shared__ double red[256];
  const int j = blockIdx.x;
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

// descending order: rank_j = #{i : s_i > s_j} + #{i < j : s_i == s_j}
__global__ void __launch_bounds__(256)
rank_sort_kernel(const double* sigma, int n, int* perm, double* sorted) {
  const int j = blockIdx.x * 256 + threadIdx.x;
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

// k = #{ j : sqrt(sum_{i>=j} s_i^2) > epsrel * s_0 }  (single workgroup)
__global__ void __launch_bounds__(1024)
truncate_kernel(const double* sorted, int n, double epsrel, SvdInfo* info,
                int max_rank) {
  __shared__ double part[1024];
  __shared__ int cnt[1024];
  const int tid = threadIdx.x;
  const int per = (n + 1023) / 1024;
  const int lo = min(n, tid * per), hi = min(n, lo + per);
  double s = 0.0;
  for (int i = lo; i < hi; ++i) s += sorted[i] * sorted[i];
  part[tid] = s;
  __syncthreads();
  // suffix sums of the per-thread partials (serial over 1024 values by thread 0)
  if (tid == 0) {
    double acc = 0.0;
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
    double d = 0.0;
    for (int i = n - 1; i >= k; --i) d += sorted[i] * sorted[i];
    info->discarded = sqrt(d);
  }
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
  const long total = (long)a.A * a.B * k;
  for (long e = blockIdx.x * 256L + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    const int b = e % a.B;
    const long t = e / a.B;
    const int aa = t % a.A;
    const int j = t / a.A;
    double2 v = a.src[((long)aa * a.B + b) + (long)a.perm[j] * a.lds];
    if (a.conj) v.y = -v.y;
    a.dst[j * dj + aa * da + b * db] = v;
  }
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
