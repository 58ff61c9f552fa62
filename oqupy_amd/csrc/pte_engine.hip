// Host side of the MI355X process-tensor engine: device-resident tensor chains,
// the PT-TEMPO and TEMPO step logic, compute_dynamics, and the C ABI declared
// in include/ptengine.h.
//
// The step logic restates the reference backends
// (oqupy/backends/pt_tempo_backend.py, oqupy/backends/tempo_backend.py,
// oqupy/backends/node_array.py) on explicit [left, physical, right] site
// tensors and [west, in, out, east] influence nodes; every contraction,
// factorisation and truncation runs in the kernels of pte_kernels.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>
#include <exception>
#include <thread>
#include <condition_variable>
#include <mutex>
#include <cstdio>
#include <cstdlib>

#include "ptengine.h"
#include "pte_device.h"
#include "pte_kernels.h"
#include "pte_rank.h"
#include "pte_eta.h"

namespace pte {

using cplx = std::complex<double>;

static thread_local std::string g_last_error;

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIP_CHECK(x)                                                         \
  do {                                                                       \
    hipError_t e_ = (x);                                                     \
    if (e_ != hipSuccess)                                                    \
      throw Error(PTE_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

static void require(bool ok, int code, const std::string& msg) {
  if (!ok) throw Error(code, msg);
}

static int grid_for(long total, int per_block = 256) {
  long g = (total + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > 65535L * 8) g = 65535L * 8;
  return (int)g;
}

// ---------------------------------------------------------------------------
// device buffers
// ---------------------------------------------------------------------------
struct DevBuf {
  double2* p = nullptr;
  size_t cap = 0;  // complex entries
  bool fixed = false;  // carved from a caller's memory: never reallocated
  void ensure(size_t n, hipStream_t s) {
    if (n <= cap) return;
    if (fixed)
      throw Error(PTE_ERR_CAPACITY, "matrix exceeds the allocation-free workspace");
    if (p) HIP_CHECK(hipFreeAsync(p, s));
    size_t c = std::max<size_t>(n, cap * 3 / 2);
    c = std::max<size_t>(c, 256);
    HIP_CHECK(hipMallocAsync((void**)&p, c * sizeof(double2), s));
    cap = c;
  }
  void release(hipStream_t s) {
    if (p && !fixed) (void)hipFreeAsync(p, s);
    p = nullptr;
    cap = 0;
  }
  // point at `count` entries of fixed memory at *base, advancing *base
  void carve(char** base, size_t count) {
    p = reinterpret_cast<double2*>(*base);
    cap = count;
    fixed = true;
    *base += (count * sizeof(double2) + 255) & ~(size_t)255;
  }
};

// A tensor on the device with up to four dimensions (row-major).
struct DTensor {
  double2* p = nullptr;
  int d[4] = {1, 1, 1, 1};
  long size() const { return (long)d[0] * d[1] * d[2] * d[3]; }
};

static DTensor dalloc(int a, int b, int c, int e, hipStream_t s) {
  DTensor t;
  t.d[0] = a; t.d[1] = b; t.d[2] = c; t.d[3] = e;
  long n = std::max<long>(1, t.size());
  HIP_CHECK(hipMallocAsync((void**)&t.p, n * sizeof(double2), s));
  return t;
}
static void dfree(DTensor& t, hipStream_t s) {
  if (t.p) (void)hipFreeAsync(t.p, s);
  t.p = nullptr;
}
static DTensor dupload(const std::vector<cplx>& h, int a, int b, int c, int e,
                       hipStream_t s) {
  DTensor t = dalloc(a, b, c, e, s);
  require((long)h.size() == t.size(), PTE_ERR_ARG, "internal: upload size mismatch");
  HIP_CHECK(hipMemcpyAsync(t.p, h.data(), h.size() * sizeof(double2),
                           hipMemcpyHostToDevice, s));
  return t;
}

// An influence node [west][in][out][east]; middle nodes of the delta form
// W[w,p,s,v] = delta(w,v) delta(p,s) I[w,p] keep only I (west-major).
struct Node {
  DTensor dense;   // [wl][n][s][wr] (valid when kind == NODE_DENSE)
  DTensor delta;   // [wl][p] weights (valid when kind == NODE_DELTA)
  int kind = NODE_DENSE;
  int wl = 1, n = 1, s = 1, wr = 1;
};

static Node make_dense(const std::vector<cplx>& h, int wl, int n, int s, int wr,
                       hipStream_t st) {
  Node nd;
  nd.kind = NODE_DENSE;
  nd.wl = wl; nd.n = n; nd.s = s; nd.wr = wr;
  nd.dense = dupload(h, wl, n, s, wr, st);
  return nd;
}
// I is given as [w][p] (w: the bond index carried along the chain)
static Node make_delta(const std::vector<cplx>& Iwp, int w, int p, hipStream_t st) {
  Node nd;
  nd.kind = NODE_DELTA;
  nd.wl = w; nd.n = p; nd.s = p; nd.wr = w;
  nd.delta = dupload(Iwp, w, p, 1, 1, st);
  return nd;
}
static void free_node(Node& nd, hipStream_t s) {
  dfree(nd.dense, s);
  dfree(nd.delta, s);
}
// dense host copy of a node (for closing legs on the host)
static std::vector<cplx> node_dense_host(const Node& nd, hipStream_t s) {
  std::vector<cplx> out((size_t)nd.wl * nd.n * nd.s * nd.wr, cplx(0.0, 0.0));
  if (nd.kind == NODE_DENSE) {
    HIP_CHECK(hipMemcpyAsync(out.data(), nd.dense.p, out.size() * sizeof(double2),
                             hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
  } else {
    std::vector<cplx> I((size_t)nd.wl * nd.n);
    HIP_CHECK(hipMemcpyAsync(I.data(), nd.delta.p, I.size() * sizeof(double2),
                             hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    for (int w = 0; w < nd.wl; ++w)
      for (int p = 0; p < nd.n; ++p)
        out[(((size_t)w * nd.n + p) * nd.s + p) * nd.wr + w] = I[(size_t)w * nd.n + p];
  }
  return out;
}

static std::vector<cplx> to_cplx(const double* data, long count) {
  std::vector<cplx> v(count);
  for (long i = 0; i < count; ++i) {
    v[i] = cplx(data[2 * i], data[2 * i + 1]);
    require(std::isfinite(data[2 * i]) && std::isfinite(data[2 * i + 1]),
            PTE_ERR_NUMERIC, "non-finite value in an input tensor");
  }
  return v;
}

// ---------------------------------------------------------------------------
// workspace and the truncated SVD
// ---------------------------------------------------------------------------
// Longest column for the QR-preconditioned SVD: qrcp_kernel holds the step's
// reflector and apply_q_kernel one column of Q in LDS, 16 bytes per row
// (144 KiB of the 160 KiB of LDS of a compute unit, beside their reduction
// buffers). Taller matrices take the unpreconditioned Jacobi.
constexpr int kQrMaxRows = 9216;

// largest rank a truncated SVD of an m x n matrix can keep
static inline int k_bound(int m, int n) { return std::max(1, std::min(m, n)); }

struct Batcher;

// Learned truncation-rank form: per-bond history on the device, the
// predictor's parameters and the statistics of the capped SVDs.
struct RankState {
  int mode = 0;                 // 0 exact, 1 previous steps + oversample, 2 MLP,
                                // 3 adaptive (stop at the acceptance threshold)
  float* params = nullptr;      // device, kRankParams (mode 2)
  float alpha = 0.f, temperature = 0.f;
  int dkmax = 1, num_steps = 1;
  int oversample = 8;
  float margin = 0.1f;
  double theta = 0.02;          // accept when residual <= theta * epsrel * |beta_0|
  float step_frac = 0.f;
  float* hist = nullptr;
  int* hist_rank = nullptr;
  int* cap_d = nullptr;
  double* out_d = nullptr;      // stop summary of the last pivoted QR (4 doubles)
  double* out_h = nullptr;
  float* rec = nullptr;
  int* rec_count = nullptr;
  int rec_cap = 0;
  long svds = 0, capped = 0, fallbacks = 0, qr_steps = 0;

  void init(hipStream_t st) {
    HIP_CHECK(hipMalloc((void**)&hist, sizeof(float) * kRankKeys * kRankHist));
    HIP_CHECK(hipMalloc((void**)&hist_rank, sizeof(int) * kRankKeys));
    HIP_CHECK(hipMalloc((void**)&cap_d, sizeof(int)));
    HIP_CHECK(hipMalloc((void**)&out_d, 4 * sizeof(double)));
    HIP_CHECK(hipHostMalloc((void**)&out_h, 4 * sizeof(double), hipHostMallocDefault));
    HIP_CHECK(hipMalloc((void**)&rec_count, sizeof(int)));
    HIP_CHECK(hipMemsetAsync(hist_rank, 0, sizeof(int) * kRankKeys, st));
    HIP_CHECK(hipMemsetAsync(rec_count, 0, sizeof(int), st));
    HIP_CHECK(hipMemsetAsync(out_d, 0, 4 * sizeof(double), st));
    HIP_CHECK(hipStreamSynchronize(st));
  }
  void release() {
    (void)hipFree(params); (void)hipFree(hist); (void)hipFree(hist_rank);
    (void)hipFree(cap_d); (void)hipFree(out_d); (void)hipHostFree(out_h);
    (void)hipFree(rec); (void)hipFree(rec_count);
    params = nullptr; hist = nullptr; hist_rank = nullptr; cap_d = nullptr;
    out_d = nullptr; out_h = nullptr; rec = nullptr; rec_count = nullptr;
  }
  RankQuery query(int key, int m, int n, float epsrel) const {
    RankQuery q;
    q.key = key;
    q.step_frac = step_frac;
    q.kind = (float)(key / kRankKeysPerKind);
    q.rel_frac = (float)(key % kRankKeysPerKind) / (float)std::max(dkmax, 1);
    q.log_n = std::log2((float)std::max(1, std::min(m, n))) / 8.0f;
    q.alpha = alpha;
    q.temperature = temperature;
    q.log_eps = std::log10(epsrel) / 10.0f;
    q.m = m;
    q.n = n;
    return q;
  }
};

// history key of a bond: kind 0 zip-up, 1 sweep; rel = distance of the
// site from the newest one
static int rank_key(int kind, int rel) {
  return kind * kRankKeysPerKind + std::max(0, std::min(rel, kRankKeysPerKind - 1));
}

struct Workspace {
  hipStream_t stream = nullptr;
  // stream of the SVD phases: `stream`, or the batch stream while a batched
  // launch serves this workspace
  hipStream_t svd_stream = nullptr;
  Batcher* batcher = nullptr;
  hipEvent_t ev_ready = nullptr;
  DevBuf G, X, V, carry, carry2, svh, gpart, Q;
  // pivoted-QR state of the last svd_qrcp call
  double2* qr_tau = nullptr;
  double2* qr_vscale = nullptr;
  double* qr_beta = nullptr;
  int* qr_piv = nullptr;
  int* qr_done = nullptr;
  size_t qr_cap = 0;
  QrcpCand* qr_cand = nullptr;
  int* qr_rank_d = nullptr;
  int* qr_rank_h = nullptr;
  int qr_m = 0, qr_n = 0, qr_r = 0;
  int num_cus = 0;
  double* sigma = nullptr;
  double* sorted = nullptr;
  int* perm = nullptr;
  size_t nsv_cap = 0;
  unsigned* bar = nullptr;
  unsigned* jskip = nullptr;   // converged-pair bookkeeping of the Jacobi
  size_t jskip_cap = 0;        // words
  int jskip_words = 0;         // words the last prepared Jacobi uses
  int* flags = nullptr;
  double* fro2 = nullptr;
  SvdInfo* info_d = nullptr;
  SvdInfo* info_h = nullptr;
  int max_sweeps = 80;
  int inner_sweeps = 1;      // tuning: PTE_JACOBI_INNER_SWEEPS
  int jacobi_sync_words = kBarrierWords;   // words of `bar` the last prepared Jacobi uses
  int qr_blocks = 256;       // tuning: PTE_QR_BLOCKS (workgroups of the pivoted QR)
  int slice_rows = 64;       // tuning: PTE_JACOBI_SLICE_ROWS (min rows per slice)
  // LDS-tile Jacobi (one workgroup per compute unit): enabled unless
  // PTE_JACOBI_TILE=0; slices keep at least tile_rows rows of X
  // (PTE_JACOBI_TILE_ROWS); a slice's X and V rows together fit kTileMaxRows
  bool tile_enabled = true;
  int tile_rows = 64;
  int tile_coop_blocks = 0;
  bool small_enabled = true; // tuning: PTE_SMALL_SVD=0 sends every SVD to the grid path
  int max_coop_blocks = 0;
  double epsrel = 1e-7;
  int chi_cap = 512;
  int precision = PTE_PRECISION_C128;
  // statistics of the current step
  int n_svd = 0, sweeps = 0, max_bond = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // accumulated timing of the Jacobi SVD kernel (HIP events on this stream)
  hipEvent_t evj0 = nullptr, evj1 = nullptr;
  hipEvent_t evq0 = nullptr, evq1 = nullptr;
  long jac_count = 0, qr_count = 0, small_count = 0;
  double qr_ms = 0.0, small_ms = 0.0;
  // optional per-SVD log (environment variable PTE_SVD_TRACE = file path):
  // m, n, pivoted-QR rank (0 without QR), kept rank, sweeps, QR ms, Jacobi ms
  FILE* trace = nullptr;
  // learned truncation-rank form (null when off)
  RankState* rank = nullptr;
  const double* trunc_extra2 = nullptr;   // residual folded into the cut (capped QR)
  int tr_m = 0, tr_n = 0, tr_r = 0;
  float tr_qr_ms = 0.f;
  double jac_ms = 0.0, jac_flops = 0.0, jac_bytes = 0.0;
  bool external = false;     // buffers carved from caller memory (init_external)
  // optional log of the truncated SVDs (pte_pt_svd_log): per SVD the shape of
  // the matrix in the reference's orientation, the kept rank and the kept
  // singular values (read back from `sorted` after the cut); recording stops
  // once log_cap values are held
  long log_cap = 0;
  std::vector<int> log_shapes;
  std::vector<double> log_values;

  void log_svd(int m, int n, int k) {
    if (log_cap <= 0 || (long)log_values.size() + k > log_cap) return;
    const size_t off = log_values.size();
    log_values.resize(off + (size_t)k);
    HIP_CHECK(hipMemcpyAsync(log_values.data() + off, sorted, (size_t)k * sizeof(double),
                             hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    log_shapes.push_back(m);
    log_shapes.push_back(n);
    log_shapes.push_back(k);
  }

  void init(double eps, int cap) {
    epsrel = eps;
    chi_cap = cap;
    HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    svd_stream = stream;
    HIP_CHECK(hipEventCreateWithFlags(&ev_ready, hipEventDisableTiming));
    hipMemPool_t pool;
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
      uint64_t thr = ~0ull;
      (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
    }
    HIP_CHECK(hipMalloc((void**)&bar, sizeof(unsigned) * kSyncWords));
    HIP_CHECK(hipMalloc((void**)&flags, sizeof(int) * max_sweeps));
    HIP_CHECK(hipMalloc((void**)&fro2, sizeof(double)));
    HIP_CHECK(hipMalloc((void**)&info_d, sizeof(SvdInfo)));
    HIP_CHECK(hipHostMalloc((void**)&info_h, sizeof(SvdInfo), hipHostMallocDefault));
    int per_cu = 0;
    HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void*>(jacobi_kernel), 256, 0));
    hipDeviceProp_t prop;
    HIP_CHECK(hipGetDeviceProperties(&prop, dev));
    max_coop_blocks = per_cu * prop.multiProcessorCount;
    num_cus = prop.multiProcessorCount;
    require(prop.sharedMemPerBlock >= kSmallLdsMax + 1024, PTE_ERR_DEVICE,
            "the engine needs 160 KiB of LDS per workgroup (gfx950)");
    init_tile_kernel();
    HIP_CHECK(hipMalloc((void**)&qr_cand, sizeof(QrcpCand) * 2 * 256));
    HIP_CHECK(hipMalloc((void**)&qr_rank_d, sizeof(int)));
    HIP_CHECK(hipHostMalloc((void**)&qr_rank_h, sizeof(int), hipHostMallocDefault));
    require(max_coop_blocks > 0, PTE_ERR_DEVICE, "Jacobi kernel cannot be resident");
    HIP_CHECK(hipEventCreate(&ev0));
    HIP_CHECK(hipEventCreate(&ev1));
    HIP_CHECK(hipEventCreate(&evj0));
    HIP_CHECK(hipEventCreate(&evj1));
    HIP_CHECK(hipEventCreate(&evq0));
    HIP_CHECK(hipEventCreate(&evq1));
    if (const char* v = std::getenv("PTE_JACOBI_INNER_SWEEPS"))
      inner_sweeps = std::max(1, std::min(8, std::atoi(v)));
    if (const char* v = std::getenv("PTE_SMALL_SVD")) small_enabled = std::atoi(v) != 0;
    HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(small_svd_kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)kSmallLdsMax));
    HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(small_svd_batch_kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)kSmallLdsMax));
    if (const char* v = std::getenv("PTE_JACOBI_SLICE_ROWS"))
      slice_rows = std::max(32, std::min(4096, std::atoi(v)));
    if (const char* v = std::getenv("PTE_QR_BLOCKS"))
      qr_blocks = std::max(1, std::min(256, std::atoi(v)));
    if (const char* path = std::getenv("PTE_SVD_TRACE")) {
      if (*path) trace = std::fopen(path, "a");
    }
  }

  // Timing events are optional (null in the allocation-free workspace).
  void mark(hipEvent_t ev) {
    if (ev) HIP_CHECK(hipEventRecord(ev, svd_stream));
  }
  float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (a && b) HIP_CHECK(hipEventElapsedTime(&ms, a, b));
    return ms;
  }

  // A workspace for one truncated SVD of at most max_m x max_n (and
  // max_n x max_m) in `bytes` of caller-provided device memory at `base`,
  // on the null stream: no stream, event or memory is created, and a larger
  // matrix is a capacity error. Host-side summaries go to host_info.
  void init_external(double eps, int cap, char* base, size_t bytes, int max_m, int max_n,
                     SvdInfo* host_info, int* host_rank) {
    external = true;
    epsrel = eps;
    chi_cap = cap;
    stream = nullptr;
    svd_stream = nullptr;
    char* const begin = base;
    const size_t mn = (size_t)max_m * max_n;
    const size_t big = std::max(max_m, max_n), small = std::min(max_m, max_n);
    auto take = [&](size_t b) {
      char* p = base;
      base += (b + 255) & ~(size_t)255;
      return p;
    };
    bar = reinterpret_cast<unsigned*>(take(sizeof(unsigned) * kSyncWords));
    {
      const size_t nbmax = (size_t)jacobi_blocks(std::max(max_m, max_n));
      jskip_cap = nbmax + nbmax * (nbmax - 1) / 2;
      jskip = reinterpret_cast<unsigned*>(take(sizeof(unsigned) * jskip_cap));
    }
    flags = reinterpret_cast<int*>(take(sizeof(int) * max_sweeps));
    fro2 = reinterpret_cast<double*>(take(sizeof(double)));
    info_d = reinterpret_cast<SvdInfo*>(take(sizeof(SvdInfo)));
    qr_cand = reinterpret_cast<QrcpCand*>(take(sizeof(QrcpCand) * 2 * 256));
    qr_rank_d = reinterpret_cast<int*>(take(sizeof(int)));
    qr_tau = reinterpret_cast<double2*>(take(sizeof(double2) * big));
    qr_vscale = reinterpret_cast<double2*>(take(sizeof(double2) * big));
    qr_beta = reinterpret_cast<double*>(take(sizeof(double) * big));
    qr_piv = reinterpret_cast<int*>(take(sizeof(int) * big));
    qr_done = reinterpret_cast<int*>(take(sizeof(int) * big));
    qr_cap = big;
    sigma = reinterpret_cast<double*>(take(sizeof(double) * big));
    sorted = reinterpret_cast<double*>(take(sizeof(double) * big));
    perm = reinterpret_cast<int*>(take(sizeof(int) * big));
    nsv_cap = big;
    Q.carve(&base, mn);
    X.carve(&base, mn);
    V.carve(&base, big * big);
    G.carve(&base, mn);       // the input matrix, row-major
    carry.carve(&base, mn);   // left factor
    svh.carve(&base, mn);     // S V^H
    const int npairs = jacobi_blocks((int)big) / 2;
    gpart.carve(&base, (size_t)2 * npairs * 16 * 1024);
    (void)small;
    require((size_t)(base - begin) <= bytes, PTE_ERR_CAPACITY,
            "internal: allocation-free workspace larger than its memory");
    info_h = host_info;
    qr_rank_h = host_rank;
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    int per_cu = 0;
    HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void*>(jacobi_kernel), 256, 0));
    hipDeviceProp_t prop;
    HIP_CHECK(hipGetDeviceProperties(&prop, dev));
    max_coop_blocks = per_cu * prop.multiProcessorCount;
    num_cus = prop.multiProcessorCount;
    require(prop.sharedMemPerBlock >= kSmallLdsMax + 1024, PTE_ERR_DEVICE,
            "the engine needs 160 KiB of LDS per workgroup (gfx950)");
    init_tile_kernel();
    require(max_coop_blocks > 0, PTE_ERR_DEVICE, "Jacobi kernel cannot be resident");
    HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(small_svd_kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)kSmallLdsMax));
  }

  void destroy() {
    if (external) return;
    if (!stream) return;
    G.release(stream); X.release(stream); V.release(stream);
    carry.release(stream); carry2.release(stream); svh.release(stream);
    gpart.release(stream); Q.release(stream);
    (void)hipStreamSynchronize(stream);
    (void)hipFree(qr_tau); (void)hipFree(qr_vscale); (void)hipFree(qr_beta);
    (void)hipFree(qr_piv); (void)hipFree(qr_done); (void)hipFree(qr_cand);
    (void)hipFree(qr_rank_d); (void)hipHostFree(qr_rank_h);
    (void)hipFree(sigma); (void)hipFree(sorted); (void)hipFree(perm);
    (void)hipFree(bar); (void)hipFree(jskip); (void)hipFree(flags); (void)hipFree(fro2);
    (void)hipFree(info_d);
    (void)hipHostFree(info_h);
    (void)hipEventDestroy(ev0); (void)hipEventDestroy(ev1);
    (void)hipEventDestroy(evj0); (void)hipEventDestroy(evj1);
    (void)hipEventDestroy(evq0); (void)hipEventDestroy(evq1);
    (void)hipEventDestroy(ev_ready);
    if (trace) std::fclose(trace);
    trace = nullptr;
    (void)hipStreamDestroy(stream);
    stream = nullptr;
  }

  void ensure_sv(size_t n) {
    if (n <= nsv_cap) return;
    require(!external, PTE_ERR_CAPACITY, "matrix exceeds the allocation-free workspace");
    HIP_CHECK(hipStreamSynchronize(stream));
    HIP_CHECK(hipStreamSynchronize(svd_stream));
    (void)hipFree(sigma); (void)hipFree(sorted); (void)hipFree(perm);
    size_t c = std::max<size_t>(n, 1024);
    HIP_CHECK(hipMalloc((void**)&sigma, c * sizeof(double)));
    HIP_CHECK(hipMalloc((void**)&sorted, c * sizeof(double)));
    HIP_CHECK(hipMalloc((void**)&perm, c * sizeof(int)));
    nsv_cap = c;
  }

  // ---- phases of the truncated SVD, shared by the single and the batched
  // launches (svd_stream is the stream they are enqueued on) ----

  // Launch limits of the LDS-tile Jacobi (its LDS allows one workgroup per
  // compute unit) and its tuning variables.
  void init_tile_kernel() {
    // LDS of the QR path: one column of up to kQrMaxRows rows
    for (const void* f : {reinterpret_cast<const void*>(qrcp_kernel),
                          reinterpret_cast<const void*>(qrcp_batch_kernel),
                          reinterpret_cast<const void*>(apply_q_kernel)})
      HIP_CHECK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)(kQrMaxRows * sizeof(double2))));
    HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(jacobi_tile_kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)tile_lds_bytes(kTileMaxRows)));
    int per_cu = 0;
    HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void*>(jacobi_tile_kernel), 256,
        tile_lds_bytes(kTileMaxRows)));
    tile_coop_blocks = per_cu * num_cus;
    if (const char* v = std::getenv("PTE_JACOBI_TILE")) tile_enabled = std::atoi(v) != 0;
    if (const char* v = std::getenv("PTE_JACOBI_TILE_ROWS"))
      tile_rows = std::max(32, std::min(kTileMaxRows, std::atoi(v)));
  }

  static size_t tile_lds_bytes(int rows) {
    return (size_t)rows * kJacobiTileLd * sizeof(double2);
  }

  // Row slices of the LDS-tile Jacobi on X (m x n): the most slices (up to
  // 16, at least tile_rows rows of X each) that one co-resident grid holds;
  // 0 when the path is off or a slice's rows do not fit the tile.
  int jacobi_tile_slices(int m, int n) const {
    if (!tile_enabled || tile_coop_blocks <= 0) return 0;
    const int npairs = jacobi_blocks(n) / 2;
    if (npairs > tile_coop_blocks) return 0;
    const int R = std::max(1, std::min({tile_coop_blocks / npairs,
                                        (m + tile_rows - 1) / tile_rows, 16}));
    auto round32 = [](int v) { return ((v + 31) / 32) * 32; };
    const int rows = round32((m + R - 1) / R) + round32((n + R - 1) / R);
    return rows <= kTileMaxRows ? R : 0;
  }

  // Buffers and arguments of the Jacobi on X (m x n, column-major, ld = m);
  // at most max_grid workgroups. *grid_out receives the grid size.
  JacobiArgs jacobi_prepare(int m, int n, int max_grid, int* grid_out, int slices = 0) {
    require(m >= 1 && n >= 1, PTE_ERR_ARG, "internal: empty SVD");
    V.ensure((size_t)n * n, svd_stream);
    ensure_sv((size_t)n);
    const int nb = jacobi_blocks(n);
    const int npairs = nb / 2;
    require(nb <= (int)kSyncMaxBlocks, PTE_ERR_CAPACITY,
            "SVD of " + std::to_string(n) + " columns exceeds the Jacobi's block limit");
    require(npairs <= max_grid, PTE_ERR_CAPACITY,
            "SVD of " + std::to_string(n) + " columns exceeds the co-resident grid");
    // row slices per block pair: each slice keeps at least slice_rows rows of
    // X, at most 16 slices, within the grid
    const int R = slices > 0 ? slices
                             : std::max(1, std::min({max_grid / npairs,
                                                     (m + slice_rows - 1) / slice_rows, 16}));
    auto round32 = [](int v) { return ((v + 31) / 32) * 32; };
    const int xrows = round32((m + R - 1) / R), vrows = round32((n + R - 1) / R);
    if (R > 1) gpart.ensure((size_t)2 * npairs * R * 1024, svd_stream);
    jacobi_sync_words = kBarrierWords + 2 * nb;
    jskip_words = nb + nb * (nb - 1) / 2;
    ensure_jskip((size_t)jskip_words);
    JacobiArgs ja;
    ja.X = X.p; ja.m = m; ja.n = n; ja.ldx = m;
    ja.V = V.p; ja.ldv = n;
    ja.nb = nb;
    ja.slices = R; ja.xrows_per = xrows; ja.vrows_per = vrows;
    ja.gpart = gpart.p;
    ja.tol = std::max(m, 16) * 2.220446049250313e-16;
    ja.abs_tol_scale = 2.220446049250313e-16;
    ja.max_sweeps = max_sweeps;
    ja.inner_sweeps = inner_sweeps;
    ja.bar = bar; ja.flags = flags; ja.fro2 = fro2;
    ja.info = &info_d->sweeps;
    ja.skip = jskip;
    *grid_out = npairs * R;
    return ja;
  }

  void ensure_jskip(size_t words) {
    if (words <= jskip_cap) return;
    require(!external, PTE_ERR_CAPACITY, "matrix exceeds the allocation-free workspace");
    HIP_CHECK(hipStreamSynchronize(stream));
    HIP_CHECK(hipStreamSynchronize(svd_stream));
    (void)hipFree(jskip);
    const size_t c = std::max<size_t>(words, 4096);
    HIP_CHECK(hipMalloc((void**)&jskip, c * sizeof(unsigned)));
    jskip_cap = c;
  }

  static int jacobi_blocks(int n) {
    int nb = (n + kJacobiBlock - 1) / kJacobiBlock;
    if (nb % 2) ++nb;
    return std::max(nb, 2);
  }

  // natural grid of a Jacobi on m x n (before any budget)
  int jacobi_natural_grid(int m, int n) const {
    const int npairs = jacobi_blocks(n) / 2;
    const int R = std::max(1, std::min({std::min(max_coop_blocks, 1024) / std::max(npairs, 1),
                                        (m + slice_rows - 1) / slice_rows, 16}));
    return npairs * R;
  }

  void jacobi_reset() {
    HIP_CHECK(hipMemsetAsync(bar, 0, sizeof(unsigned) * jacobi_sync_words, svd_stream));
    HIP_CHECK(hipMemsetAsync(jskip, 0, sizeof(unsigned) * jskip_words, svd_stream));
    HIP_CHECK(hipMemsetAsync(flags, 0, sizeof(int) * max_sweeps, svd_stream));
    HIP_CHECK(hipMemsetAsync(fro2, 0, sizeof(double), svd_stream));
  }

  // singular values, descending order, cut; the summary goes to info_h
  void jacobi_finish(int m, int n) {
    column_norm_kernel<<<n, 256, 0, svd_stream>>>(X.p, m, n, m, sigma);
    rank_sort_kernel<<<(n + 255) / 256, 256, 0, svd_stream>>>(sigma, n, perm, sorted);
    truncate_kernel<<<1, 1024, 0, svd_stream>>>(sorted, n, epsrel, info_d, chi_cap,
                                                 trunc_extra2);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(info_h, info_d, sizeof(SvdInfo), hipMemcpyDeviceToHost,
                             svd_stream));
  }

  // after synchronisation: checks, statistics, kept rank
  // `small`: the SVD ran in small_svd_kernel (its time is kept apart from
  // the Jacobi kernel's statistics)
  int jacobi_check(int m, int n, float ms, bool small = false) {
    require(info_h->sweeps > 0, PTE_ERR_NUMERIC,
            "Jacobi SVD did not converge (" + std::to_string(m) + "x" +
                std::to_string(n) + ")");
    require(!info_h->overflow, PTE_ERR_CAPACITY,
            "bond dimension " + std::to_string(info_h->rank) +
                " exceeds the capacity chi_cap=" + std::to_string(chi_cap));
    require(info_h->rank >= 1, PTE_ERR_NUMERIC, "SVD of a zero matrix");
    // algorithmic work of a thin complex SVD of an M x N matrix (R-SVD count
    // of Golub & Van Loan, times 4 for complex arithmetic) and the HBM bytes
    // of reading and writing M once plus writing V
    const double mm = std::max(m, n), nn = std::min(m, n);
    if (small) {
      small_ms += ms;
      small_count += 1;
    } else {
      jac_ms += ms;
      jac_count += 1;
      jac_flops += 4.0 * (6.0 * mm * nn * nn + 20.0 * nn * nn * nn);
      jac_bytes += 16.0 * (2.0 * (double)m * n + (double)n * n);
    }
    if (trace) {
      if (tr_r == 0) { tr_m = n; tr_n = m; tr_qr_ms = 0.f; }
      std::fprintf(trace, "%d,%d,%d,%d,%d,%.4f,%.4f\n", tr_m, tr_n, tr_r, info_h->rank,
                   info_h->sweeps, tr_qr_ms, ms);
      tr_r = 0;
    }
    n_svd += 1;
    sweeps += info_h->sweeps;
    max_bond = std::max(max_bond, info_h->rank);
    return info_h->rank;
  }

  // Jacobi SVD of X (m x n, column-major, ld = m, already in X) with V (n x n)
  // accumulated; sorts and truncates. Returns the kept rank.
  int svd(int m, int n) {
    if (batcher) return batched_svd(false, m, n);
    int grid = 0;
    const int tile_slices = jacobi_tile_slices(m, n);
    JacobiArgs ja = jacobi_prepare(m, n, std::min(max_coop_blocks, 1024), &grid, tile_slices);
    jacobi_reset();
    void* kargs[] = {&ja};
    mark(evj0);
    if (tile_slices > 0) {
      HIP_CHECK(hipLaunchCooperativeKernel(
          reinterpret_cast<const void*>(jacobi_tile_kernel), dim3(grid), dim3(256), kargs,
          tile_lds_bytes(ja.xrows_per + ja.vrows_per), svd_stream));
    } else {
      HIP_CHECK(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(jacobi_kernel),
                                           dim3(grid), dim3(256), kargs, 0, svd_stream));
    }
    mark(evj1);
    jacobi_finish(m, n);
    HIP_CHECK(hipStreamSynchronize(svd_stream));
    return jacobi_check(m, n, elapsed(evj0, evj1));
  }

  void ensure_qr(size_t n) {
    if (n <= qr_cap) return;
    require(!external, PTE_ERR_CAPACITY, "matrix exceeds the allocation-free workspace");
    HIP_CHECK(hipStreamSynchronize(stream));
    HIP_CHECK(hipStreamSynchronize(svd_stream));
    (void)hipFree(qr_tau); (void)hipFree(qr_vscale); (void)hipFree(qr_beta);
    (void)hipFree(qr_piv); (void)hipFree(qr_done);
    const size_t c = std::max<size_t>(n, 1024);
    HIP_CHECK(hipMalloc((void**)&qr_tau, c * sizeof(double2)));
    HIP_CHECK(hipMalloc((void**)&qr_vscale, c * sizeof(double2)));
    HIP_CHECK(hipMalloc((void**)&qr_beta, c * sizeof(double)));
    HIP_CHECK(hipMalloc((void**)&qr_piv, c * sizeof(int)));
    HIP_CHECK(hipMalloc((void**)&qr_done, c * sizeof(int)));
    qr_cap = c;
  }

  // Truncated SVD of M (m x n, column-major in Q, ld = m, m <= kQrMaxRows):
  // pivoted QR  M P = Q R  (stopped once the unfactored block is below
  // eps * ||M||_F, r steps), then the Jacobi SVD of Y = R^H (n x r) in X,
  // Y W = Z. With perm the descending order of the column norms of Z:
  //   left singular vector j = Q W[:, perm[j]]            (emit_left_q)
  //   row j of S V^H at column piv[s] = conj(Z[s, perm[j]])  (emit_svh)
  // Returns the kept rank k.
  // `key` >= 0 names the bond (rank_key) for the learned form: its history
  // is updated and, with a predictor, the pivoted QR is capped at the
  // predicted step count + oversample. The capped factorisation is accepted
  // when its residual is at most theta * epsrel * |beta_0| (|beta_0|, the
  // largest column norm, is a lower bound of s0), and the residual then
  // counts as discarded weight in the cut; otherwise the pivoted QR resumes
  // where it stopped and runs to the exact stop. The step count the history
  // and the trace keep is the first one whose residual met that threshold
  // (mode 3 stops there without a prediction).
  int svd_qrcp(int m, int n, int key = -1) {
    if (batcher) return batched_svd(true, m, n);
    if (small_ok(m, n)) {
      const bool adaptive = rank != nullptr && key >= 0 && rank->mode == 3;
      const int k = svd_small(m, n, adaptive);
      if (adaptive) {
        rank->svds += 1;
        rank->capped += 1;
        rank->qr_steps += info_h->pad;
      }
      return k;
    }
    const bool learn = rank != nullptr && key >= 0;
    const bool capped = learn && rank->mode > 0;
    const RankQuery q = learn ? rank->query(key, m, n, (float)epsrel) : RankQuery();
    if (capped && rank->mode != 3) {
      RankPredictArgs pa;
      pa.params = rank->mode == 2 ? rank->params : nullptr;
      pa.hist = rank->hist;
      pa.hist_rank = rank->hist_rank;
      pa.q = q;
      pa.oversample = rank->oversample;
      pa.margin = rank->margin;
      pa.nmax = std::min(m, n);
      pa.cap = rank->cap_d;
      rank_predict_kernel<<<1, kRankWidth, 0, svd_stream>>>(pa);
    }
    float ms = 0.f;
    const bool predicted = capped && rank->mode != 3;
    int r = qrcp_run(m, n, predicted ? rank->cap_d : nullptr, 0, 0.0, &ms,
                     capped && rank->mode == 3);
    bool folded = false;
    if (capped) {
      rank->capped += 1;
      const double* o = rank->out_h;   // resid2, ||M||_F^2, |beta_0|, mark
      const double thr = rank->theta * epsrel * o[2];
      if (o[0] <= thr * thr) {
        folded = o[0] > 0.0;
      } else {
        rank->fallbacks += 1;
        const double eps = 2.220446049250313e-16;
        float ms2 = 0.f;
        r = qrcp_run(m, n, nullptr, r, eps * eps * o[1], &ms2);
        ms += ms2;
      }
    }
    if (rank) rank->qr_steps += r;
    qr_finish_host(m, n, r, ms);
    build_rh_kernel<<<grid_for((long)n * r), 256, 0, svd_stream>>>(Q.p, m, n, qr_piv,
                                                                   qr_beta, qr_rank_d, X.p);
    trunc_extra2 = folded ? rank->out_d : nullptr;
    int k;
    try {
      k = svd(n, r);
    } catch (...) {
      trunc_extra2 = nullptr;
      throw;
    }
    trunc_extra2 = nullptr;
    if (learn) {
      rank->svds += 1;
      RankRecordArgs ra;
      ra.hist = rank->hist;
      ra.hist_rank = rank->hist_rank;
      ra.q = q;
      ra.sorted = sorted;
      ra.info = info_d;
      ra.need = rank->out_d + 3;
      ra.rec = rank->rec;
      ra.rec_count = rank->rec_count;
      ra.rec_cap = rank->rec_cap;
      rank_record_kernel<<<1, kRankWidth, 0, svd_stream>>>(ra);
      HIP_CHECK(hipGetLastError());
    }
    return k;
  }

  // one launch of the pivoted QR of Q (m x n): capped at *max_steps when
  // given, resumed after `start` steps when start > 0 (stop threshold
  // drop_abs2 on the squared residual). Returns the steps taken in total;
  // the stop summary goes to rank->out_h when the learned form is on.
  int qrcp_run(int m, int n, const int* max_steps, int start, double drop_abs2, float* ms,
               bool stop_at_mark = false) {
    const int nblk = qr_natural_grid(n);
    QrcpArgs qa = qr_prepare(m, n, nblk);
    qa.max_steps = max_steps;
    qa.start = start;
    qa.drop_abs2 = drop_abs2;
    qa.out = rank ? rank->out_d : nullptr;
    qa.mark_rel = rank ? rank->theta * epsrel : 0.0;
    qa.stop_at_mark = stop_at_mark ? 1 : 0;
    HIP_CHECK(hipMemsetAsync(bar, 0, sizeof(unsigned) * kBarrierWords, svd_stream));
    void* kargs[] = {&qa};
    mark(evq0);
    HIP_CHECK(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(qrcp_kernel),
                                         dim3(nblk), dim3(256), kargs,
                                         (size_t)m * sizeof(double2), svd_stream));
    mark(evq1);
    HIP_CHECK(hipMemcpyAsync(qr_rank_h, qr_rank_d, sizeof(int), hipMemcpyDeviceToHost,
                             svd_stream));
    if (rank)
      HIP_CHECK(hipMemcpyAsync(rank->out_h, rank->out_d, 4 * sizeof(double),
                               hipMemcpyDeviceToHost, svd_stream));
    HIP_CHECK(hipStreamSynchronize(svd_stream));
    *ms = elapsed(evq0, evq1);
    return *qr_rank_h;
  }

  // ---- small SVDs: one workgroup, LDS resident (small_svd_kernel) ----
  // dynamic LDS of small_svd_kernel: 160 KiB less 1 KiB for its static part
  static constexpr size_t kSmallLdsMax = 159 * 1024;
  // the QR phase needs M and 64 bytes per column; the Jacobi phase Y
  // (n x r) and, when it fits, W (r x r), r <= min(m, n); the epilogue Y and
  // two r-vectors
  static size_t small_qr_lds(int m, int n) { return (size_t)m * n * 16 + (size_t)n * 64; }
  static size_t small_lds(int m, int n) {
    const size_t r = (size_t)std::min(m, n);
    return std::max({small_qr_lds(m, n), (size_t)n * r * 16 + r * 16,
                     std::min(kSmallLdsMax, ((size_t)n * r + r * r) * 16)});
  }
  bool small_ok(int m, int n) const {
    return small_enabled && n <= kSmallMaxCols && small_qr_lds(m, n) <= kSmallLdsMax;
  }

  // buffers and arguments of small_svd_kernel on Q (m x n)
  SmallSvdArgs small_prepare(int m, int n, SvdInfo* copy) {
    require(m >= 1 && n >= 1, PTE_ERR_ARG, "internal: empty SVD");
    const int rmax = std::min(m, n);
    ensure_qr((size_t)n);
    ensure_sv((size_t)rmax);
    X.ensure((size_t)n * rmax, svd_stream);
    V.ensure((size_t)rmax * rmax, svd_stream);
    SmallSvdArgs sa;
    sa.A = Q.p; sa.m = m; sa.n = n;
    sa.tau = qr_tau; sa.vscale = qr_vscale; sa.beta = qr_beta; sa.piv = qr_piv;
    sa.qrank = qr_rank_d;
    sa.Y = X.p; sa.W = V.p;
    sa.sigma = sigma; sa.perm = perm; sa.sorted = sorted; sa.info = info_d;
    sa.epsrel = epsrel; sa.max_rank = chi_cap; sa.max_sweeps = max_sweeps;
    sa.copy = copy;
    sa.stop_rel = 0.0;
    sa.lds_bytes = (long)small_lds(m, n);
    return sa;
  }

  // after synchronisation (summary in info_h): the same checks, statistics
  // and state as the pivoted-QR + Jacobi path
  int small_finish(int m, int n, float ms) {
    const int r = info_h->pad;
    require(r >= 1, PTE_ERR_NUMERIC, "SVD of a zero matrix");
    qr_count += 1;
    tr_m = m; tr_n = n; tr_r = r; tr_qr_ms = 0.f;
    qr_m = m; qr_n = n; qr_r = r;
    return jacobi_check(n, r, ms, true);
  }

  // `adaptive`: the learned form's adaptive stop of the QR (its residual
  // counted in the cut)
  int svd_small(int m, int n, bool adaptive = false) {
    SmallSvdArgs sa = small_prepare(m, n, nullptr);
    if (adaptive) sa.stop_rel = rank->theta * epsrel;
    // timed (two event records on the stream) only for the SVD trace
    const bool timed = trace != nullptr;
    if (timed) mark(evj0);
    small_svd_kernel<<<1, kSmallThreads, small_lds(m, n), svd_stream>>>(sa);
    if (timed) mark(evj1);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(info_h, info_d, sizeof(SvdInfo), hipMemcpyDeviceToHost,
                             svd_stream));
    HIP_CHECK(hipStreamSynchronize(svd_stream));
    return small_finish(m, n, timed ? elapsed(evj0, evj1) : 0.f);
  }

  int qr_natural_grid(int n) const { return std::max(1, std::min({n, qr_blocks, num_cus})); }
  static int qr_min_grid(int n) { return std::max(1, (n + kQrcpMaxLocal - 1) / kQrcpMaxLocal); }

  // buffers and arguments of the pivoted QR of Q (m x n) on nblk workgroups
  QrcpArgs qr_prepare(int m, int n, int nblk) {
    require(m >= 1 && n >= 1 && m <= kQrMaxRows, PTE_ERR_ARG, "internal: bad QR shape");
    ensure_qr((size_t)n);
    require((n + nblk - 1) / nblk <= kQrcpMaxLocal, PTE_ERR_CAPACITY,
            "matrix of " + std::to_string(n) + " columns exceeds the pivoted QR");
    QrcpArgs qa;
    qa.A = Q.p; qa.m = m; qa.n = n; qa.lda = m;
    qa.drop = 2.220446049250313e-16;
    qa.tau = qr_tau; qa.vscale = qr_vscale; qa.beta = qr_beta;
    qa.piv = qr_piv; qa.done = qr_done; qa.cand = qr_cand; qa.rank = qr_rank_d;
    qa.bar = bar;
    qa.max_steps = nullptr;
    qa.out = nullptr;
    qa.start = 0;
    qa.drop_abs2 = 0.0;
    qa.mark_rel = 0.0;
    qa.stop_at_mark = 0;
    return qa;
  }

  // after synchronisation: rank r of the pivoted QR (in *qr_rank_h),
  // statistics, X sized for Y = R^H; returns the arguments of building Y
  RhArgs qr_finish_host(int m, int n, int r, float ms) {
    require(r >= 1, PTE_ERR_NUMERIC, "SVD of a zero matrix");
    qr_ms += ms;
    qr_count += 1;
    tr_m = m; tr_n = n; tr_r = r; tr_qr_ms = ms;
    X.ensure((size_t)n * r, svd_stream);
    qr_m = m; qr_n = n; qr_r = r;
    RhArgs ra;
    ra.A = Q.p; ra.lda = m; ra.n = n; ra.pad = 0;
    ra.piv = qr_piv; ra.beta = qr_beta; ra.rank = qr_rank_d; ra.Y = X.p;
    return ra;
  }

  int qr_finish(int m, int n, float ms) {
    const int r = *qr_rank_h;
    qr_finish_host(m, n, r, ms);
    build_rh_kernel<<<grid_for((long)n * r), 256, 0, svd_stream>>>(Q.p, m, n, qr_piv,
                                                                   qr_beta, qr_rank_d, X.p);
    return r;
  }

  SvdPostArgs post_args(int m, int n) const {
    SvdPostArgs pa;
    pa.X = X.p; pa.m = m; pa.n = n;
    pa.sigma = sigma; pa.perm = perm; pa.sorted = sorted; pa.info = info_d;
    pa.epsrel = epsrel; pa.max_rank = chi_cap; pa.pad = 0;
    return pa;
  }

  int batched_svd(bool qr, int m, int n);

  // dst[j*dj + a*da + b*db] = (Q W[:, perm[j]])[a*B + b] for j < k, with
  // every stride d = d0 + dk * k
  void emit_left_q(int B, long dj0, long djk, long da0, long dak, long db0, long dbk,
                   double2* dst, int k) {
    ApplyQArgs aq;
    aq.A = Q.p; aq.m = qr_m; aq.lda = qr_m;
    aq.tau = qr_tau; aq.vscale = qr_vscale; aq.piv = qr_piv; aq.qrank = qr_rank_d;
    aq.W = V.p; aq.ldw = qr_r; aq.perm = perm; aq.info = info_d;
    aq.B = B;
    aq.dj0 = dj0; aq.djk = djk; aq.da0 = da0; aq.dak = dak; aq.db0 = db0; aq.dbk = dbk;
    aq.dst = dst;
    if (qr_m <= kApplyQWaveRows)
      apply_q_wave_kernel<<<(k + 3) / 4, 256, 0, stream>>>(aq);
    else
      apply_q_kernel<<<std::min(k, 1024), 256, (size_t)qr_m * sizeof(double2), stream>>>(aq);
  }

  // S V^H rows of the last svd_qrcp: column c = a*B + b of M goes to
  // dst[j*dj + a*da + b*db], conjugated
  void emit_svh(int B, long dj0, long djk, long da0, long dak, long db0, long dbk,
                double2* dst, int k) {
    emit(X.p, qr_n, qr_n / B, B, dj0, djk, da0, dak, db0, dbk, 1, dst, k, qr_piv);
  }

  void emit(const double2* src, long lds, int A, int B, long dj0, long djk,
            long da0, long dak, long db0, long dbk, int conj, double2* dst, int k,
            const int* rowmap = nullptr) {
    EmitArgs ea;
    ea.src = src; ea.lds = lds; ea.perm = perm; ea.info = info_d;
    ea.A = A; ea.B = B;
    ea.dj0 = dj0; ea.djk = djk; ea.da0 = da0; ea.dak = dak; ea.db0 = db0; ea.dbk = dbk;
    ea.conj = conj; ea.rowmap = rowmap; ea.dst = dst;
    emit_kernel<<<grid_for((long)A * B * k), 256, 0, stream>>>(ea);
  }

  void gemm(int M, int N, int K, const double2* A, long sai, long sak,
            const double2* B, long sbk, long sbj, double2* C, long sci, long scj) {
    GemmArgs g;
    g.M = M; g.N = N; g.K = K;
    g.A = A; g.sai = sai; g.sak = sak;
    g.B = B; g.sbk = sbk; g.sbj = sbj;
    g.C = C; g.sci = sci; g.scj = scj;
    g.conjA = 0; g.conjB = 0;
    g.alpha = make_double2(1.0, 0.0);
    g.beta_zero = 1;
    dim3 grid((N + 63) / 64, (M + 63) / 64);
    zgemm_mfma_kernel<<<grid, 256, 0, stream>>>(g);
  }

  // The network contraction of a zip-up site: complex64 on the matrix cores
  // in the precision form, complex128 otherwise. (The sweeps' absorption of
  // S V^H into the neighbouring site is complex128 in every form: it carries
  // the full dynamic range of the kept singular values into the state, and
  // rounding it to complex64 at every site of every step set the form's
  // error floor.)
  void contract(int M, int N, int K, const double2* A, long sai, long sak,
                const double2* B, long sbk, long sbj, double2* C, long sci, long scj) {
    if (precision != PTE_PRECISION_C64) {
      gemm(M, N, K, A, sai, sak, B, sbk, sbj, C, sci, scj);
      return;
    }
    GemmArgs g;
    g.M = M; g.N = N; g.K = K;
    g.A = A; g.sai = sai; g.sak = sak;
    g.B = B; g.sbk = sbk; g.sbj = sbj;
    g.C = C; g.sci = sci; g.scj = scj;
    g.conjA = 0; g.conjB = 0;
    g.alpha = make_double2(1.0, 0.0);
    g.beta_zero = 1;
    dim3 grid((N + 63) / 64, (M + 63) / 64);
    cgemm_mfma_kernel<<<grid, 256, 0, stream>>>(g);
  }
};

// ---------------------------------------------------------------------------
// Batched truncated SVDs for many engines stepped together (one host thread
// per engine). An engine thread that reaches an SVD posts a request and
// waits; when every running thread has posted (or left), the last one to
// arrive serves all requests at once on the batch stream: one cooperative
// launch of qrcp_batch_kernel for the pivoted QRs, one of
// jacobi_batch_kernel for the Jacobi SVDs, each workgroup working on one
// request with that request's own barrier counter. Requests that do not fit
// the co-resident grid together go into several launches.
// ---------------------------------------------------------------------------
struct SvdRequest {
  Workspace* ws = nullptr;
  bool qr = false;
  int m = 0, n = 0;       // shape handed to the SVD (Q for qr, X otherwise)
  int jm = 0, jn = 0;     // shape of the Jacobi
  int k = 0;
  bool done = false;
  bool failed = false;
  std::exception_ptr err;
};

struct Batcher {
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::mutex mu;
  std::condition_variable cv;
  int running = 0;
  std::vector<SvdRequest*> pending;
  int max_jacobi_grid = 0;
  int tile_jacobi_grid = 0;   // co-resident grid of jacobi_tile_batch_kernel (0: off)
  int num_cus = 0;
  // statistics: served batches, requests, launches, host wall time serving
  long n_batches = 0, n_requests = 0, n_launches = 0;
  double serve_ms = 0.0, qr_ms_total = 0.0, jac_ms_total = 0.0;
  // argument arrays: pinned staging and device copies
  char* stage_h = nullptr;
  char* stage_d = nullptr;
  size_t stage_cap = 0;

  void init() {
    HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    HIP_CHECK(hipEventCreate(&ev0));
    HIP_CHECK(hipEventCreate(&ev1));
    int dev = 0, per_cu = 0;
    HIP_CHECK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIP_CHECK(hipGetDeviceProperties(&prop, dev));
    num_cus = prop.multiProcessorCount;
    HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void*>(jacobi_batch_kernel), 256, 0));
    max_jacobi_grid = std::min(per_cu * num_cus, 1024);
    HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(small_svd_batch_kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)Workspace::kSmallLdsMax));
    require(max_jacobi_grid > 0, PTE_ERR_DEVICE, "batched Jacobi kernel cannot be resident");
    // the LDS-tile variant: one workgroup per compute unit
    const size_t tile_lds = Workspace::tile_lds_bytes(kTileMaxRows);
    HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(jacobi_tile_batch_kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)tile_lds));
    HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void*>(jacobi_tile_batch_kernel), 256, tile_lds));
    tile_jacobi_grid = std::min(per_cu * num_cus, 1024);
    if (const char* v = std::getenv("PTE_JACOBI_TILE"))
      if (std::atoi(v) == 0) tile_jacobi_grid = 0;
  }

  void destroy() {
    if (!stream) return;
    (void)hipStreamSynchronize(stream);
    (void)hipFree(stage_d);
    (void)hipHostFree(stage_h);
    (void)hipHostFree(ranks_h); (void)hipHostFree(infos_h);
    (void)hipFree(ranks_d); (void)hipFree(infos_d);
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    (void)hipStreamDestroy(stream);
    stream = nullptr;
  }

  void ensure_stage(size_t bytes) {
    if (bytes <= stage_cap) return;
    HIP_CHECK(hipStreamSynchronize(stream));
    (void)hipFree(stage_d);
    (void)hipHostFree(stage_h);
    const size_t c = std::max<size_t>(bytes, 1 << 16);
    HIP_CHECK(hipHostMalloc((void**)&stage_h, c, hipHostMallocDefault));
    HIP_CHECK(hipMalloc((void**)&stage_d, c));
    stage_cap = c;
  }

  void enter(int count) {
    std::lock_guard<std::mutex> lk(mu);
    running += count;
  }

  // an engine thread is done with its step
  void leave() {
    std::unique_lock<std::mutex> lk(mu);
    running -= 1;
    serve_if_complete(lk);
  }

  int submit(Workspace* ws, bool qr, int m, int n) {
    SvdRequest rq;
    rq.ws = ws; rq.qr = qr; rq.m = m; rq.n = n;
    HIP_CHECK(hipEventRecord(ws->ev_ready, ws->stream));
    std::unique_lock<std::mutex> lk(mu);
    pending.push_back(&rq);
    serve_if_complete(lk);
    cv.wait(lk, [&] { return rq.done; });
    if (rq.failed) std::rethrow_exception(rq.err);
    return rq.k;
  }

  // with the lock held: when every running thread waits, serve them
  void serve_if_complete(std::unique_lock<std::mutex>& lk) {
    if (pending.empty() || (int)pending.size() < running) return;
    std::vector<SvdRequest*> batch;
    batch.swap(pending);
    lk.unlock();
    const auto t0 = std::chrono::steady_clock::now();
    n_batches += 1;
    n_requests += (long)batch.size();
    try {
      execute(batch);
    } catch (...) {
      for (auto* r : batch) {
        r->ws->svd_stream = r->ws->stream;
        if (!r->failed) { r->failed = true; r->err = std::current_exception(); }
      }
    }
    serve_ms += std::chrono::duration<double, std::milli>(
                    std::chrono::steady_clock::now() - t0).count();
    lk.lock();
    for (auto* r : batch) r->done = true;
    cv.notify_all();
  }

  // greedy chunks of consecutive requests whose minimum grids fit `cap`
  template <class MinGrid>
  static std::vector<std::pair<size_t, size_t>> chunks(const std::vector<SvdRequest*>& rs,
                                                       int cap, MinGrid min_grid) {
    std::vector<std::pair<size_t, size_t>> out;
    size_t i = 0;
    while (i < rs.size()) {
      size_t j = i;
      long used = 0;
      while (j < rs.size() && (j == i || used + min_grid(rs[j]) <= cap)) {
        used += min_grid(rs[j]);
        ++j;
      }
      out.emplace_back(i, j);
      i = j;
    }
    return out;
  }

  // host image of several argument arrays, copied to the device at once
  struct Pack {
    std::vector<char> h;
    size_t add(const void* p, size_t bytes) {
      const size_t off = (h.size() + 255) / 256 * 256;
      h.resize(off + bytes);
      if (bytes) std::memcpy(h.data() + off, p, bytes);
      return off;
    }
  };

  // copy a pack to the device (after the previous one has been consumed);
  // returns the device base address
  char* upload(const Pack& pk) {
    ensure_stage(pk.h.size());
    HIP_CHECK(hipStreamSynchronize(stream));
    std::memcpy(stage_h, pk.h.data(), pk.h.size());
    HIP_CHECK(hipMemcpyAsync(stage_d, stage_h, pk.h.size(), hipMemcpyHostToDevice, stream));
    return stage_d;
  }

  // starts and workgroup->problem map of a batched cooperative launch
  static void layout(const std::vector<int>& grids, std::vector<int>* starts,
                     std::vector<int>* map) {
    starts->assign(grids.size() + 1, 0);
    for (size_t p = 0; p < grids.size(); ++p) (*starts)[p + 1] = (*starts)[p] + grids[p];
    map->assign((*starts)[grids.size()], 0);
    for (size_t p = 0; p < grids.size(); ++p)
      for (int b = (*starts)[p]; b < (*starts)[p + 1]; ++b) (*map)[b] = (int)p;
  }

  // workgroups per request: its minimum, plus a share of what is left of
  // `cap` proportional to what it could still use (sum <= cap when the
  // minima fit)
  static std::vector<int> allocate(const std::vector<int>& nat, const std::vector<int>& mn,
                                   int cap) {
    long smin = 0, sextra = 0;
    for (size_t i = 0; i < nat.size(); ++i) {
      smin += mn[i];
      sextra += std::max(0, nat[i] - mn[i]);
    }
    const long rem = std::max(0L, (long)cap - smin);
    std::vector<int> g(nat.size());
    for (size_t i = 0; i < nat.size(); ++i) {
      const long extra = std::max(0, nat[i] - mn[i]);
      const long add = (sextra <= rem) ? extra : (sextra > 0 ? extra * rem / sextra : 0);
      g[i] = mn[i] + (int)add;
    }
    return g;
  }

  static void fail(SvdRequest* r) {
    r->failed = true;
    r->err = std::current_exception();
  }

  // pinned host arrays for the gathered QR ranks and SVD summaries
  int* ranks_h = nullptr;
  SvdInfo* infos_h = nullptr;
  int* ranks_d = nullptr;
  SvdInfo* infos_d = nullptr;
  size_t gather_cap = 0;

  void ensure_gather(size_t n) {
    if (n <= gather_cap) return;
    HIP_CHECK(hipStreamSynchronize(stream));
    (void)hipHostFree(ranks_h); (void)hipHostFree(infos_h);
    (void)hipFree(ranks_d); (void)hipFree(infos_d);
    const size_t c = std::max<size_t>(n, 256);
    HIP_CHECK(hipHostMalloc((void**)&ranks_h, c * sizeof(int), hipHostMallocDefault));
    HIP_CHECK(hipHostMalloc((void**)&infos_h, c * sizeof(SvdInfo), hipHostMallocDefault));
    HIP_CHECK(hipMalloc((void**)&ranks_d, c * sizeof(int)));
    HIP_CHECK(hipMalloc((void**)&infos_d, c * sizeof(SvdInfo)));
    gather_cap = c;
  }

  // small SVDs: one workgroup each, all in one launch
  void run_small(std::vector<SvdRequest*>& rs) {
    ensure_gather(rs.size());
    std::vector<SmallSvdArgs> args;
    std::vector<SvdRequest*> in;
    size_t lds = 0;
    for (auto* r : rs) {
      try {
        args.push_back(r->ws->small_prepare(r->m, r->n, infos_d + in.size()));
        lds = std::max(lds, Workspace::small_lds(r->m, r->n));
        in.push_back(r);
      } catch (...) {
        fail(r);
      }
    }
    if (in.empty()) return;
    for (auto& sa : args) sa.lds_bytes = (long)lds;
    Pack pk;
    const size_t oa = pk.add(args.data(), args.size() * sizeof(SmallSvdArgs));
    char* base = upload(pk);
    HIP_CHECK(hipEventRecord(ev0, stream));
    small_svd_batch_kernel<<<(unsigned)in.size(), kSmallBatchThreads, lds, stream>>>(
        reinterpret_cast<const SmallSvdArgs*>(base + oa));
    HIP_CHECK(hipEventRecord(ev1, stream));
    HIP_CHECK(hipGetLastError());
    n_launches += 1;
    HIP_CHECK(hipMemcpyAsync(infos_h, infos_d, in.size() * sizeof(SvdInfo),
                             hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
    jac_ms_total += ms;
    for (size_t i = 0; i < in.size(); ++i) {
      SvdRequest* r = in[i];
      try {
        *r->ws->info_h = infos_h[i];
        r->k = r->ws->small_finish(r->m, r->n, ms);
      } catch (...) {
        fail(r);
      }
    }
  }

  void run_qr(std::vector<SvdRequest*>& qrs) {
    int max_m = 0;
    for (auto* r : qrs) max_m = std::max(max_m, r->m);
    const size_t lds = (size_t)max_m * sizeof(double2);
    int per_cu = 0;
    HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void*>(qrcp_batch_kernel), 256, lds));
    const int cap = std::min(per_cu * num_cus, 1024);
    require(cap > 0, PTE_ERR_DEVICE, "batched QR kernel cannot be resident");
    auto ming = [](SvdRequest* r) { return Workspace::qr_min_grid(r->n); };
    for (auto& ch : chunks(qrs, cap, ming)) {
      std::vector<int> nat, mn;
      for (size_t i = ch.first; i < ch.second; ++i) {
        nat.push_back(qrs[i]->ws->qr_natural_grid(qrs[i]->n));
        mn.push_back(Workspace::qr_min_grid(qrs[i]->n));
      }
      const std::vector<int> alloc = allocate(nat, mn, cap);
      std::vector<QrcpArgs> args;
      std::vector<ZeroItem> zeros;
      std::vector<int> grids;
      std::vector<SvdRequest*> in;
      for (size_t i = ch.first; i < ch.second; ++i) {
        SvdRequest* r = qrs[i];
        try {
          const int g = alloc[i - ch.first];
          args.push_back(r->ws->qr_prepare(r->m, r->n, g));
          zeros.push_back(ZeroItem{r->ws->bar, (int)kBarrierWords, 0});
          grids.push_back(g);
          in.push_back(r);
        } catch (...) {
          fail(r);
        }
      }
      if (in.empty()) continue;
      std::vector<int> starts, map;
      layout(grids, &starts, &map);
      const int total = starts.back();
      require(total <= cap, PTE_ERR_CAPACITY, "internal: batched QR grid too large");
      ensure_gather(in.size());
      Pack pk;
      const size_t oa = pk.add(args.data(), args.size() * sizeof(QrcpArgs));
      const size_t os = pk.add(starts.data(), starts.size() * sizeof(int));
      const size_t om = pk.add(map.data(), map.size() * sizeof(int));
      const size_t oz = pk.add(zeros.data(), zeros.size() * sizeof(ZeroItem));
      char* base = upload(pk);
      const QrcpArgs* da = reinterpret_cast<const QrcpArgs*>(base + oa);
      const int* ds = reinterpret_cast<const int*>(base + os);
      const int* dm = reinterpret_cast<const int*>(base + om);
      zero_items_kernel<<<(int)zeros.size(), 256, 0, stream>>>(
          reinterpret_cast<const ZeroItem*>(base + oz));
      void* kargs[] = {&da, &ds, &dm};
      HIP_CHECK(hipEventRecord(ev0, stream));
      HIP_CHECK(hipLaunchCooperativeKernel(
          reinterpret_cast<const void*>(qrcp_batch_kernel), dim3(total), dim3(256), kargs,
          lds, stream));
      n_launches += 1;
      HIP_CHECK(hipEventRecord(ev1, stream));
      gather_qr_rank_kernel<<<1, 256, 0, stream>>>(da, (int)in.size(), ranks_d);
      HIP_CHECK(hipMemcpyAsync(ranks_h, ranks_d, in.size() * sizeof(int),
                               hipMemcpyDeviceToHost, stream));
      HIP_CHECK(hipStreamSynchronize(stream));
      float ms = 0.f;
      HIP_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
      qr_ms_total += ms;
      std::vector<RhArgs> rh;
      long max_elems = 0;
      for (size_t i = 0; i < in.size(); ++i) {
        SvdRequest* r = in[i];
        try {
          const int rk = ranks_h[i];
          rh.push_back(r->ws->qr_finish_host(r->m, r->n, rk, ms));
          r->jm = r->n;
          r->jn = rk;
          max_elems = std::max(max_elems, (long)r->n * rk);
        } catch (...) {
          fail(r);
        }
      }
      if (rh.empty()) continue;
      Pack pr;
      const size_t orh = pr.add(rh.data(), rh.size() * sizeof(RhArgs));
      char* b2 = upload(pr);
      const int gx = std::max(1, std::min<int>((int)((max_elems + 255) / 256), 1024));
      build_rh_batch_kernel<<<dim3(gx, (unsigned)rh.size()), 256, 0, stream>>>(
          reinterpret_cast<const RhArgs*>(b2 + orh));
      HIP_CHECK(hipGetLastError());
    }
  }

  void run_jacobi(std::vector<SvdRequest*>& jr) {
    // the LDS-tile kernel when its grid holds the batch's block pairs; a
    // chunk whose slices do not all fit the tile takes the streaming kernel
    // (its grid within the same cap)
    const bool tile_try = tile_jacobi_grid > 0;
    const int cap = tile_try ? std::min(tile_jacobi_grid, max_jacobi_grid) : max_jacobi_grid;
    auto ming = [](SvdRequest* r) { return Workspace::jacobi_blocks(r->jn) / 2; };
    for (auto& ch : chunks(jr, cap, ming)) {
      std::vector<int> nat, mn;
      for (size_t i = ch.first; i < ch.second; ++i) {
        nat.push_back(jr[i]->ws->jacobi_natural_grid(jr[i]->jm, jr[i]->jn));
        mn.push_back(Workspace::jacobi_blocks(jr[i]->jn) / 2);
      }
      const std::vector<int> alloc = allocate(nat, mn, cap);
      std::vector<JacobiArgs> args;
      std::vector<SvdPostArgs> post;
      std::vector<ZeroItem> zeros;
      std::vector<int> grids;
      std::vector<SvdRequest*> in;
      int max_n = 1;
      for (size_t i = ch.first; i < ch.second; ++i) {
        SvdRequest* r = jr[i];
        try {
          int g = 0;
          args.push_back(r->ws->jacobi_prepare(r->jm, r->jn, alloc[i - ch.first], &g));
          post.push_back(r->ws->post_args(r->jm, r->jn));
          Workspace* w = r->ws;
          zeros.push_back(ZeroItem{w->bar, w->jacobi_sync_words, 0});
          zeros.push_back(ZeroItem{w->jskip, w->jskip_words, 0});
          zeros.push_back(ZeroItem{reinterpret_cast<unsigned*>(w->flags), w->max_sweeps, 0});
          zeros.push_back(ZeroItem{reinterpret_cast<unsigned*>(w->fro2), 2, 0});
          grids.push_back(g);
          in.push_back(r);
          max_n = std::max(max_n, r->jn);
        } catch (...) {
          fail(r);
        }
      }
      if (in.empty()) continue;
      std::vector<int> starts, map;
      layout(grids, &starts, &map);
      const int total = starts.back();
      require(total <= cap, PTE_ERR_CAPACITY, "internal: batched Jacobi grid too large");
      ensure_gather(in.size());
      Pack pk;
      const size_t oa = pk.add(args.data(), args.size() * sizeof(JacobiArgs));
      const size_t os = pk.add(starts.data(), starts.size() * sizeof(int));
      const size_t om = pk.add(map.data(), map.size() * sizeof(int));
      const size_t oz = pk.add(zeros.data(), zeros.size() * sizeof(ZeroItem));
      const size_t op = pk.add(post.data(), post.size() * sizeof(SvdPostArgs));
      char* base = upload(pk);
      const JacobiArgs* da = reinterpret_cast<const JacobiArgs*>(base + oa);
      const int* ds = reinterpret_cast<const int*>(base + os);
      const int* dm = reinterpret_cast<const int*>(base + om);
      const SvdPostArgs* dp = reinterpret_cast<const SvdPostArgs*>(base + op);
      zero_items_kernel<<<(int)zeros.size(), 256, 0, stream>>>(
          reinterpret_cast<const ZeroItem*>(base + oz));
      void* kargs[] = {&da, &ds, &dm};
      int tile_rows = 0;
      for (const JacobiArgs& ja : args) tile_rows = std::max(tile_rows, ja.xrows_per + ja.vrows_per);
      const bool tile = tile_try && tile_rows <= kTileMaxRows;
      HIP_CHECK(hipEventRecord(ev0, stream));
      if (tile) {
        HIP_CHECK(hipLaunchCooperativeKernel(
            reinterpret_cast<const void*>(jacobi_tile_batch_kernel), dim3(total), dim3(256),
            kargs, Workspace::tile_lds_bytes(tile_rows), stream));
      } else {
        HIP_CHECK(hipLaunchCooperativeKernel(
            reinterpret_cast<const void*>(jacobi_batch_kernel), dim3(total), dim3(256), kargs,
            0, stream));
      }
      n_launches += 1;
      HIP_CHECK(hipEventRecord(ev1, stream));
      const unsigned np = (unsigned)in.size();
      column_norm_batch_kernel<<<dim3(max_n, np), 256, 0, stream>>>(dp);
      rank_sort_batch_kernel<<<dim3((max_n + 255) / 256, np), 256, 0, stream>>>(dp);
      truncate_batch_kernel<<<np, 1024, 0, stream>>>(dp, infos_d);
      HIP_CHECK(hipGetLastError());
      HIP_CHECK(hipMemcpyAsync(infos_h, infos_d, in.size() * sizeof(SvdInfo),
                               hipMemcpyDeviceToHost, stream));
      HIP_CHECK(hipStreamSynchronize(stream));
      float ms = 0.f;
      HIP_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
      jac_ms_total += ms;
      for (size_t i = 0; i < in.size(); ++i) {
        SvdRequest* r = in[i];
        try {
          *r->ws->info_h = infos_h[i];
          r->k = r->ws->jacobi_check(r->jm, r->jn, ms);
        } catch (...) {
          fail(r);
        }
      }
    }
  }

  void execute(std::vector<SvdRequest*>& batch) {
    for (auto* r : batch) {
      HIP_CHECK(hipStreamWaitEvent(stream, r->ws->ev_ready, 0));
      r->ws->svd_stream = stream;
    }
    std::vector<SvdRequest*> qrs, smalls;
    for (auto* r : batch) {
      if (r->qr && r->ws->small_ok(r->m, r->n)) smalls.push_back(r);
      else if (r->qr) qrs.push_back(r);
      else { r->jm = r->m; r->jn = r->n; }
    }
    if (!smalls.empty()) run_small(smalls);
    if (!qrs.empty()) run_qr(qrs);
    std::vector<SvdRequest*> jr;
    for (auto* r : batch)
      if (!r->failed && !(r->qr && r->ws->small_ok(r->m, r->n))) jr.push_back(r);
    if (!jr.empty()) run_jacobi(jr);
    for (auto* r : batch) r->ws->svd_stream = r->ws->stream;
  }
};

int Workspace::batched_svd(bool qr, int m, int n) { return batcher->submit(this, qr, m, n); }

// ---------------------------------------------------------------------------
// chain operations (node_array.py) on device sites [l][p][r]
// ---------------------------------------------------------------------------
struct Chain {
  Workspace* ws = nullptr;
  std::vector<DTensor> sites;  // d[0]=left, d[1]=phys, d[2]=right

  hipStream_t st() const { return ws->stream; }

  void set_site(int i, DTensor t) {
    dfree(sites[i], st());
    sites[i] = t;
  }

  // Truncating SVD sweep left to right from `start` to `stop` (exclusive of
  // the SVD at `stop`) -- NodeArray.svd_sweep, node_array.py:251-273.
  void sweep_right(int start, int stop) {
    for (int i = start; i < stop; ++i) {
      DTensor& A = sites[i];
      const int Dl = A.d[0], P = A.d[1], Dr = A.d[2];
      const long cnt = (long)Dl * P * Dr;
      const int m = Dl * P;
      int k;
      DTensor U;
      ws->svh.ensure((size_t)k_bound(m, Dr) * Dr, st());
      if (m <= kQrMaxRows) {
        // M = A as (Dl*P) x Dr, column-major copy for the pivoted QR
        ws->Q.ensure(cnt, st());
        transpose_kernel<<<grid_for(cnt), 256, 0, st()>>>(A.p, m, Dr, ws->Q.p);
        k = ws->svd_qrcp(m, Dr, rank_key(1, (int)sites.size() - 1 - i));
        // new site [l][p][j]: row i = (l,p) -> i*k + j
        U = dalloc(Dl, P, k, 1, st());
        ws->emit_left_q(1, /*dj*/ 1, 0, /*da*/ 0, 1, 0, 0, U.p, k);
        // svh[j][r]
        ws->emit_svh(1, /*dj*/ Dr, 0, /*da*/ 1, 0, 0, 0, ws->svh.p, k);
      } else {
        // M = A as (Dl*P) x Dr; Jacobi on X = M^H (Dr x Dl*P),
        // X[r + c*Dr] = conj(A[c*Dr + r])
        ws->X.ensure(cnt, st());
        conj_copy_kernel<<<grid_for(cnt), 256, 0, st()>>>(A.p, cnt, ws->X.p);
        k = ws->svd(Dr, m);
        // new site [l][p][j] = V[(l*P+p) + perm[j]*ldv]
        U = dalloc(Dl, P, k, 1, st());
        ws->emit(ws->V.p, m, m, 1, /*dj*/ 1, 0, /*da*/ 0, 1, 0, 0, 0, U.p, k);
        // svh[j][r] = conj(Z[r + perm[j]*Dr])
        ws->emit(ws->X.p, Dr, Dr, 1, /*dj*/ Dr, 0, /*da*/ 1, 0, 0, 0, 1, ws->svh.p, k);
      }
      ws->log_svd(m, Dr, k);
      // next site: [j][q][s] = sum_r svh[j][r] next[r][q][s]
      DTensor& Bn = sites[i + 1];
      require(Bn.d[0] == Dr, PTE_ERR_ARG, "internal: bond mismatch in sweep");
      const int Pq = Bn.d[1], Ds = Bn.d[2];
      DTensor Nn = dalloc(k, Pq, Ds, 1, st());
      // the absorption of S V^H stays complex128 in every precision form
      ws->gemm(k, Pq * Ds, Dr, ws->svh.p, Dr, 1, Bn.p, (long)Pq * Ds, 1, Nn.p,
               (long)Pq * Ds, 1);
      set_site(i, U);
      set_site(i + 1, Nn);
    }
  }

  // Truncating SVD sweep right to left from `start` down to `stop`
  // (exclusive) -- NodeArray.svd_sweep, node_array.py:274-296.
  void sweep_left(int start, int stop) {
    for (int i = start; i > stop; --i) {
      DTensor& A = sites[i];
      const int Dl = A.d[0], P = A.d[1], Dr = A.d[2];
      const long cnt = (long)Dl * P * Dr;
      const int m = P * Dr;
      int k;
      DTensor U;
      ws->svh.ensure((size_t)k_bound(m, Dl) * Dl, st());
      if (m <= kQrMaxRows) {
        // M[(p,r), l] = A[l][p][r] is column-major (P*Dr) x Dl in A's memory
        ws->Q.ensure(cnt, st());
        HIP_CHECK(hipMemcpyAsync(ws->Q.p, A.p, cnt * sizeof(double2),
                                 hipMemcpyDeviceToDevice, st()));
        k = ws->svd_qrcp(m, Dl);
        // new site [j][p][r]: row i = (p,r) -> j*m + i
        U = dalloc(k, P, Dr, 1, st());
        ws->emit_left_q(1, /*dj*/ m, 0, /*da*/ 1, 0, 0, 0, U.p, k);
        // svh[j][l]
        ws->emit_svh(1, /*dj*/ Dl, 0, /*da*/ 1, 0, 0, 0, ws->svh.p, k);
      } else {
        // X = M^H (Dl x P*Dr): X[l + c*Dl] = conj(A[l*(P*Dr) + c])
        ws->X.ensure(cnt, st());
        conj_transpose_kernel<<<grid_for(cnt), 256, 0, st()>>>(A.p, Dl, m, ws->X.p);
        k = ws->svd(Dl, m);
        // new site [j][p][r] = V[(p*Dr + r) + perm[j]*ldv]
        U = dalloc(k, P, Dr, 1, st());
        ws->emit(ws->V.p, m, m, 1, /*dj*/ m, 0, /*da*/ 1, 0, 0, 0, 0, U.p, k);
        // svh[j][l] = conj(Z[l + perm[j]*Dl])
        ws->emit(ws->X.p, Dl, Dl, 1, /*dj*/ Dl, 0, /*da*/ 1, 0, 0, 0, 1, ws->svh.p, k);
      }
      ws->log_svd(m, Dl, k);
      // previous site: [l'][q][j] = sum_l prev[l'][q][l] svh[j][l]
      DTensor& Bp = sites[i - 1];
      require(Bp.d[2] == Dl, PTE_ERR_ARG, "internal: bond mismatch in sweep");
      const int Dl2 = Bp.d[0], Pq = Bp.d[1];
      DTensor Np = dalloc(Dl2, Pq, k, 1, st());
      ws->gemm(Dl2 * Pq, k, Dl, Bp.p, Dl, 1, ws->svh.p, 1, Dl, Np.p, k, 1);
      set_site(i, U);
      set_site(i - 1, Np);
    }
  }

  // Zip `mpo` into the last mpo.size() sites, right to left, with a truncated
  // SVD at every site but the leftmost -- NodeArray.zip_up, node_array.py:413-554
  // with direction 'left'.
  void zip_left(std::vector<Node>& mpo) {
    const int n = (int)mpo.size();
    const int first = (int)sites.size() - n;
    require(first >= 0, PTE_ERR_ARG, "internal: MPO longer than the MPS");
    // identity carry over (right bond of the last site, east leg of the last node)
    int Dx, Cr, Cv;  // carry dims: [r][x][v]
    {
      const DTensor& A = sites[first + n - 1];
      const Node& W = mpo[n - 1];
      Cr = A.d[2]; Cv = W.wr; Dx = Cr * Cv;
      ws->carry.ensure((size_t)Cr * Dx * Cv, st());
      identity_carry_kernel<<<grid_for((long)Cr * Dx * Cv), 256, 0, st()>>>(
          ws->carry.p, Cr, Cv);
    }
    for (int j = n - 1; j >= 0; --j) {
      const int ia = first + j;
      DTensor& A = sites[ia];
      Node& W = mpo[j];
      const int Dl = A.d[0], P = A.d[1], Dr = A.d[2];
      require(Dr == Cr && W.wr == Cv && W.n == P, PTE_ERR_ARG,
              "internal: leg mismatch in zip-up");
      const int Wl = W.wl, S = W.s;
      // G[(l,p),(x,v)] = sum_r A[(l,p), r] carry[r, (x,v)]
      ws->G.ensure((size_t)Dl * P * Dx * Cv, st());
      ws->contract(Dl * P, Dx * Cv, Dr, A.p, Dr, 1, ws->carry.p, (long)Dx * Cv, 1, ws->G.p,
               (long)Dx * Cv, 1);
      MpoApplyArgs ma;
      ma.G = ws->G.p;
      ma.W = W.dense.p;
      ma.I = W.delta.p;
      ma.kind = W.kind;
      ma.Dl = Dl; ma.P = P; ma.Dx = Dx; ma.V = Cv; ma.Wl = Wl; ma.S = S;
      const long total = (long)Dx * S * Dl * Wl;
      if (j == 0) {
        require(Wl == 1, PTE_ERR_ARG, "internal: first node has a west leg");
        DTensor T = dalloc(Dl, S, Dx, 1, st());
        ma.mode = MPO_OUT_SITE;
        ma.out = T.p;
        ma.ldx = 0;
        mpo_apply_kernel<<<grid_for(total), 256, 0, st()>>>(ma);
        set_site(ia, T);
        break;
      }
      const int mrows = Dl * Wl, ncols = Dx * S;
      int k;
      DTensor U;
      ws->carry2.ensure((size_t)Dl * k_bound(mrows, ncols) * Wl, st());
      if (ncols <= kQrMaxRows) {
        // M[(x,s),(l,w)] column-major (Dx*S) x (Dl*Wl) for the pivoted QR
        ws->Q.ensure((size_t)mrows * ncols, st());
        ma.mode = MPO_OUT_M;
        ma.out = ws->Q.p;
        ma.ldx = ncols;
        mpo_apply_kernel<<<grid_for(total), 256, 0, st()>>>(ma);
        k = ws->svd_qrcp(ncols, mrows, rank_key(0, (int)sites.size() - 1 - ia));
        // new site [j][s][x]: row i = x*S + s (a = x, b = s)
        U = dalloc(k, S, Dx, 1, st());
        ws->emit_left_q(S, /*dj*/ (long)S * Dx, 0, /*da*/ 1, 0, /*db*/ Dx, 0, U.p, k);
        // next carry [l][j][w]: column c = l*Wl + w of M (a = l, b = w)
        ws->emit_svh(Wl, /*dj*/ Wl, 0, /*da*/ 0, Wl, /*db*/ 1, 0, ws->carry2.p, k);
      } else {
        // X = M^H with M[(x,s),(l,w)]: X is (Dl*Wl) x (Dx*S)
        ws->X.ensure((size_t)mrows * ncols, st());
        ma.mode = MPO_OUT_SVD_H;
        ma.out = ws->X.p;
        ma.ldx = mrows;
        mpo_apply_kernel<<<grid_for(total), 256, 0, st()>>>(ma);
        k = ws->svd(mrows, ncols);
        // new site [j][s][x] = V[(x*S + s) + perm[j]*ldv]   (a = x, b = s)
        U = dalloc(k, S, Dx, 1, st());
        ws->emit(ws->V.p, ncols, Dx, S, /*dj*/ (long)S * Dx, 0, /*da*/ 1, 0, /*db*/ Dx, 0,
                 0, U.p, k);
        // next carry [l][j][w] = conj(Z[(l*Wl + w) + perm[j]*mrows])   (a = l, b = w)
        ws->emit(ws->X.p, mrows, Dl, Wl, /*dj*/ Wl, 0, /*da*/ 0, Wl, /*db*/ 1, 0, 1,
                 ws->carry2.p, k);
      }
      ws->log_svd(ncols, mrows, k);
      std::swap(ws->carry, ws->carry2);
      Cr = Dl; Dx = k; Cv = Wl;
      set_site(ia, U);
    }
  }

  // Zip `mpo` (same length as the chain) left to right with a truncated SVD
  // at every site but the rightmost -- NodeArray.zip_up with direction
  // 'right', left_index=0, right_index=-1 (TEMPO, tempo_backend.py:539-547).
  void zip_right(std::vector<Node>& mpo) {
    const int n = (int)mpo.size();
    require((int)sites.size() == n, PTE_ERR_ARG, "internal: zip lengths differ");
    // carry [(x,w)][l] with x = (l0, w0): identity
    int Dx, Cl, Cw;
    {
      const DTensor& A = sites[0];
      const Node& W = mpo[0];
      Cl = A.d[0]; Cw = W.wl; Dx = Cl * Cw;
      // carry'[(x*Cw + w)*Cl + l] = delta(x, l*Cw + w): same pattern as the
      // identity carry with (R, V) = (Cl, Cw) read as [(x,w)][l]
      std::vector<cplx> h((size_t)Dx * Cw * Cl, cplx(0.0, 0.0));
      for (int x = 0; x < Dx; ++x)
        for (int w = 0; w < Cw; ++w)
          for (int l = 0; l < Cl; ++l)
            if (x == l * Cw + w) h[((size_t)x * Cw + w) * Cl + l] = 1.0;
      ws->carry.ensure(h.size(), st());
      HIP_CHECK(hipMemcpyAsync(ws->carry.p, h.data(), h.size() * sizeof(double2),
                               hipMemcpyHostToDevice, st()));
      HIP_CHECK(hipStreamSynchronize(st()));
    }
    for (int j = 0; j < n; ++j) {
      DTensor& A = sites[j];
      Node& W = mpo[j];
      const int Dl = A.d[0], P = A.d[1], Dr = A.d[2];
      require(Dl == Cl && W.wl == Cw && W.n == P, PTE_ERR_ARG,
              "internal: leg mismatch in zip-up (right)");
      const int E = W.wr, S = W.s;
      // G[(x,w),(p,r)] = sum_l carry[(x,w), l] A[l, (p,r)]
      ws->G.ensure((size_t)Dx * Cw * P * Dr, st());
      ws->contract(Dx * Cw, P * Dr, Dl, ws->carry.p, Dl, 1, A.p, (long)P * Dr, 1, ws->G.p,
               (long)P * Dr, 1);
      MpoApplyArgs ma;
      ma.G = ws->G.p;
      ma.W = W.dense.p;
      ma.I = W.delta.p;
      ma.kind = W.kind;
      ma.Dl = Dr; ma.P = P; ma.Dx = Dx; ma.V = Cw; ma.Wl = E; ma.S = S;
      const long total = (long)Dx * S * Dr * E;
      if (j == n - 1) {
        require(Dr == 1, PTE_ERR_ARG, "internal: last site has a right bond");
        DTensor T = dalloc(Dx, S, E, 1, st());
        ma.mode = MPO_OUT_SITE;
        ma.out = T.p;
        ma.ldx = 0;
        mpo_apply_right_kernel<<<grid_for(total), 256, 0, st()>>>(ma);
        set_site(j, T);
        break;
      }
      const int mrows = Dr * E, ncols = Dx * S;
      int k;
      DTensor U;
      ws->carry2.ensure((size_t)k_bound(mrows, ncols) * E * Dr, st());
      if (ncols <= kQrMaxRows) {
        // M[(x,s),(r,e)] column-major (Dx*S) x (Dr*E) for the pivoted QR
        ws->Q.ensure((size_t)mrows * ncols, st());
        ma.mode = MPO_OUT_M;
        ma.out = ws->Q.p;
        ma.ldx = ncols;
        mpo_apply_right_kernel<<<grid_for(total), 256, 0, st()>>>(ma);
        k = ws->svd_qrcp(ncols, mrows);
        // new site [x][s][j]: row i = x*S + s -> i*k + j
        U = dalloc(Dx, S, k, 1, st());
        ws->emit_left_q(1, /*dj*/ 1, 0, /*da*/ 0, 1, 0, 0, U.p, k);
        // carry [(j,e)][r]: column c = r*E + e of M (a = r, b = e)
        ws->emit_svh(E, /*dj*/ (long)E * Dr, 0, /*da*/ 1, 0, /*db*/ Dr, 0, ws->carry2.p, k);
      } else {
        // X = M^H with M[(x,s),(r,e)]: X is (Dr*E) x (Dx*S)
        ws->X.ensure((size_t)mrows * ncols, st());
        ma.mode = MPO_OUT_SVD_H;
        ma.out = ws->X.p;
        ma.ldx = mrows;
        mpo_apply_right_kernel<<<grid_for(total), 256, 0, st()>>>(ma);
        k = ws->svd(mrows, ncols);
        // new site [x][s][j] = V[(x*S+s) + perm[j]*ldv]  (a = x*S+s)
        U = dalloc(Dx, S, k, 1, st());
        ws->emit(ws->V.p, ncols, Dx * S, 1, /*dj*/ 1, 0, /*da*/ 0, 1, 0, 0, 0, U.p, k);
        // carry [(j,e)][r] = conj(Z[(r*E + e) + perm[j]*mrows])  (a = r, b = e)
        ws->emit(ws->X.p, mrows, Dr, E, /*dj*/ (long)E * Dr, 0, /*da*/ 1, 0, /*db*/ Dr, 0,
                 1, ws->carry2.p, k);
      }
      ws->log_svd(ncols, mrows, k);
      std::swap(ws->carry, ws->carry2);
      Cl = Dr; Cw = E; Dx = k;
      set_site(j, U);
    }
  }

  void clear() {
    for (auto& s : sites) dfree(s, st());
    sites.clear();
  }
};

// ---------------------------------------------------------------------------
// host construction helpers (oqupy/util.py create_delta layouts)
// ---------------------------------------------------------------------------
struct HostInfl {
  std::vector<cplx> v;
  int rows = 0, cols = 0;  // cols == 0: vector
  bool set = false;
};

static cplx hget(const HostInfl& h, int i, int j) { return h.v[(size_t)i * h.cols + j]; }

// ---------------------------------------------------------------------------
// PT-TEMPO engine (oqupy/backends/pt_tempo_backend.py)
// ---------------------------------------------------------------------------
struct PtEngine {
  pte_pt_config cfg;
  Workspace ws;
  Chain mps;
  std::vector<Node> mpo;
  std::vector<HostInfl> infl;
  std::vector<int> north_map, west_map;
  int num_infl = 0;
  int step = 0;  // 0: not initialised
  std::vector<cplx> sum_north_scaled;
  // caps (device) after pte_pt_compute_caps
  std::vector<DTensor> caps;
  double last_ms = 0.0;

  hipStream_t st() { return ws.stream; }

  void init_engine() {
    ws.init(cfg.epsrel, cfg.chi_cap);
    ws.precision = cfg.precision;
    mps.ws = &ws;
    num_infl = std::min(cfg.num_steps, cfg.dkmax + 1);
    infl.resize(num_infl);
  }

  // pt_tempo_backend.py:105-190
  void initialize() {
    const int d = cfg.dimension, d2 = d * d;
    for (int i = 0; i < num_infl; ++i)
      require(infl[i].set, PTE_ERR_STATE, "influence " + std::to_string(i) + " not set");
    sum_north_scaled.assign(cfg.n_north, cplx((double)d, 0.0));
    const bool uniq = !north_map.empty();
    int n_north = 0, n_west = 0;
    if (uniq) {
      n_north = *std::max_element(north_map.begin(), north_map.end()) + 1;
      n_west = *std::max_element(west_map.begin(), west_map.end()) + 1;
    }
    mps.sites.assign(num_infl, DTensor());
    mpo.assign(num_infl, Node());
    for (int i = 0; i < num_infl; ++i) {
      const HostInfl& h = infl[i];
      if (i == 0) {
        if (uniq) {
          require(h.cols == 0 && h.rows == n_north, PTE_ERR_ARG,
                  "influence 0 must be a vector of the north classes");
          std::vector<cplx> node((size_t)n_west * d2 * n_north, 0.0);
          std::vector<cplx> site((size_t)d2 * n_north, 0.0);
          for (int i1 = 0; i1 < d2; ++i1) {
            const cplx val = h.v[north_map[i1]] / (double)d;
            node[((size_t)west_map[i1] * d2 + i1) * n_north + north_map[i1]] = val;
            site[(size_t)i1 * n_north + north_map[i1]] = val / (double)d;
          }
          mpo[0] = make_dense(node, 1, n_west, d2, n_north, st());
          mps.sites[0] = dupload(site, 1, d2, n_north, 1, st());
        } else {
          require(h.rows == d2 && h.cols == d2, PTE_ERR_ARG, "influence 0 must be d^2 x d^2");
          // create_delta(I0/d, [1,1,0]): W[j][j][i] = I0[i][j]/d
          std::vector<cplx> node((size_t)d2 * d2 * d2, 0.0);
          std::vector<cplx> site((size_t)d2 * d2, 0.0);
          for (int a = 0; a < d2; ++a)
            for (int b = 0; b < d2; ++b) {
              const cplx val = hget(h, a, b) / (double)d;
              node[((size_t)b * d2 + b) * d2 + a] = val;
              site[(size_t)a * d2 + b] = hget(h, b, a) / (double)d / (double)d;  // (I0/d).T/d
            }
          mpo[0] = make_dense(node, 1, d2, d2, d2, st());
          mps.sites[0] = dupload(site, 1, d2, d2, 1, st());
        }
        if (num_infl == 1) {
          // a single node keeps its open right leg, nothing else to build
        }
      } else if (i == num_infl - 1) {
        require(h.cols > 0, PTE_ERR_ARG, "influence must be a matrix");
        // add_singleton(1), add_singleton(3): W[l][0][s][0] = I[l][s]
        mpo[i] = make_dense(h.v, h.rows, 1, h.cols, 1, st());
        mps.sites[i] = dupload(h.v, h.rows, h.cols, 1, 1, st());
      } else {
        require(h.cols > 0, PTE_ERR_ARG, "influence must be a matrix");
        // create_delta(I, [0,1,1,0]): W[l][n][n][l] = I[l][n]
        mpo[i] = make_delta(h.v, h.rows, h.cols, st());
        // create_delta(I/d, [0,1,0]): A[l][p][l] = I[l][p]/d
        std::vector<cplx> site((size_t)h.rows * h.cols * h.rows, 0.0);
        for (int l = 0; l < h.rows; ++l)
          for (int p = 0; p < h.cols; ++p)
            site[((size_t)l * h.cols + p) * h.rows + l] = hget(h, l, p) / (double)d;
        mps.sites[i] = dupload(site, h.rows, h.cols, h.rows, 1, st());
      }
    }
    begin_stats();
    const int n = (int)mps.sites.size();
    mps.sweep_left(n - 1, 0);
    mps.sweep_right(0, n - 1);
    end_stats();
    step = 1;
  }

  void begin_stats() {
    ws.n_svd = 0; ws.sweeps = 0; ws.max_bond = 0;
    HIP_CHECK(hipEventRecord(ws.ev0, st()));
  }
  void end_stats() {
    HIP_CHECK(hipEventRecord(ws.ev1, st()));
    HIP_CHECK(hipEventSynchronize(ws.ev1));
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, ws.ev0, ws.ev1));
    last_ms = ms;
  }

  // pt_tempo_backend.py:192-282
  bool compute_step(const double* last, int rows, int cols) {
    require(step >= 1, PTE_ERR_STATE, "engine not initialised");
    require(step < cfg.num_steps, PTE_ERR_STATE, "process tensor already complete");
    step += 1;
    const bool end_phase = step > cfg.num_steps - num_infl + 1;
    if (ws.rank) ws.rank->step_frac = (float)step / (float)cfg.num_steps;
    begin_stats();
    if (end_phase) {
      require(mpo.size() >= 2, PTE_ERR_STATE, "internal: MPO too short to shrink");
      // drop the last node, close the new last node's east leg with sum_north*d
      free_node(mpo.back(), st());
      mpo.pop_back();
      Node& nd = mpo.back();
      std::vector<cplx> dense = node_dense_host(nd, st());
      require((int)sum_north_scaled.size() == nd.wr, PTE_ERR_ARG,
              "sum_north length does not match the MPO bond");
      std::vector<cplx> closed((size_t)nd.wl * nd.n * nd.s, 0.0);
      for (size_t a = 0; a < closed.size(); ++a) {
        cplx acc = 0.0;
        for (int v = 0; v < nd.wr; ++v) acc += dense[a * nd.wr + v] * sum_north_scaled[v];
        closed[a] = acc;
      }
      const int wl = nd.wl, nn = nd.n, ss = nd.s;
      free_node(nd, st());
      nd = make_dense(closed, wl, nn, ss, 1, st());
    } else {
      if (last != nullptr) {
        std::vector<cplx> v = to_cplx(last, (long)rows * cols);
        Node& nd = mpo.back();
        require(rows == nd.wl && cols == nd.s, PTE_ERR_ARG,
                "replacement influence has the wrong shape");
        const int wl = nd.wl;
        free_node(nd, st());
        nd = make_dense(v, wl, 1, cols, 1, st());
      }
      // join the [[1.0]] site
      std::vector<cplx> one(1, cplx(1.0, 0.0));
      mps.sites.push_back(dupload(one, 1, 1, 1, 1, st()));
    }
    mps.zip_left(mpo);
    mps.sweep_right(step - 2, (int)mps.sites.size() - 1);
    end_stats();
    return step < cfg.num_steps;
  }

  // pt_tempo_backend.py:284-307: [l][p][r] -> [l][r][p] * d
  void get_mpo_tensor(int k, double* out, long cap, int* shape3) {
    require(step >= cfg.num_steps, PTE_ERR_STATE, "process tensor not complete");
    require((int)mps.sites.size() == cfg.num_steps, PTE_ERR_STATE,
            "internal: MPS length differs from the number of steps");
    require(k >= 0 && k < cfg.num_steps, PTE_ERR_ARG, "step index out of range");
    const DTensor& s = mps.sites[k];
    const int Dl = s.d[0], P = s.d[1], Dr = s.d[2];
    shape3[0] = Dl; shape3[1] = Dr; shape3[2] = P;
    require(cap >= (long)Dl * P * Dr, PTE_ERR_ARG, "output buffer too small");
    std::vector<cplx> h((size_t)Dl * P * Dr);
    HIP_CHECK(hipMemcpyAsync(h.data(), s.p, h.size() * sizeof(double2),
                             hipMemcpyDeviceToHost, st()));
    HIP_CHECK(hipStreamSynchronize(st()));
    const double d = cfg.dimension;
    for (int l = 0; l < Dl; ++l)
      for (int r = 0; r < Dr; ++r)
        for (int p = 0; p < P; ++p) {
          const cplx v = h[((size_t)l * P + p) * Dr + r] * d;
          const size_t o = ((size_t)l * Dr + r) * P + p;
          out[2 * o] = v.real();
          out[2 * o + 1] = v.imag();
        }
  }

  // raw MPS site i, [left][physical][right], at any step
  void get_site(int i, double* out, long cap, int* shape3) {
    require(step >= 1, PTE_ERR_STATE, "engine not initialised");
    require(i >= 0 && i < (int)mps.sites.size(), PTE_ERR_ARG, "site index out of range");
    const DTensor& s = mps.sites[i];
    for (int a = 0; a < 3; ++a) shape3[a] = s.d[a];
    require(cap >= s.size(), PTE_ERR_ARG, "output buffer too small");
    HIP_CHECK(hipMemcpyAsync(out, s.p, (size_t)s.size() * sizeof(double2),
                             hipMemcpyDeviceToHost, st()));
    HIP_CHECK(hipStreamSynchronize(st()));
  }

  // process_tensor.py:380-406, with trace_square = vec(1)/d; sites keep the
  // [l][p][r] layout and the factor d of the export:
  // cap_k[a] = sum_{p,b} d*A_k[a][p][b] cap_{k+1}[b] tr2[p]
  void compute_caps() {
    require(step >= cfg.num_steps, PTE_ERR_STATE, "process tensor not complete");
    const int N = cfg.num_steps, d = cfg.dimension;
    for (auto& c : caps) dfree(c, st());
    caps.assign(N + 1, DTensor());
    std::vector<cplx> one(1, cplx(1.0, 0.0));
    caps[N] = dupload(one, 1, 1, 1, 1, st());
    for (int k = N - 1; k >= 0; --k) {
      const DTensor& s = mps.sites[k];
      const int Dl = s.d[0], P = s.d[1], Dr = s.d[2];
      require(P == d * d, PTE_ERR_ARG, "internal: PT site physical leg is not d^2");
      // tmp[a][p] = sum_b A[a][p][b] cap[b]
      DTensor tmp = dalloc(Dl, P, 1, 1, st());
      ws.gemm(Dl * P, 1, Dr, s.p, Dr, 1, caps[k + 1].p, 1, 1, tmp.p, 1, 1);
      // cap[a] = sum_p tmp[a][p] * d * tr2[p]   (d * tr2 = 1 on the diagonal)
      std::vector<cplx> vec(P, 0.0);
      for (int i = 0; i < d; ++i) vec[i * d + i] = 1.0;
      DTensor dv = dupload(vec, P, 1, 1, 1, st());
      DTensor c = dalloc(Dl, 1, 1, 1, st());
      leg_vector_kernel<<<grid_for(Dl), 256, 0, st()>>>(tmp.p, Dl, P, 1, dv.p, c.p);
      dfree(tmp, st());
      dfree(dv, st());
      caps[k] = c;
    }
    HIP_CHECK(hipStreamSynchronize(st()));
  }

  void destroy() {
    for (auto& c : caps) dfree(c, st());
    mps.clear();
    for (auto& nd : mpo) free_node(nd, st());
    mpo.clear();
    if (ws.rank) {
      (void)hipStreamSynchronize(st());
      ws.rank->release();
      delete ws.rank;
      ws.rank = nullptr;
    }
    ws.destroy();
  }

  RankState& rank_state() {
    if (!ws.rank) {
      ws.rank = new RankState();
      ws.rank->init(st());
      ws.rank->dkmax = std::max(cfg.dkmax, 1);
      ws.rank->num_steps = cfg.num_steps;
    }
    return *ws.rank;
  }
};

// ---------------------------------------------------------------------------
// TEMPO engine (oqupy/backends/tempo_backend.py:314-626)
// ---------------------------------------------------------------------------
struct TempoEngine {
  pte_tempo_config cfg;
  Workspace ws;
  Chain mps;
  std::vector<HostInfl> infl;   // dk >= 0
  std::vector<int> north_map, west_map;
  std::vector<Node> mpo;        // reversed: [dk=K ... dk=0]
  int step = -1;                // -1: not initialised
  std::vector<cplx> sum_north, sum_west;
  DTensor sum_north_d;

  hipStream_t st() { return ws.stream; }

  void init_engine() {
    ws.init(cfg.epsrel, cfg.chi_cap);
    ws.precision = cfg.precision;
    mps.ws = &ws;
  }

  HostInfl& slot(int dk) {
    if ((int)infl.size() <= dk) infl.resize(dk + 1);
    return infl[dk];
  }

  // create_delta(I, [1,0,0,1]): W[w=j][n=i][s=i][e=j] = I[i][j] -> delta form
  // with weights Iw[w][p] = I[p][w]
  Node four_legs(const HostInfl& h) {
    std::vector<cplx> Iw((size_t)h.cols * h.rows);
    for (int i = 0; i < h.rows; ++i)
      for (int j = 0; j < h.cols; ++j) Iw[(size_t)j * h.rows + i] = hget(h, i, j);
    Node nd = make_delta(Iw, h.cols, h.rows, st());
    return nd;
  }

  void initialize(const double* rho0, const double* unitary) {
    const int d = cfg.dimension, d2 = d * d;
    std::vector<cplx> U = to_cplx(unitary, (long)d * d);
    // super_u = kron(U, conj(U)), super_u_dagg = kron(U^H, U^T)
    std::vector<cplx> su((size_t)d2 * d2), sud((size_t)d2 * d2);
    for (int a = 0; a < d; ++a)
      for (int b = 0; b < d; ++b)
        for (int c = 0; c < d; ++c)
          for (int e = 0; e < d; ++e) {
            const size_t r = (size_t)a * d + c, col = (size_t)b * d + e;
            su[r * d2 + col] = U[a * d + b] * std::conj(U[c * d + e]);
            sud[r * d2 + col] = std::conj(U[b * d + a]) * U[e * d + c];
          }
    sum_north.assign(cfg.n_north, cplx(1.0, 0.0));
    sum_west.assign(cfg.n_west, cplx(1.0, 0.0));
    sum_north_d = dupload(sum_north, cfg.n_north, 1, 1, 1, st());
    const int count = (cfg.dkmax < 0) ? 1 : cfg.dkmax + 1;
    std::vector<Node> nodes;
    for (int i = 0; i < count; ++i) {
      require((int)infl.size() > i && infl[i].set, PTE_ERR_STATE,
              "influence " + std::to_string(i) + " not set");
      const HostInfl& h = infl[i];
      if (i == 0) {
        if (!north_map.empty()) {
          const int n_north = *std::max_element(north_map.begin(), north_map.end()) + 1;
          const int n_west = *std::max_element(west_map.begin(), west_map.end()) + 1;
          require(h.cols == 0 && h.rows == n_north, PTE_ERR_ARG,
                  "influence 0 must be a vector of the north classes");
          std::vector<cplx> node((size_t)n_west * d2 * n_north * d2, 0.0);
          for (int i1 = 0; i1 < d2; ++i1)
            node[(((size_t)west_map[i1] * d2 + i1) * n_north + north_map[i1]) * d2 + i1] =
                h.v[north_map[i1]];
          nodes.push_back(make_dense(node, n_west, d2, n_north, d2, st()));
        } else {
          require(h.rows == d2 && h.cols == d2, PTE_ERR_ARG, "influence 0 must be d^2 x d^2");
          // W0[w][n][s][e] = delta(w,n,s,e) I0[n][n], then the basis change
          // W'[w][m][s][f] = sum_{n,e} W0[w][n][s][e] sud[n][m] su[f][e]
          std::vector<cplx> node((size_t)d2 * d2 * d2 * d2, 0.0);
          for (int w = 0; w < d2; ++w)
            for (int m = 0; m < d2; ++m)
              for (int s = 0; s < d2; ++s)
                for (int f = 0; f < d2; ++f) {
                  // only n = e = w = s contributes
                  if (w != s) continue;
                  const cplx val = hget(h, w, w) * sud[(size_t)w * d2 + m] *
                                   su[(size_t)f * d2 + w];
                  node[(((size_t)w * d2 + m) * d2 + s) * d2 + f] = val;
                }
          nodes.push_back(make_dense(node, d2, d2, d2, d2, st()));
        }
      } else {
        nodes.push_back(four_legs(h));
      }
    }
    mpo.assign(nodes.rbegin(), nodes.rend());
    std::vector<cplx> r0 = to_cplx(rho0, d2);
    mps.clear();
    mps.sites.push_back(dupload(r0, 1, d2, 1, 1, st()));
    step = 0;
  }

  // a node whose west leg is closed with sum_west (apply_vector, left=True)
  Node close_west(const Node& nd) {
    std::vector<cplx> dense = node_dense_host(nd, st());
    require((int)sum_west.size() == nd.wl, PTE_ERR_ARG,
            "sum_west length does not match the MPO bond");
    std::vector<cplx> closed((size_t)nd.n * nd.s * nd.wr, 0.0);
    for (int w = 0; w < nd.wl; ++w)
      for (size_t a = 0; a < closed.size(); ++a)
        closed[a] += dense[(size_t)w * closed.size() + a] * sum_west[w];
    return make_dense(closed, 1, nd.n, nd.s, nd.wr, st());
  }

  // tempo_backend.py:439-575
  void compute_step(const double* p1, const double* p2, const double* extra, int rows,
                    int cols, double* state_out) {
    require(step >= 0, PTE_ERR_STATE, "engine not initialised");
    step += 1;
    const int d = cfg.dimension, d2 = d * d;
    std::vector<cplx> P1 = to_cplx(p1, (long)d2 * d2), P2 = to_cplx(p2, (long)d2 * d2);
    // window of the MPO (borrowed nodes except the replaced/closed ones)
    std::vector<Node> window;
    std::vector<Node> owned;
    const int K = cfg.dkmax;
    if (K < 0) {
      window = mpo;
      // grow the MPO with influence(len(window)) for the next step
      const int dk = (int)window.size();
      require((int)infl.size() > dk && infl[dk].set, PTE_ERR_STATE,
              "influence " + std::to_string(dk) + " not set");
      mpo.insert(mpo.begin(), four_legs(infl[dk]));
    } else if (step <= K) {
      window.assign(mpo.end() - step, mpo.end());
    } else {
      window = mpo;
      if (extra != nullptr) {
        HostInfl h;
        h.v = to_cplx(extra, (long)rows * cols);
        h.rows = rows; h.cols = cols; h.set = true;
        owned.push_back(four_legs(h));
        window[0] = owned.back();
      }
    }
    owned.push_back(close_west(window[0]));
    window[0] = owned.back();
    // first half-step on the newest site: A[l][q][r] = sum_p P1[q][p] A[l][p][r]
    {
      DTensor& A = mps.sites.back();
      const int Dl = A.d[0], P = A.d[1], Dr = A.d[2];
      require(P == d2, PTE_ERR_ARG, "internal: newest site is not d^2");
      DTensor Pd = dupload(P1, d2, d2, 1, 1, st());
      DTensor N = dalloc(Dl, P, Dr, 1, st());
      for (int l = 0; l < Dl; ++l)
        ws.gemm(P, Dr, P, Pd.p, P, 1, A.p + (long)l * P * Dr, Dr, 1,
                N.p + (long)l * P * Dr, Dr, 1);
      dfree(Pd, st());
      mps.set_site((int)mps.sites.size() - 1, N);
    }
    // trace out the oldest site once the window is full
    if (mps.sites.size() != window.size()) {
      DTensor& A0 = mps.sites[0];
      const int Dl = A0.d[0], P = A0.d[1], Dr = A0.d[2];
      require(Dl == 1 && P == cfg.n_north, PTE_ERR_ARG, "internal: oldest site shape");
      DTensor tr = dalloc(1, Dr, 1, 1, st());
      leg_vector_kernel<<<grid_for(Dr), 256, 0, st()>>>(A0.p, 1, P, Dr, sum_north_d.p,
                                                        tr.p);
      DTensor& A1 = mps.sites[1];
      const int P1d = A1.d[1], Dr1 = A1.d[2];
      DTensor N = dalloc(1, P1d, Dr1, 1, st());
      ws.gemm(1, P1d * Dr1, Dr, tr.p, Dr, 1, A1.p, (long)P1d * Dr1, 1, N.p,
              (long)P1d * Dr1, 1);
      dfree(tr, st());
      mps.set_site(1, N);
      dfree(mps.sites[0], st());
      mps.sites.erase(mps.sites.begin());
    }
    ws.n_svd = 0; ws.sweeps = 0; ws.max_bond = 0;
    mps.zip_right(window);
    const int n = (int)mps.sites.size();
    mps.sweep_left(n - 1, 0);
    for (auto& nd : owned) free_node(nd, st());
    // join prop_2: site[e][p] = P2[p][e]
    {
      std::vector<cplx> site((size_t)d2 * d2);
      for (int e = 0; e < d2; ++e)
        for (int p = 0; p < d2; ++p) site[(size_t)e * d2 + p] = P2[(size_t)p * d2 + e];
      require(mps.sites.back().d[2] == d2, PTE_ERR_ARG, "internal: east leg is not d^2");
      mps.sites.push_back(dupload(site, d2, d2, 1, 1, st()));
    }
    // read out: sum every physical leg but the newest
    DevBuf vec, vec2;
    std::vector<cplx> one(1, cplx(1.0, 0.0));
    vec.ensure(1, st());
    HIP_CHECK(hipMemcpyAsync(vec.p, one.data(), sizeof(double2), hipMemcpyHostToDevice,
                             st()));
    int len = 1;
    for (size_t i = 0; i + 1 < mps.sites.size(); ++i) {
      const DTensor& A = mps.sites[i];
      const int Dl = A.d[0], P = A.d[1], Dr = A.d[2];
      require(Dl == len && P == cfg.n_north, PTE_ERR_ARG, "internal: readout shape");
      // tmp[p][r] = sum_l vec[l] A[l][p][r]; vec'[r] = sum_p tmp[p][r]
      ws.G.ensure((size_t)P * Dr, st());
      ws.gemm(1, P * Dr, Dl, vec.p, Dl, 1, A.p, (long)P * Dr, 1, ws.G.p, (long)P * Dr, 1);
      vec2.ensure(Dr, st());
      leg_vector_kernel<<<grid_for(Dr), 256, 0, st()>>>(ws.G.p, 1, P, Dr, sum_north_d.p,
                                                        vec2.p);
      std::swap(vec, vec2);
      len = Dr;
    }
    const DTensor& L = mps.sites.back();
    require(L.d[0] == len, PTE_ERR_ARG, "internal: readout shape (last)");
    DevBuf outv;
    outv.ensure(d2, st());
    ws.gemm(1, d2, len, vec.p, len, 1, L.p, d2, 1, outv.p, d2, 1);
    std::vector<cplx> h(d2);
    HIP_CHECK(hipMemcpyAsync(h.data(), outv.p, d2 * sizeof(double2),
                             hipMemcpyDeviceToHost, st()));
    HIP_CHECK(hipStreamSynchronize(st()));
    vec.release(st()); vec2.release(st()); outv.release(st());
    for (int i = 0; i < d2; ++i) {
      state_out[2 * i] = h[i].real();
      state_out[2 * i + 1] = h[i].imag();
    }
  }

  void destroy() {
    mps.clear();
    for (auto& nd : mpo) free_node(nd, st());
    mpo.clear();
    dfree(sum_north_d, st());
    ws.destroy();
  }
};

// ---------------------------------------------------------------------------
// compute_dynamics kernels (system_dynamics.py:631-700), for S systems
// (grid y) through one process tensor; v, w are [S][bond][d2], propagators
// [S][N][2][d2][d2]
// ---------------------------------------------------------------------------

// w[a][i] = sum_j (P1 v[a])[j] Tin[j][i]   (Tin = identity when null)
__global__ void __launch_bounds__(256)
dyn_pre_kernel(const double2* v, int A, int d2, long vstride, const double2* P1,
               long pstride, const double2* tin, double2* w) {
  const int sys = blockIdx.y;
  v += sys * vstride;
  w += sys * vstride;
  P1 += sys * pstride;
  const long total = (long)A * d2;
  for (long e = blockIdx.x * 256L + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    const int i = e % d2;
    const long a = e / d2;
    double2 acc = make_double2(0.0, 0.0);
    for (int j = 0; j < d2; ++j) {
      double2 pj = make_double2(0.0, 0.0);
      for (int q = 0; q < d2; ++q) cfma(pj, P1[j * d2 + q], v[a * d2 + q]);
      if (tin) cfma(acc, pj, tin[j * d2 + i]);
      else if (j == i) acc = pj;
    }
    w[e] = acc;
  }
}

// u[b][i] = sum_a w[a][i] T[a][i][b] * scale; v'[b][o] = sum_i u[b][i] Tout[i][o]
// then v''[b][o2] = sum_o P2[o2][o] v'[b][o].  One thread per b; the site is
// read in its [a][i][b] layout (coalesced over b).
__global__ void __launch_bounds__(256)
dyn_mpo_kernel(const double2* w, const double2* T, int A, int d2, int B, long vstride,
               double scale, const double2* tout, const double2* P2, long pstride,
               double2* vout) {
  const int b = blockIdx.x * 256 + threadIdx.x;
  if (b >= B) return;
  const int sys = blockIdx.y;
  w += sys * vstride;
  vout += sys * vstride;
  P2 += sys * pstride;
  double2 u[16];
  for (int i = 0; i < d2; ++i) u[i] = make_double2(0.0, 0.0);
  for (int a = 0; a < A; ++a)
    for (int i = 0; i < d2; ++i)
      cfma(u[i], w[(long)a * d2 + i], T[((long)a * d2 + i) * B + b]);
  double2 vp[16];
  for (int o = 0; o < d2; ++o) {
    double2 acc = make_double2(0.0, 0.0);
    for (int i = 0; i < d2; ++i) {
      if (tout) cfma(acc, u[i], tout[i * d2 + o]);
      else if (i == o) acc = u[i];
    }
    vp[o] = cscale(acc, scale);
  }
  for (int o2 = 0; o2 < d2; ++o2) {
    double2 acc = make_double2(0.0, 0.0);
    for (int o = 0; o < d2; ++o) cfma(acc, P2[o2 * d2 + o], vp[o]);
    vout[(long)b * d2 + o2] = acc;
  }
}

// rho[o] = sum_a cap[a] v[a][o]   (one workgroup per system)
__global__ void __launch_bounds__(256)
dyn_readout_kernel(const double2* v, const double2* cap, int A, int d2, long vstride,
                   double2* rho, long rstride) {
  __shared__ double2 red[256];
  const int sys = blockIdx.y;
  v += sys * vstride;
  rho += sys * rstride;
  for (int o = 0; o < d2; ++o) {
    double2 acc = make_double2(0.0, 0.0);
    for (int a = threadIdx.x; a < A; a += 256) cfma(acc, cap[a], v[(long)a * d2 + o]);
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (threadIdx.x < s) red[threadIdx.x] = cadd(red[threadIdx.x], red[threadIdx.x + s]);
      __syncthreads();
    }
    if (threadIdx.x == 0) rho[o] = red[0];
    __syncthreads();
  }
}

// ---- one system coupled to several environments ----
// The state carries one bond leg per process tensor: v[b_0 .. b_{E-1}][i],
// flattened with the last environment fastest and the Liouville index i
// innermost.

constexpr int kMaxEnvironments = 8;

// w[r][o] = sum_i M[o][i] v[r][i]  (M = P, or M = T^T when trans: w = v T)
__global__ void __launch_bounds__(256)
dyn_rows_kernel(const double2* v, long rows, int d2, const double2* M, int trans, double2* w) {
  const long total = rows * d2;
  for (long e = blockIdx.x * 256L + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    const int o = e % d2;
    const long r = e / d2;
    double2 acc = make_double2(0.0, 0.0);
    for (int i = 0; i < d2; ++i)
      cfma(acc, trans ? M[i * d2 + o] : M[o * d2 + i], v[r * d2 + i]);
    w[e] = acc;
  }
}

// Environment e with the legs before it already advanced:
//   out[l][b][q][o] = scale * sum_i (sum_a v[l][a][q][i] T[a][i][b]) Tout[i][o]
// (Tout = identity when null); one thread per (l, b, q), T read in its
// [a][i][b] layout.
__global__ void __launch_bounds__(256)
dyn_env_kernel(const double2* v, long L, int A, long Q, int d2, const double2* T, int B,
               const double2* tout, double scale, double2* out) {
  const long total = L * B * Q;
  for (long e = blockIdx.x * 256L + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    const long q = e % Q;
    const long t = e / Q;
    const int b = t % B;
    const long l = t / B;
    double2 acc[16];
    for (int i = 0; i < d2; ++i) acc[i] = make_double2(0.0, 0.0);
    for (int a = 0; a < A; ++a) {
      const double2* vr = v + ((l * A + a) * Q + q) * d2;
      for (int i = 0; i < d2; ++i) cfma(acc[i], vr[i], T[((long)a * d2 + i) * B + b]);
    }
    double2* orow = out + ((l * B + b) * Q + q) * d2;
    for (int o = 0; o < d2; ++o) {
      double2 y = make_double2(0.0, 0.0);
      for (int i = 0; i < d2; ++i) {
        if (tout) cfma(y, acc[i], tout[i * d2 + o]);
        else if (i == o) y = acc[i];
      }
      orow[o] = cscale(y, scale);
    }
  }
}

struct CapSet {
  const double2* cap[kMaxEnvironments];
  int bond[kMaxEnvironments];
  int E;
};

// rho[o] = sum_{a_0..a_{E-1}} prod_e cap_e[a_e] v[a_0..a_{E-1}][o]  (one workgroup)
__global__ void __launch_bounds__(256)
dyn_readout_multi_kernel(const double2* v, CapSet cs, int d2, double2* rho) {
  __shared__ double2 red[256];
  long rows = 1;
  for (int e = 0; e < cs.E; ++e) rows *= cs.bond[e];
  for (int o = 0; o < d2; ++o) {
    double2 acc = make_double2(0.0, 0.0);
    for (long f = threadIdx.x; f < rows; f += 256) {
      long rem = f;
      double2 wgt = make_double2(1.0, 0.0);
      for (int e = cs.E - 1; e >= 0; --e) {
        const int x = rem % cs.bond[e];
        rem /= cs.bond[e];
        wgt = cmul(wgt, cs.cap[e][x]);
      }
      cfma(acc, wgt, v[f * d2 + o]);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (threadIdx.x < s) red[threadIdx.x] = cadd(red[threadIdx.x], red[threadIdx.x + s]);
      __syncthreads();
    }
    if (threadIdx.x == 0) rho[o] = red[0];
    __syncthreads();
  }
}

// Device-resident process tensor of one environment, as the dynamics reads it.
struct DynEnv {
  std::vector<const double2*> sites;   // N sites [a][i][b]
  std::vector<int> bonds;              // N + 1
  std::vector<const double2*> caps;    // N + 1
  double scale;                        // factor applied to every site
  const double* tin;                   // host d2 x d2 or null
  const double* tout;
};

// compute_dynamics of one system through E environments
// (oqupy/system_dynamics.py:119-171, 654-700): per step the readout with the
// product of the caps, P1, each environment's site with its transforms, P2.
static void run_dynamics_multi(hipStream_t st, int d2, int N, const std::vector<DynEnv>& env,
                               const double* props, const double* rho0, double* states_out) {
  const int E = (int)env.size();
  require(d2 >= 1 && d2 <= 16, PTE_ERR_ARG, "Liouville dimension must be 1..16");
  require(E >= 1 && E <= kMaxEnvironments, PTE_ERR_ARG, "1 to 8 process tensors");
  for (long i = 0; i < 2L * N * d2 * d2 * 2; ++i)
    require(std::isfinite(props[i]), PTE_ERR_NUMERIC, "non-finite propagator");
  auto bond = [&](int e, int k) { return env[e].bonds[k]; };
  // largest state: while environment e is applied at step k the legs before
  // e are at k + 1 and the others at k
  long maxrows = 1;
  for (int k = 0; k < N; ++k)
    for (int e = 0; e <= E; ++e) {
      long r = 1;
      for (int f = 0; f < E; ++f) r *= bond(f, f < e ? k + 1 : k);
      maxrows = std::max(maxrows, r);
    }
  require(maxrows * d2 <= (1L << 31), PTE_ERR_CAPACITY,
          "product of the bond dimensions too large for the dynamics state");
  auto up = [&](const double* h, long cnt) {
    double2* p = nullptr;
    HIP_CHECK(hipMallocAsync((void**)&p, std::max<long>(cnt, 1) * sizeof(double2), st));
    HIP_CHECK(hipMemcpyAsync(p, h, cnt * sizeof(double2), hipMemcpyHostToDevice, st));
    return p;
  };
  double2* dprops = up(props, 2L * N * d2 * d2);
  std::vector<double2*> tin(E, nullptr), tout(E, nullptr);
  for (int e = 0; e < E; ++e) {
    if (env[e].tin) tin[e] = up(env[e].tin, (long)d2 * d2);
    if (env[e].tout) tout[e] = up(env[e].tout, (long)d2 * d2);
  }
  double2 *v = nullptr, *w = nullptr, *rho = nullptr;
  HIP_CHECK(hipMallocAsync((void**)&v, (size_t)maxrows * d2 * sizeof(double2), st));
  HIP_CHECK(hipMallocAsync((void**)&w, (size_t)maxrows * d2 * sizeof(double2), st));
  HIP_CHECK(hipMallocAsync((void**)&rho, (size_t)(N + 1) * d2 * sizeof(double2), st));
  HIP_CHECK(hipMemcpyAsync(v, rho0, d2 * sizeof(double2), hipMemcpyHostToDevice, st));
  auto caps_at = [&](int k) {
    CapSet cs;
    cs.E = E;
    for (int e = 0; e < kMaxEnvironments; ++e) { cs.cap[e] = nullptr; cs.bond[e] = 1; }
    for (int e = 0; e < E; ++e) { cs.cap[e] = env[e].caps[k]; cs.bond[e] = bond(e, k); }
    return cs;
  };
  for (int k = 0; k < N; ++k) {
    dyn_readout_multi_kernel<<<1, 256, 0, st>>>(v, caps_at(k), d2, rho + (long)k * d2);
    const double2* P1 = dprops + (long)k * 2 * d2 * d2;
    const double2* P2 = P1 + d2 * d2;
    long rows = 1;
    for (int e = 0; e < E; ++e) rows *= bond(e, k);
    dyn_rows_kernel<<<grid_for(rows * d2), 256, 0, st>>>(v, rows, d2, P1, 0, w);
    std::swap(v, w);
    for (int e = 0; e < E; ++e) {
      long L = 1, Q = 1;
      for (int f = 0; f < e; ++f) L *= bond(f, k + 1);
      for (int f = e + 1; f < E; ++f) Q *= bond(f, k);
      const int A = bond(e, k), B = bond(e, k + 1);
      if (tin[e]) {
        dyn_rows_kernel<<<grid_for(L * A * Q * d2), 256, 0, st>>>(v, L * A * Q, d2, tin[e], 1,
                                                                  w);
        std::swap(v, w);
      }
      dyn_env_kernel<<<grid_for(L * B * Q), 256, 0, st>>>(v, L, A, Q, d2, env[e].sites[k], B,
                                                          tout[e], env[e].scale, w);
      std::swap(v, w);
    }
    rows = 1;
    for (int e = 0; e < E; ++e) rows *= bond(e, k + 1);
    dyn_rows_kernel<<<grid_for(rows * d2), 256, 0, st>>>(v, rows, d2, P2, 0, w);
    std::swap(v, w);
  }
  dyn_readout_multi_kernel<<<1, 256, 0, st>>>(v, caps_at(N), d2, rho + (long)N * d2);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpyAsync(states_out, rho, (size_t)(N + 1) * d2 * sizeof(double2),
                           hipMemcpyDeviceToHost, st));
  HIP_CHECK(hipStreamSynchronize(st));
  (void)hipFreeAsync(v, st); (void)hipFreeAsync(w, st); (void)hipFreeAsync(rho, st);
  (void)hipFreeAsync(dprops, st);
  for (int e = 0; e < E; ++e) {
    if (tin[e]) (void)hipFreeAsync(tin[e], st);
    if (tout[e]) (void)hipFreeAsync(tout[e], st);
  }
  HIP_CHECK(hipStreamSynchronize(st));
}

// Runs the dynamics of S systems over device-resident sites [a][i][b]
// (scaled by `scale`) and caps; props: S x N x 2 x d2 x d2, rho0: S x d2,
// states_out: S x (N+1) x d2 (interleaved complex).
static void run_dynamics(hipStream_t st, int d2, int N, const std::vector<const double2*>& sites,
                         const std::vector<int>& bonds, const std::vector<const double2*>& caps,
                         const double* props, const double* tin_h, const double* tout_h,
                         const double* rho0, double scale, double* states_out, int S = 1) {
  require(d2 >= 1 && d2 <= 16, PTE_ERR_ARG, "Liouville dimension must be 1..16");
  require(S >= 1 && S <= 65535, PTE_ERR_ARG, "number of systems must be 1..65535");
  const long pstride = 2L * N * d2 * d2;      // per system, complex entries
  auto up = [&](const double* h, long count) {
    double2* p = nullptr;
    HIP_CHECK(hipMallocAsync((void**)&p, std::max<long>(count, 1) * sizeof(double2), st));
    HIP_CHECK(hipMemcpyAsync(p, h, count * sizeof(double2), hipMemcpyHostToDevice, st));
    return p;
  };
  for (long i = 0; i < (long)S * pstride * 2; ++i)
    require(std::isfinite(props[i]), PTE_ERR_NUMERIC, "non-finite propagator");
  double2* dprops = up(props, (long)S * pstride);
  double2* tin = tin_h ? up(tin_h, (long)d2 * d2) : nullptr;
  double2* tout = tout_h ? up(tout_h, (long)d2 * d2) : nullptr;
  int maxb = 1;
  for (int b : bonds) maxb = std::max(maxb, b);
  const long vstride = (long)maxb * d2, rstride = (long)(N + 1) * d2;
  double2 *v = nullptr, *v2 = nullptr, *w = nullptr, *rho = nullptr;
  HIP_CHECK(hipMallocAsync((void**)&v, (size_t)S * vstride * sizeof(double2), st));
  HIP_CHECK(hipMallocAsync((void**)&v2, (size_t)S * vstride * sizeof(double2), st));
  HIP_CHECK(hipMallocAsync((void**)&w, (size_t)S * vstride * sizeof(double2), st));
  HIP_CHECK(hipMallocAsync((void**)&rho, (size_t)S * rstride * sizeof(double2), st));
  HIP_CHECK(hipMemcpy2DAsync(v, vstride * sizeof(double2), rho0, d2 * sizeof(double2),
                             d2 * sizeof(double2), S, hipMemcpyHostToDevice, st));
  for (int k = 0; k < N; ++k) {
    const int A = bonds[k], B = bonds[k + 1];
    dyn_readout_kernel<<<dim3(1, S), 256, 0, st>>>(v, caps[k], A, d2, vstride,
                                                   rho + (long)k * d2, rstride);
    const double2* P1 = dprops + (long)k * 2 * d2 * d2;
    const double2* P2 = P1 + d2 * d2;
    dyn_pre_kernel<<<dim3(grid_for((long)A * d2), S), 256, 0, st>>>(v, A, d2, vstride, P1,
                                                                    pstride, tin, w);
    dyn_mpo_kernel<<<dim3((B + 255) / 256, S), 256, 0, st>>>(w, sites[k], A, d2, B, vstride,
                                                             scale, tout, P2, pstride, v2);
    std::swap(v, v2);
  }
  dyn_readout_kernel<<<dim3(1, S), 256, 0, st>>>(v, caps[N], bonds[N], d2, vstride,
                                                 rho + (long)N * d2, rstride);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpyAsync(states_out, rho, (size_t)S * rstride * sizeof(double2),
                           hipMemcpyDeviceToHost, st));
  HIP_CHECK(hipStreamSynchronize(st));
  (void)hipFreeAsync(v, st); (void)hipFreeAsync(v2, st); (void)hipFreeAsync(w, st); (void)hipFreeAsync(rho, st);
  (void)hipFreeAsync(dprops, st);
  if (tin) (void)hipFreeAsync(tin, st);
  if (tout) (void)hipFreeAsync(tout, st);
  HIP_CHECK(hipStreamSynchronize(st));
}

template <class F>
static int guarded(F&& f) {
  try {
    f();
    return PTE_OK;
  } catch (const Error& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return PTE_ERR_DEVICE;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return PTE_ERR_DEVICE;
  }
}

// device buffers and stream of one stateless C-ABI call (the rank
// predictor's network, eta), freed on exit
struct CallBuffers {
  hipStream_t st = nullptr;
  std::vector<void*> owned;
  template <class T>
  T* alloc(size_t n) {
    T* p = nullptr;
    HIP_CHECK(hipMallocAsync((void**)&p, std::max<size_t>(n, 1) * sizeof(T), st));
    owned.push_back(p);
    return p;
  }
  template <class T>
  T* upload(const T* h, size_t n) {
    T* p = alloc<T>(n);
    HIP_CHECK(hipMemcpyAsync(p, h, n * sizeof(T), hipMemcpyHostToDevice, st));
    return p;
  }
  CallBuffers() { HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking)); }
  ~CallBuffers() {
    (void)hipStreamSynchronize(st);
    for (void* p : owned) (void)hipFreeAsync(p, st);
    (void)hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
  }
};

}  // namespace pte

using namespace pte;

struct pte_pt_engine { PtEngine impl; };
struct pte_tempo_engine { TempoEngine impl; };
struct pte_pt_batch {
  std::vector<pte_pt_engine*> engines;
  Batcher batcher;
};

static void store_infl(HostInfl& h, const double* data, int rows, int cols) {
  require(data != nullptr && rows > 0 && cols >= 0, PTE_ERR_ARG, "bad influence shape");
  const long count = (cols == 0) ? rows : (long)rows * cols;
  h.v = to_cplx(data, count);
  h.rows = rows;
  h.cols = cols;
  h.set = true;
}

extern "C" {

const char* pte_last_error(void) { return g_last_error.c_str(); }
const char* pte_version(void) { return "ptengine 0.1 gfx950"; }

int pte_pt_create(const pte_pt_config* cfg, pte_pt_engine** out) {
  return guarded([&] {
    require(cfg && out, PTE_ERR_ARG, "null argument");
    require(cfg->dimension >= 1 && cfg->dimension <= 4, PTE_ERR_ARG,
            "dimension must be 1..4");
    require(cfg->num_steps >= 2, PTE_ERR_ARG, "num_steps must be >= 2");
    require(cfg->dkmax >= 1, PTE_ERR_ARG, "dkmax must be >= 1");
    require(cfg->epsrel > 0.0 && std::isfinite(cfg->epsrel), PTE_ERR_ARG,
            "epsrel must be positive");
    require(cfg->precision == PTE_PRECISION_C128 || cfg->precision == PTE_PRECISION_C64,
            PTE_ERR_ARG, "precision must be PTE_PRECISION_C128 or PTE_PRECISION_C64");
    require(cfg->chi_cap >= 1, PTE_ERR_ARG, "chi_cap must be positive");
    require(cfg->n_north >= 1, PTE_ERR_ARG, "n_north must be positive");
    auto* e = new pte_pt_engine();
    e->impl.cfg = *cfg;
    try {
      e->impl.init_engine();
    } catch (...) {
      delete e;
      throw;
    }
    *out = e;
  });
}

int pte_pt_set_influence(pte_pt_engine* e, int dk, const double* data, int rows, int cols) {
  return guarded([&] {
    require(e != nullptr, PTE_ERR_ARG, "null engine");
    require(dk >= 0 && dk < e->impl.num_infl, PTE_ERR_ARG, "dk out of range");
    store_infl(e->impl.infl[dk], data, rows, cols);
  });
}

int pte_pt_set_degeneracy_maps(pte_pt_engine* e, const int* nmap, const int* wmap, int length) {
  return guarded([&] {
    require(e && nmap && wmap, PTE_ERR_ARG, "null argument");
    const int d2 = e->impl.cfg.dimension * e->impl.cfg.dimension;
    require(length == d2, PTE_ERR_ARG, "degeneracy maps must have length d^2");
    e->impl.north_map.assign(nmap, nmap + length);
    e->impl.west_map.assign(wmap, wmap + length);
    for (int i = 0; i < length; ++i)
      require(nmap[i] >= 0 && nmap[i] < d2 && wmap[i] >= 0 && wmap[i] < d2, PTE_ERR_ARG,
              "degeneracy map entry out of range");
  });
}

int pte_pt_initialize(pte_pt_engine* e) {
  return guarded([&] {
    require(e != nullptr, PTE_ERR_ARG, "null engine");
    require(e->impl.step == 0, PTE_ERR_STATE, "already initialised");
    e->impl.initialize();
  });
}

int pte_pt_step(pte_pt_engine* e, const double* last, int rows, int cols, int* more) {
  return guarded([&] {
    require(e && more, PTE_ERR_ARG, "null argument");
    *more = e->impl.compute_step(last, rows, cols) ? 1 : 0;
  });
}

int pte_pt_get_step(pte_pt_engine* e, int* step) {
  return guarded([&] {
    require(e && step, PTE_ERR_ARG, "null argument");
    *step = e->impl.step;
  });
}

int pte_pt_num_sites(pte_pt_engine* e, int* n) {
  return guarded([&] {
    require(e && n, PTE_ERR_ARG, "null argument");
    *n = (int)e->impl.mps.sites.size();
  });
}

int pte_pt_bond_dims(pte_pt_engine* e, int* out, int n) {
  return guarded([&] {
    require(e && out, PTE_ERR_ARG, "null argument");
    const auto& s = e->impl.mps.sites;
    require(n == (int)s.size() - 1, PTE_ERR_ARG, "n must be n_sites - 1");
    for (int i = 0; i < n; ++i) out[i] = s[i].d[2];
  });
}

int pte_pt_get_mpo_tensor(pte_pt_engine* e, int step, double* out, long cap, int* shape3) {
  return guarded([&] {
    require(e && out && shape3, PTE_ERR_ARG, "null argument");
    e->impl.get_mpo_tensor(step, out, cap, shape3);
  });
}

int pte_pt_get_site(pte_pt_engine* e, int i, double* out, long cap, int* shape3) {
  return guarded([&] {
    require(e != nullptr && out != nullptr && shape3 != nullptr, PTE_ERR_ARG,
            "null argument");
    e->impl.get_site(i, out, cap, shape3);
  });
}

int pte_pt_compute_caps(pte_pt_engine* e) {
  return guarded([&] {
    require(e != nullptr, PTE_ERR_ARG, "null engine");
    e->impl.compute_caps();
  });
}

int pte_pt_get_cap(pte_pt_engine* e, int k, double* out, int cap, int* length) {
  return guarded([&] {
    require(e && out && length, PTE_ERR_ARG, "null argument");
    auto& caps = e->impl.caps;
    require(!caps.empty(), PTE_ERR_STATE, "caps not computed");
    require(k >= 0 && k < (int)caps.size(), PTE_ERR_ARG, "cap index out of range");
    const int n = caps[k].d[0];
    *length = n;
    require(cap >= n, PTE_ERR_ARG, "output buffer too small");
    HIP_CHECK(hipMemcpyAsync(out, caps[k].p, n * sizeof(double2), hipMemcpyDeviceToHost,
                             e->impl.st()));
    HIP_CHECK(hipStreamSynchronize(e->impl.st()));
  });
}

int pte_pt_last_step_stats(pte_pt_engine* e, int* n_svd, int* sweeps, int* max_bond,
                           double* ms) {
  return guarded([&] {
    require(e != nullptr, PTE_ERR_ARG, "null engine");
    if (n_svd) *n_svd = e->impl.ws.n_svd;
    if (sweeps) *sweeps = e->impl.ws.sweeps;
    if (max_bond) *max_bond = e->impl.ws.max_bond;
    if (ms) *ms = e->impl.last_ms;
  });
}

int pte_pt_svd_log(pte_pt_engine* e, long max_values) {
  return guarded([&] {
    require(e != nullptr && max_values >= 0, PTE_ERR_ARG, "bad engine or log size");
    Workspace& ws = e->impl.ws;
    ws.log_cap = max_values;
    ws.log_shapes.clear();
    ws.log_values.clear();
  });
}

int pte_pt_get_svd_log(pte_pt_engine* e, int* shapes, int max_svds, double* values,
                       long max_values, int* n_svds, long* n_values) {
  return guarded([&] {
    require(e != nullptr && n_svds != nullptr && n_values != nullptr, PTE_ERR_ARG,
            "null argument");
    const Workspace& ws = e->impl.ws;
    const int ns = (int)(ws.log_shapes.size() / 3);
    const long nv = (long)ws.log_values.size();
    *n_svds = ns;
    *n_values = nv;
    if (shapes != nullptr)
      std::memcpy(shapes, ws.log_shapes.data(),
                  sizeof(int) * 3 * (size_t)std::max(0, std::min(ns, max_svds)));
    if (values != nullptr)
      std::memcpy(values, ws.log_values.data(),
                  sizeof(double) * (size_t)std::max(0L, std::min(nv, max_values)));
  });
}

int pte_pt_svd_kernel_stats(pte_pt_engine* e, long* count, double* ms, double* flops,
                            double* bytes) {
  return guarded([&] {
    require(e != nullptr, PTE_ERR_ARG, "null engine");
    if (count) *count = e->impl.ws.jac_count;
    if (ms) *ms = e->impl.ws.jac_ms;
    if (flops) *flops = e->impl.ws.jac_flops;
    if (bytes) *bytes = e->impl.ws.jac_bytes;
  });
}

int pte_pt_set_rank_mode(pte_pt_engine* e, int mode, const float* params, double alpha,
                         double temperature, int oversample, double margin, double theta) {
  return guarded([&] {
    require(e != nullptr, PTE_ERR_ARG, "null engine");
    require(mode >= 0 && mode <= 3, PTE_ERR_ARG, "rank mode must be 0..3");
    require(mode != 2 || params != nullptr, PTE_ERR_ARG, "the MLP mode needs parameters");
    require(oversample >= 0 && margin >= 0.0 && theta > 0.0 && theta <= 1.0, PTE_ERR_ARG,
            "bad oversample, margin or theta");
    require(e->impl.cfg.precision == PTE_PRECISION_C128, PTE_ERR_ARG,
            "the learned rank form runs with complex128 contractions");
    RankState& r = e->impl.rank_state();
    if (mode == 2) {
      for (int i = 0; i < kRankParams; ++i)
        require(std::isfinite(params[i]), PTE_ERR_NUMERIC, "non-finite MLP parameter");
      if (!r.params) HIP_CHECK(hipMalloc((void**)&r.params, sizeof(float) * kRankParams));
      HIP_CHECK(hipMemcpyAsync(r.params, params, sizeof(float) * kRankParams,
                               hipMemcpyHostToDevice, e->impl.st()));
      HIP_CHECK(hipStreamSynchronize(e->impl.st()));
    }
    r.mode = mode;
    r.alpha = (float)alpha;
    r.temperature = (float)temperature;
    r.oversample = oversample;
    r.margin = (float)margin;
    r.theta = theta;
  });
}

int pte_pt_rank_trace(pte_pt_engine* e, int capacity) {
  return guarded([&] {
    require(e != nullptr && capacity >= 0, PTE_ERR_ARG, "bad arguments");
    RankState& r = e->impl.rank_state();
    hipStream_t st = e->impl.st();
    HIP_CHECK(hipStreamSynchronize(st));
    if (r.rec) HIP_CHECK(hipFree(r.rec));
    r.rec = nullptr;
    r.rec_cap = capacity;
    if (capacity > 0)
      HIP_CHECK(hipMalloc((void**)&r.rec, sizeof(float) * (size_t)capacity * kRankRecord));
    HIP_CHECK(hipMemsetAsync(r.rec_count, 0, sizeof(int), st));
    HIP_CHECK(hipStreamSynchronize(st));
  });
}

int pte_pt_get_rank_trace(pte_pt_engine* e, float* out, int max_records, int* count) {
  return guarded([&] {
    require(e != nullptr && count != nullptr, PTE_ERR_ARG, "null argument");
    RankState* r = e->impl.ws.rank;
    if (!r || !r->rec) {
      *count = 0;
      return;
    }
    hipStream_t st = e->impl.st();
    int c = 0;
    HIP_CHECK(hipMemcpyAsync(&c, r->rec_count, sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    *count = c;
    if (out && max_records > 0) {
      const int n = std::min(c, max_records);
      HIP_CHECK(hipMemcpyAsync(out, r->rec, sizeof(float) * (size_t)n * kRankRecord,
                               hipMemcpyDeviceToHost, st));
      HIP_CHECK(hipStreamSynchronize(st));
    }
  });
}

int pte_pt_rank_stats(pte_pt_engine* e, long* svds, long* capped, long* fallbacks,
                      long* qr_steps) {
  return guarded([&] {
    require(e && svds && capped && fallbacks && qr_steps, PTE_ERR_ARG, "null argument");
    const RankState* r = e->impl.ws.rank;
    *svds = r ? r->svds : 0;
    *capped = r ? r->capped : 0;
    *fallbacks = r ? r->fallbacks : 0;
    *qr_steps = r ? r->qr_steps : 0;
  });
}

// ---- the rank predictor's network on its own (training, inference) ----

int pte_eta_batch(int n_baths, const double* bath_params, const int* cutoff_types,
                  const double* gl_nodes, const double* gl_weights, int n_tau,
                  const int* tau_bath, const double* taus, double* eta_out) {
  return guarded([&] {
    require(n_baths >= 1 && n_tau >= 1 && bath_params && cutoff_types && gl_nodes &&
                gl_weights && tau_bath && taus && eta_out,
            PTE_ERR_ARG, "bad arguments");
    std::vector<EtaBath> baths(n_baths);
    std::vector<double> tmax(n_baths, 0.0);
    for (int i = 0; i < n_tau; ++i) {
      require(tau_bath[i] >= 0 && tau_bath[i] < n_baths, PTE_ERR_ARG, "bad bath index");
      require(std::isfinite(taus[i]), PTE_ERR_ARG, "non-finite time");
      tmax[tau_bath[i]] = std::max(tmax[tau_bath[i]], std::fabs(taus[i]));
    }
    for (int b = 0; b < n_baths; ++b) {
      EtaBath& e = baths[b];
      e.alpha = bath_params[4 * b];
      e.zeta = bath_params[4 * b + 1];
      e.wc = bath_params[4 * b + 2];
      e.temp = bath_params[4 * b + 3];
      e.cutoff = cutoff_types[b];
      require(e.cutoff >= 0 && e.cutoff <= 2, PTE_ERR_ARG, "unknown cutoff type");
      require(std::isfinite(e.alpha) && e.zeta >= 1.0 && e.wc > 0.0 && e.temp >= 0.0,
              PTE_ERR_ARG, "bath parameters: zeta >= 1, w_c > 0, T >= 0 required");
      const double top = e.wc * (e.cutoff == ETA_CUTOFF_EXPONENTIAL ? 40.0
                                 : e.cutoff == ETA_CUTOFF_GAUSSIAN ? 6.5 : 1.0);
      double width = e.wc / 8.0;
      if (tmax[b] > 0.0) width = std::min(width, 1.0 / tmax[b]);
      const double np = std::ceil(top / width);
      require(np <= 1e7, PTE_ERR_CAPACITY, "time range too long for the eta quadrature");
      e.npan = std::max(1, (int)np);
      e.width = top / e.npan;
    }
    CallBuffers buf;   // device buffers freed on exit
    EtaBath* db = buf.upload(baths.data(), baths.size());
    int* dtb = buf.upload(tau_bath, n_tau);
    double* dt = buf.upload(taus, n_tau);
    double* gx = buf.upload(gl_nodes, kEtaNodes);
    double* gw = buf.upload(gl_weights, kEtaNodes);
    double2* out = buf.alloc<double2>(n_tau);
    eta_kernel<<<n_tau, 256, 0, buf.st>>>(db, dtb, dt, n_tau, gx, gw, out);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(eta_out, out, sizeof(double2) * n_tau, hipMemcpyDeviceToHost,
                             buf.st));
    HIP_CHECK(hipStreamSynchronize(buf.st));
  });
}

int pte_rank_param_count(void) { return kRankParams; }
int pte_rank_feature_count(void) { return kRankFeatures; }

int pte_rank_mlp_predict(const float* params, const float* X, int n, float* yhat) {
  return guarded([&] {
    require(params && X && yhat && n >= 1, PTE_ERR_ARG, "bad arguments");
    CallBuffers b;
    float* P = b.upload(params, kRankParams);
    float* dX = b.upload(X, (size_t)n * kRankFeatures);
    float* H = b.alloc<float>((size_t)n * 3 * kRankWidth);
    float* Y = b.alloc<float>(n);
    mlp_forward_kernel<<<n, kRankWidth, 0, b.st>>>(P, dX, nullptr, n, H, Y);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(yhat, Y, sizeof(float) * n, hipMemcpyDeviceToHost, b.st));
    HIP_CHECK(hipStreamSynchronize(b.st));
  });
}

int pte_rank_mlp_grad(const float* params, const float* X, const float* y, int n, float* grad,
                      float* loss) {
  return guarded([&] {
    require(params && X && y && grad && loss && n >= 1, PTE_ERR_ARG, "bad arguments");
    CallBuffers b;
    float* P = b.upload(params, kRankParams);
    float* dX = b.upload(X, (size_t)n * kRankFeatures);
    float* dy = b.upload(y, n);
    float* H = b.alloc<float>((size_t)n * 3 * kRankWidth);
    float* D = b.alloc<float>((size_t)n * 3 * kRankWidth);
    float* Y = b.alloc<float>(n);
    float* d4 = b.alloc<float>(n);
    float* G = b.alloc<float>(kRankParams);
    float* L = b.alloc<float>(1);
    mlp_forward_kernel<<<n, kRankWidth, 0, b.st>>>(P, dX, nullptr, n, H, Y);
    mlp_backward_kernel<<<n, kRankWidth, 0, b.st>>>(P, H, Y, dy, nullptr, n, D, d4);
    mlp_grad_kernel<<<(kRankParams + 255) / 256, 256, 0, b.st>>>(dX, nullptr, H, D, d4, n, G);
    mlp_loss_kernel<<<1, 256, 0, b.st>>>(Y, dy, nullptr, n, L);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(grad, G, sizeof(float) * kRankParams, hipMemcpyDeviceToHost,
                             b.st));
    HIP_CHECK(hipMemcpyAsync(loss, L, sizeof(float), hipMemcpyDeviceToHost, b.st));
    HIP_CHECK(hipStreamSynchronize(b.st));
  });
}

int pte_rank_mlp_train(float* params, const float* X, const float* y, int n, int epochs,
                       int batch, float lr, unsigned seed, float* epoch_loss) {
  return guarded([&] {
    require(params && X && y && n >= 1 && epochs >= 1 && batch >= 1 && lr > 0.f,
            PTE_ERR_ARG, "bad arguments");
    for (long i = 0; i < (long)n * kRankFeatures; ++i)
      require(std::isfinite(X[i]), PTE_ERR_NUMERIC, "non-finite feature");
    for (int i = 0; i < n; ++i)
      require(std::isfinite(y[i]), PTE_ERR_NUMERIC, "non-finite target");
    batch = std::min(batch, n);
    CallBuffers b;
    float* P = b.upload(params, kRankParams);
    float* dX = b.upload(X, (size_t)n * kRankFeatures);
    float* dy = b.upload(y, n);
    float* H = b.alloc<float>((size_t)batch * 3 * kRankWidth);
    float* D = b.alloc<float>((size_t)batch * 3 * kRankWidth);
    float* Y = b.alloc<float>(batch);
    float* d4 = b.alloc<float>(batch);
    float* G = b.alloc<float>(kRankParams);
    float* M = b.alloc<float>(kRankParams);
    float* V = b.alloc<float>(kRankParams);
    const int nb = (n + batch - 1) / batch;
    float* L = b.alloc<float>((size_t)epochs * nb);
    int* idx = b.alloc<int>(n);
    HIP_CHECK(hipMemsetAsync(M, 0, sizeof(float) * kRankParams, b.st));
    HIP_CHECK(hipMemsetAsync(V, 0, sizeof(float) * kRankParams, b.st));
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    unsigned long long state = seed * 6364136223846793005ull + 1442695040888963407ull;
    auto next = [&]() {   // 64-bit LCG (Knuth's MMIX constants), high bits
      state = state * 6364136223846793005ull + 1442695040888963407ull;
      return (unsigned)(state >> 33);
    };
    int t = 0;
    for (int ep = 0; ep < epochs; ++ep) {
      for (int i = n - 1; i > 0; --i) std::swap(order[i], order[next() % (unsigned)(i + 1)]);
      // the host order is copied before the next epoch reshuffles it
      HIP_CHECK(hipMemcpyAsync(idx, order.data(), sizeof(int) * n, hipMemcpyHostToDevice,
                               b.st));
      HIP_CHECK(hipStreamSynchronize(b.st));
      for (int bi = 0; bi < nb; ++bi) {
        const int off = bi * batch, cnt = std::min(batch, n - off);
        mlp_forward_kernel<<<cnt, kRankWidth, 0, b.st>>>(P, dX, idx + off, cnt, H, Y);
        mlp_loss_kernel<<<1, 256, 0, b.st>>>(Y, dy, idx + off, cnt, L + (long)ep * nb + bi);
        mlp_backward_kernel<<<cnt, kRankWidth, 0, b.st>>>(P, H, Y, dy, idx + off, cnt, D, d4);
        mlp_grad_kernel<<<(kRankParams + 255) / 256, 256, 0, b.st>>>(dX, idx + off, H, D, d4,
                                                                    cnt, G);
        ++t;
        mlp_adam_kernel<<<(kRankParams + 255) / 256, 256, 0, b.st>>>(
            P, G, M, V, kRankParams, lr, 0.9f, 0.999f, 1e-8f, t);
      }
      HIP_CHECK(hipGetLastError());
    }
    std::vector<float> lh((size_t)epochs * nb);
    HIP_CHECK(hipMemcpyAsync(lh.data(), L, sizeof(float) * lh.size(), hipMemcpyDeviceToHost,
                             b.st));
    HIP_CHECK(hipMemcpyAsync(params, P, sizeof(float) * kRankParams, hipMemcpyDeviceToHost,
                             b.st));
    HIP_CHECK(hipStreamSynchronize(b.st));
    if (epoch_loss)
      for (int ep = 0; ep < epochs; ++ep) {
        double acc = 0.0, w = 0.0;
        for (int bi = 0; bi < nb; ++bi) {
          const int cnt = std::min(batch, n - bi * batch);
          acc += (double)lh[(size_t)ep * nb + bi] * cnt;
          w += cnt;
        }
        epoch_loss[ep] = (float)(acc / w);
      }
  });
}

void pte_pt_destroy(pte_pt_engine* e) {
  if (!e) return;
  guarded([&] { e->impl.destroy(); });
  delete e;
}

int pte_pt_batch_create(pte_pt_engine* const* engines, int count, pte_pt_batch** out) {
  return guarded([&] {
    require(engines && out && count >= 1, PTE_ERR_ARG, "bad batch arguments");
    for (int i = 0; i < count; ++i) {
      require(engines[i] != nullptr, PTE_ERR_ARG, "null engine in batch");
      for (int j = 0; j < i; ++j)
        require(engines[j] != engines[i], PTE_ERR_ARG, "engine listed twice in batch");
    }
    auto* b = new pte_pt_batch();
    b->engines.assign(engines, engines + count);
    try {
      b->batcher.init();
    } catch (...) {
      delete b;
      throw;
    }
    *out = b;
  });
}

int pte_pt_batch_run(pte_pt_batch* b, int max_steps, int* more) {
  return guarded([&] {
    require(b != nullptr && more != nullptr, PTE_ERR_ARG, "null argument");
    require(max_steps >= 1, PTE_ERR_ARG, "max_steps must be positive");
    std::vector<PtEngine*> active;
    for (auto* e : b->engines) {
      require(e->impl.step >= 1, PTE_ERR_STATE, "engine in batch not initialised");
      if (e->impl.step < e->impl.cfg.num_steps) active.push_back(&e->impl);
    }
    std::vector<std::exception_ptr> errs(active.size());
    if (!active.empty()) {
      b->batcher.enter((int)active.size());
      std::vector<std::thread> threads;
      threads.reserve(active.size());
      for (size_t i = 0; i < active.size(); ++i) {
        threads.emplace_back([&, i] {
          PtEngine* eng = active[i];
          eng->ws.batcher = &b->batcher;
          try {
            for (int t = 0; t < max_steps && eng->step < eng->cfg.num_steps; ++t)
              eng->compute_step(nullptr, 0, 0);
          } catch (...) {
            errs[i] = std::current_exception();
          }
          eng->ws.batcher = nullptr;
          b->batcher.leave();
        });
      }
      for (auto& t : threads) t.join();
    }
    for (size_t i = 0; i < b->engines.size(); ++i) {
      const PtEngine& e = b->engines[i]->impl;
      more[i] = e.step < e.cfg.num_steps ? 1 : 0;
    }
    for (auto& er : errs)
      if (er) std::rethrow_exception(er);
  });
}

int pte_pt_batch_stats(pte_pt_batch* b, long* batches, long* requests, long* launches,
                       double* serve_ms, double* qr_ms, double* jacobi_ms) {
  return guarded([&] {
    require(b && batches && requests && launches && serve_ms && qr_ms && jacobi_ms,
            PTE_ERR_ARG, "null argument");
    *batches = b->batcher.n_batches;
    *requests = b->batcher.n_requests;
    *launches = b->batcher.n_launches;
    *serve_ms = b->batcher.serve_ms;
    *qr_ms = b->batcher.qr_ms_total;
    *jacobi_ms = b->batcher.jac_ms_total;
  });
}

void pte_pt_batch_destroy(pte_pt_batch* b) {
  if (!b) return;
  guarded([&] { b->batcher.destroy(); });
  delete b;
}

int pte_tempo_create(const pte_tempo_config* cfg, pte_tempo_engine** out) {
  return guarded([&] {
    require(cfg && out, PTE_ERR_ARG, "null argument");
    require(cfg->dimension >= 1 && cfg->dimension <= 4, PTE_ERR_ARG,
            "dimension must be 1..4");
    require(cfg->dkmax >= -1, PTE_ERR_ARG, "dkmax must be >= 0 or -1");
    require(cfg->epsrel > 0.0 && std::isfinite(cfg->epsrel), PTE_ERR_ARG,
            "epsrel must be positive");
    require(cfg->precision == PTE_PRECISION_C128 || cfg->precision == PTE_PRECISION_C64,
            PTE_ERR_ARG, "precision must be PTE_PRECISION_C128 or PTE_PRECISION_C64");
    require(cfg->chi_cap >= 1 && cfg->n_north >= 1 && cfg->n_west >= 1, PTE_ERR_ARG,
            "bad capacity or vector lengths");
    auto* e = new pte_tempo_engine();
    e->impl.cfg = *cfg;
    try {
      e->impl.init_engine();
    } catch (...) {
      delete e;
      throw;
    }
    *out = e;
  });
}

int pte_tempo_set_influence(pte_tempo_engine* e, int dk, const double* data, int rows,
                            int cols) {
  return guarded([&] {
    require(e != nullptr, PTE_ERR_ARG, "null engine");
    require(dk >= 0, PTE_ERR_ARG, "dk must be non-negative");
    store_infl(e->impl.slot(dk), data, rows, cols);
  });
}

int pte_tempo_set_degeneracy_maps(pte_tempo_engine* e, const int* nmap, const int* wmap,
                                  int length) {
  return guarded([&] {
    require(e && nmap && wmap, PTE_ERR_ARG, "null argument");
    const int d2 = e->impl.cfg.dimension * e->impl.cfg.dimension;
    require(length == d2, PTE_ERR_ARG, "degeneracy maps must have length d^2");
    e->impl.north_map.assign(nmap, nmap + length);
    e->impl.west_map.assign(wmap, wmap + length);
  });
}

int pte_tempo_initialize(pte_tempo_engine* e, const double* rho0, const double* unitary) {
  return guarded([&] {
    require(e && rho0 && unitary, PTE_ERR_ARG, "null argument");
    require(e->impl.step < 0, PTE_ERR_STATE, "already initialised");
    e->impl.initialize(rho0, unitary);
  });
}

int pte_tempo_step(pte_tempo_engine* e, const double* p1, const double* p2,
                   const double* extra, int rows, int cols, double* state_out) {
  return guarded([&] {
    require(e && p1 && p2 && state_out, PTE_ERR_ARG, "null argument");
    e->impl.compute_step(p1, p2, extra, rows, cols, state_out);
  });
}

int pte_tempo_get_step(pte_tempo_engine* e, int* step) {
  return guarded([&] {
    require(e && step, PTE_ERR_ARG, "null argument");
    *step = e->impl.step;
  });
}

int pte_tempo_bond_dims(pte_tempo_engine* e, int* out, int n, int* n_sites) {
  return guarded([&] {
    require(e && n_sites, PTE_ERR_ARG, "null argument");
    const auto& s = e->impl.mps.sites;
    *n_sites = (int)s.size();
    if (out) {
      require(n >= (int)s.size() - 1, PTE_ERR_ARG, "output too small");
      for (size_t i = 0; i + 1 < s.size(); ++i) out[i] = s[i].d[2];
    }
  });
}

void pte_tempo_destroy(pte_tempo_engine* e) {
  if (!e) return;
  guarded([&] { e->impl.destroy(); });
  delete e;
}

int pte_contract(int precision, int m, int n, int k, const double* A, const double* B,
                 double* C) {
  return guarded([&] {
    require(m >= 1 && n >= 1 && k >= 1 && A && B && C, PTE_ERR_ARG, "bad arguments");
    require(precision == PTE_PRECISION_C128 || precision == PTE_PRECISION_C64, PTE_ERR_ARG,
            "unknown precision");
    Workspace ws;
    ws.init(1e-7, 1);
    ws.precision = precision;
    try {
      DTensor da = dupload(to_cplx(A, (long)m * k), m, k, 1, 1, ws.stream);
      DTensor db = dupload(to_cplx(B, (long)k * n), k, n, 1, 1, ws.stream);
      DTensor dc = dalloc(m, n, 1, 1, ws.stream);
      ws.contract(m, n, k, da.p, k, 1, db.p, n, 1, dc.p, n, 1);
      HIP_CHECK(hipGetLastError());
      HIP_CHECK(hipMemcpyAsync(C, dc.p, (size_t)m * n * sizeof(double2),
                               hipMemcpyDeviceToHost, ws.stream));
      HIP_CHECK(hipStreamSynchronize(ws.stream));
      dfree(da, ws.stream);
      dfree(db, ws.stream);
      dfree(dc, ws.stream);
    } catch (...) {
      ws.destroy();
      throw;
    }
    ws.destroy();
  });
}

int pte_split_svd(int m, int n, const double* M, double epsrel, int chi_cap, double* U,
                  double* S, double* SVh, int* k_out, int* sweeps_out) {
  return guarded([&] {
    require(m >= 1 && n >= 1 && M && U && S && SVh && k_out, PTE_ERR_ARG,
            "bad arguments");
    require(epsrel > 0.0 && chi_cap >= 1, PTE_ERR_ARG, "bad epsrel or capacity");
    Workspace ws;
    ws.init(epsrel, chi_cap);
    try {
      const long cnt = (long)m * n;
      std::vector<cplx> h = to_cplx(M, cnt);
      DTensor dm = dupload(h, m, n, 1, 1, ws.stream);
      int k;
      DevBuf u, svh;
      if (m <= kQrMaxRows) {
        // column-major copy of M for the pivoted QR
        ws.Q.ensure(cnt, ws.stream);
        transpose_kernel<<<grid_for(cnt), 256, 0, ws.stream>>>(dm.p, m, n, ws.Q.p);
        k = ws.svd_qrcp(m, n);
        u.ensure((size_t)m * k, ws.stream);
        svh.ensure((size_t)k * n, ws.stream);
        ws.emit_left_q(1, 1, 0, 0, 1, 0, 0, u.p, k);
        ws.emit_svh(1, n, 0, 1, 0, 0, 0, svh.p, k);
      } else {
        ws.X.ensure(cnt, ws.stream);
        // X = M^H (n x m, column-major) is the elementwise conjugate of M
        conj_copy_kernel<<<grid_for(cnt), 256, 0, ws.stream>>>(dm.p, cnt, ws.X.p);
        k = ws.svd(n, m);
        u.ensure((size_t)m * k, ws.stream);
        svh.ensure((size_t)k * n, ws.stream);
        ws.emit(ws.V.p, m, m, 1, 1, 0, 0, 1, 0, 0, 0, u.p, k);
        ws.emit(ws.X.p, n, n, 1, n, 0, 1, 0, 0, 0, 1, svh.p, k);
      }
      HIP_CHECK(hipMemcpyAsync(U, u.p, (size_t)m * k * sizeof(double2),
                               hipMemcpyDeviceToHost, ws.stream));
      HIP_CHECK(hipMemcpyAsync(SVh, svh.p, (size_t)k * n * sizeof(double2),
                               hipMemcpyDeviceToHost, ws.stream));
      HIP_CHECK(hipMemcpyAsync(S, ws.sorted, (size_t)k * sizeof(double),
                               hipMemcpyDeviceToHost, ws.stream));
      HIP_CHECK(hipStreamSynchronize(ws.stream));
      *k_out = k;
      if (sweeps_out) *sweeps_out = ws.info_h->sweeps;
      u.release(ws.stream);
      svh.release(ws.stream);
      dfree(dm, ws.stream);
    } catch (...) {
      ws.destroy();
      throw;
    }
    ws.destroy();
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Allocation-free truncated SVD (pte_split_svd_noalloc): the workspace is this
// file-scope device array, one copy per GPU, held until the process ends.
// Host code never uses the array's name as an address: each device's address
// comes from hipGetSymbolAddress on its first call.
// ---------------------------------------------------------------------------
constexpr int kNoAllocMaxDim = 256;
constexpr size_t kNoAllocBytes = (size_t)12 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

static std::mutex g_noalloc_mutex;
constexpr int kNoAllocMaxDevices = 64;
static char* g_noalloc_ptr[kNoAllocMaxDevices] = {};

// device address of this GPU's scratch (caller holds g_noalloc_mutex)
static char* noalloc_scratch() {
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  require(dev >= 0 && dev < kNoAllocMaxDevices, PTE_ERR_DEVICE, "device index out of range");
  if (g_noalloc_ptr[dev] == nullptr) {
    void* p = nullptr;
    HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    require(p != nullptr, PTE_ERR_DEVICE, "no address for the SVD scratch");
    g_noalloc_ptr[dev] = static_cast<char*>(p);
  }
  return g_noalloc_ptr[dev];
}

extern "C" {

int pte_split_svd_noalloc(int m, int n, const double* M, double epsrel, int chi_cap,
                          double* U, double* S, double* SVh, int* k_out, int* sweeps_out) {
  return guarded([&] {
    require(m >= 1 && n >= 1 && M && U && S && SVh && k_out, PTE_ERR_ARG,
            "bad arguments");
    require(epsrel > 0.0 && chi_cap >= 1, PTE_ERR_ARG, "bad epsrel or capacity");
    require(m <= kNoAllocMaxDim && n <= kNoAllocMaxDim, PTE_ERR_CAPACITY,
            "pte_split_svd_noalloc takes matrices of at most " +
                std::to_string(kNoAllocMaxDim) + " x " + std::to_string(kNoAllocMaxDim));
    const long cnt = (long)m * n;
    for (long i = 0; i < 2 * cnt; ++i)
      require(std::isfinite(M[i]), PTE_ERR_NUMERIC, "non-finite value in an input tensor");
    std::lock_guard<std::mutex> lock(g_noalloc_mutex);
    char* base = noalloc_scratch();
    SvdInfo info_host{};
    int rank_host = 0;
    Workspace ws;
    ws.init_external(epsrel, chi_cap, base, kNoAllocBytes, kNoAllocMaxDim, kNoAllocMaxDim,
                     &info_host, &rank_host);
    HIP_CHECK(hipMemcpyAsync(ws.G.p, M, (size_t)cnt * sizeof(double2), hipMemcpyHostToDevice,
                             ws.stream));
    // column-major copy of M for the pivoted QR (m <= kQrMaxRows here)
    transpose_kernel<<<grid_for(cnt), 256, 0, ws.stream>>>(ws.G.p, m, n, ws.Q.p);
    const int k = ws.svd_qrcp(m, n);
    ws.emit_left_q(1, 1, 0, 0, 1, 0, 0, ws.carry.p, k);
    ws.emit_svh(1, n, 0, 1, 0, 0, 0, ws.svh.p, k);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(U, ws.carry.p, (size_t)m * k * sizeof(double2),
                             hipMemcpyDeviceToHost, ws.stream));
    HIP_CHECK(hipMemcpyAsync(SVh, ws.svh.p, (size_t)k * n * sizeof(double2),
                             hipMemcpyDeviceToHost, ws.stream));
    HIP_CHECK(hipMemcpyAsync(S, ws.sorted, (size_t)k * sizeof(double),
                             hipMemcpyDeviceToHost, ws.stream));
    HIP_CHECK(hipStreamSynchronize(ws.stream));
    *k_out = k;
    if (sweeps_out) *sweeps_out = info_host.sweeps;
  });
}

int pte_dynamics_run(int d2, int num_steps, const int* bonds, const double* tensors,
                     const double* caps, const double* props, const double* tin,
                     const double* tout, const double* rho0, double* states_out) {
  return guarded([&] {
    require(bonds && tensors && caps && props && rho0 && states_out, PTE_ERR_ARG,
            "null argument");
    require(num_steps >= 1, PTE_ERR_ARG, "num_steps must be >= 1");
    require(bonds[0] == 1, PTE_ERR_ARG, "the first bond must be 1");
    hipStream_t st;
    HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    std::vector<int> b(bonds, bonds + num_steps + 1);
    std::vector<double2*> dsites(num_steps), dcaps(num_steps + 1);
    long off = 0, coff = 0;
    for (int k = 0; k < num_steps; ++k) {
      require(b[k] >= 1 && b[k + 1] >= 1, PTE_ERR_ARG, "bond dims must be positive");
      // tensors come as [a][b][i]; the kernel reads [a][i][b]
      const long cnt = (long)b[k] * b[k + 1] * d2;
      std::vector<cplx> h(cnt);
      for (int a = 0; a < b[k]; ++a)
        for (int bb = 0; bb < b[k + 1]; ++bb)
          for (int i = 0; i < d2; ++i) {
            const long src = off + ((long)a * b[k + 1] + bb) * d2 + i;
            h[((long)a * d2 + i) * b[k + 1] + bb] = cplx(tensors[2 * src], tensors[2 * src + 1]);
          }
      HIP_CHECK(hipMallocAsync((void**)&dsites[k], cnt * sizeof(double2), st));
      HIP_CHECK(hipMemcpyAsync(dsites[k], h.data(), cnt * sizeof(double2),
                               hipMemcpyHostToDevice, st));
      HIP_CHECK(hipStreamSynchronize(st));
      off += cnt;
    }
    require(b[num_steps] >= 1, PTE_ERR_ARG, "bond dims must be positive");
    for (int k = 0; k <= num_steps; ++k) {
      HIP_CHECK(hipMallocAsync((void**)&dcaps[k], b[k] * sizeof(double2), st));
      HIP_CHECK(hipMemcpyAsync(dcaps[k], caps + 2 * coff, b[k] * sizeof(double2),
                               hipMemcpyHostToDevice, st));
      coff += b[k];
    }
    std::vector<const double2*> cs(dsites.begin(), dsites.end());
    std::vector<const double2*> cc(dcaps.begin(), dcaps.end());
    run_dynamics(st, d2, num_steps, cs, b, cc, props, tin, tout, rho0, 1.0, states_out);
    for (auto p : dsites) (void)hipFreeAsync(p, st);
    for (auto p : dcaps) (void)hipFreeAsync(p, st);
    HIP_CHECK(hipStreamSynchronize(st));
    (void)hipStreamDestroy(st);
  });
}

int pte_dynamics_run_engine(pte_pt_engine* e, const double* props, const double* tin,
                            const double* tout, const double* rho0, double* states_out) {
  return guarded([&] {
    require(e && props && rho0 && states_out, PTE_ERR_ARG, "null argument");
    PtEngine& pe = e->impl;
    require(pe.step >= pe.cfg.num_steps, PTE_ERR_STATE, "process tensor not complete");
    require(!pe.caps.empty(), PTE_ERR_STATE, "caps not computed");
    const int N = pe.cfg.num_steps, d2 = pe.cfg.dimension * pe.cfg.dimension;
    std::vector<int> b(N + 1);
    std::vector<const double2*> sites(N), caps(N + 1);
    for (int k = 0; k < N; ++k) {
      b[k] = pe.mps.sites[k].d[0];
      sites[k] = pe.mps.sites[k].p;
      require(pe.mps.sites[k].d[1] == d2, PTE_ERR_ARG, "internal: site leg is not d^2");
    }
    b[N] = pe.mps.sites[N - 1].d[2];
    for (int k = 0; k <= N; ++k) caps[k] = pe.caps[k].p;
    run_dynamics(pe.st(), d2, N, sites, b, caps, props, tin, tout, rho0,
                 (double)pe.cfg.dimension, states_out);
  });
}

int pte_dynamics_run_engine_batch(pte_pt_engine* e, int nsys, const double* props,
                                  const double* tin, const double* tout, const double* rho0,
                                  double* states_out) {
  return guarded([&] {
    require(e && props && rho0 && states_out && nsys >= 1, PTE_ERR_ARG, "bad arguments");
    PtEngine& pe = e->impl;
    require(pe.step >= pe.cfg.num_steps, PTE_ERR_STATE, "process tensor not complete");
    require(!pe.caps.empty(), PTE_ERR_STATE, "caps not computed");
    const int N = pe.cfg.num_steps, d2 = pe.cfg.dimension * pe.cfg.dimension;
    std::vector<int> b(N + 1);
    std::vector<const double2*> sites(N), caps(N + 1);
    for (int k = 0; k < N; ++k) {
      b[k] = pe.mps.sites[k].d[0];
      sites[k] = pe.mps.sites[k].p;
      require(pe.mps.sites[k].d[1] == d2, PTE_ERR_ARG, "internal: site leg is not d^2");
    }
    b[N] = pe.mps.sites[N - 1].d[2];
    for (int k = 0; k <= N; ++k) caps[k] = pe.caps[k].p;
    run_dynamics(pe.st(), d2, N, sites, b, caps, props, tin, tout, rho0,
                 (double)pe.cfg.dimension, states_out, nsys);
  });
}

int pte_dynamics_run_engines(pte_pt_engine* const* engines, int count, const double* props,
                             const double* const* transforms_in,
                             const double* const* transforms_out, const double* rho0,
                             double* states_out) {
  return guarded([&] {
    require(engines && props && rho0 && states_out && transforms_in && transforms_out,
            PTE_ERR_ARG, "null argument");
    require(count >= 1 && count <= kMaxEnvironments, PTE_ERR_ARG, "1 to 8 process tensors");
    require(engines[0] != nullptr, PTE_ERR_ARG, "null engine");
    const int N = engines[0]->impl.cfg.num_steps, d = engines[0]->impl.cfg.dimension;
    std::vector<DynEnv> env(count);
    for (int e = 0; e < count; ++e) {
      require(engines[e] != nullptr, PTE_ERR_ARG, "null engine");
      PtEngine& pe = engines[e]->impl;
      require(pe.step >= pe.cfg.num_steps, PTE_ERR_STATE, "process tensor not complete");
      require(!pe.caps.empty(), PTE_ERR_STATE, "caps not computed");
      require(pe.cfg.num_steps == N && pe.cfg.dimension == d, PTE_ERR_ARG,
              "process tensors of different lengths or dimensions");
      DynEnv& en = env[e];
      en.sites.resize(N);
      en.bonds.resize(N + 1);
      en.caps.resize(N + 1);
      for (int k = 0; k < N; ++k) {
        require(pe.mps.sites[k].d[1] == d * d, PTE_ERR_ARG, "internal: site leg is not d^2");
        en.sites[k] = pe.mps.sites[k].p;
        en.bonds[k] = pe.mps.sites[k].d[0];
      }
      en.bonds[N] = pe.mps.sites[N - 1].d[2];
      for (int k = 0; k <= N; ++k) en.caps[k] = pe.caps[k].p;
      en.scale = (double)d;
      en.tin = transforms_in[e];
      en.tout = transforms_out[e];
      // the sites are read on the first engine's stream
      if (e > 0) HIP_CHECK(hipStreamSynchronize(pe.st()));
    }
    run_dynamics_multi(engines[0]->impl.st(), d * d, N, env, props, rho0, states_out);
  });
}

int pte_dynamics_run_multi(int d2, int num_steps, int count, const int* bonds,
                           const double* tensors, const double* caps, const double* props,
                           const double* const* transforms_in,
                           const double* const* transforms_out, const double* rho0,
                           double* states_out) {
  return guarded([&] {
    require(bonds && tensors && caps && props && rho0 && states_out && transforms_in &&
                transforms_out, PTE_ERR_ARG, "null argument");
    require(num_steps >= 1, PTE_ERR_ARG, "num_steps must be >= 1");
    require(count >= 1 && count <= kMaxEnvironments, PTE_ERR_ARG, "1 to 8 process tensors");
    hipStream_t st;
    HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    std::vector<double2*> owned;
    std::vector<DynEnv> env(count);
    long off = 0, coff = 0;
    try {
      for (int e = 0; e < count; ++e) {
        DynEnv& en = env[e];
        en.bonds.assign(bonds + (long)e * (num_steps + 1), bonds + (long)(e + 1) * (num_steps + 1));
        require(en.bonds[0] == 1, PTE_ERR_ARG, "the first bond must be 1");
        for (int k = 0; k <= num_steps; ++k)
          require(en.bonds[k] >= 1, PTE_ERR_ARG, "bond dims must be positive");
        en.sites.resize(num_steps);
        en.caps.resize(num_steps + 1);
        for (int k = 0; k < num_steps; ++k) {
          // tensors come as [a][b][i]; the kernels read [a][i][b]
          const int A = en.bonds[k], B = en.bonds[k + 1];
          const long cnt = (long)A * B * d2;
          std::vector<cplx> h(cnt);
          for (int x = 0; x < A; ++x)
            for (int y = 0; y < B; ++y)
              for (int i = 0; i < d2; ++i) {
                const long src = off + ((long)x * B + y) * d2 + i;
                h[((long)x * d2 + i) * B + y] = cplx(tensors[2 * src], tensors[2 * src + 1]);
              }
          double2* p = nullptr;
          HIP_CHECK(hipMallocAsync((void**)&p, cnt * sizeof(double2), st));
          owned.push_back(p);
          HIP_CHECK(hipMemcpyAsync(p, h.data(), cnt * sizeof(double2), hipMemcpyHostToDevice,
                                   st));
          HIP_CHECK(hipStreamSynchronize(st));
          en.sites[k] = p;
          off += cnt;
        }
        for (int k = 0; k <= num_steps; ++k) {
          double2* p = nullptr;
          HIP_CHECK(hipMallocAsync((void**)&p, en.bonds[k] * sizeof(double2), st));
          owned.push_back(p);
          HIP_CHECK(hipMemcpyAsync(p, caps + 2 * coff, en.bonds[k] * sizeof(double2),
                                   hipMemcpyHostToDevice, st));
          en.caps[k] = p;
          coff += en.bonds[k];
        }
        en.scale = 1.0;
        en.tin = transforms_in[e];
        en.tout = transforms_out[e];
      }
      run_dynamics_multi(st, d2, num_steps, env, props, rho0, states_out);
    } catch (...) {
      (void)hipStreamSynchronize(st);
      for (auto p : owned) (void)hipFreeAsync(p, st);
      (void)hipStreamSynchronize(st);
      (void)hipStreamDestroy(st);
      throw;
    }
    for (auto p : owned) (void)hipFreeAsync(p, st);
    HIP_CHECK(hipStreamSynchronize(st));
    HIP_CHECK(hipStreamDestroy(st));
  });
}

}  // extern "C"
