// Device-side helpers shared by the process-tensor engine kernels (gfx950).
//
// Complex numbers are double2 (x = real, y = imaginary), interleaved as in
// numpy's complex128, so host arrays are copied without conversion.
#pragma once

#include <hip/hip_runtime.h>

namespace pte {

__device__ __forceinline__ double2 cadd(double2 a, double2 b) {
  return make_double2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ double2 csub(double2 a, double2 b) {
  return make_double2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 cscale(double2 a, double s) {
  return make_double2(a.x * s, a.y * s);
}
// acc += a * b
__device__ __forceinline__ void cfma(double2& acc, double2 a, double2 b) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(-a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(a.y, b.x, acc.y);
}
// acc += conj(a) * b
__device__ __forceinline__ void cfmac(double2& acc, double2 a, double2 b) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(-a.y, b.x, acc.y);
}
__device__ __forceinline__ double cabs2(double2 a) { return a.x * a.x + a.y * a.y; }

// Cross-lane moves of a double by DPP (two 32-bit halves): lane i reads the
// lane given by the control word (quad permutes, row mirrors).
constexpr int kDppXor1 = 0xB1;        // quad_perm [1,0,3,2]
constexpr int kDppXor2 = 0x4E;        // quad_perm [2,3,0,1]
constexpr int kDppHalfMirror = 0x141; // lane i <-> 7 - i within 8
constexpr int kDppMirror = 0x140;     // lane i <-> 15 - i within 16

template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  const long long b = __double_as_longlong(v);
  int lo = (int)(b & 0xffffffffll), hi = (int)(b >> 32);
  lo = __builtin_amdgcn_update_dpp(0, lo, CTRL, 0xF, 0xF, false);
  hi = __builtin_amdgcn_update_dpp(0, hi, CTRL, 0xF, 0xF, false);
  return __longlong_as_double((long long)(unsigned)lo | ((long long)hi << 32));
}
template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}

// Sums over aligned groups of G lanes (G a power of two <= 64) of four
// values at once. Every step adds a partner's value symmetrically (x + y on
// one lane, y + x on the other), so all lanes of a group end with the same
// bits.
__device__ __forceinline__ void group_sum4(double& a, double& b, double& c, double& d, int G) {
  if (G >= 2) {
    a += dpp_d<kDppXor1>(a); b += dpp_d<kDppXor1>(b);
    c += dpp_d<kDppXor1>(c); d += dpp_d<kDppXor1>(d);
  }
  if (G >= 4) {
    a += dpp_d<kDppXor2>(a); b += dpp_d<kDppXor2>(b);
    c += dpp_d<kDppXor2>(c); d += dpp_d<kDppXor2>(d);
  }
  if (G >= 8) {
    a += dpp_d<kDppHalfMirror>(a); b += dpp_d<kDppHalfMirror>(b);
    c += dpp_d<kDppHalfMirror>(c); d += dpp_d<kDppHalfMirror>(d);
  }
  if (G >= 16) {
    a += dpp_d<kDppMirror>(a); b += dpp_d<kDppMirror>(b);
    c += dpp_d<kDppMirror>(c); d += dpp_d<kDppMirror>(d);
  }
  if (G >= 32) {
    a += __shfl_xor(a, 16, 64); b += __shfl_xor(b, 16, 64);
    c += __shfl_xor(c, 16, 64); d += __shfl_xor(d, 16, 64);
  }
  if (G >= 64) {
    a += __shfl_xor(a, 32, 64); b += __shfl_xor(b, 32, 64);
    c += __shfl_xor(c, 32, 64); d += __shfl_xor(d, 32, 64);
  }
}

__device__ __forceinline__ double group_sum1(double a, int G) {
  double b = 0.0, c = 0.0, d = 0.0;
  group_sum4(a, b, c, d, G);
  return a;
}

// Sum over one 64-lane wavefront; every lane gets the same bits.
__device__ __forceinline__ double wave_sum(double v) { return group_sum1(v, 64); }

// Sum over a 256-thread workgroup with one __syncthreads: wavefront sums go
// to slot[parity][wave]; alternating the parity between calls keeps a fast
// wavefront from overwriting slots that another is still reading. Every
// thread returns the same value.
__device__ __forceinline__ double block_sum256(double v, double (*slot)[4], int& parity) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) slot[parity][threadIdx.x >> 6] = v;
  __syncthreads();
  const double r = (slot[parity][0] + slot[parity][1]) + (slot[parity][2] + slot[parity][3]);
  parity ^= 1;
  return r;
}

__device__ __forceinline__ double2 block_sum256c(double2 v, double (*slot)[4],
                                                 double (*sloti)[4], int& parity) {
  v.x = wave_sum(v.x);
  v.y = wave_sum(v.y);
  if ((threadIdx.x & 63) == 0) {
    slot[parity][threadIdx.x >> 6] = v.x;
    sloti[parity][threadIdx.x >> 6] = v.y;
  }
  __syncthreads();
  const double2 r = make_double2(
      (slot[parity][0] + slot[parity][1]) + (slot[parity][2] + slot[parity][3]),
      (sloti[parity][0] + sloti[parity][1]) + (sloti[parity][2] + sloti[parity][3]));
  parity ^= 1;
  return r;
}

// Agent-scope (device-coherent) accesses for data that workgroups of one
// cooperative launch exchange. They bypass the per-XCD L2 copies, so the grid
// barrier below needs no L2 writeback or invalidation; every read of such
// data after a barrier and every write before one goes through these.
__device__ __forceinline__ double ld_agent(const double* p) {
  return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int ld_agent(const int* p) {
  return __hip_atomic_load(const_cast<int*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned ld_agent(const unsigned* p) {
  return __hip_atomic_load(const_cast<unsigned*>(p), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(unsigned* p, unsigned v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double2 ld_agent(const double2* p) {
  return make_double2(ld_agent(&p->x), ld_agent(&p->y));
}
__device__ __forceinline__ void st_agent(double2* p, double2 v) {
  st_agent(&p->x, v.x);
  st_agent(&p->y, v.y);
}

// The same agent-scope accesses for a whole complex entry in one 16-byte
// buffer instruction (the pair of 8-byte atomics above takes two). The
// buffer resource covers `bytes` bytes from `base` (accesses beyond it load
// zero and store nothing); the cache policy is sc1, the bits the compiler
// gives the agent-scope atomics on gfx950.
constexpr int kBufAgentPolicy = 16;                  // sc1
constexpr int kBufResourceWord3 = 0x00020000;        // raw 32-bit buffer (gfx9)
constexpr long kBufMaxEntries = 1L << 27;            // complex entries in 2^31 bytes

__device__ __forceinline__ __amdgpu_buffer_rsrc_t agent_rsrc(const void* base, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)bytes,
                                           kBufResourceWord3);
}
typedef int int4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ double2 ld_agent16(__amdgpu_buffer_rsrc_t r, unsigned byte_off) {
  const int4v d = __builtin_amdgcn_raw_buffer_load_b128(r, (int)byte_off, 0, kBufAgentPolicy);
  double2 v;
  __builtin_memcpy(&v, &d, sizeof(v));
  return v;
}
__device__ __forceinline__ void st_agent16(__amdgpu_buffer_rsrc_t r, unsigned byte_off,
                                           double2 v) {
  int4v d;
  __builtin_memcpy(&d, &v, sizeof(v));
  __builtin_amdgcn_raw_buffer_store_b128(d, r, (int)byte_off, 0, kBufAgentPolicy);
}

// Barrier for a group of co-resident workgroups of a cooperative launch.
// bar[0] is zeroed before the launch and only grows; the barrier of
// generation `epoch` waits until epoch * nblocks arrivals. Every wavefront
// first waits for its outstanding memory operations (s_waitcnt 0), so its
// agent-scope stores are performed before its workgroup arrives.
__device__ __forceinline__ void grid_barrier(unsigned* bar, unsigned nblocks,
                                             unsigned& epoch) {
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  epoch += 1;
  if (threadIdx.x == 0) {
    __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned target = epoch * nblocks;
    while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target)
      __builtin_amdgcn_s_sleep(1);
  }
  __syncthreads();
}

// Grid-wide barrier over all nblocks workgroups of a launch (or of one
// problem of a batched launch; bid is the workgroup's index in it), in two
// levels: workgroup bid arrives on the counter of group bid % kBarrierGroups,
// the last arrival of a group on the top counter bar[0], and every
// workgroup waits on the top counter. With one counter, the arrivals of 256
// workgroups on one word serialise (measured 3.7 us per barrier on MI355X
// against 2.2 us in two levels with eight groups). The counters are
// kBarrierStride words apart; all kBarrierWords are zeroed before the
// launch. Small grids use the single counter.
constexpr unsigned kBarrierGroups = 8;
constexpr unsigned kBarrierStride = 64;
constexpr unsigned kBarrierWords = kBarrierStride * (kBarrierGroups + 1);
// words of the synchronisation buffer: the grid counters, then (Jacobi) a
// pair barrier and a pair completion counter per block pair and a version per
// block
constexpr unsigned kSyncMaxBlocks = 8192;
constexpr unsigned kSyncWords = kBarrierWords + 2 * kSyncMaxBlocks;

__device__ __forceinline__ void grid_barrier_all(unsigned* bar, unsigned nblocks,
                                                 unsigned bid, unsigned& epoch) {
  if (nblocks < 4 * kBarrierGroups) {
    grid_barrier(bar, nblocks, epoch);
    return;
  }
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  epoch += 1;
  if (threadIdx.x == 0) {
    const unsigned g = bid % kBarrierGroups;
    const unsigned gsize = nblocks / kBarrierGroups + (g < nblocks % kBarrierGroups ? 1u : 0u);
    unsigned* grp = bar + kBarrierStride * (g + 1);
    const unsigned old =
        __hip_atomic_fetch_add(grp, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (old == epoch * gsize - 1)
      __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned target = epoch * kBarrierGroups;
    while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target)
      __builtin_amdgcn_s_sleep(1);
  }
  __syncthreads();
}

}  // namespace pte
