// gp_scan.h -- device-wide exclusive prefix sum for the counting sort and
// the pair offsets: 64-lane wavefront scans (__shfl_up) inside 256-thread
// tiles of 2048 elements, tile totals scanned recursively, then added back.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gp {

constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    T t = __shfl_up(v, d, 64);
    if (lane >= d) v += t;
  }
  return v;
}

// exclusive scan of in[0..n) (of type Tin, widened to T) into out[0..n);
// tile totals go to sums[blockIdx.x]
template <typename Tin, typename T>
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_tiles(const Tin* __restrict__ in,
                                                           T* __restrict__ out,
                                                           T* __restrict__ sums, int64_t n) {
  __shared__ T wsum[SCAN_BLOCK / 64];
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
  T v[SCAN_ITEMS];
  T tot = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    int64_t idx = base + k;
    v[k] = (idx < n) ? (T)in[idx] : (T)0;
    tot += v[k];
  }
  T incl = wave_incl_scan(tot);
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 63) wsum[wid] = incl;
  __syncthreads();
  T woff = 0;
  for (int w = 0; w < wid; ++w) woff += wsum[w];
  T run = woff + incl - tot;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    int64_t idx = base + k;
    if (idx < n) out[idx] = run;
    run += v[k];
  }
  if (threadIdx.x == SCAN_BLOCK - 1) sums[blockIdx.x] = woff + incl;
}

template <typename T>
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_add(T* __restrict__ out,
                                                         const T* __restrict__ sums, int64_t n) {
  const T add = sums[blockIdx.x];
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
  for (int k = threadIdx.x; k < SCAN_TILE; k += SCAN_BLOCK) {
    int64_t idx = base + k;
    if (idx < n) out[idx] += add;
  }
}

template <typename T>
__global__ void k_copy1(T* dst, const T* src, int zero) {
  dst[0] = zero ? (T)0 : src[0];
}

// ---------------------------------------------------------------------------
// Two-launch exclusive scan of int counts (reduce, then scan): the first
// launch writes each tile's total; in the second every workgroup sums the
// totals of the tiles before its own (at most a few thousand, read from L2)
// and scans its tile from there.  No recursion and no separate add or copy
// launches; the last tile writes the grand total to out[n].  The sums must
// fit 31 bits (cell counts).  (A single-pass decoupled look-back was slower
// here: its look-back reads must reach a point of coherence shared by all
// XCDs, and the inclusive prefixes propagate 64 tiles per round trip.)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(SCAN_BLOCK) void k_tile_sums(const int* __restrict__ in,
                                                          int* __restrict__ sums, int n) {
  __shared__ int wsum[SCAN_BLOCK / 64];
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
  int tot = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    const int64_t idx = base + k;
    tot += (idx < n) ? in[idx] : 0;
  }
  for (int o = 32; o > 0; o >>= 1) tot += __shfl_xor(tot, o);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = tot;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < SCAN_BLOCK / 64; ++w) t += wsum[w];
    sums[blockIdx.x] = t;
  }
}

// tpre: the tiles' exclusive prefixes, when scanned beforehand (else each
// tile sums the totals of the tiles before it)
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_from_sums(const int* __restrict__ in,
                                                               int* __restrict__ out,
                                                               const int* __restrict__ sums,
                                                               int n, int ntiles,
                                                               const int* __restrict__ tpre) {
  __shared__ int wsum[SCAN_BLOCK / 64];
  __shared__ int s_excl;
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // the tile's own inputs are requested first, so their loads overlap the
  // walk over the preceding totals
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
  int v[SCAN_ITEMS];
  int tot = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    const int64_t idx = base + k;
    v[k] = (idx < n) ? in[idx] : 0;
    tot += v[k];
  }
  // exclusive prefix of this tile: the totals of tiles [0, blockIdx.x)
  if (tpre) {
    if (threadIdx.x == 0) s_excl = tpre[blockIdx.x];
  } else {
    int pre = 0;
    for (int t = threadIdx.x; t < (int)blockIdx.x; t += SCAN_BLOCK) pre += sums[t];
    for (int o = 32; o > 0; o >>= 1) pre += __shfl_xor(pre, o);
    if (lane == 0) wsum[wid] = pre;
    __syncthreads();
    if (threadIdx.x == 0) {
      int e = 0;
      for (int w = 0; w < SCAN_BLOCK / 64; ++w) e += wsum[w];
      s_excl = e;
    }
  }
  __syncthreads();
  const int incl = wave_incl_scan(tot);
  if (lane == 63) wsum[wid] = incl;   // wsum reuse: s_excl was read after the barrier above
  __syncthreads();
  int woff = 0, agg = 0;
  for (int w = 0; w < SCAN_BLOCK / 64; ++w) {
    if (w < wid) woff += wsum[w];
    agg += wsum[w];
  }
  int run = s_excl + woff + incl - tot;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    const int64_t idx = base + k;
    if (idx < n) out[idx] = run;
    run += v[k];
  }
  if ((int)blockIdx.x == ntiles - 1 && threadIdx.x == 0) out[n] = s_excl + agg;
}

// Short inputs (the cell counts of a typical grid, a few ten thousand) in
// one workgroup and one launch: passes of SCAN1_CHUNK inputs are staged in
// LDS with coalesced loads, each thread scans its SCAN1_ITEMS consecutive
// values there, the thread totals are scanned across the workgroup and
// carried from pass to pass, and the results leave coalesced.  Above
// SCAN1_MAX inputs the two-launch scan's parallelism wins.
constexpr int SCAN1_BLOCK = 1024;
constexpr int SCAN1_ITEMS = 8;
constexpr int SCAN1_CHUNK = SCAN1_BLOCK * SCAN1_ITEMS;   // 32 KiB of LDS
constexpr int SCAN1_MAX = 4 * SCAN1_CHUNK;

__global__ __launch_bounds__(SCAN1_BLOCK) void k_scan_onewg(const int* __restrict__ in,
                                                            int* __restrict__ out, int n) {
  __shared__ int buf[SCAN1_CHUNK];
  __shared__ int wsum[SCAN1_BLOCK / 64];
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int carry = 0;
  for (int c0 = 0; c0 < n; c0 += SCAN1_CHUNK) {
    for (int k = threadIdx.x; k < SCAN1_CHUNK; k += SCAN1_BLOCK)
      buf[k] = (c0 + k < n) ? in[c0 + k] : 0;
    __syncthreads();
    int v[SCAN1_ITEMS];
    int tot = 0;
#pragma unroll
    for (int k = 0; k < SCAN1_ITEMS; ++k) {
      v[k] = buf[threadIdx.x * SCAN1_ITEMS + k];
      tot += v[k];
    }
    const int incl = wave_incl_scan(tot);
    if (lane == 63) wsum[wid] = incl;
    __syncthreads();
    int woff = 0, agg = 0;
    for (int w = 0; w < SCAN1_BLOCK / 64; ++w) {
      if (w < wid) woff += wsum[w];
      agg += wsum[w];
    }
    int run = carry + woff + incl - tot;
#pragma unroll
    for (int k = 0; k < SCAN1_ITEMS; ++k) {
      buf[threadIdx.x * SCAN1_ITEMS + k] = run;
      run += v[k];
    }
    __syncthreads();
    for (int k = threadIdx.x; k < SCAN1_CHUNK; k += SCAN1_BLOCK)
      if (c0 + k < n) out[c0 + k] = buf[k];
    carry += agg;
    __syncthreads();   // buf and wsum are rewritten by the next pass
  }
  if (threadIdx.x == 0) out[n] = carry;
}

// out[i] = sum_{k<i} in[k] for i < n, out[n] = total; sums holds two ints
// per tile of SCAN_TILE inputs, plus one
inline int64_t scan2_sums(int64_t n) { return (n + SCAN_TILE - 1) / SCAN_TILE; }
constexpr int SCAN3_MIN_TILES = 2048;

inline void scan_exclusive_int(const int* in, int* out, int n, int* sums, hipStream_t st) {
  const int ntiles = (int)scan2_sums(n);
  if (n <= 0) {
    hipLaunchKernelGGL((k_copy1<int>), dim3(1), dim3(1), 0, st, out, (const int*)out, 1);
    return;
  }
  if (n <= SCAN1_MAX) {
    hipLaunchKernelGGL(k_scan_onewg, dim3(1), dim3(SCAN1_BLOCK), 0, st, in, out, n);
    return;
  }
  hipLaunchKernelGGL(k_tile_sums, dim3((unsigned)ntiles), dim3(SCAN_BLOCK), 0, st, in, sums, n);
  // many tiles: their totals scanned in one workgroup first, so each tile
  // reads its prefix instead of summing every total before it (a walk that
  // grows with the square of the tile count)
  const int* tpre = nullptr;
  if (ntiles >= SCAN3_MIN_TILES && ntiles <= SCAN1_MAX) {
    hipLaunchKernelGGL(k_scan_onewg, dim3(1), dim3(SCAN1_BLOCK), 0, st, sums, sums + ntiles,
                       ntiles);
    tpre = sums + ntiles;
  }
  hipLaunchKernelGGL(k_scan_from_sums, dim3((unsigned)ntiles), dim3(SCAN_BLOCK), 0, st, in, out,
                     sums, n, ntiles, tpre);
}

// scratch elements needed by scan_exclusive for n inputs
inline int64_t scan_scratch(int64_t n) {
  const int64_t nt = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (nt <= 1) return 1;
  return 2 * nt + 1 + scan_scratch(nt);
}

// out[i] = sum_{k<i} in[k] for i < n, and out[n] = total (out has n+1 slots)
template <typename Tin, typename T>
void scan_exclusive(const Tin* in, T* out, int64_t n, T* scratch, hipStream_t st) {
  const int64_t nt = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (n <= 0) {
    hipLaunchKernelGGL((k_copy1<T>), dim3(1), dim3(1), 0, st, out, (const T*)out, 1);
    return;
  }
  T* sums = scratch;
  hipLaunchKernelGGL((k_scan_tiles<Tin, T>), dim3((unsigned)nt), dim3(SCAN_BLOCK), 0, st, in,
                     out, sums, n);
  if (nt > 1) {
    T* sums_ex = scratch + nt;   // nt + 1 slots, its total lands at sums_ex[nt]
    scan_exclusive<T, T>(sums, sums_ex, nt, scratch + 2 * nt + 1, st);
    hipLaunchKernelGGL((k_scan_add<T>), dim3((unsigned)nt), dim3(SCAN_BLOCK), 0, st, out,
                       sums_ex, n);
    hipLaunchKernelGGL((k_copy1<T>), dim3(1), dim3(1), 0, st, out + n, (const T*)(sums_ex + nt), 0);
  } else {
    hipLaunchKernelGGL((k_copy1<T>), dim3(1), dim3(1), 0, st, out + n, (const T*)sums, 0);
  }
}

}  // namespace gp
