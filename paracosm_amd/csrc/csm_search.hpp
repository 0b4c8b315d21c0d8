// Search primitives over ascending u32 arrays (adjacency segments in global
// memory, the leaf filter in LDS) used by the search kernel.  Plain C++: with
// CSM_SEARCH_HOST defined the device qualifiers compile away, so that the CPU
// tests check them against exhaustive searches.
#pragma once

#include <stdint.h>

#ifdef CSM_SEARCH_HOST
#define __device__
#define __forceinline__ inline
#endif

namespace csm {

// version word ins<<16 | del: visible to update positions ins..del
__device__ __forceinline__ bool ts_visible(uint32_t ts, uint32_t pos) {
  return (ts >> 16) <= pos && pos <= (ts & 0xFFFFu);
}

// first index i in [lo, hi) with a[i] >= x (hi if none).  The halvings of
// the binary search are unchanged, but while at least 3 entries remain the
// 3 pivots the next two halvings can visit are loaded together and the two
// comparisons made afterwards: one memory round trip per two levels (the
// loads are independent, so they overlap).
__device__ __forceinline__ uint32_t lower_bound(const uint32_t* __restrict__ a, uint32_t lo,
                                                uint32_t hi, uint32_t x) {
  uint32_t n = hi - lo;
  while (n >= 3) {
    // [lo, lo + n1) | m | [m + 1, m + 1 + n2), both halves non-empty
    const uint32_t n1 = n >> 1, n2 = n - n1 - 1, m = lo + n1;
    const uint32_t vm = a[m], vA = a[lo + (n1 >> 1)], vB = a[m + 1 + (n2 >> 1)];
    const bool c1 = vm < x;
    const uint32_t l2 = c1 ? m + 1 : lo, k2 = c1 ? n2 : n1;
    const bool c2 = (c1 ? vB : vA) < x;
    const uint32_t h2 = k2 >> 1;
    lo = c2 ? l2 + h2 + 1 : l2;
    n = c2 ? k2 - h2 - 1 : h2;
  }
  while (n > 0) {
    uint32_t half = n >> 1;
    uint32_t mid = lo + half;
    bool less = a[mid] < x;
    lo = less ? mid + 1 : lo;
    n = less ? n - half - 1 : half;
  }
  return lo;
}

// is packed value x visible in [lo, hi) at update position pos
__device__ __forceinline__ bool contains(const uint32_t* __restrict__ adj,
                                         const uint32_t* __restrict__ ats, uint32_t lo,
                                         uint32_t hi, uint32_t x, bool versioned, uint32_t pos) {
  uint32_t i = lower_bound(adj, lo, hi, x);
  if (!versioned) return i < hi && adj[i] == x;
  for (; i < hi && adj[i] == x; ++i)
    if (ts_visible(ats[i], pos)) return true;
  return false;
}

// contains() for a segment [lo, lo + n) that is the same for every lane of
// the wave (only x differs): the halvings depend on n only, so all lanes
// step together, two halvings per round trip (3 pivots loaded together); the
// first pivots are one address for the whole wave.  Finds the last entry
// <= x, then (versioned) scans back over equal entries for a visible one.
__device__ __forceinline__ bool contains_u(const uint32_t* __restrict__ adj,
                                           const uint32_t* __restrict__ ats, uint32_t lo,
                                           uint32_t n, uint32_t x, bool versioned, uint32_t pos) {
  if (n == 0) return false;
  uint32_t b = lo, m = n;
  while (m > 2) {
    const uint32_t h1 = m >> 1, m1 = m - h1;
    const uint32_t h2 = m1 >> 1;
    const uint32_t v1 = adj[b + h1], vA = adj[b + h2], vB = adj[b + h1 + h2];
    const bool c1 = v1 <= x;
    const bool c2 = (c1 ? vB : vA) <= x;
    b = c1 ? b + h1 : b;
    b = c2 ? b + h2 : b;
    m = m1 - h2;
  }
  if (m == 2) b = adj[b + 1] <= x ? b + 1 : b;
  if (!versioned) return adj[b] == x;
  for (uint32_t i = b + 1; i-- > lo && adj[i] == x;)
    if (ts_visible(ats[i], pos)) return true;
  return false;
}

// is x in the ascending LDS array a[0, n): the halvings depend on n only, so
// every lane runs the same number of steps (no divergent loop exits)
__device__ __forceinline__ bool lds_contains(const uint32_t* a, uint32_t n, uint32_t x) {
  if (n == 0) return false;
  uint32_t b = 0;
  for (uint32_t m = n; m > 1;) {
    const uint32_t half = m >> 1;
    b = a[b + half] <= x ? b + half : b;
    m -= half;
  }
  return a[b] == x;
}

// Runs of the sorted insertion keys (vertex << 32 | entry, padded with ~0):
// end of the run of vertex v that starts at bs, i.e. the first index after
// bs whose key is not v's (a larger vertex or the padding), or n.
// Galloping, so a run of b keys costs O(log b) loads.
__device__ __forceinline__ uint32_t run_end(const uint64_t* __restrict__ keys, uint32_t bs,
                                            uint32_t n, uint32_t v) {
  const uint64_t bound = ((uint64_t)v + 1) << 32;
  uint32_t lo = bs + 1;  // every index below lo is in the run
  uint32_t hi = n;       // n, or an index known to be past the run
  for (uint64_t step = 1;; step <<= 1) {
    const uint64_t probe = lo + step - 1;
    if (probe >= n) break;
    if (keys[probe] >= bound) {
      hi = (uint32_t)probe;
      break;
    }
    lo = (uint32_t)probe + 1;
  }
  while (lo < hi) {
    const uint32_t m = lo + ((hi - lo) >> 1);
    if (keys[m] >= bound) hi = m;
    else lo = m + 1;
  }
  return lo;
}

}  // namespace csm
