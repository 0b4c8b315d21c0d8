// SymBi index on the device: d1/d2 bitmaps and the n1/np1/n2/nc2 counters of
// the reference's DCS/DCG (ParaCOSM/CSM/matching/SymBi/symbi.cpp:246-466),
// one bitmap/counter row per (query, query vertex), per DAG edge (n1) and per
// directed query edge (n2), each row indexed by data vertex id.  The DCS
// lists themselves are the label-partitioned segments of the device graph.
//
// Definitions (fix point over the BFS DAG of the query):
//   n1[u<-p][v]  = #{w in S(v, L(p), l(u,p)) : d1[p][w]}           (p parent of u)
//   np1[u][v]    = #{parents p : n1[u<-p][v] > 0},  d1[u][v] = np1 == |parents|
//   n2[u->c][v]  = #{w in S(v, L(c), l(u,c)) : d2[c][w]}           (any neighbour c)
//   nc2[u][v]    = #{children c : n2[u->c][v] > 0}, d2[u][v] = d1 && nc2 == |children|
//
// A batch changes the index in two bulk phases around the search: insertions
// before it (the index then describes the union of the graph before the batch
// and the batch's insertions -- a superset, so d2 is still a sound filter for
// every update's exact search) and deletions after it.  Each phase is a
// snapshot step over the changed edges followed by frontier propagation, d1
// first, then d2.  A counter transition (0 -> 1, P -1 -> P, ...) is observed by
// exactly one atomic, and flag writes happen in their own kernels, so the
// result is the fix point whatever the order of the atomics.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "csm_kernels.hpp"

namespace csm {

constexpr uint32_t ERR_INDEX_OVERFLOW = 32u;

// One row per (query, query vertex u).
struct SxRow {
  uint32_t label;                 // dense label of u
  uint32_t np, nc;                // parents, children
  uint32_t pad;
  uint32_t par_row[MAXNB];        // flag row of each parent p
  uint32_t par_key[MAXNB];        // key of S(v, L(p), l(u,p))
  uint32_t par_n1[MAXNB];         // n1 row of (u <- p)
  uint32_t par_n2up[MAXNB];       // n2 row of (u -> p)
  uint32_t par_n2dn[MAXNB];       // n2 row of (p -> u)
  uint32_t ch_row[MAXNB];
  uint32_t ch_key[MAXNB];         // key of S(v, L(c), l(u,c))
  uint32_t ch_n1[MAXNB];          // n1 row of (c <- u)
  uint32_t ch_n2dn[MAXNB];        // n2 row of (u -> c)
  uint32_t ch_n2up[MAXNB];        // n2 row of (c -> u)
};

struct SxView {
  const SxRow* __restrict__ rows;
  uint32_t* d1;
  uint32_t* d2;
  uint32_t* np1;
  uint32_t* nc2;
  uint32_t* n1;
  uint32_t* n2;
  uint32_t words;      // bitmap words per row
  uint32_t nvc;        // counter entries per row (vertex capacity)
};

struct SxList {
  unsigned long long* buf;   // (row << 32) | v
  uint32_t* count;
  uint32_t cap;
};

__device__ __forceinline__ bool bit_get(const uint32_t* b, uint32_t words, uint32_t row, uint32_t v) {
  return (b[(size_t)row * words + (v >> 5)] >> (v & 31)) & 1u;
}
__device__ __forceinline__ bool bit_set(uint32_t* b, uint32_t words, uint32_t row, uint32_t v) {
  uint32_t m = 1u << (v & 31);
  return !(atomicOr(&b[(size_t)row * words + (v >> 5)], m) & m);
}
__device__ __forceinline__ bool bit_clear(uint32_t* b, uint32_t words, uint32_t row, uint32_t v) {
  uint32_t m = 1u << (v & 31);
  return (atomicAnd(&b[(size_t)row * words + (v >> 5)], ~m) & m) != 0;
}
__device__ __forceinline__ uint32_t* cnt_at(uint32_t* base, uint32_t nvc, uint32_t row, uint32_t v) {
  return base + (size_t)row * nvc + v;
}
__device__ __forceinline__ void push(const SxList& l, uint32_t row, uint32_t v, uint32_t* err) {
  uint32_t i = atomicAdd(l.count, 1u);
  if (i < l.cap) l.buf[i] = ((unsigned long long)row << 32) | v;
  else atomicOr(err, ERR_INDEX_OVERFLOW);
}

// entry i of a segment [lo, hi) present in the union view (graph before the
// batch plus its insertions: the first non-hidden version of each value) or
// the end view (alive after the batch)
__device__ __forceinline__ bool present(const GraphView& g, uint32_t i, uint32_t lo, bool dirty,
                                        bool end_view) {
  if (!dirty) return true;
  uint32_t ts = g.ats[i];
  if (ts == TS_DEAD) return false;
  if (end_view) return (ts & 0xFFFFu) == 0xFFFFu;
  return i == lo || g.adj[i - 1] != g.adj[i];
}

// ------------------------------------------------------------------ build
// mode 0: top-down rows (n1, np1, d1) of flag row r
// mode 1: bottom-up rows (n2 of children, nc2, d2)
// mode 2: n2 of parents (after all d2 are final)
__global__ void k_sx_build(GraphView g, SxView x, uint32_t r, int mode) {
  const SxRow* R = x.rows + r;
  const uint32_t lab = R->label;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < g.nv; v += gridDim.x * blockDim.x) {
    if (g.vlab[v] != lab) continue;
    uint2 rec = g.vrec[v];
    uint32_t off = rec.x, len = rec.y & LEN_MASK;
    bool dirty = (rec.y & DIRTY) != 0;
    if (mode == 0) {
      uint32_t np = 0;
      for (uint32_t j = 0; j < R->np; ++j) {
        uint2 sg = seg_at(g, v, off, len, R->par_key[j]);
        uint32_t c = 0, pr = R->par_row[j];
        for (uint32_t i = sg.x; i < sg.y; ++i)
          if (present(g, i, sg.x, dirty, true) && bit_get(x.d1, x.words, pr, g.adj[i] & g.vmask)) ++c;
        *cnt_at(x.n1, x.nvc, R->par_n1[j], v) = c;
        np += c > 0;
      }
      *cnt_at(x.np1, x.nvc, r, v) = np;
      if (np == R->np) bit_set(x.d1, x.words, r, v);
    } else if (mode == 1) {
      uint32_t nc = 0;
      for (uint32_t j = 0; j < R->nc; ++j) {
        uint2 sg = seg_at(g, v, off, len, R->ch_key[j]);
        uint32_t c = 0, cr = R->ch_row[j];
        for (uint32_t i = sg.x; i < sg.y; ++i)
          if (present(g, i, sg.x, dirty, true) && bit_get(x.d2, x.words, cr, g.adj[i] & g.vmask)) ++c;
        *cnt_at(x.n2, x.nvc, R->ch_n2dn[j], v) = c;
        nc += c > 0;
      }
      *cnt_at(x.nc2, x.nvc, r, v) = nc;
      if (nc == R->nc && bit_get(x.d1, x.words, r, v)) bit_set(x.d2, x.words, r, v);
    } else {
      for (uint32_t j = 0; j < R->np; ++j) {
        uint2 sg = seg_at(g, v, off, len, R->par_key[j]);
        uint32_t c = 0, pr = R->par_row[j];
        for (uint32_t i = sg.x; i < sg.y; ++i)
          if (present(g, i, sg.x, dirty, true) && bit_get(x.d2, x.words, pr, g.adj[i] & g.vmask)) ++c;
        *cnt_at(x.n2, x.nvc, R->par_n2up[j], v) = c;
      }
    }
  }
}

// TurboFlux's explicit-path counts (CountDownwards, parallel_turboflux.cpp:
// 469-491), in closed form over one tree edge u -> c at a time, parents before
// children.  calls_u[v] is the number of times CountDownwards(u, v) runs: 1
// for every data vertex with the root's label (GenerateMatchingOrder calls it
// once per key of DCS_[root -> first child], :536-541), and for a child the
// number of distinct parent vertices whose walk reached it (the walk of a
// (c, parent vertex) pair happens on its first run only; later runs add its
// memoised count).  Each run of (u, v) adds to paths[c] the number of
// d2-valid children w in S(v, L(c), l(u,c)); each first walk adds 1 to
// calls_c[w].
__global__ void k_tf_root_calls(const uint32_t* __restrict__ vlab, uint32_t lab,
                                uint32_t* __restrict__ calls, uint32_t nv) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += gridDim.x * blockDim.x)
    calls[v] = vlab[v] == lab ? 1u : 0u;
}

__global__ void k_tf_paths(GraphView g, SxView x, uint32_t key, uint32_t crow,
                           const uint32_t* __restrict__ calls_u, uint32_t* __restrict__ calls_c,
                           unsigned long long* __restrict__ paths) {
  unsigned long long acc = 0;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < g.nv; v += gridDim.x * blockDim.x) {
    const uint32_t runs = calls_u[v];
    if (runs == 0) continue;
    const uint2 rec = g.vrec[v];
    const bool dirty = (rec.y & DIRTY) != 0;
    const uint2 sg = seg_at(g, v, rec.x, rec.y & LEN_MASK, key);
    uint32_t n = 0;
    for (uint32_t i = sg.x; i < sg.y; ++i) {
      if (!present(g, i, sg.x, dirty, true)) continue;
      const uint32_t w = g.adj[i] & g.vmask;
      if (!bit_get(x.d2, x.words, crow, w)) continue;
      ++n;
      atomicAdd(&calls_c[w], 1u);
    }
    acc += (unsigned long long)runs * n;
  }
  if (acc) atomicAdd(paths, acc);
}

// a new vertex: root rows get d1, rows without children also d2
__global__ void k_sx_add_vertex(SxView x, uint32_t nrows, uint32_t v, uint32_t lab) {
  for (uint32_t r = threadIdx.x; r < nrows; r += blockDim.x) {
    const SxRow* R = x.rows + r;
    if (R->label != lab) continue;
    if (R->np == 0) {
      bit_set(x.d1, x.words, r, v);
      if (R->nc == 0) bit_set(x.d2, x.words, r, v);
    }
  }
}

// Vertex deletion: v leaves every index row.  Its incident edges are already
// deleted, so what remains are the unconditional d1/d2 bits of rows without
// parents (and children) and counters that are zero.
__global__ void k_sx_remove_vertex(SxView x, uint32_t nrows, uint32_t n1rows, uint32_t n2rows,
                                   uint32_t v) {
  for (uint32_t r = threadIdx.x; r < nrows; r += blockDim.x) {
    bit_clear(x.d1, x.words, r, v);
    bit_clear(x.d2, x.words, r, v);
    *cnt_at(x.np1, x.nvc, r, v) = 0;
    *cnt_at(x.nc2, x.nvc, r, v) = 0;
  }
  for (uint32_t r = threadIdx.x; r < n1rows; r += blockDim.x) *cnt_at(x.n1, x.nvc, r, v) = 0;
  for (uint32_t r = threadIdx.x; r < n2rows; r += blockDim.x) *cnt_at(x.n2, x.nvc, r, v) = 0;
}

// ---------------------------------------------------- changed-edge marks
// Per update: 1 = insertion of an edge absent before the batch (first version
// of its value in the list), 2 = deletion leaving the edge absent after it
// (its version is the last visible one), 0 otherwise.
__global__ void k_sx_mark(BatchView b, GraphView g, const uint32_t* __restrict__ slab,
                          uint8_t* __restrict__ mark) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < b.n; i += gridDim.x * blockDim.x) {
    uint8_t mk = 0;
    uint32_t l = slab[i];
    uint32_t x = b.v1[i], y = b.v2[i];
    if (b.type[i] == 'e' && l != NO_LABEL && x < g.nv && y < g.nv && g.vlab[y] != NO_LABEL) {
      uint32_t pos = i + 1;
      uint2 r = g.vrec[x];
      uint32_t off = r.x, end = r.x + (r.y & LEN_MASK);
      uint32_t val = ((g.vlab[y] * g.nel + l) << g.vbits) | y;
      uint32_t j = lower_bound(g.adj, off, end, val);
      if (j < end && g.adj[j] == val) {
        if (b.add[i]) {
          uint32_t ts = g.ats[j];
          if (ts != TS_DEAD && (ts >> 16) == pos) mk = 1;
        } else {
          uint32_t last = NONE;
          for (uint32_t t = j; t < end && g.adj[t] == val; ++t)
            if (g.ats[t] != TS_DEAD) last = t;
          if (last != NONE && (g.ats[last] & 0xFFFFu) == pos) mk = 2;
        }
      }
    }
    mark[i] = mk;
  }
}

struct SxSeedArgs {
  BatchView b;
  GraphView g;
  SxView x;
  const uint32_t* __restrict__ slab;
  const uint8_t* __restrict__ mark;
  const uint32_t* __restrict__ tab_off;   // bucket (L(p), L(c), l) -> DAG edges
  const uint32_t* __restrict__ tab_ent;   // (parent row << 8) | child index
  uint32_t nlab;
  uint32_t want;                          // 1 insert phase, 2 delete phase
  int stage;                              // 0: d1 part, 1: d2 part
  SxList out;
  uint32_t* err;
};

// Snapshot step over the changed edges: for every DAG edge (p -> c) whose
// labels match the edge in either orientation, the increments/decrements of
// symbi.cpp:683-700 (insertion) and :904-912 (deletion) with the flags read
// before any propagation of this phase.
__global__ void k_sx_seed(SxSeedArgs a) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < a.b.n; i += gridDim.x * blockDim.x) {
    if (a.mark[i] != a.want) continue;
    const uint32_t l = a.slab[i];
    const bool ins = a.want == 1;
    for (int o = 0; o < 2; ++o) {
      uint32_t x1 = o ? a.b.v2[i] : a.b.v1[i], x2 = o ? a.b.v1[i] : a.b.v2[i];
      uint32_t l1 = a.g.vlab[x1], l2 = a.g.vlab[x2];
      if (l1 >= a.nlab || l2 >= a.nlab) continue;
      uint32_t bucket = (l1 * a.nlab + l2) * a.g.nel + l;
      for (uint32_t e = a.tab_off[bucket]; e < a.tab_off[bucket + 1]; ++e) {
        uint32_t pr = a.tab_ent[e] >> 8, j = a.tab_ent[e] & 0xFFu;
        const SxRow* P = a.x.rows + pr;
        uint32_t cr = P->ch_row[j];
        const SxRow* C = a.x.rows + cr;
        if (a.stage == 0) {
          if (!bit_get(a.x.d1, a.x.words, pr, x1)) continue;
          uint32_t* n1 = cnt_at(a.x.n1, a.x.nvc, P->ch_n1[j], x2);
          if (ins) {
            if (atomicAdd(n1, 1u) == 0u)
              if (atomicAdd(cnt_at(a.x.np1, a.x.nvc, cr, x2), 1u) + 1u == C->np)
                push(a.out, cr, x2, a.err);
          } else {
            if (atomicSub(n1, 1u) == 1u)
              if (atomicSub(cnt_at(a.x.np1, a.x.nvc, cr, x2), 1u) == C->np)
                push(a.out, cr, x2, a.err);
          }
        } else {
          if (bit_get(a.x.d2, a.x.words, cr, x2)) {
            uint32_t* n2 = cnt_at(a.x.n2, a.x.nvc, P->ch_n2dn[j], x1);
            if (ins) {
              if (atomicAdd(n2, 1u) == 0u)
                if (atomicAdd(cnt_at(a.x.nc2, a.x.nvc, pr, x1), 1u) + 1u == P->nc &&
                    bit_get(a.x.d1, a.x.words, pr, x1))
                  push(a.out, pr, x1, a.err);
            } else {
              if (atomicSub(n2, 1u) == 1u)
                if (atomicSub(cnt_at(a.x.nc2, a.x.nvc, pr, x1), 1u) == P->nc)
                  push(a.out, pr, x1, a.err);
            }
          }
          if (bit_get(a.x.d2, a.x.words, pr, x1)) {
            uint32_t* n2 = cnt_at(a.x.n2, a.x.nvc, P->ch_n2up[j], x2);
            if (ins) atomicAdd(n2, 1u); else atomicSub(n2, 1u);
          }
        }
      }
    }
  }
}

// Flag step: set (or clear) the flag of every candidate; those that changed
// form the frontier; `also` (nullable) records every changed entry.
__global__ void k_sx_flags(SxList cand, uint32_t* bits, uint32_t words, bool set, SxList front,
                           SxList also, uint32_t* err) {
  uint32_t n = *cand.count;
  if (n > cand.cap) n = cand.cap;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    unsigned long long e = cand.buf[i];
    uint32_t r = (uint32_t)(e >> 32), v = (uint32_t)e;
    bool changed = set ? bit_set(bits, words, r, v) : bit_clear(bits, words, r, v);
    if (!changed) continue;
    push(front, r, v, err);
    if (also.buf) push(also, r, v, err);
  }
}

// Newly d1 entries with every child supported become d2 candidates
// (the d2 part of InsertionTopDown, symbi.cpp:410-416); entries that lost d1
// lose d2 too (symbi.cpp:440-447).
__global__ void k_sx_from_d1(SxList src, SxView x, bool ins, SxList out, uint32_t* err) {
  uint32_t n = *src.count;
  if (n > src.cap) n = src.cap;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    unsigned long long e = src.buf[i];
    uint32_t r = (uint32_t)(e >> 32), v = (uint32_t)e;
    if (!ins || *cnt_at(x.nc2, x.nvc, r, v) == x.rows[r].nc) push(out, r, v, err);
  }
}

// Propagation from a frontier, one wavefront per entry (u, v) scanning the
// DCS segments of v:
//   stage 0 (d1 changed): children c, w in S(v, L(c), l): n1[c<-u][w] +-1,
//            transitions of np1[c][w] give d1 candidates (symbi.cpp:718-726,
//            :920-928)
//   stage 1 (d2 changed): parents p, w in S(v, L(p), l): n2[p->u][w] +-1,
//            transitions of nc2[p][w] give d2 candidates; children c:
//            n2[c->u][w] +-1 (symbi.cpp:728-746, :930-942)
__global__ void k_sx_propagate(GraphView g, SxView x, SxList front, bool ins, bool end_view,
                               int stage, SxList out, uint32_t* err) {
  const uint32_t lane = lane_id();
  const uint32_t wid = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
  const uint32_t nw = (gridDim.x * blockDim.x) / WAVE;
  uint32_t n = *front.count;
  if (n > front.cap) n = front.cap;
  for (uint32_t it = wid; it < n; it += nw) {
    unsigned long long e = front.buf[it];
    const uint32_t r = (uint32_t)(e >> 32), v = (uint32_t)e;
    const SxRow* R = x.rows + r;
    uint2 rec = g.vrec[v];
    const uint32_t off = rec.x, len = rec.y & LEN_MASK;
    const bool dirty = (rec.y & DIRTY) != 0;
    if (stage == 0) {
      for (uint32_t j = 0; j < R->nc; ++j) {
        uint2 sg = seg_at(g, v, off, len, R->ch_key[j]);
        const uint32_t cr = R->ch_row[j], nprow = R->ch_n1[j];
        const uint32_t need = x.rows[cr].np;
        for (uint32_t i = sg.x + lane; i < sg.y; i += WAVE) {
          if (!present(g, i, sg.x, dirty, end_view)) continue;
          uint32_t w = g.adj[i] & g.vmask;
          uint32_t* c1 = cnt_at(x.n1, x.nvc, nprow, w);
          if (ins) {
            if (atomicAdd(c1, 1u) == 0u)
              if (atomicAdd(cnt_at(x.np1, x.nvc, cr, w), 1u) + 1u == need) push(out, cr, w, err);
          } else {
            if (atomicSub(c1, 1u) == 1u)
              if (atomicSub(cnt_at(x.np1, x.nvc, cr, w), 1u) == need) push(out, cr, w, err);
          }
        }
      }
    } else {
      for (uint32_t j = 0; j < R->np; ++j) {
        uint2 sg = seg_at(g, v, off, len, R->par_key[j]);
        const uint32_t pr = R->par_row[j], n2row = R->par_n2dn[j];
        const uint32_t need = x.rows[pr].nc;
        for (uint32_t i = sg.x + lane; i < sg.y; i += WAVE) {
          if (!present(g, i, sg.x, dirty, end_view)) continue;
          uint32_t w = g.adj[i] & g.vmask;
          uint32_t* c2 = cnt_at(x.n2, x.nvc, n2row, w);
          if (ins) {
            if (atomicAdd(c2, 1u) == 0u)
              if (atomicAdd(cnt_at(x.nc2, x.nvc, pr, w), 1u) + 1u == need &&
                  bit_get(x.d1, x.words, pr, w))
                push(out, pr, w, err);
          } else {
            if (atomicSub(c2, 1u) == 1u)
              if (atomicSub(cnt_at(x.nc2, x.nvc, pr, w), 1u) == need) push(out, pr, w, err);
          }
        }
      }
      for (uint32_t j = 0; j < R->nc; ++j) {
        uint2 sg = seg_at(g, v, off, len, R->ch_key[j]);
        const uint32_t n2row = R->ch_n2up[j];
        for (uint32_t i = sg.x + lane; i < sg.y; i += WAVE) {
          if (!present(g, i, sg.x, dirty, end_view)) continue;
          uint32_t w = g.adj[i] & g.vmask;
          uint32_t* c2 = cnt_at(x.n2, x.nvc, n2row, w);
          if (ins) atomicAdd(c2, 1u); else atomicSub(c2, 1u);
        }
      }
    }
  }
}

// propagation must have drained within the launched levels
__global__ void k_sx_check(const uint32_t* count, uint32_t* err) {
  if (threadIdx.x == 0 && *count) atomicOr(err, ERR_INDEX_OVERFLOW);
}

__global__ void k_sx_zero(uint32_t* counts, uint32_t n) {
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) counts[i] = 0;
}

}  // namespace csm
