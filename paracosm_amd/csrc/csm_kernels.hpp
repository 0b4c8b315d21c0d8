// gfx950 kernels of the incremental match engine.
//
// Device graph layout (HBM):
//   vrec[v]  = {off, len | DIRTY}: the neighbour list of v is adj[off, off+len)
//   adj[i]   = packed neighbour entry  (key << vbits) | nbr,  key = L(nbr)*nel + elabel
//              so each list is sorted by (neighbour label, edge label, id) and
//              the label-filtered neighbour set S(v, L, l) -- the reference's
//              DCS list (ParaCOSM/CSM/matching/SymBi/symbi.cpp:404-529) and the
//              label test of the GraphFlow scan (graphflow.cpp:212-218) -- is
//              one contiguous segment of it.
//   ats[i]   = version word of adj[i]: (ins << 16) | del, positions inside the
//              current batch; an entry is visible to the search of the update
//              at position p iff ins <= p <= del.  Untouched entries hold
//              TS_BASE (always visible); only lists flagged DIRTY hold others.
//   vlab[v]  = dense vertex label (NO_LABEL when the vertex does not exist).
//   dslot[v] = key directory slot of v (NONE: none); dir[slot*dstride + b] is
//              the first index, relative to the list start, whose key is >= b
//              (b = 0..nkeys, the last entry is the list length), so a segment
//              lookup costs two dependent loads instead of two binary searches
//              over the whole list.  Lists of at least a threshold length get
//              a slot; every physical rewrite of a list with a slot rebuilds
//              its directory before the next search.
//
// Lists are never edited in place: a list that changes is rewritten into
// fresh space taken from the pool with one atomic bump, and the pool is
// repacked when it runs low.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "csm_search.hpp"

namespace csm {

constexpr uint32_t DIRTY = 0x80000000u;
constexpr uint32_t LEN_MASK = 0x7fffffffu;
constexpr uint32_t TS_BASE = 0x0000FFFFu;   // ins 0, never deleted
constexpr uint32_t TS_DEAD = 0xFFFF0000u;   // never visible
constexpr uint32_t POS_MAX = 0xFFFEu;       // largest update position in a batch
constexpr uint32_t NO_LABEL = 0xFFFFFFFFu;
constexpr uint32_t NONE = 0xFFFFFFFFu;
constexpr int MAXK = 32;                    // query vertices supported by the search
constexpr int MAXNB = MAXK - 1;
constexpr int WAVE = 64;
constexpr int SEARCH_WAVES = 4;             // waves per search workgroup (MK <= 16)
// waves per search workgroup for queries of up to MK vertices: the per-wave
// LDS state grows with MK^2, so the largest layout runs two waves per group
template <int MK>
constexpr int search_waves() { return MK > 16 ? 2 : SEARCH_WAVES; }
constexpr uint32_t FCAP = 512;              // LDS leaf filter entries per wave
constexpr uint32_t LEAF_SPLIT = 8192;       // a parent with more leaves becomes a task

// error bits in Ctrl::err
constexpr uint32_t ERR_DELETE_MISSING = 1u;
constexpr uint32_t ERR_SEED_OVERFLOW = 2u;
constexpr uint32_t ERR_CHUNK_OVERFLOW = 4u;
constexpr uint32_t ERR_POOL_OVERFLOW = 8u;
constexpr uint32_t ERR_BAD_UPDATE = 16u;
constexpr uint32_t ERR_VERTEX_UPDATE = 64u;   // a vertex record inside a device batch

constexpr uint32_t ROUND_GROUP_MAX = 16;   // search rounds enqueued per host read-back
constexpr uint32_t NXCD = 8;           // MI355X: 8 XCDs, workgroups dispatched round-robin
constexpr uint32_t XHEAD_STRIDE = 32;  // u32 words between two XCD chunk heads

struct Ctrl {
  uint32_t pool_top;      // bump pointer of the current pool (entries)
  uint32_t nseeds;        // seeds emitted this batch
  uint32_t nchunks;       // search tasks (seed chunks) this batch
  uint32_t chunk_head;    // next chunk to run
  uint32_t err;           // ERR_* bits
  uint32_t err_at;        // first failing update index (min)
  uint32_t nruns;         // vertices with inserted entries this batch
  uint32_t ndirty_next;   // dirty vertices recorded for the next batch
  uint32_t ndirty_cur;    // dirty vertices to compact at the start of this batch
  uint32_t abort;         // host-set: stop searching
  uint32_t nspill;        // tasks spilled by the current search round
  uint32_t nspill_chunks; // their chunks
  uint32_t dir_top;       // key directory slots handed out
  uint32_t rounds;        // spilled rounds that had chunks, this batch
  uint32_t pad[2];
  unsigned long long live;        // live entries (sum of list lengths)
  unsigned long long stat_tasks;
  unsigned long long stat_cands;
  unsigned long long stat_bytes;
  unsigned long long prof[32];    // per-phase wave cycles, calls and sizes (CSM_PHASE_PROF builds)
  // next chunk of each XCD's share of the search round, one 128 B line each
  uint32_t xhead[NXCD * XHEAD_STRIDE];
  // tasks and chunk ids each XCD has spilled into its region of the spill
  // buffers this round, and the chunk count of each region for this round
  uint32_t xspill[NXCD * XHEAD_STRIDE];
  uint32_t xspill_chunks[NXCD * XHEAD_STRIDE];
  uint32_t xn[NXCD];
  // rounds enqueued back to back: slot i is set by the i-th launch of the
  // group once a wave of it ran a task (the host counts only busy launches)
  uint32_t round_busy[ROUND_GROUP_MAX];
};

// One search plan per (query, matching order): positions 0..k-1.
struct PlanRec {
  uint32_t k;
  uint32_t q;                    // query id (result column)
  uint32_t symbi;                // 1: filter candidates by the d2 bitmap
  // search settings of this plan's query size (the autotuner's choice):
  // candidates per chunk when a range is cut, and the work units a task
  // may use before it spills
  uint16_t chunk, budget;
  uint32_t qv[MAXK];             // query vertex at position p
  uint32_t nb[MAXK];             // backward neighbours of position p
  uint32_t bw[MAXK][MAXNB];      // (backward position) | (segment key << 8)
  uint32_t d2row[MAXK];          // row of position p's query vertex in the d2 bitmap
};

// Time limit, result cap and work counters of a search (SearchArgs::lim).
struct SearchLimits {
  // time limit: a host-written word in fine-grained host memory; once it
  // reads non-zero, waves stop their tasks and drain the remaining chunks,
  // flagging each affected update with USTAT_TIME_LIMIT and its query in qtl
  const uint32_t* habort;
  uint32_t* ustatus;          // per update of the batch: USTAT_* bits
  uint32_t* qtl;              // per query: 1 once the time limit dropped some of its work
  unsigned long long* uwork;  // per update: candidate entries tested
  unsigned long long ucap;    // per (update, query) result cap
  // embedding rows (--print-enum): row r at emb + r * emb_w holds the
  // update, the query and the data vertex of each query vertex; emb_n counts
  // every embedding found, rows past emb_cap are not written
  uint32_t* emb;
  unsigned long long* emb_n;
  unsigned long long emb_cap;
  uint32_t emb_w;
};
constexpr uint32_t LIM_TIME = 1u, LIM_CAP = 2u, LIM_WORK = 4u, LIM_ENUM = 8u;


// A search task: update t, plan qo, the first `depth` positions fixed in m[],
// and the candidate range of position `depth` (absolute pool offsets into the
// segment of backward neighbour `jmin`; hi == NONE: the whole segment).  The
// seeds of an update are depth-2 tasks; tasks of deeper prefixes come from
// spills.  Chunk c of a task covers [lo + c*chunk, lo + (c+1)*chunk).
struct Task {
  uint32_t t;
  uint32_t qo;
  uint32_t depth;
  uint32_t jmin;
  uint32_t lo, hi;
  uint32_t base;     // first chunk id
  uint32_t csize;    // candidates per chunk
  uint32_t m[MAXK];
};
constexpr uint32_t TASK_WORDS = sizeof(Task) / 4;
constexpr uint32_t MAX_TASK_CHUNKS = 64;   // a task is cut into at most this many chunks

// chunks of a range of n candidates: at least `chunk` each, at most MAX_TASK_CHUNKS
__device__ __forceinline__ uint2 chunking(uint32_t n, uint32_t chunk) {
  uint32_t nch = (n + chunk - 1) / chunk;
  if (nch > MAX_TASK_CHUNKS) nch = MAX_TASK_CHUNKS;
  if (nch == 0) nch = 1;
  return make_uint2(nch, (n + nch - 1) / nch);
}

struct GraphView {
  const uint2* __restrict__ vrec;
  const uint32_t* __restrict__ vlab;
  const uint32_t* __restrict__ adj;
  const uint32_t* __restrict__ ats;
  const uint32_t* __restrict__ dslot;
  const uint32_t* __restrict__ dir;
  uint32_t dstride;
  uint32_t vbits;
  uint32_t vmask;
  uint32_t nel;
  uint32_t nv;
};

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

// segment of key `key` inside the list [off, off+len)
__device__ __forceinline__ uint2 segment(const uint32_t* __restrict__ adj, uint32_t off,
                                         uint32_t len, uint32_t key, uint32_t vbits) {
  uint32_t end = off + len;
  uint32_t lo = lower_bound(adj, off, end, key << vbits);
  uint64_t hk = (uint64_t)(key + 1) << vbits;
  uint32_t hi = hk > 0xFFFFFFFFull ? end : lower_bound(adj, lo, end, (uint32_t)hk);
  return make_uint2(lo, hi);
}

// segment of key `key` in v's list [off, off+len), through v's key directory
// when it has one
__device__ __forceinline__ uint2 seg_at(const GraphView& g, uint32_t v, uint32_t off,
                                        uint32_t len, uint32_t key) {
  const uint32_t s = g.dslot[v];
  if (s != NONE) {
    const uint32_t* __restrict__ D = g.dir + (size_t)s * g.dstride + key;
    return make_uint2(off + D[0], off + D[1]);
  }
  return segment(g.adj, off, len, key, g.vbits);
}

// the same from v's record; *rec_y receives the list's length word (DIRTY in
// bit 31)
__device__ __forceinline__ uint2 seg_of(const GraphView& g, uint32_t v, uint32_t key,
                                        uint32_t* rec_y) {
  const uint2 r = g.vrec[v];
  *rec_y = r.y;
  return seg_at(g, v, r.x, r.y & LEN_MASK, key);
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & (WAVE - 1); }

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ uint32_t popc64(uint64_t x) { return (uint32_t)__popcll(x); }

__device__ __forceinline__ uint32_t uniform(uint32_t x) {
  return __builtin_amdgcn_readfirstlane(x);
}

template <typename T>
__device__ __forceinline__ T wave_sum(T x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
  return x;
}

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    uint64_t y = __shfl_xor(x, o);
    x = y < x ? y : x;
  }
  return x;
}

// (least x, its lane) over lanes [0, n) (ties to the lower lane; value
// 0xFFFFFFFF and lane 0 when every value is 0xFFFFFFFF or n == 0): the same
// as wave_min_u64((x << 32) | lane) when the lanes from n on hold
// 0xFFFFFFFF, with n scalar reads instead of a 64-lane reduction
__device__ __forceinline__ uint64_t lanes_min_u32(uint32_t x, uint32_t n) {
  uint32_t best = 0xFFFFFFFFu, bl = 0;
  for (uint32_t j = 0; j < n; ++j) {
    const uint32_t v = __builtin_amdgcn_readlane(x, j);
    if (v < best) { best = v; bl = j; }
  }
  return ((uint64_t)best << 32) | bl;
}

__device__ __forceinline__ bool d2_bit(const uint32_t* __restrict__ d2, uint32_t row,
                                       uint32_t words_per_row, uint32_t v) {
  return (d2[(size_t)row * words_per_row + (v >> 5)] >> (v & 31)) & 1u;
}

// ------------------------------------------------------------ graph build
// Directed entries of an edge list, keyed (src << 32 | nbr) for the duplicate
// pass.  Self-loops are dropped.
__global__ void k_build_keys(const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst,
                             uint64_t n, uint64_t* __restrict__ keys,
                             uint32_t* __restrict__ vals) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t a = src[i], b = dst[i];
    bool loop = a == b;
    keys[2 * i] = loop ? ~0ull : ((uint64_t)a << 32) | b;
    keys[2 * i + 1] = loop ? ~0ull : ((uint64_t)b << 32) | a;
    vals[2 * i] = (uint32_t)i;
    vals[2 * i + 1] = (uint32_t)i;
  }
}

// After a stable sort by (src, nbr): keep the first copy of each pair (the
// earliest input edge, as Graph::AddEdge ignores later duplicates,
// graph.cpp:63-64) and re-key it by (src << 32 | packed entry).
__global__ void k_build_pack(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                             uint64_t n, const uint32_t* __restrict__ vlab,
                             const uint32_t* __restrict__ elab_dense, uint32_t nel,
                             uint32_t vbits, uint64_t* __restrict__ out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t k = keys[i];
    bool dup = (k == ~0ull) || (i > 0 && keys[i - 1] == k);
    if (dup) { out[i] = ~0ull; continue; }
    uint32_t a = (uint32_t)(k >> 32), b = (uint32_t)k;
    // an edge to a vertex without a label (never declared by a 'v' record)
    // takes part in no match and has no segment key: it is left out
    if (vlab[a] == NO_LABEL || vlab[b] == NO_LABEL) { out[i] = ~0ull; continue; }
    uint32_t key = vlab[b] * nel + elab_dense[vals[i]];
    out[i] = ((uint64_t)a << 32) | ((key << vbits) | b);
  }
}

// Sorted (src, packed) entries -> pool + vertex records.
__global__ void k_build_lists(const uint64_t* __restrict__ keys, uint64_t n,
                              uint32_t* __restrict__ adj, uint32_t* __restrict__ ats,
                              uint2* __restrict__ vrec, uint32_t* __restrict__ vend) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t k = keys[i];
    if (k == ~0ull) continue;
    uint32_t a = (uint32_t)(k >> 32);
    adj[i] = (uint32_t)k;
    ats[i] = TS_BASE;
    if (i == 0 || (uint32_t)(keys[i - 1] >> 32) != a) vrec[a].x = (uint32_t)i;
    if (i + 1 == n || keys[i + 1] == ~0ull || (uint32_t)(keys[i + 1] >> 32) != a)
      vend[a] = (uint32_t)(i + 1);
  }
}

__global__ void k_build_lens(uint2* __restrict__ vrec, const uint32_t* __restrict__ vend,
                             uint32_t nv) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += gridDim.x * blockDim.x) {
    uint32_t e = vend[v];
    vrec[v].y = e ? e - vrec[v].x : 0u;
    if (!e) vrec[v].x = 0u;
  }
}

// ------------------------------------------------------------ batch apply
struct BatchView {
  const uint8_t* __restrict__ type;   // 'e' / 'v'
  const uint8_t* __restrict__ add;
  const uint32_t* __restrict__ v1;
  const uint32_t* __restrict__ v2;
  const uint32_t* __restrict__ label; // original edge labels (inserts)
  uint32_t n;
};

// Unpack csm_update records (16 B) into structure-of-arrays.
struct UpdateRec { uint8_t type, add; uint16_t pad; uint32_t id1, id2, label; };

// Vertex records are applied on the host (the host entry points turn them into
// no-op records of type 0); one reaching a device batch is an error.
__global__ void k_unpack(const UpdateRec* __restrict__ u, uint32_t n, uint8_t* __restrict__ type,
                         uint8_t* __restrict__ add, uint32_t* __restrict__ v1,
                         uint32_t* __restrict__ v2, uint32_t* __restrict__ lab,
                         const uint32_t* __restrict__ elab_map, uint32_t elab_map_n,
                         Ctrl* __restrict__ ctrl) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    UpdateRec r = u[i];
    if (r.type == 'v') {
      atomicOr(&ctrl->err, ERR_VERTEX_UPDATE);
      atomicMin(&ctrl->err_at, i);
    }
    type[i] = r.type;
    add[i] = r.add;
    v1[i] = r.id1;
    v2[i] = r.id2;
    // dense edge label for insertions (NO_LABEL if unknown)
    uint32_t l = NO_LABEL;
    if (r.type == 'e' && r.add) {
      // elab_map: sorted original labels; dense id = index
      uint32_t lo = 0, hi = elab_map_n;
      while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (elab_map[mid] < r.label) lo = mid + 1; else hi = mid;
      }
      if (lo < elab_map_n && elab_map[lo] == r.label) l = lo;
    }
    lab[i] = l;
  }
}

// Directed entries of the batch's insertions keyed (src << 32 | packed),
// valued by update position; other records give a sentinel key.
__global__ void k_batch_keys(BatchView b, const uint32_t* __restrict__ vlab, uint32_t nel,
                             uint32_t vbits, uint32_t nv, uint64_t* __restrict__ keys,
                             uint32_t* __restrict__ vals, uint32_t* __restrict__ ctrl_err,
                             uint32_t* __restrict__ ctrl_err_at) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < b.n; i += gridDim.x * blockDim.x) {
    uint64_t k0 = ~0ull, k1 = ~0ull;
    if (b.type[i] == 'e' && b.add[i]) {
      uint32_t a = b.v1[i], c = b.v2[i], l = b.label[i];
      if (a < nv && c < nv && a != c && l != NO_LABEL && vlab[a] != NO_LABEL &&
          vlab[c] != NO_LABEL) {
        k0 = ((uint64_t)a << 32) | (((vlab[c] * nel + l) << vbits) | c);
        k1 = ((uint64_t)c << 32) | (((vlab[a] * nel + l) << vbits) | a);
      } else {
        atomicOr(ctrl_err, ERR_BAD_UPDATE);
        atomicMin(ctrl_err_at, i);
      }
    }
    keys[2 * i] = k0;
    keys[2 * i + 1] = k1;
    vals[2 * i] = i + 1;
    vals[2 * i + 1] = i + 1;
  }
}

// Heads of the sorted insertion entries -> runs (one per touched source).
__global__ void k_runs(const uint64_t* __restrict__ keys, uint32_t n, uint32_t* __restrict__ run_start,
                       uint32_t* __restrict__ nruns) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    uint64_t k = keys[i];
    if (k == ~0ull) continue;
    if (i == 0 || (keys[i - 1] >> 32) != (k >> 32)) run_start[atomicAdd(nruns, 1u)] = i;
  }
}

// Compaction of lists dirtied by the previous batch: drop entries deleted in
// that batch, reset the version of the rest to TS_BASE, clear DIRTY.  One
// workgroup per vertex.  A list without tombstones is fixed in place.
template <int BS>
__global__ __launch_bounds__(BS) void k_compact(uint2* __restrict__ vrec, uint32_t* __restrict__ adj,
                                                uint32_t* __restrict__ ats,
                                                const uint32_t* __restrict__ list,
                                                const uint32_t* __restrict__ count,
                                                Ctrl* __restrict__ ctrl, uint32_t pool_cap) {
  __shared__ uint32_t s_cnt[BS / WAVE];
  __shared__ uint32_t s_total, s_off;
  const uint32_t n = *count;
  for (uint32_t it = blockIdx.x; it < n; it += gridDim.x) {
    uint32_t v = list[it];
    uint2 r = vrec[v];
    uint32_t off = r.x, len = r.y & LEN_MASK;
    // pass 1: tombstones?
    uint32_t dead = 0;
    for (uint32_t i = threadIdx.x; i < len; i += BS) dead += (ats[off + i] & 0xFFFFu) != 0xFFFFu;
    dead = wave_sum(dead);
    if (lane_id() == 0) s_cnt[threadIdx.x / WAVE] = dead;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t t = 0;
      for (int w = 0; w < BS / WAVE; ++w) t += s_cnt[w];
      s_total = t;
      uint32_t keep = len - t;
      if (t) {
        uint32_t o = atomicAdd(&ctrl->pool_top, keep);
        if ((uint64_t)o + keep > pool_cap) { atomicOr(&ctrl->err, ERR_POOL_OVERFLOW); o = NONE; }
        s_off = o;
        atomicAdd(&ctrl->live, (unsigned long long)0 - (unsigned long long)t);
      } else {
        s_off = off;
      }
    }
    __syncthreads();
    uint32_t total_dead = s_total, noff = s_off;
    if (noff == NONE) { __syncthreads(); continue; }
    if (total_dead == 0) {
      for (uint32_t i = threadIdx.x; i < len; i += BS) ats[off + i] = TS_BASE;
    } else {
      // pass 2: ordered stream compaction in chunks of BS
      uint32_t written = 0;
      for (uint32_t c = 0; c < len; c += BS) {
        uint32_t i = c + threadIdx.x;
        bool keep = i < len && (ats[off + i] & 0xFFFFu) == 0xFFFFu;
        uint64_t m = ballot(keep);
        uint32_t before = popc64(m & ((1ull << lane_id()) - 1ull));
        __syncthreads();
        if (lane_id() == 0) s_cnt[threadIdx.x / WAVE] = popc64(m);
        __syncthreads();
        uint32_t wbase = 0, ctot = 0;
        for (int w = 0; w < BS / WAVE; ++w) {
          if (w < (int)(threadIdx.x / WAVE)) wbase += s_cnt[w];
          ctot += s_cnt[w];
        }
        if (keep) {
          adj[noff + written + wbase + before] = adj[off + i];
          ats[noff + written + wbase + before] = TS_BASE;
        }
        written += ctot;
      }
    }
    if (threadIdx.x == 0) vrec[v] = make_uint2(noff, len - total_dead);
    __syncthreads();
  }
}

// Merge each touched vertex's sorted new entries into a fresh copy of its
// (clean) list.  Runs are taken in groups of MERGE_GROUP.  The lanes of wave 0
// size the group's runs, and one atomic per counter reserves the pool space,
// the live count and the dirty-list slots of the whole group: one single-lane
// atomic per run and counter serialises on the one line of Ctrl.  Then the
// workgroup merges the group's runs one after another.  New entries carry
// their update position as `ins`; the list is flagged DIRTY and recorded for
// compaction at the start of the next batch.
constexpr uint32_t MERGE_GROUP = 16;
template <int BS>
__global__ __launch_bounds__(BS) void k_merge(uint2* __restrict__ vrec, uint32_t* __restrict__ adj,
                                              uint32_t* __restrict__ ats,
                                              const uint64_t* __restrict__ keys,
                                              const uint32_t* __restrict__ vals, uint32_t nkeys,
                                              const uint32_t* __restrict__ run_start,
                                              Ctrl* __restrict__ ctrl, uint32_t* __restrict__ dirty_next,
                                              uint32_t pool_cap) {
  static_assert(BS % WAVE == 0 && MERGE_GROUP <= WAVE, "group sized by wave 0");
  __shared__ uint32_t s_v[MERGE_GROUP], s_bs[MERGE_GROUP], s_b[MERGE_GROUP];
  __shared__ uint32_t s_off[MERGE_GROUP], s_a[MERGE_GROUP], s_noff[MERGE_GROUP];
  __shared__ uint32_t s_slot[MERGE_GROUP];
  const uint32_t nr = ctrl->nruns;
  for (uint32_t g0 = blockIdx.x * MERGE_GROUP; g0 < nr; g0 += gridDim.x * MERGE_GROUP) {
    const uint32_t cnt = min(MERGE_GROUP, nr - g0);
    if (threadIdx.x < WAVE) {
      const uint32_t t = threadIdx.x;
      uint32_t v = 0, bs = 0, b = 0, off = 0, a = 0, fresh = 0;
      if (t < cnt) {
        bs = run_start[g0 + t];
        v = (uint32_t)(keys[bs] >> 32);
        b = run_end(keys, bs, nkeys, v) - bs;
        const uint2 r = vrec[v];
        off = r.x;
        a = r.y & LEN_MASK;
        fresh = (r.y & DIRTY) ? 0u : 1u;
      }
      // inclusive prefix sums over lanes 0 .. MERGE_GROUP-1
      const unsigned long long ab = (unsigned long long)a + b;
      unsigned long long iab = ab, ib = b;
      uint32_t ifr = fresh;
#pragma unroll
      for (uint32_t o = 1; o < MERGE_GROUP; o <<= 1) {
        const unsigned long long yab = __shfl_up(iab, o), yb = __shfl_up(ib, o);
        const uint32_t yfr = __shfl_up(ifr, o);
        if (t >= o) { iab += yab; ib += yb; ifr += yfr; }
      }
      const unsigned long long tot_ab = __shfl(iab, MERGE_GROUP - 1);
      const unsigned long long tot_b = __shfl(ib, MERGE_GROUP - 1);
      const uint32_t tot_fr = __shfl(ifr, MERGE_GROUP - 1);
      uint32_t o = NONE, od = 0;
      if (t == 0) {
        if (tot_ab <= pool_cap) {
          o = atomicAdd(&ctrl->pool_top, (uint32_t)tot_ab);
          if ((uint64_t)o + tot_ab > pool_cap) o = NONE;
        }
        if (o == NONE) atomicOr(&ctrl->err, ERR_POOL_OVERFLOW);
        else if (tot_fr) od = atomicAdd(&ctrl->ndirty_next, tot_fr);
        atomicAdd(&ctrl->live, tot_b);
      }
      o = __shfl(o, 0);
      od = __shfl(od, 0);
      if (t < cnt) {
        s_v[t] = v;
        s_bs[t] = bs;
        s_b[t] = b;
        s_off[t] = off;
        s_a[t] = a;
        s_noff[t] = o == NONE ? NONE : o + (uint32_t)(iab - ab);
        s_slot[t] = (o != NONE && fresh) ? od + (ifr - fresh) : NONE;
      }
    }
    __syncthreads();
    for (uint32_t q = 0; q < cnt; ++q) {
      const uint32_t noff = s_noff[q];
      if (noff == NONE) continue;
      const uint32_t v = s_v[q], bs = s_bs[q], b = s_b[q], off = s_off[q], a = s_a[q];
      // old entries: out = i + #(new < old[i])
      for (uint32_t i = threadIdx.x; i < a; i += BS) {
        uint32_t x = adj[off + i];
        uint32_t lo = 0, n = b;
        while (n > 0) {
          uint32_t h = n >> 1;
          bool less = (uint32_t)keys[bs + lo + h] < x;
          lo = less ? lo + h + 1 : lo;
          n = less ? n - h - 1 : h;
        }
        adj[noff + i + lo] = x;
        ats[noff + i + lo] = ats[off + i];
      }
      // new entries: out = j + #(old <= new[j])
      for (uint32_t j = threadIdx.x; j < b; j += BS) {
        uint32_t x = (uint32_t)keys[bs + j];
        uint32_t lo = 0, n = a;
        while (n > 0) {
          uint32_t h = n >> 1;
          bool le = adj[off + lo + h] <= x;
          lo = le ? lo + h + 1 : lo;
          n = le ? n - h - 1 : h;
        }
        adj[noff + j + lo] = x;
        ats[noff + j + lo] = (vals[bs + j] << 16) | 0xFFFFu;
      }
      // v's record is read only in the group pass above, so it can be
      // rewritten while other waves still merge this run
      if (threadIdx.x == 0) {
        vrec[v] = make_uint2(noff, (a + b) | DIRTY);
        if (s_slot[q] != NONE) dirty_next[s_slot[q]] = v;
      }
    }
    __syncthreads();
  }
}

// Rebuild the key directories of rewritten lists.  One wave per vertex; the
// vertices are list[0 .. *count) (DIR_LIST), the heads of the merge runs
// (DIR_RUNS) or every vertex below n (DIR_ALL).  A vertex without a slot gets
// one when its list has at least `threshold` entries and slots are left;
// otherwise it keeps using binary search.
constexpr int DIR_LIST = 0, DIR_RUNS = 1, DIR_ALL = 2;
template <int SRC>
__global__ void k_dir_build(const uint2* __restrict__ vrec, const uint32_t* __restrict__ adj,
                            uint32_t* __restrict__ dslot, uint32_t* __restrict__ dir,
                            uint32_t dstride, uint32_t vbits, uint32_t threshold,
                            const uint32_t* __restrict__ list, const uint32_t* __restrict__ count,
                            const uint64_t* __restrict__ keys, uint32_t n,
                            Ctrl* __restrict__ ctrl, uint32_t dir_cap) {
  const uint32_t lane = lane_id();
  const uint32_t wid = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
  const uint32_t nw = (gridDim.x * blockDim.x) / WAVE;
  const uint32_t total = SRC == DIR_LIST ? *count : SRC == DIR_RUNS ? ctrl->nruns : n;
  const uint32_t nk = dstride - 1;
  for (uint32_t it = wid; it < total; it += nw) {
    uint32_t v = SRC == DIR_LIST ? list[it]
               : SRC == DIR_RUNS ? (uint32_t)(keys[list[it]] >> 32) : it;
    const uint2 r = vrec[v];
    const uint32_t off = r.x, len = r.y & LEN_MASK;
    uint32_t s = uniform(dslot[v]);
    if (s == NONE) {
      if (len < threshold) continue;
      if (lane == 0) {
        s = atomicAdd(&ctrl->dir_top, 1u);
        if (s >= dir_cap) s = NONE;
        else dslot[v] = s;
      }
      s = uniform(__shfl(s, 0));
      if (s == NONE) continue;
    }
    uint32_t* __restrict__ D = dir + (size_t)s * dstride;
    if (len >= 16u * dstride) {
      // a long list (a hub): one binary search per key instead of a walk of
      // every entry; D[b] is the first entry whose key is at least b
      for (uint32_t b = lane; b <= nk; b += WAVE)
        D[b] = b == nk ? len : lower_bound(adj, off, off + len, b << vbits) - off;
      continue;
    }
    // entry i opens the keys (key(i-1), key(i)]
    for (uint32_t i = lane; i < len; i += WAVE) {
      const uint32_t kc = adj[off + i] >> vbits;
      const uint32_t kb = i ? (adj[off + i - 1] >> vbits) + 1u : 0u;
      for (uint32_t b = kb; b <= kc && b <= nk; ++b) D[b] = i;
    }
    const uint32_t kt = len ? (adj[off + len - 1] >> vbits) + 1u : 0u;
    for (uint32_t b = kt + lane; b <= nk; b += WAVE) D[b] = len;
  }
}

// Locate the live version of (a -> b) for the update at position pos: the
// entry for neighbour b (any edge label) with the largest ins < pos
// (mode 0), or any version visible just before pos (mode 1).  Returns the
// pool index or NONE, and the dense edge label in *el.
__device__ __forceinline__ uint32_t find_version(const GraphView& g, uint32_t a, uint32_t b,
                                                 uint32_t pos, int mode, uint32_t* el) {
  uint2 r = g.vrec[a];
  uint32_t off = r.x, end = r.x + (r.y & LEN_MASK);
  uint32_t lb_ = g.vlab[b];
  uint32_t best = NONE, best_ins = 0;
  for (uint32_t l = 0; l < g.nel; ++l) {
    uint32_t x = ((lb_ * g.nel + l) << g.vbits) | b;
    for (uint32_t i = lower_bound(g.adj, off, end, x); i < end && g.adj[i] == x; ++i) {
      uint32_t ts = g.ats[i];
      uint32_t ins = ts >> 16, del = ts & 0xFFFFu;
      if (mode == 0) {
        if (ins < pos && (best == NONE || ins >= best_ins)) { best = i; best_ins = ins; *el = l; }
      } else {
        if (ins < pos && del >= pos) { *el = l; return i; }
      }
    }
  }
  return best;
}

// Deletions: mark the live version of both directed entries deleted at the
// update's position and record the edge label for the seeds.
__global__ void k_delete_mark(BatchView b, GraphView g, uint32_t* __restrict__ ats,
                              uint2* __restrict__ vrec, uint32_t* __restrict__ seed_label,
                              Ctrl* __restrict__ ctrl, uint32_t* __restrict__ dirty_next) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < b.n; i += gridDim.x * blockDim.x) {
    if (b.type[i] != 'e') { seed_label[i] = NO_LABEL; continue; }
    if (b.add[i]) { seed_label[i] = b.label[i]; continue; }
    uint32_t pos = i + 1, x = b.v1[i], y = b.v2[i];
    uint32_t el0 = NO_LABEL, el1 = NO_LABEL;
    uint32_t i0 = NONE, i1 = NONE;
    if (x < g.nv && y < g.nv && g.vlab[x] != NO_LABEL && g.vlab[y] != NO_LABEL) {
      i0 = find_version(g, x, y, pos, 0, &el0);
      i1 = find_version(g, y, x, pos, 0, &el1);
    }
    if (i0 == NONE || i1 == NONE) {
      atomicOr(&ctrl->err, ERR_DELETE_MISSING);
      atomicMin(&ctrl->err_at, i);
      seed_label[i] = NO_LABEL;
      continue;
    }
    // the chosen versions must still be alive; a second deletion of the same
    // version is a deletion of an absent edge
    uint32_t t0 = ats[i0], t1 = ats[i1];
    uint32_t e0 = (t0 & 0xFFFF0000u) | 0xFFFFu, e1 = (t1 & 0xFFFF0000u) | 0xFFFFu;
    bool ok0 = atomicCAS(&ats[i0], e0, (t0 & 0xFFFF0000u) | pos) == e0;
    bool ok1 = atomicCAS(&ats[i1], e1, (t1 & 0xFFFF0000u) | pos) == e1;
    if (!ok0 || !ok1) {
      atomicOr(&ctrl->err, ERR_DELETE_MISSING);
      atomicMin(&ctrl->err_at, i);
      seed_label[i] = NO_LABEL;
      continue;
    }
    seed_label[i] = el0;
    uint32_t* y0 = reinterpret_cast<uint32_t*>(&vrec[x]) + 1;
    uint32_t* y1 = reinterpret_cast<uint32_t*>(&vrec[y]) + 1;
    if (!(atomicOr(y0, DIRTY) & DIRTY)) dirty_next[atomicAdd(&ctrl->ndirty_next, 1u)] = x;
    if (!(atomicOr(y1, DIRTY) & DIRTY)) dirty_next[atomicAdd(&ctrl->ndirty_next, 1u)] = y;
  }
}

// Insertions of an edge that is already present just before the insertion
// leave the graph unchanged (graph.cpp:63-64): hide the new version.  The
// update is still searched, as the reference's AddEdge does.
__global__ void k_dedup(BatchView b, GraphView g, uint32_t* __restrict__ ats) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < b.n; i += gridDim.x * blockDim.x) {
    if (b.type[i] != 'e' || !b.add[i] || b.label[i] == NO_LABEL) continue;
    uint32_t pos = i + 1;
    for (int dir = 0; dir < 2; ++dir) {
      uint32_t x = dir ? b.v2[i] : b.v1[i], y = dir ? b.v1[i] : b.v2[i];
      if (x >= g.nv || y >= g.nv || x == y) continue;
      uint32_t el;
      if (find_version(g, x, y, pos, 1, &el) == NONE) continue;
      // hide this update's own version(s) of (x -> y)
      uint2 r = g.vrec[x];
      uint32_t off = r.x, end = r.x + (r.y & LEN_MASK);
      uint32_t xv = ((g.vlab[y] * g.nel + b.label[i]) << g.vbits) | y;
      for (uint32_t j = lower_bound(g.adj, off, end, xv); j < end && g.adj[j] == xv; ++j)
        if ((g.ats[j] >> 16) == pos) ats[j] = TS_DEAD;
    }
  }
}

// ------------------------------------------------------------------ seeds
// per-update status bits of a batch (csm_update_status)
constexpr uint32_t USTAT_TIME_LIMIT = 1u;  // search stopped: counts are lower bounds
constexpr uint32_t USTAT_CAPPED = 2u;      // a count reached the per-update result cap
// passed the label filter but certified before its search to change no count:
// for every seed, some query neighbour of a seed endpoint has no candidate
constexpr uint32_t USTAT_CERTIFIED = 4u;

// Reserve one row per lane with `ok` set (one atomic per wave) and write
// each lane's embedding: vertex_of(i) is the data vertex of query vertex i.
template <class F>
__device__ __forceinline__ void emit_rows(const SearchLimits* lim, bool ok, uint32_t t,
                                          uint32_t q, uint32_t k, F vertex_of) {
  const uint32_t lane = lane_id();
  const uint64_t m = ballot(ok);
  if (m == 0) return;
  unsigned long long base = 0;
  if (lane == 0) base = atomicAdd(lim->emb_n, (unsigned long long)popc64(m));
  base = __shfl(base, 0);
  if (!ok) return;
  const unsigned long long r = base + popc64(m & ((1ull << lane) - 1ull));
  if (r >= lim->emb_cap) return;
  uint32_t* row = lim->emb + r * lim->emb_w;
  row[0] = t;
  row[1] = q;
  for (uint32_t i = 0; i < k; ++i) row[2 + i] = vertex_of(i);
}

struct SeedArgs {
  GraphView g;
  BatchView b;
  const uint32_t* __restrict__ seed_label;   // dense edge label per update
  const uint32_t* __restrict__ tab_off;      // [nlab*nlab*nel + 1]
  const uint32_t* __restrict__ tab_ent;      // (qo << 1) | swap
  const PlanRec* __restrict__ plans;
  const uint32_t* __restrict__ d2;           // SymBi d2 bitmap (nullable)
  uint32_t d2_words;
  uint32_t nlab;
  uint32_t nq;
  uint32_t versioned;
  uint32_t homo;
  uint32_t want_hash;
  Task* seeds;
  uint32_t seed_cap;
  uint32_t* chunk_owner;
  uint32_t chunk_cap;
  Ctrl* ctrl;
  unsigned long long* counts;
  unsigned long long* hashes;
  uint8_t* safe;
  uint32_t* ustatus;          // USTAT_CERTIFIED for updates no seed survives (nullable)
  const SearchLimits* lim;    // embedding rows of 2-vertex queries (enumerate != 0)
  uint32_t enumerate;
  // seeds of the search-time order rule: every seed that passes d2 becomes a
  // task (no neighbourhood certificate, and a task even when the static
  // depth-2 segment is empty), so that the search tree is the reference's
  uint32_t dyn;
};

// Emit the depth-2 seeds of the lanes with `have` set (one task each, cut into
// ch.x chunks) with one reservation per wave.  Chunk ids reserved past the
// capacity are flagged; ids below it whose task did not fit are marked
// ownerless (NONE), so every id the search can reach has been written.
__device__ __forceinline__ void emit_seed_wave(bool have, uint32_t t, uint32_t qo, uint32_t m0,
                                               uint32_t m1, uint32_t bj, uint32_t blo,
                                               uint32_t bhi, uint2 ch, Task* seeds,
                                               uint32_t seed_cap, uint32_t* chunk_owner,
                                               uint32_t chunk_cap, Ctrl* ctrl) {
  const uint32_t lane = lane_id();
  const uint64_t hm = ballot(have);
  if (!hm) return;
  const uint32_t nt = popc64(hm), ti = popc64(hm & ((1ull << lane) - 1ull));
  uint32_t incl = have ? ch.x : 0u;
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    uint32_t y = __shfl_up(incl, o);
    if ((int)lane >= o) incl += y;
  }
  const uint32_t nc = uniform(__shfl(incl, WAVE - 1));
  uint32_t tb = 0, cb = 0;
  if (lane == 0) {
    tb = atomicAdd(&ctrl->nseeds, nt);
    cb = atomicAdd(&ctrl->nchunks, nc);
    if ((uint64_t)tb + nt > seed_cap) atomicOr(&ctrl->err, ERR_SEED_OVERFLOW);
    if ((uint64_t)cb + nc > chunk_cap) atomicOr(&ctrl->err, ERR_CHUNK_OVERFLOW);
  }
  tb = uniform(__shfl(tb, 0));
  cb = uniform(__shfl(cb, 0));
  if (!have) return;
  const uint32_t sidx = tb + ti;
  const bool fits = (uint64_t)sidx < seed_cap;
  const uint32_t first = cb + incl - ch.x;
  if (fits) {
    Task* tk = seeds + sidx;
    tk->t = t; tk->qo = qo; tk->depth = 2; tk->jmin = bj; tk->lo = blo; tk->hi = bhi;
    tk->base = first; tk->csize = ch.y; tk->m[0] = m0; tk->m[1] = m1;
  }
  for (uint32_t c = 0, id = first; c < ch.x; ++c, ++id)
    if ((uint64_t)id < chunk_cap) chunk_owner[id] = fits ? sidx : NONE;
}

// One wavefront per update, one lane per seed template: fix m[0], m[1], find
// the depth-2 candidate segment (the smaller label-filtered segment of the
// two endpoints) and emit the seed as a task cut into chunks.  Updates with
// no template are "safe" (the label-triple filter of
// parallel_graphflow.cpp:1990-2016).
__global__ void k_seeds(SeedArgs a) {
  const uint32_t lane = lane_id();
  const uint32_t wid = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
  const uint32_t nw = (gridDim.x * blockDim.x) / WAVE;
  for (uint32_t i = wid; i < a.b.n; i += nw) {
    uint32_t l = a.seed_label[i];
    uint32_t x = a.b.v1[i], y = a.b.v2[i];
    bool edge = a.b.type[i] == 'e' && l != NO_LABEL && x < a.g.nv && y < a.g.nv;
    uint32_t s = 0, e = 0;
    if (edge) {
      uint32_t lx = a.g.vlab[x], ly = a.g.vlab[y];
      if (lx != NO_LABEL && ly != NO_LABEL && lx < a.nlab && ly < a.nlab) {
        uint32_t bucket = (lx * a.nlab + ly) * a.g.nel + l;
        s = a.tab_off[bucket];
        e = a.tab_off[bucket + 1];
      }
    }
    s = uniform(s);
    e = uniform(e);
    if (lane == 0 && a.safe) a.safe[i] = (s == e);
    bool survived = false;      // some seed of this update passed every check
    for (uint32_t j0 = s; j0 < e; j0 += WAVE) {
      const uint32_t j = j0 + lane;
      bool have = false;
      uint32_t qo = 0, m0 = 0, m1 = 0, blo = 0, bhi = 0, bj = 0;
      uint2 ch = make_uint2(0u, 0u);
      if (j < e) {
        uint32_t ent = a.tab_ent[j];
        qo = ent >> 1;
        const PlanRec* P = a.plans + qo;
        m0 = (ent & 1) ? y : x;
        m1 = (ent & 1) ? x : y;
        bool ok = !(P->symbi && a.d2 &&
                    !(d2_bit(a.d2, P->d2row[0], a.d2_words, m0) &&
                      d2_bit(a.d2, P->d2row[1], a.d2_words, m1)));
        // neighbourhood certificate: every other query neighbour of the two
        // seed endpoints needs a candidate among their neighbours with its
        // label and edge label; an empty segment (in the pool, so at every
        // position of the batch) rules the seed out
        for (uint32_t p = 2; ok && !a.dyn && p < P->k; ++p)
          for (uint32_t jj = 0; ok && jj < P->nb[p]; ++jj) {
            const uint32_t w = P->bw[p][jj];
            if ((w & 0xFFu) > 1u) continue;
            uint32_t ry;
            const uint2 sg = seg_of(a.g, (w & 0xFFu) == 0 ? m0 : m1, w >> 8, &ry);
            ok = sg.y > sg.x;
          }
        if (ok && P->k == 2) {
          survived = true;
          atomicAdd(&a.counts[(size_t)i * a.nq + P->q], 1ull);
          if (a.want_hash) {
            uint32_t mq[2];
            mq[P->qv[0]] = m0;
            mq[P->qv[1]] = m1;
            uint64_t h = mix64(mix64((uint64_t)mq[0] + 1) + (uint64_t)mq[1] + 1);
            atomicAdd(&a.hashes[(size_t)i * a.nq + P->q], (unsigned long long)h);
          }
          if (a.enumerate) {
            const unsigned long long r = atomicAdd(a.lim->emb_n, 1ull);
            if (r < a.lim->emb_cap) {
              uint32_t* row = a.lim->emb + r * a.lim->emb_w;
              row[0] = i;
              row[1] = P->q;
              row[2 + P->qv[0]] = m0;
              row[2 + P->qv[1]] = m1;
            }
          }
          ok = false;
        }
        if (ok) {
          // depth-2 segment: smallest label-filtered partner segment
          uint32_t best = NONE;
          for (uint32_t jj = 0; jj < P->nb[2]; ++jj) {
            uint32_t w = P->bw[2][jj];
            uint32_t vx = (w & 0xFFu) == 0 ? m0 : m1;
            uint32_t ry;
            uint2 sg = seg_of(a.g, vx, w >> 8, &ry);
            uint32_t sz = sg.y - sg.x;
            if (best == NONE || sz < best) { best = sz; blo = sg.x; bhi = sg.y; bj = jj; }
          }
          if (best != 0 && best != NONE) {
            have = true;
            ch = chunking(best, P->chunk);
          } else if (a.dyn) {
            have = true;
            ch = chunking(0u, P->chunk);
          }
        }
      }
      survived = survived || have;
      emit_seed_wave(have, i, qo, m0, m1, bj, blo, bhi, ch, a.seeds, a.seed_cap, a.chunk_owner,
                     a.chunk_cap, a.ctrl);
    }
    // passed the label filter, yet no seed can grow into an embedding
    if (a.ustatus && s < e && ballot(survived) == 0 && lane == 0)
      a.ustatus[i] |= USTAT_CERTIFIED;
  }
}

// ----------------------------------------------------------------- search
struct SearchArgs {
  GraphView g;
  const PlanRec* __restrict__ plans;
  const Task* __restrict__ tasks;
  const uint32_t* __restrict__ chunk_owner;   // NONE: reserved id without a task
  uint32_t chunk_cap;                         // ids beyond it were never written
  Task* spill;                // next round's tasks (nullable: never spill)
  uint32_t* spill_owner;      // next round's chunk owners
  uint32_t spill_cap, spill_chunk_cap;
  uint32_t xregions;          // this round's chunks sit in per-XCD regions (spilled rounds)
  Ctrl* ctrl;
  unsigned long long* counts;
  unsigned long long* hashes;
  const uint32_t* __restrict__ d2;
  uint32_t d2_words;
  uint32_t nq;
  uint32_t homo;
  uint32_t pos_offset;        // position of update index 0 minus 1 (0 in a batch)
  unsigned long long* qcount; // per-query running match count (nullable)
  uint32_t* qabort;           // per-query abort flags (nullable)
  unsigned long long budget;  // per-query match budget in this launch (0: none)
  // LIM_* bits: which of the rarely used features below are on; their
  // pointers and values sit in device memory (lim) so that the hot loop
  // keeps one word of them in registers
  uint32_t lflags;
  const SearchLimits* __restrict__ lim;
  uint32_t round_slot;        // this launch's slot in Ctrl::round_busy
  // small batches leave the GPU partly idle, so their rounds are kept short:
  // plan budgets of at least 256 units are divided by 2^budget_shift (not
  // below 64)
  uint32_t budget_shift;
};

// real-time counter ticks (100 MHz) between two polls of the host abort word
constexpr uint64_t ABORT_POLL_TICKS = 2000;

// Per-wave DFS state in LDS, sized for queries of up to MK vertices.
template <int MK>
struct WaveState {
  uint32_t m[MK];             // data vertex at each position
  uint32_t mq[MK];            // data vertex of each query vertex (hash)
  uint32_t cur[MK], end[MK];
  uint32_t bn[MK], bi[MK];
  uint32_t jmin[MK], pdirty[MK], udirty[MK];
  uint32_t plo[MK][MK - 1], phi[MK][MK - 1];
  uint32_t cand[MK][WAVE];
  // fused last position: per-parent segments of the current parent buffer
  uint32_t olo[WAVE], ohi[WAVE], pref[WAVE + 1];
  uint32_t src[WAVE];
  uint32_t ljf, ljp, lfsize, lC;
  // F: the last position's valid vertices with respect to all fixed
  // neighbours (ascending); lF = |F|, or NONE when F was not built
  uint32_t lF;
  uint32_t lCeff;             // C without the vertex at position k-3
  uint32_t fvalid;            // F still holds for the current prefix
  uint32_t F[FCAP];
  // data vertex whose segment plo/phi[p][j] currently hold (NONE: none); the
  // graph does not change during a launch, so a hit needs no lookup
  uint32_t pvx[MK][MK - 1];
  // the current plan's per-position fields (plan index pqo)
  uint32_t pqo;
  uint32_t pnb[MK], pd2[MK], pqv[MK];
  uint32_t pchunk, pbudget;   // the plan's split chunk and spill budget
  uint32_t pbw[MK][MK - 1];
  // rarely used per-wave state, kept out of registers: the time-limit word
  // as last read and when, and the candidate count at the task's start
  uint32_t aborted;
  uint32_t pad_;
  unsigned long long last_poll, cands0;
};

// One wavefront runs one depth-2 chunk of a seed at a time: an iterative DFS
// over positions 2..k-2 whose per-position candidate buffers (one wave-width
// each) live in LDS.  Each refill tests 64 consecutive entries of the u_min
// segment in parallel: version visibility, injectivity against m[], the SymBi
// d2 bit, and a binary search of every other backward neighbour's segment
// (the join of graphflow.cpp:222-238).
//
// The last position is never descended into.  Its backward neighbours split
// into "fixed" ones (positions < k-2, constant while position k-2 iterates)
// and at most one "parent" (position k-2).  Fixed segments are found once per
// entry of position k-2; then every refill of position k-2 (up to 64 parents)
// is finished in one go:
//   * no parent neighbour: the leaf set S is the same for all parents, so its
//     valid size C is counted once and each parent adds C - [parent in S];
//   * otherwise each lane finds its parent's own segment, the smaller of the
//     own and the fixed segment is the candidate source, and the wave walks
//     the concatenation of all parents' sources 64 entries at a time.
// Optional per-phase cycle accounting of the search (built with
// -DCSM_PHASE_PROF): wave cycles spent in each phase, summed over waves,
// and counters (printed as "counter i" with CSM_TRACE_ROUNDS set):
//   0 leaf entries walked            1 parents handed to the leaf loop
//   2 parents walking their own segment
//   3 / 5 leaf entries walked without / with F in LDS
//   4 / 7 wave cycles of leaf loops without / with F in LDS
//   6 sum of the parents' own segment lengths
//   8 backward neighbours and 9 candidates at refills of position k-2
//  10 entries of position k-2 without a parent neighbour
//  11 smallest fixed segment lengths at entries of position k-2
//  12 wave lifetime (100 MHz ticks)    13 waves
#ifdef CSM_PHASE_PROF
#define PROF_BEGIN(id) const uint64_t prof_t0_##id = __builtin_readcyclecounter()
#define PROF_END(id, i) \
  do { prof_acc[i] += __builtin_readcyclecounter() - prof_t0_##id; ++prof_n[i]; } while (0)
#else
#define PROF_BEGIN(id) (void)0
#define PROF_END(id, i) (void)0
#endif
#ifdef CSM_PHASE_PROF
#define PROF_ADD(i, x) (prof_x[i] += (x))
#else
#define PROF_ADD(i, x) (void)0
#endif

// the search's lambdas are always inlined: an outlined one is called with its
// captures spilled to scratch memory
#define CSM_AINL __attribute__((always_inline))

#ifndef CSM_XCD_SPLIT
#define CSM_XCD_SPLIT 1
#endif
#ifndef CSM_XCD_REGIONS
#define CSM_XCD_REGIONS 1
#endif
// spilled tasks stay on the XCD that spilled them (needs the XCD shares)
constexpr bool XCD_REGIONS = CSM_XCD_REGIONS && CSM_XCD_SPLIT;
#ifndef CSM_CHUNK_GRAB
#define CSM_CHUNK_GRAB 1
#endif
// six waves per SIMD for queries of up to 8 vertices: the kernel fits in 80
// VGPRs only when the compiler is held to it (left at five it takes 96)
#ifndef CSM_SEARCH_WAVES_PER_EU
#define CSM_SEARCH_WAVES_PER_EU 6
#endif
template <int MK, bool VERSIONED, bool HASH, bool SYMBI>
__global__ __launch_bounds__(search_waves<MK>() * WAVE)
__attribute__((amdgpu_waves_per_eu(MK <= 8 ? CSM_SEARCH_WAVES_PER_EU : MK <= 10 ? 5 : MK <= 12 ? 4 : MK <= 16 ? 3 : MK <= 24 ? 2 : 1, 8))) void k_search(SearchArgs a) {
  __shared__ WaveState<MK> S_all[search_waves<MK>()];
  const uint32_t lane = lane_id();
  const uint64_t lt_mask = (1ull << lane) - 1ull;
  WaveState<MK>& S = S_all[threadIdx.x / WAVE];
  if (lane_id() == 0) {
    S.pqo = NONE;
    S.fvalid = 0;
    S.aborted = 0;
    S.last_poll = 0;
  }
  for (uint32_t i = lane; i < MK * (MK - 1); i += WAVE) S.pvx[i / (MK - 1)][i % (MK - 1)] = NONE;
  wave_sync();
  const uint32_t* __restrict__ adj = a.g.adj;
  const uint32_t* __restrict__ ats = a.g.ats;
  const uint32_t vbits = a.g.vbits, vmask = a.g.vmask;
  // n_tasks and n_cands are the same in every lane; n_bytes (algorithmic
  // bytes of the segments a DFS node reads) is kept per lane and summed once
  // at the end
  unsigned long long n_tasks = 0, n_cands = 0, n_bytes = 0;
#ifdef CSM_PHASE_PROF
  unsigned long long prof_acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  unsigned long long prof_n[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  unsigned long long prof_x[16] = {0};
  const uint64_t prof_kt0 = __builtin_amdgcn_s_memrealtime();
#endif

  // Each XCD (workgroups are dispatched to the 8 XCDs round-robin) first
  // takes chunks from its own eighth of the round's chunk ids, so that the
  // consecutive chunks of a task run under one L2; a wave whose share is
  // used up takes from the next XCD's share.  A wave claims CSM_CHUNK_GRAB
  // consecutive ids per atomic (1 measures best once the shares give the L2
  // locality; fewer per claim when a round has few chunks per wave).
  uint32_t c_next = 0, c_end = 0;
  uint32_t xdone = 0;                       // XCD shares seen used up
  const uint32_t xcd = blockIdx.x % NXCD;
  // spilled rounds: XCD y's share is region y of the spill buffers, where
  // XCD y's waves wrote the tasks of the previous round
  const uint32_t xn_lane = (XCD_REGIONS && a.xregions && lane < NXCD) ? a.ctrl->xn[lane] : 0u;
  const uint32_t nwaves = gridDim.x * (uint32_t)search_waves<MK>();
  // the host's time-limit word, polled at most every ABORT_POLL_TICKS
  auto poll_abort = [&]() CSM_AINL -> bool {
    if (!(a.lflags & LIM_TIME)) return false;
    if (S.aborted) return true;
    const uint64_t now = __builtin_amdgcn_s_memrealtime();
    if (now - S.last_poll < ABORT_POLL_TICKS) return false;
    uint32_t v = 0;
    if (lane == 0) {
      v = __hip_atomic_load(a.lim->habort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      S.last_poll = now;
      S.aborted = v;
    }
    return __builtin_amdgcn_readlane(v, 0) != 0;
  };
  while (true) {
    PROF_BEGIN(fetch);
    uint32_t nch = uniform(__atomic_load_n(&a.ctrl->nchunks, __ATOMIC_RELAXED));
    nch = nch < a.chunk_cap ? nch : a.chunk_cap;
    if (c_next >= c_end) {
      uint32_t grab = CSM_CHUNK_GRAB;
      while (grab > 1 && (uint64_t)nch < 8ull * grab * nwaves) grab >>= 1;
#if CSM_XCD_SPLIT
      bool got = false;
      for (uint32_t xi = 0; xi < NXCD && !got; ++xi) {
        const uint32_t y = (xcd + xi) % NXCD;
        if ((xdone >> y) & 1u) continue;
        uint32_t lo, hi;
        if (XCD_REGIONS && a.xregions) {
          lo = y * (a.chunk_cap / NXCD);
          hi = lo + __builtin_amdgcn_readlane(xn_lane, y);
        } else {
          lo = (uint32_t)((uint64_t)nch * y / NXCD);
          hi = (uint32_t)((uint64_t)nch * (y + 1) / NXCD);
        }
        uint32_t g = 0;
        if (lane == 0 && lo < hi) g = atomicAdd(&a.ctrl->xhead[y * XHEAD_STRIDE], grab);
        g = __builtin_amdgcn_readlane(g, 0);
        if (lo < hi && g < hi - lo) {
          c_next = lo + g;
          c_end = hi - c_next < grab ? hi : c_next + grab;
          got = true;
        } else {
          xdone |= 1u << y;
        }
      }
      if (!got) break;
#else
      uint32_t c0 = 0;
      if (lane == 0) c0 = atomicAdd(&a.ctrl->chunk_head, grab);
      c_next = __builtin_amdgcn_readlane(c0, 0);
      c_end = c_next + grab;
#endif
    }
    const uint32_t c = c_next++;
#if !CSM_XCD_SPLIT
    if (c >= nch) break;
#endif
    if (__atomic_load_n(&a.ctrl->abort, __ATOMIC_RELAXED)) break;
    const uint32_t sidx = uniform(a.chunk_owner[c]);
    if (sidx == NONE) continue;
    const Task* __restrict__ TK = a.tasks + sidx;
    const uint32_t t = uniform(TK->t), qo = uniform(TK->qo);
    const uint32_t d0 = uniform(TK->depth);
    const PlanRec* __restrict__ P = a.plans + qo;
    const uint32_t k = uniform(P->k);
    const uint32_t pos = t + 1 + a.pos_offset;
    const uint32_t q = uniform(P->q);
    if (a.qabort && __atomic_load_n(&a.qabort[q], __ATOMIC_RELAXED)) continue;
    if (poll_abort()) {
      // time limit: this chunk is not searched
      if (lane == 0) {
        atomicOr(&a.lim->ustatus[t], USTAT_TIME_LIMIT);
        atomicOr(&a.lim->qtl[q], 1u);
      }
      continue;
    }
    if (a.lflags & LIM_CAP) {
      // the count of (t, q) has reached the cap: nothing left to find
      uint32_t full = 0;
      if (lane == 0)
        full = __hip_atomic_load(&a.counts[(size_t)t * a.nq + q], __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT) >= a.lim->ucap;
      if (__builtin_amdgcn_readlane(full, 0)) continue;
    }
    ++n_tasks;
    if ((a.lflags & LIM_WORK) && lane == 0) S.cands0 = n_cands;
    if (qo != S.pqo) {
      if (lane < MK) {
        S.pnb[lane] = P->nb[lane];
        if (lane == 0) {
          S.pchunk = P->chunk;
          const uint32_t b = P->budget;
          S.pbudget = b >= 256u ? max(64u, b >> a.budget_shift) : b;
        }
        S.pd2[lane] = P->d2row[lane];
        S.pqv[lane] = P->qv[lane];
      }
      for (uint32_t i = lane; i < MK * (MK - 1); i += WAVE) {
        S.pbw[i / (MK - 1)][i % (MK - 1)] = P->bw[i / (MK - 1)][i % (MK - 1)];
        S.pvx[i / (MK - 1)][i % (MK - 1)] = NONE;
      }
      if (lane == 0) S.pqo = qo;
      wave_sync();
    }
    if (lane < d0) {
      const uint32_t v = TK->m[lane];
      S.m[lane] = v;
      if (HASH) S.mq[S.pqv[lane]] = v;
    }
    if (lane == 0) S.fvalid = 0;
    wave_sync();
    unsigned long long cnt = 0, hsum = 0, flushed = 0;
    uint32_t ticks = 0;
    uint32_t work = 0;          // budget units used by this task (one per 64 candidates)
    bool may_spill = false;
    bool stop = false;
    bool lost = false;          // stopped by the time limit: work was dropped
    const uint32_t L1 = k - 1, L2 = k - 2;   // last and second-to-last positions

    // publish the running count; stop if the query's budget is exhausted or
    // the update's count has reached the result cap
    auto flush = [&](bool final_) CSM_AINL {
      unsigned long long d = cnt - flushed;
      if (lane == 0 && d) {
        if (!(a.lflags & LIM_CAP)) atomicAdd(&a.counts[(size_t)t * a.nq + q], d);
        if (a.qcount) {
          unsigned long long tot = atomicAdd(&a.qcount[q], d) + d;
          if (a.budget && tot > a.budget) atomicOr(&a.qabort[q], 1u);
        }
      }
      if (a.lflags & LIM_CAP) {
        // add the count and see the update's running total (the cap stops
        // the task once it is reached)
        uint32_t capped = 0;
        if (lane == 0) {
          unsigned long long* c = &a.counts[(size_t)t * a.nq + q];
          const unsigned long long tot_u =
              d ? atomicAdd(c, d) + d
                : __hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          capped = tot_u >= a.lim->ucap;
        }
        if (!final_ && __builtin_amdgcn_readlane(capped, 0)) stop = true;
      }
      flushed = cnt;
      if (!final_ && a.qabort)
        stop = uniform(__atomic_load_n(&a.qabort[q], __ATOMIC_RELAXED)) != 0;
    };

    // test a candidate entry e of position lv's u_min segment (joins with the
    // other backward neighbours of lv)
    auto test = [&](uint32_t lv, uint32_t e, uint32_t end, uint32_t& vout) CSM_AINL -> bool {
      bool ok = e < end;
      uint32_t v = 0;
      // the plan fields and cached segment bounds read below are the same
      // for every lane: held in scalar registers
      if (ok) {
        v = adj[e] & vmask;
        if (VERSIONED && uniform(S.udirty[lv])) ok = ts_visible(ats[e], pos);
      }
      if (ok && !a.homo) {
        for (uint32_t i = 0; i < lv; ++i) ok = ok && (S.m[i] != v);
      }
      if (SYMBI && ok) ok = d2_bit(a.d2, uniform(S.pd2[lv]), a.d2_words, v);
      const uint32_t nb = uniform(S.pnb[lv]);
      const uint32_t jm = uniform(S.jmin[lv]);
      const uint32_t pdm = VERSIONED ? uniform(S.pdirty[lv]) : 0u;
      for (uint32_t j = 0; j < nb; ++j) {
        if (j == jm) continue;
        const uint32_t w = uniform(S.pbw[lv][j]);
        const uint32_t slo = uniform(S.plo[lv][j]), shi = uniform(S.phi[lv][j]);
        if (!ok) continue;
        const uint32_t x = ((w >> 8) << vbits) | v;
        ok = contains_u(adj, ats, slo, shi - slo, x, (pdm >> j) & 1u, pos);
      }
      vout = v;
      return ok;
    };

    // is w a valid last-position vertex w.r.t. the fixed neighbours (all but
    // `skip`), injectivity against m[0..ninj-1] and the d2 bit
    auto last_fixed_ok = [&](uint32_t w, uint32_t skip, uint32_t ninj) CSM_AINL -> bool {
      bool ok = true;
      if (!a.homo)
        for (uint32_t i = 0; i < ninj; ++i) ok = ok && (S.m[i] != w);
      if (SYMBI && ok) ok = d2_bit(a.d2, S.pd2[L1], a.d2_words, w);
      const uint32_t nb = uniform(S.pnb[L1]);
      for (uint32_t j = 0; j < nb; ++j) {
        if (j == skip || j == S.ljp || !ok) continue;
        uint32_t bw = S.pbw[L1][j];
        uint32_t x = ((bw >> 8) << vbits) | w;
        bool vd = VERSIONED && ((S.pdirty[L1] >> j) & 1u);
        const uint32_t slo = uniform(S.plo[L1][j]);
        ok = contains_u(adj, ats, slo, uniform(S.phi[L1][j]) - slo, x, vd, pos);
      }
      return ok;
    };

    // fixed neighbours of the last position, once per entry of position k-2;
    // false if no leaf can complete (an empty fixed segment or an empty F).
    // F is built without the injectivity test against position k-3, which the
    // leaves make themselves, so it is reused across the entries of position
    // k-2 that share the fixed vertices and the prefix m[0..k-4].
    auto prepare_last = [&]() CSM_AINL -> bool {
      PROF_BEGIN(pl);
      const uint32_t nb = uniform(S.pnb[L1]);
      uint32_t my_lo = 0, my_hi = 0, my_dirty = 0, my_size = 0xFFFFFFFFu, is_par = 0, miss = 0;
      if (lane < nb) {
        uint32_t w = S.pbw[L1][lane];
        uint32_t bp = w & 0xFFu;
        if (bp == L2) {
          is_par = 1;
          // keep the parent slot's cached dirty bit (the slot belongs to
          // tasks that start at the last position)
          my_dirty = VERSIONED ? ((S.pdirty[L1] >> lane) & 1u) : 0u;
        } else {
          const uint32_t vx = S.m[bp];
          if (vx == S.pvx[L1][lane]) {
            my_lo = S.plo[L1][lane];
            my_hi = S.phi[L1][lane];
            my_dirty = VERSIONED ? ((S.pdirty[L1] >> lane) & 1u) : 0u;
          } else {
            miss = 1;
            uint32_t ry;
            uint2 sg = seg_of(a.g, vx, w >> 8, &ry);
            my_lo = sg.x;
            my_hi = sg.y;
            my_dirty = VERSIONED ? (ry >> 31) : 0u;
            S.plo[L1][lane] = my_lo;
            S.phi[L1][lane] = my_hi;
            S.pvx[L1][lane] = vx;
          }
          my_size = my_hi - my_lo;
        }
      }
      const bool reuse = S.fvalid != 0 && ballot(miss != 0) == 0;
      uint64_t pm = ballot(is_par != 0);
      uint32_t jp = pm ? (uint32_t)__builtin_ctzll(pm) : NONE;
      uint64_t mn = lanes_min_u32(my_size, nb);
      uint32_t jf = (uint32_t)(mn >> 32) == 0xFFFFFFFFu ? NONE : uniform((uint32_t)mn);
      uint32_t fsize = uniform((uint32_t)(mn >> 32));
      const uint32_t dmask = VERSIONED ? (uint32_t)ballot(my_dirty != 0) : 0u;
      n_bytes += 4ull * (my_size == 0xFFFFFFFFu ? 0u : my_size);
      wave_sync();
      if (lane == 0) {
        S.ljp = jp;
        S.ljf = jf;
        S.lfsize = fsize;
        S.pdirty[L1] = dmask;
        if (!reuse) S.fvalid = 0;
      }
      wave_sync();
      if (jf != NONE && fsize == 0) return false;
      PROF_ADD(10, jp == NONE ? 1u : 0u);
      PROF_ADD(11, jf == NONE ? 0u : fsize);
      const uint32_t mk3 = S.m[L2 - 1];
      // C without the vertex at position k-3 (one of the C leaves at most
      // once): looked up in F, or checked in the fixed segments when F is
      // not in LDS
      auto c_eff = [&](uint32_t C, bool kept) CSM_AINL -> uint32_t {
        if (a.homo || C == 0) return C;
        if (kept) return lds_contains(S.F, C, mk3) ? C - 1u : C;
        const bool fd = VERSIONED && ((S.pdirty[L1] >> jf) & 1u);
        const uint32_t x = ((S.pbw[L1][jf] >> 8) << vbits) | mk3;
        const uint32_t slo = uniform(S.plo[L1][jf]);
        const bool in = contains_u(adj, ats, slo, uniform(S.phi[L1][jf]) - slo, x, fd,
                                   pos) &&
                        last_fixed_ok(mk3, jf, L2 - 1);
        return in ? C - 1u : C;
      };
      if (reuse) {
        const uint32_t Ce = c_eff(S.lC, S.lF != NONE);
        if (lane == 0) S.lCeff = Ce;
        wave_sync();
        PROF_END(pl, 2);
        return Ce != 0;
      }
      if (lane == 0) S.lF = NONE;
      wave_sync();
      if (jf != NONE && (fsize <= FCAP || (jp == NONE && !HASH))) {
        // walk the smallest fixed segment once: F (when it fits in LDS) and
        // C, the number of valid leaves shared by every parent, both without
        // the injectivity test against position k-3, so that they hold for
        // every entry of position k-2 under the same fixed vertices
        const uint32_t lo = __builtin_amdgcn_readlane(my_lo, jf), hi = __builtin_amdgcn_readlane(my_hi, jf);
        const bool fd = VERSIONED && ((dmask >> jf) & 1u);
        const bool keep = fsize <= FCAP;
        const uint32_t ninj = L2 - 1;
        uint32_t C = 0;
        for (uint32_t b = lo; b < hi; b += WAVE) {
          uint32_t e = b + lane;
          bool ok = e < hi;
          uint32_t w = 0;
          if (ok) {
            w = adj[e] & vmask;
            if (fd) ok = ts_visible(ats[e], pos);
          }
          ok = ok && last_fixed_ok(w, jf, ninj);
          const uint64_t m = ballot(ok);
          if (keep && ok) S.F[C + popc64(m & lt_mask)] = w;
          C += popc64(m);
          n_cands += (hi - b) < WAVE ? (hi - b) : WAVE;
        }
        work += (hi - lo) / WAVE;
        wave_sync();
        const uint32_t Ce = c_eff(C, keep);
        if (lane == 0) {
          S.lC = C;
          S.lCeff = Ce;
          if (keep) S.lF = C;
          S.fvalid = 1;
        }
        wave_sync();
        if (Ce == 0) return false;   // no leaf can complete under this prefix
      }
      PROF_END(pl, 2);
      return true;
    };

    // enter position lv: partner segments, u_min, candidate range
    auto enter = [&](uint32_t lv, uint32_t rlo, uint32_t rhi, int forced_j) CSM_AINL -> bool {
      const uint32_t nb = uniform(S.pnb[lv]);
      uint32_t my_lo = 0, my_hi = 0, my_dirty = 0, my_size = 0xFFFFFFFFu;
      if (lane < nb) {
        uint32_t w = S.pbw[lv][lane];
        uint32_t vx = S.m[w & 0xFFu];
        if (vx == S.pvx[lv][lane]) {
          my_lo = S.plo[lv][lane];
          my_hi = S.phi[lv][lane];
          my_dirty = VERSIONED ? ((S.pdirty[lv] >> lane) & 1u) : 0u;
        } else {
          uint32_t ry;
          uint2 sg = seg_of(a.g, vx, w >> 8, &ry);
          my_lo = sg.x;
          my_hi = sg.y;
          my_dirty = VERSIONED ? (ry >> 31) : 0u;
          S.plo[lv][lane] = my_lo;
          S.phi[lv][lane] = my_hi;
          S.pvx[lv][lane] = vx;
        }
        my_size = my_hi - my_lo;
      }
      uint64_t mn = (forced_j >= 0) ? (((uint64_t)__builtin_amdgcn_readlane(my_size, forced_j) << 32) | (uint32_t)forced_j)
                                    : lanes_min_u32(my_size, nb);
      uint32_t jm = uniform((uint32_t)mn);
      uint32_t msz = uniform((uint32_t)(mn >> 32));
      uint32_t dmask = VERSIONED ? (uint32_t)ballot(my_dirty != 0) : 0u;
      n_bytes += 4ull * (lane < nb ? my_size : 0u);
      // the cached slots' dirty bits and F's validity, before any early return
      S.pdirty[lv] = dmask;
      if (lv < L2) S.fvalid = 0;   // the prefix that F excludes has changed
      wave_sync();
      if (msz == 0) return false;
      uint32_t lo = __builtin_amdgcn_readlane(my_lo, jm), hi = __builtin_amdgcn_readlane(my_hi, jm);
      lo = lo > rlo ? lo : rlo;
      hi = hi < rhi ? hi : rhi;
      S.cur[lv] = lo;
      S.end[lv] = hi;
      S.bn[lv] = 0;
      S.bi[lv] = 0;
      S.jmin[lv] = jm;
      S.udirty[lv] = (dmask >> jm) & 1u;
      if (lo >= hi) return false;
      if (lv == L2) return prepare_last();
      return true;
    };

    // reserve n_t spill tasks and n_c chunk ids with one atomic each; when
    // they do not fit, the chunk ids below the cap are marked ownerless and
    // the caller keeps the work
    auto reserve_spill = [&](uint32_t n_t, uint32_t n_c, uint32_t& tb, uint32_t& cb) CSM_AINL -> bool {
      tb = 0;
      cb = 0;
      if (XCD_REGIONS) {
        // this XCD's region: an eighth of each spill buffer
        const uint32_t tcap = a.spill_cap / NXCD, ccap = a.spill_chunk_cap / NXCD;
        if (lane == 0) {
          tb = atomicAdd(&a.ctrl->xspill[xcd * XHEAD_STRIDE], n_t);
          cb = atomicAdd(&a.ctrl->xspill_chunks[xcd * XHEAD_STRIDE], n_c);
          atomicAdd(&a.ctrl->nspill, n_t);
        }
        tb = __builtin_amdgcn_readlane(tb, 0);
        cb = __builtin_amdgcn_readlane(cb, 0);
        const uint32_t tb0 = xcd * tcap, cb0 = xcd * ccap;
        if ((uint64_t)tb + n_t <= tcap && (uint64_t)cb + n_c <= ccap) {
          tb += tb0;
          cb += cb0;
          return true;
        }
        for (uint64_t id = (uint64_t)cb + lane; id < (uint64_t)cb + n_c && id < ccap; id += WAVE)
          a.spill_owner[cb0 + id] = NONE;
        return false;
      }
      if (lane == 0) {
        tb = atomicAdd(&a.ctrl->nspill, n_t);
        cb = atomicAdd(&a.ctrl->nspill_chunks, n_c);
      }
      tb = __builtin_amdgcn_readlane(tb, 0);
      cb = __builtin_amdgcn_readlane(cb, 0);
      if ((uint64_t)tb + n_t <= a.spill_cap && (uint64_t)cb + n_c <= a.spill_chunk_cap) return true;
      for (uint64_t id = (uint64_t)cb + lane; id < (uint64_t)cb + n_c && id < a.spill_chunk_cap; id += WAVE)
        a.spill_owner[id] = NONE;
      return false;
    };

    // write spill task ti: positions 0..depth-2 from m[], position depth-1 =
    // extra, candidate range [lo, hi) of backward neighbour jm in chunks from ci
    auto emit_task = [&](uint32_t& ti, uint32_t& ci, uint32_t depth, uint32_t jm, uint32_t lo,
                         uint32_t hi, uint2 ch, uint32_t extra) CSM_AINL {
      const uint32_t nch = ch.x;
      Task* o = a.spill + ti;
      uint32_t word = 0;
      if (lane < TASK_WORDS) {
        if (lane == 0) word = t;
        else if (lane == 1) word = qo;
        else if (lane == 2) word = depth;
        else if (lane == 3) word = jm;
        else if (lane == 4) word = lo;
        else if (lane == 5) word = hi;
        else if (lane == 6) word = ci;
        else if (lane == 7) word = ch.y;
        else {
          uint32_t p = lane - 8;
          word = p + 1 < depth ? S.m[p] : (p + 1 == depth ? extra : 0u);
        }
        reinterpret_cast<uint32_t*>(o)[lane] = word;
      }
      for (uint32_t j = lane; j < nch; j += WAVE) a.spill_owner[ci + j] = ti;
      ++ti;
      ci += nch;
    };

    // the leaves of parent v (candidate range [lo, hi) of the last position's
    // backward neighbour jm) as a task of the next round
    auto spill_leaves = [&](uint32_t v, uint32_t jm, uint32_t lo, uint32_t hi) CSM_AINL -> bool {
      const uint2 ch = chunking(hi - lo, S.pchunk);
      uint32_t tb, cb;
      if (!reserve_spill(1u, ch.x, tb, cb)) return false;
      emit_task(tb, cb, L1, jm, lo, hi, ch, v);
      return true;
    };

    // all leaves under the bn parents now in cand[k-2]
    auto finish_parents = [&](uint32_t bn) CSM_AINL {
      PROF_BEGIN(fp);
      const uint32_t jp = uniform(S.ljp), jf = uniform(S.ljf);
      const uint32_t nF = uniform(S.lF);   // F in LDS (NONE: not built)
      const bool useF = nF != NONE;
      const uint32_t mk3 = S.m[L2 - 1];
      uint32_t v = lane < bn ? S.cand[L2][lane] : 0u;
      if (jp == NONE && !HASH) {
        // C - [parent is itself a valid leaf]
        bool self = false;
        if (lane < bn && !a.homo) {
          if (useF) {
            self = lds_contains(S.F, nF, v);
          } else {
            const bool fd = VERSIONED && ((S.pdirty[L1] >> jf) & 1u);
            uint32_t bw = S.pbw[L1][jf];
            uint32_t x = ((bw >> 8) << vbits) | v;
            const uint32_t slo = uniform(S.plo[L1][jf]);
            self = contains_u(adj, ats, slo, uniform(S.phi[L1][jf]) - slo, x, fd, pos) &&
                   last_fixed_ok(v, jf, L2);
          }
        }
        cnt += (unsigned long long)bn * S.lCeff - popc64(ballot(self));
        if (lane == 0) n_bytes += 4ull * bn * S.lfsize;
        PROF_END(fp, 5);
        return;
      }
      // per-parent source: own segment S(v, key_jp), or F / the smallest
      // fixed segment
      const uint32_t fsz = useF ? nF : S.lfsize;
      uint32_t olo = 0, ohi = 0, odirty = 0, size = 0, src = 0;
      if (lane < bn) {
        if (jp != NONE) {
          uint32_t ry;
          uint2 sg = seg_of(a.g, v, S.pbw[L1][jp] >> 8, &ry);
          olo = sg.x;
          ohi = sg.y;
          odirty = VERSIONED ? (ry >> 31) : 0u;
          uint32_t osz = sg.y - sg.x;
          if (osz == 0) size = 0;
          else if (jf != NONE && fsz < osz) { size = fsz; src = 0; }
          else { size = osz; src = 1; }
          n_bytes += 4ull * osz;
        } else {
          size = fsz;
          src = 0;
        }
        S.olo[lane] = olo;
        S.ohi[lane] = ohi;
        S.src[lane] = src | (odirty << 1);
      }
      // parents with very many leaves go to the next round as tasks of their
      // own, so that no chunk runs unboundedly long
      if (may_spill) {
        uint64_t big = ballot(lane < bn && size > LEAF_SPLIT);
        while (big) {
          const uint32_t p = (uint32_t)__builtin_ctzll(big);
          big &= big - 1;
          const bool own = (__builtin_amdgcn_readlane(src, p) & 1u) != 0;
          const uint32_t lo = own ? __builtin_amdgcn_readlane(olo, p) : S.plo[L1][jf];
          const uint32_t hi = own ? __builtin_amdgcn_readlane(ohi, p) : S.phi[L1][jf];
          if (!spill_leaves(S.cand[L2][p], own ? jp : jf, lo, hi)) {
            may_spill = false;
            break;
          }
          if (lane == p) size = 0;
        }
      }
      // exclusive prefix of sizes over the parents
      uint32_t incl = size;
#pragma unroll
      for (int o = 1; o < WAVE; o <<= 1) {
        uint32_t y = __shfl_up(incl, o);
        if ((int)lane >= o) incl += y;
      }
      const uint32_t total = __builtin_amdgcn_readlane(incl, WAVE - 1);
      work += total / WAVE;
      PROF_ADD(0, total);
      PROF_ADD(1, bn);
      PROF_ADD(2, popc64(ballot(lane < bn && src == 1)));
      PROF_ADD(useF ? 5 : 3, total);
      PROF_ADD(6, uniform(popc64(ballot(lane < bn)) ? (uint32_t)wave_sum((unsigned long long)(lane < bn ? (ohi - olo) : 0u)) : 0u));
      S.pref[lane] = incl - size;
      if (lane == 0) S.pref[WAVE] = total;
      // parent of each window entry: the parents whose leaves start inside
      // the window mark their start in pmap (tagged with the window number,
      // so a slot needs no clearing between windows), and each entry takes
      // the last mark at or
      // before it, or the parent that ran on from the previous window.  The
      // last position's candidate buffer is free: it is never descended into.
      uint32_t* pmap = S.cand[L1];
      pmap[lane] = NONE;
      wave_sync();
      const uint32_t fsum_lo = jf != NONE ? S.plo[L1][jf] : 0u;
      const uint32_t pstart = incl - size;
      const uint64_t upto_mask = lane == WAVE - 1 ? ~0ull : (2ull << lane) - 1ull;
      uint32_t pcarry = 0;
      for (uint32_t b = 0, wid = 0; b < total; b += WAVE, ++wid) {
        const uint32_t g = b + lane;
        bool ok = g < total, need = false;
        uint32_t w = 0, p;
        // a mark is 0x800000 | (window mod 2^20) above the parent index
        // (never NONE >> 8); the slots are cleared every 2^20 windows
        const uint32_t tag = 0x800000u | (wid & 0xFFFFFu);
        if ((wid & 0xFFFFFu) == 0 && wid) pmap[lane] = NONE;
        if (lane < bn && size > 0 && pstart >= b && pstart - b < WAVE)
          pmap[pstart - b] = (tag << 8) | lane;
        wave_sync();
        {
          const uint64_t marks = ballot((pmap[lane] >> 8) == tag) & upto_mask;
          p = marks ? (pmap[63 - __builtin_clzll(marks)] & 0xFFu) : pcarry;
          pcarry = __builtin_amdgcn_readlane(p, WAVE - 1);
        }
        if (ok) {
          const uint32_t sp = S.src[p];
          const uint32_t off = g - S.pref[p];
          if ((sp & 1u) || !useF) {
            uint32_t e;
            bool dirty;
            if (sp & 1u) { e = S.olo[p] + off; dirty = (sp >> 1) & 1u; }
            else { e = fsum_lo + off; dirty = VERSIONED && ((S.pdirty[L1] >> jf) & 1u); }
            w = adj[e] & vmask;
            if (VERSIONED && dirty) ok = ts_visible(ats[e], pos);
          } else {
            w = S.F[off];   // already valid for the fixed neighbours
          }
          const uint32_t vp = S.cand[L2][p];
          if (ok && !a.homo) ok = w != vp;
          if (!useF) {
            ok = ok && last_fixed_ok(w, (sp & 1u) ? NONE : jf, L2);
          } else {
            // F leaves the test against position k-3 to the leaves
            if (ok && !a.homo) ok = w != mk3;
            if (ok && (sp & 1u)) ok = lds_contains(S.F, nF, w);
          }
          need = ok && jp != NONE && !(sp & 1u);
        }
        // leaves drawn from F or a fixed segment: join with the parent's own
        // segment; a window whose joins all fall under one parent searches
        // that segment with wave-uniform halvings
        const uint64_t nm = ballot(need);
        if (nm) {
          const uint32_t x = ((S.pbw[L1][jp] >> 8) << vbits) | w;
          const uint32_t p0 = __builtin_amdgcn_readlane(p, (int)__builtin_ctzll(nm));
          if (ballot(need && p != p0) == 0) {
            const uint32_t lo0 = uniform(S.olo[p0]), hi0 = uniform(S.ohi[p0]);
            const bool d0 = VERSIONED && ((uniform(S.src[p0]) >> 1) & 1u);
            if (need) ok = contains_u(adj, ats, lo0, hi0 - lo0, x, d0, pos);
          } else if (need) {
            ok = contains(adj, ats, S.olo[p], S.ohi[p], x, VERSIONED && ((S.src[p] >> 1) & 1u), pos);
          }
        }
        cnt += popc64(ballot(ok));
        n_cands += (total - b) < WAVE ? (total - b) : WAVE;
        if (HASH) {
          uint64_t h = 0;
          if (ok) {
            const uint32_t vp = S.cand[L2][p];
            for (uint32_t i = 0; i < k; ++i) {
              uint32_t mv = (i == S.pqv[L1]) ? w : (i == S.pqv[L2]) ? vp : S.mq[i];
              h = mix64(h + (uint64_t)mv + 1);
            }
          }
          hsum += wave_sum(h);
          if (a.lflags & LIM_ENUM)
            emit_rows(a.lim, ok, t, q, k, [&](uint32_t i) {
              return (i == S.pqv[L1]) ? w : (i == S.pqv[L2]) ? S.cand[L2][p] : S.mq[i];
            });
        }
      }
      if (jp == NONE && lane == 0) n_bytes += 4ull * bn * S.lfsize;
#ifdef CSM_PHASE_PROF
      PROF_ADD(useF ? 7 : 4, __builtin_readcyclecounter() - prof_t0_fp);
#endif
      PROF_END(fp, 6);
    };

    // spill the rest of this task's DFS as new tasks (one per remaining range
    // and one per buffered candidate); false if the spill buffers are full
    auto spill = [&](uint32_t lv_now) CSM_AINL -> bool {
      uint32_t n_t = 0, n_c = 0;
      for (uint32_t L = d0; L <= lv_now; ++L) {
        uint32_t cu = S.cur[L], en = S.end[L];
        if (cu < en) { ++n_t; n_c += chunking(en - cu, S.pchunk).x; }
        if (L < L2) { uint32_t r = S.bn[L] - S.bi[L]; n_t += r; n_c += r; }
      }
      n_t = uniform(n_t);
      n_c = uniform(n_c);
      if (n_t == 0) return true;
      uint32_t tb, cb;
      if (!reserve_spill(n_t, n_c, tb, cb)) return false;
      uint32_t ti = tb, ci = cb;
      for (uint32_t L = d0; L <= lv_now; ++L) {
        uint32_t cu = S.cur[L], en = S.end[L];
        if (cu < en)
          emit_task(ti, ci, L, S.jmin[L], cu, en, chunking(en - cu, S.pchunk), L > 0 ? S.m[L - 1] : 0u);
        if (L < L2)
          for (uint32_t i = S.bi[L]; i < S.bn[L]; ++i)
            emit_task(ti, ci, L + 1, NONE, 0u, NONE, make_uint2(1u, NONE), S.cand[L][i]);
      }
      return true;
    };

    const uint32_t off = c - uniform(TK->base);
    const uint32_t tlo = uniform(TK->lo), thi = uniform(TK->hi);
    uint32_t rlo = 0, rhi = NONE;
    if (thi != NONE) {
      const uint32_t cs = uniform(TK->csize);
      rlo = tlo + off * cs;
      rhi = rlo + cs;
      rhi = rhi < thi ? rhi : thi;
    }
    const uint32_t tj = uniform(TK->jmin);
    int lv = (int)d0;
    work = 0;
    may_spill = a.spill != nullptr;
    PROF_END(fetch, 0);
    PROF_BEGIN(e0);
    bool live = enter(d0, rlo, rhi, tj == NONE ? -1 : (int)tj);
    PROF_END(e0, 1);
    // a task takes one DFS step before it may spill: entering position k-2
    // can charge more than the whole budget (the walk of the fixed segment),
    // and a task spilled before any step would be handed on unchanged
    bool moved = false;
    while (live && !stop) {
      if (may_spill && ++work > S.pbudget && (uint32_t)lv < L1 && moved) {
        PROF_BEGIN(sp);
        bool ok_sp = spill((uint32_t)lv);
        PROF_END(sp, 7);
        if (ok_sp) break;
        may_spill = false;
      }
      moved = true;
      const uint32_t L = (uint32_t)lv;
      if (L == L1) {
        // k == 3: the task's range is the last position itself
        uint32_t cur = S.cur[L], end = S.end[L];
        for (; cur < end; cur += WAVE) {
          uint32_t v;
          bool ok = test(L, cur + lane, end, v);
          cnt += popc64(ballot(ok));
          n_cands += (end - cur) < WAVE ? (end - cur) : WAVE;
          if (HASH) {
            uint64_t h = 0;
            if (ok)
              for (uint32_t i = 0; i < k; ++i) {
                uint32_t mv = (i == S.pqv[L]) ? v : S.mq[i];
                h = mix64(h + (uint64_t)mv + 1);
              }
            hsum += wave_sum(h);
            if (a.lflags & LIM_ENUM)
              emit_rows(a.lim, ok, t, q, k,
                        [&](uint32_t i) { return (i == S.pqv[L]) ? v : S.mq[i]; });
          }
        }
        break;
      }
      if (++ticks == 256) {
        ticks = 0;
        if (cnt - flushed > (1ull << 20) || a.qabort || (a.lflags & LIM_CAP)) flush(false);
      }
      // every iteration, as one can walk hundreds of leaf windows (the poll
      // reads the clock, and the host word only every ABORT_POLL_TICKS)
      if (poll_abort()) {
        stop = true;
        lost = true;
      }
      uint32_t cur = S.cur[L], end = S.end[L];
      if (L == L2) {
        if (cur < end) {
          PROF_BEGIN(t2);
          PROF_ADD(8, uniform(S.pnb[L]));
          PROF_ADD(9, (end - cur) < WAVE ? (end - cur) : WAVE);
          uint32_t v;
          bool ok = test(L, cur + lane, end, v);
          uint64_t m = ballot(ok);
          n_cands += (end - cur) < WAVE ? (end - cur) : WAVE;
          if (ok) S.cand[L][popc64(m & lt_mask)] = v;
          S.cur[L] = cur + WAVE;
          wave_sync();
          PROF_END(t2, 4);
          if (m) finish_parents(popc64(m));
          continue;
        }
        --lv;
        if (lv < (int)d0) break;
        continue;
      }
      uint32_t bi = S.bi[L], bn = S.bn[L];
      if (bi < bn) {
        uint32_t v = S.cand[L][bi];
        S.bi[L] = bi + 1;
        S.m[L] = v;
        if (HASH) S.mq[S.pqv[L]] = v;
        ++lv;
        PROF_BEGIN(en);
        bool entered = enter(L + 1, 0u, 0xFFFFFFFFu, -1);
        PROF_END(en, 1);
        if (!entered) --lv;
        continue;
      }
      if (cur < end) {
        PROF_BEGIN(t1);
        uint32_t v;
        bool ok = test(L, cur + lane, end, v);
        uint64_t m = ballot(ok);
        n_cands += (end - cur) < WAVE ? (end - cur) : WAVE;
        if (ok) S.cand[L][popc64(m & lt_mask)] = v;
        S.bn[L] = popc64(m);
        S.bi[L] = 0;
        S.cur[L] = cur + WAVE;
        wave_sync();
        PROF_END(t1, 3);
        continue;
      }
      --lv;
      if (lv < (int)d0) break;
    }
    flush(true);
    if (lane == 0 && HASH && hsum) atomicAdd(&a.hashes[(size_t)t * a.nq + q], hsum);
    if (lane == 0 && lost) {
      atomicOr(&a.lim->ustatus[t], USTAT_TIME_LIMIT);
      atomicOr(&a.lim->qtl[q], 1u);
    }
    if (lane == 0 && (a.lflags & LIM_WORK) && n_cands > S.cands0)
      atomicAdd(&a.lim->uwork[t], n_cands - S.cands0);
  }
  n_bytes = wave_sum(n_bytes);
  if (lane == 0) {
    // a wave of an empty round adds nothing (and skips three contended atomics)
    if (n_tasks) {
      a.ctrl->round_busy[a.round_slot] = 1u;
      atomicAdd(&a.ctrl->stat_tasks, n_tasks);
      atomicAdd(&a.ctrl->stat_cands, n_cands);
      atomicAdd(&a.ctrl->stat_bytes, n_bytes + 4ull * n_cands);
    }
#ifdef CSM_PHASE_PROF
    for (int i = 0; i < 8; ++i) {
      atomicAdd(&a.ctrl->prof[i], prof_acc[i]);
      atomicAdd(&a.ctrl->prof[8 + i], prof_n[i]);
    }
    const uint64_t prof_kt1 = __builtin_amdgcn_s_memrealtime();
    prof_x[12] += prof_kt1 - prof_kt0;
    prof_x[13] += 1;
    for (int i = 0; i < 14; ++i) atomicAdd(&a.ctrl->prof[16 + i], prof_x[i]);
    atomicMin(&a.ctrl->prof[30], (unsigned long long)prof_kt0);
    atomicMax(&a.ctrl->prof[31], (unsigned long long)prof_kt1);
#endif
  }
}

// ------------------------------------------------------------ batch control
// counts of n updates x nq queries at or above the result cap become the cap
// (the search stops a (update, query) pair once its count reaches the cap,
// so the result is min(exact count, cap) whatever the order of the tasks)
__global__ void k_clamp_counts(unsigned long long* __restrict__ counts, uint32_t n, uint32_t nq,
                               unsigned long long cap, uint32_t* __restrict__ ustatus) {
  const uint64_t total = (uint64_t)n * nq;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * blockDim.x) {
    if (counts[i] >= cap) {
      counts[i] = cap;
      atomicOr(&ustatus[i / nq], USTAT_CAPPED);
    }
  }
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_batch_begin(Ctrl* c) {
#ifdef CSM_PHASE_PROF
  if (threadIdx.x == 0) { c->prof[30] = ~0ull; c->prof[31] = 0; }
#endif
  if (threadIdx.x == 0) {
    c->nseeds = 0;
    c->nchunks = 0;
    c->chunk_head = 0;
    for (uint32_t x = 0; x < NXCD; ++x) {
      c->xhead[x * XHEAD_STRIDE] = 0;
      c->xspill[x * XHEAD_STRIDE] = 0;
      c->xspill_chunks[x * XHEAD_STRIDE] = 0;
      c->xn[x] = 0;
    }
    c->nspill = 0;
    c->nspill_chunks = 0;
    c->nruns = 0;
    c->ndirty_next = 0;
    c->err = 0;
    c->err_at = 0xFFFFFFFFu;
    c->rounds = 0;
  }
}

// forget the seeds and spills of a search so that it can run again on the
// same batch (autotuning); the graph-update state of the batch is kept
__global__ void k_search_reset(Ctrl* c) {
  if (threadIdx.x == 0) {
    c->nseeds = 0;
    c->nchunks = 0;
    c->chunk_head = 0;
    for (uint32_t x = 0; x < NXCD; ++x) {
      c->xhead[x * XHEAD_STRIDE] = 0;
      c->xspill[x * XHEAD_STRIDE] = 0;
      c->xspill_chunks[x * XHEAD_STRIDE] = 0;
      c->xn[x] = 0;
    }
    c->nspill = 0;
    c->nspill_chunks = 0;
  }
}

// the spilled tasks of the last round become the input of the next
__global__ void k_next_round(Ctrl* c, uint32_t spill_chunk_cap) {
#ifdef CSM_PHASE_PROF
  if (threadIdx.x == 0) { c->prof[30] = ~0ull; c->prof[31] = 0; }
#endif
  if (threadIdx.x == 0) {
    if (XCD_REGIONS) {
      // region x holds XCD x's spills (ids past the region's end were never
      // handed out)
      const uint32_t ccap = spill_chunk_cap / NXCD;
      uint32_t total = 0;
      for (uint32_t x = 0; x < NXCD; ++x) {
        const uint32_t n = c->xspill_chunks[x * XHEAD_STRIDE];
        c->xn[x] = n < ccap ? n : ccap;
        total += c->xn[x];
        c->xspill[x * XHEAD_STRIDE] = 0;
        c->xspill_chunks[x * XHEAD_STRIDE] = 0;
      }
      c->nchunks = total;
    } else {
      c->nchunks = c->nspill_chunks;
    }
    if (c->nchunks) ++c->rounds;
    c->chunk_head = 0;
    for (uint32_t x = 0; x < NXCD; ++x) c->xhead[x * XHEAD_STRIDE] = 0;
    c->nspill = 0;
    c->nspill_chunks = 0;
  }
}

__global__ void k_batch_end(Ctrl* c) {
  if (threadIdx.x == 0) {
    c->ndirty_cur = c->ndirty_next;
    c->ndirty_next = 0;
  }
}

// Initial matching seeds: every directed data edge (v -> w) matching the first
// query edge of plan `qo` (each embedding maps that query edge onto exactly
// one directed data edge).  Vertices [v0, v1).
__global__ void k_initial_seeds(GraphView g, const PlanRec* __restrict__ plans, uint32_t qo,
                                uint32_t la, uint32_t key01, uint32_t v0, uint32_t v1,
                                uint32_t chunk, Task* seeds, uint32_t seed_cap,
                                uint32_t* chunk_owner, uint32_t chunk_cap, Ctrl* ctrl,
                                const uint32_t* __restrict__ d2, uint32_t d2_words, uint32_t symbi,
                                unsigned long long* counts, const SearchLimits* emb_lim) {
  const uint32_t lane = lane_id();
  const uint32_t wid = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
  const uint32_t nw = (gridDim.x * blockDim.x) / WAVE;
  const PlanRec* P = plans + qo;
  const uint32_t k = P->k, nb2 = k > 2 ? P->nb[2] : 0u;
  // one wave per first-position vertex v, one lane per entry of its segment
  for (uint32_t v = v0 + wid; v < v1; v += nw) {
    if (g.vlab[v] != la) continue;
    if (symbi && !d2_bit(d2, P->d2row[0], d2_words, v)) continue;
    uint32_t ry;
    const uint2 sg = seg_of(g, v, key01, &ry);
    const bool dirty = (ry & DIRTY) != 0;
    for (uint32_t i0 = sg.x; i0 < sg.y; i0 += WAVE) {
      const uint32_t i = i0 + lane;
      bool have = false, ok = i < sg.y;
      uint32_t w = 0, blo = 0, bhi = 0, bj = 0;
      uint2 ch = make_uint2(0u, 0u);
      if (ok) {
        if (dirty) ok = ts_visible(g.ats[i], 0xFFFFu);
        w = g.adj[i] & g.vmask;
        if (ok && symbi) ok = d2_bit(d2, P->d2row[1], d2_words, w);
      }
      if (k == 2) {
        const uint32_t n2 = popc64(ballot(ok));
        if (lane == 0 && n2) atomicAdd(&counts[P->q], (unsigned long long)n2);
        if (emb_lim)
          emit_rows(emb_lim, ok, 0u, P->q, 2u, [&](uint32_t u) { return u == P->qv[0] ? v : w; });
        continue;
      }
      if (ok) {
        uint32_t best = NONE;
        for (uint32_t jj = 0; jj < nb2; ++jj) {
          uint32_t bw = P->bw[2][jj];
          uint32_t vx = (bw & 0xFFu) == 0 ? v : w;
          uint32_t ry2;
          uint2 s2 = seg_of(g, vx, bw >> 8, &ry2);
          uint32_t sz = s2.y - s2.x;
          if (best == NONE || sz < best) { best = sz; blo = s2.x; bhi = s2.y; bj = jj; }
        }
        if (best != 0 && best != NONE) {
          have = true;
          ch = chunking(best, chunk);
        }
      }
      emit_seed_wave(have, 0u, qo, v, w, bj, blo, bhi, ch, seeds, seed_cap, chunk_owner,
                     chunk_cap, ctrl);
    }
  }
}

// ---------------------------------------------------------------- repack
__global__ void k_repack_lens(const uint2* __restrict__ vrec, uint32_t* __restrict__ lens,
                              uint32_t nv) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += gridDim.x * blockDim.x)
    lens[v] = vrec[v].y & LEN_MASK;
}

// one wavefront per list: entries and versions to their new offset
template <int BS>
__global__ __launch_bounds__(BS) void k_repack_copy(uint2* __restrict__ vrec,
                                                    const uint32_t* __restrict__ offs,
                                                    const uint32_t* __restrict__ adj_in,
                                                    const uint32_t* __restrict__ ats_in,
                                                    uint32_t* __restrict__ adj_out,
                                                    uint32_t* __restrict__ ats_out, uint32_t nv) {
  const uint32_t lane = lane_id();
  const uint32_t nw = gridDim.x * (BS / WAVE);
  for (uint32_t v = blockIdx.x * (BS / WAVE) + threadIdx.x / WAVE; v < nv; v += nw) {
    const uint2 r = vrec[v];
    const uint32_t len = r.y & LEN_MASK, no = offs[v];
    for (uint32_t i = lane; i < len; i += WAVE) {
      adj_out[no + i] = adj_in[r.x + i];
      ats_out[no + i] = ats_in[r.x + i];
    }
    if (lane == 0) vrec[v].x = no;
  }
}

}  // namespace csm
