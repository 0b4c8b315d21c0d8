// Device parsing of the reference's text formats: graph files
// (Graph::LoadFromFile, ParaCOSM/CSM/graph/graph.cpp:217-249: "t" header
// lines, "v id label", and any other leading char an edge "e a b label") and
// update streams (Graph::LoadUpdateStream, graph.cpp:251-285: "v"/"-v id
// label", and any other leading token an edge "e a b label" with is_add =
// (token == "e")), one record per line, into csm_update records (InsertUnit
// layout, utils/types.h:34-42).  The reference reads tokens across line
// breaks and stops silently at the first field that is not a number; here a
// record is one line and such a line is an error that names its number.
//
// Three passes over the bytes in HBM: newlines counted per 4 KiB tile, their
// positions written after a scan of the tile counts, then one thread per line
// decodes its record.  Records of skipped lines ("t" headers, blank lines)
// are flagged and compacted away by the caller.
#pragma once
#include "csm_kernels.hpp"

namespace csm {

constexpr int PARSE_BS = 256;
constexpr int PARSE_BYTES = 16;                       // bytes per thread and tile pass
constexpr uint32_t PARSE_TILE = PARSE_BS * PARSE_BYTES;
constexpr uint32_t PARSE_MAX_LINE = 256;              // longer lines are malformed
constexpr int PARSE_GRAPH = 0, PARSE_STREAM = 1;

// the 16 bytes of thread slot `base` (fewer at the end of the text)
__device__ __forceinline__ void load16(const uint8_t* __restrict__ text, uint64_t len,
                                       uint64_t base, uint8_t b[PARSE_BYTES]) {
  if (base + PARSE_BYTES <= len) {
    const uint4 w = *reinterpret_cast<const uint4*>(text + base);
    const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
    for (int i = 0; i < PARSE_BYTES; ++i) b[i] = (uint8_t)(ws[i >> 2] >> (8 * (i & 3)));
  } else {
#pragma unroll
    for (int i = 0; i < PARSE_BYTES; ++i) b[i] = base + i < len ? text[base + i] : 0;
  }
}

// exclusive prefix of x over the block, and the block total
__device__ __forceinline__ uint32_t block_exclusive(uint32_t x, uint32_t* total) {
  __shared__ uint32_t s_wave[PARSE_BS / WAVE];
  const uint32_t lane = lane_id(), w = threadIdx.x / WAVE;
  uint32_t incl = x;
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    uint32_t y = __shfl_up(incl, o);
    if ((int)lane >= o) incl += y;
  }
  if (lane == WAVE - 1) s_wave[w] = incl;
  __syncthreads();
  uint32_t before = 0, tot = 0;
  for (uint32_t i = 0; i < PARSE_BS / WAVE; ++i) {
    if (i < w) before += s_wave[i];
    tot += s_wave[i];
  }
  __syncthreads();
  *total = tot;
  return before + incl - x;
}

// pass 1: newlines per tile (the text's base must be 16-byte aligned)
__global__ __launch_bounds__(PARSE_BS) void k_parse_count(const uint8_t* __restrict__ text,
                                                          uint64_t len,
                                                          uint32_t* __restrict__ tile_count) {
  const uint64_t base = (uint64_t)blockIdx.x * PARSE_TILE + (uint64_t)threadIdx.x * PARSE_BYTES;
  uint8_t b[PARSE_BYTES];
  load16(text, len, base, b);
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < PARSE_BYTES; ++i) c += b[i] == '\n';
  uint32_t tot;
  block_exclusive(c, &tot);
  if (threadIdx.x == 0) tile_count[blockIdx.x] = tot;
}

// pass 2: byte offset of every newline, in order (tile_off: exclusive scan
// of the tile counts)
__global__ __launch_bounds__(PARSE_BS) void k_parse_lines(const uint8_t* __restrict__ text,
                                                          uint64_t len,
                                                          const uint32_t* __restrict__ tile_off,
                                                          uint64_t* __restrict__ line_end) {
  const uint64_t base = (uint64_t)blockIdx.x * PARSE_TILE + (uint64_t)threadIdx.x * PARSE_BYTES;
  uint8_t b[PARSE_BYTES];
  load16(text, len, base, b);
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < PARSE_BYTES; ++i) c += b[i] == '\n';
  uint32_t tot;
  uint32_t o = tile_off[blockIdx.x] + block_exclusive(c, &tot);
#pragma unroll
  for (int i = 0; i < PARSE_BYTES; ++i)
    if (b[i] == '\n') line_end[o++] = base + i;
}

__device__ __forceinline__ bool is_space(uint8_t c) { return c == ' ' || c == '\t' || c == '\r'; }

// pass 3: one thread per line.  Line i spans (line_end[i-1], line_end[i]); a
// last line without a newline ends at len.  keep[i] = 0 for skipped lines;
// a malformed line sets *err and lowers *err_line to its index.
__global__ void k_parse_records(const uint8_t* __restrict__ text, uint64_t len,
                                const uint64_t* __restrict__ line_end, uint64_t n_nl,
                                uint64_t n_lines, int mode, UpdateRec* __restrict__ rec,
                                uint8_t* __restrict__ keep, uint32_t* __restrict__ err,
                                unsigned long long* __restrict__ err_line) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_lines;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t s = i ? line_end[i - 1] + 1 : 0;
    const uint64_t e = i < n_nl ? line_end[i] : len;
    UpdateRec r = {0, 0, 0, 0, 0, 0};
    uint8_t k = 0;
    bool bad = e - s > PARSE_MAX_LINE;
    uint64_t p = s;
    while (!bad && p < e && is_space(text[p])) ++p;
    if (!bad && p < e) {
      uint32_t need = 0;
      bool first_glued = false;   // the first field may follow the type without a space
      if (mode == PARSE_GRAPH) {
        // LoadFromFile reads the type as one char (`ifs >> type`, graph.cpp:226):
        // 't' is a header, 'v' a vertex, and any other char an edge
        // (graph.cpp:228-246); the next field starts right after that char
        const uint8_t t = text[p++];
        first_glued = true;
        r.add = 1;
        if (t == 't') need = NONE;
        else if (t == 'v') { r.type = 'v'; need = 2; }
        else { r.type = 'e'; need = 3; }
      } else {
        // LoadUpdateStream reads a whitespace-delimited token (graph.cpp:263-281):
        // "v" / "-v" are vertex records, any other token is an edge record
        // with is_add = (token == "e")
        const uint64_t t0 = p;
        while (p < e && !is_space(text[p])) ++p;
        const uint64_t tl = p - t0;
        const bool is_v = (tl == 1 && text[t0] == 'v') ||
                          (tl == 2 && text[t0] == '-' && text[t0 + 1] == 'v');
        if (is_v) { r.type = 'v'; need = 2; r.add = tl == 1; }
        else { r.type = 'e'; need = 3; r.add = tl == 1 && text[t0] == 'e'; }
      }
      if (!bad && need != NONE) {
        uint32_t val[3] = {0, 0, 0};
        for (uint32_t f = 0; f < need && !bad; ++f) {
          const uint64_t q0 = p;
          while (p < e && is_space(text[p])) ++p;
          if ((p == q0 && !(f == 0 && first_glued)) || p == e) { bad = true; break; }
          uint64_t x = 0;
          const uint64_t d0 = p;
          while (p < e && text[p] >= '0' && text[p] <= '9') {
            x = x * 10 + (text[p] - '0');
            if (x > 0xFFFFFFFFull) bad = true;
            ++p;
          }
          if (p == d0 || (p < e && !is_space(text[p]))) bad = true;
          val[f] = (uint32_t)x;
        }
        while (!bad && p < e && is_space(text[p])) ++p;
        if (p != e) bad = true;   // trailing token
        if (!bad) {
          r.id1 = val[0];
          if (need == 3) { r.id2 = val[1]; r.label = val[2]; }
          else r.label = val[1];
          k = 1;
        }
      }
    }
    if (bad) {
      k = 0;
      atomicOr(err, 1u);
      atomicMin(err_line, (unsigned long long)i);
    }
    rec[i] = r;
    keep[i] = k;
  }
}

}  // namespace csm

namespace csm {

// ---------------------------------------------------------------------------
// A data graph straight from graph-file records in HBM (Graph::LoadFromFile,
// graph.cpp:217-249, with AddVertex / AddEdge, graph.cpp:23-93): vertex ids
// run to the largest id of any record; a vertex keeps the label of its first
// 'v' record (later ones are ignored, as AddVertex does) and a vertex without
// one has no label; every 'e' record is an edge in file order (duplicates are
// dropped by the CSR build, keeping the first, as AddEdge does).

// largest vertex id of any record (+1 = vertex count); *maxid starts at 0
__global__ void k_rec_maxid(const UpdateRec* __restrict__ r, uint64_t n, uint32_t* maxid,
                            unsigned long long* nedges) {
  uint32_t m = 0;
  unsigned long long ne = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const UpdateRec u = r[i];
    const uint32_t hi = u.type == 'e' ? (u.id1 > u.id2 ? u.id1 : u.id2) : u.id1;
    m = hi + 1 > m ? hi + 1 : m;
    ne += u.type == 'e';
  }
  atomicMax(maxid, m);
  if (ne) atomicAdd(nedges, ne);
}

// first 'v' record of each vertex (first[] starts at all ones)
__global__ void k_rec_first_vertex(const UpdateRec* __restrict__ r, uint64_t n,
                                   unsigned long long* __restrict__ first) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    if (r[i].type == 'v') atomicMin(&first[r[i].id1], (unsigned long long)i);
}

// label of each vertex from its first 'v' record (vl[] starts at NO_LABEL)
__global__ void k_rec_vlabels(const UpdateRec* __restrict__ r, uint64_t n,
                              const unsigned long long* __restrict__ first,
                              uint32_t* __restrict__ vl) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    if (r[i].type == 'v' && first[r[i].id1] == i) vl[r[i].id1] = r[i].label;
}

// edge label of each record (NO_LABEL for the others), for the label set
__global__ void k_rec_elabels(const UpdateRec* __restrict__ r, uint64_t n, uint32_t* __restrict__ el) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    el[i] = r[i].type == 'e' ? r[i].label : NO_LABEL;
}

// the records as the CSR build's edge arrays: an 'e' record becomes
// (id1, id2, dense edge label); any other record a self-loop, which the
// build drops.  hist (nullable) counts each edge's (L(a), L(b), label)
// triple in both orientations (the SymBi root choice, symbi.cpp:113-175).
__global__ void k_rec_edges(const UpdateRec* __restrict__ r, uint64_t n,
                            const uint32_t* __restrict__ vlab, const uint32_t* __restrict__ elmap,
                            uint32_t nel, uint32_t nlab, uint32_t* __restrict__ src,
                            uint32_t* __restrict__ dst, uint32_t* __restrict__ del,
                            unsigned long long* __restrict__ hist) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const UpdateRec u = r[i];
    if (u.type != 'e') {
      src[i] = dst[i] = 0;
      del[i] = 0;
      continue;
    }
    // dense edge label: position in the sorted label map
    uint32_t lo = 0, hi = nel;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (elmap[mid] < u.label) lo = mid + 1; else hi = mid;
    }
    src[i] = u.id1;
    dst[i] = u.id2;
    del[i] = lo;
    if (hist && u.id1 != u.id2) {
      const uint32_t la = vlab[u.id1], lb = vlab[u.id2];
      if (la < nlab && lb < nlab) {
        atomicAdd(&hist[((uint64_t)la * nlab + lb) * nel + lo], 1ull);
        atomicAdd(&hist[((uint64_t)lb * nlab + la) * nel + lo], 1ull);
      }
    }
  }
}

}  // namespace csm
