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
