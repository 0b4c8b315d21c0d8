// Search with the matching order chosen while searching: SymBi's
// ExtendableVertex rule (ParaCOSM/CSM/matching/Parallel_SymBi/
// parrallel_symbi.cpp:4217-4262, sequential SymBi/symbi.cpp:471-587).  Used by
// parallel_symbi with reference orders; the default search (k_search) runs
// static per-plan orders instead.
//
// For every unmatched query vertex u the rule keeps
//   E(u)      = min over matched neighbours w of n2[w -> u][m(w)]
//               (the d2-valid entries of DCS[w -> u][m(w)]),
//   u_min(u)  = the neighbour that gave E(u) (the first one on ties, in match
//               order), and
//   matched(u) = its matched neighbours.
// The next vertex is the unmatched one of least E among those that still
// have an unmatched neighbour ("non-isolated"), else the least among the
// isolated ones; ties go to the lower query vertex id.  Its candidates are
// the entries of DCS[u_min -> u][m(u_min)] (the label-filtered segment of
// m(u_min)'s list) that pass d2, injectivity and the other matched
// neighbours' joins.  The seed edge's endpoints are matched first, the DAG
// parent before the child.
//
// One wavefront runs one chunk of a seed: the seed's first dynamic level is
// cut into as many contiguous pieces as the seed has chunks, and chunk c
// walks piece c and everything below it.  Per-depth state (E, u_min,
// matched, the chosen vertex, its segment cursor, the join segments and a
// 64-entry candidate buffer) lives in LDS.  There are no spills: a chunk runs
// to its end, or to the time limit, the result cap or the match budget.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "csm_kernels.hpp"

namespace csm {

// Per query: the tables the order rule and the joins read.
struct DynQuery {
  uint32_t k;
  uint32_t pad[3];
  uint32_t deg[MAXK];          // query degree of u
  uint32_t nbm[MAXK];          // neighbours of u (bit w)
  uint32_t parm[MAXK];         // DAG parents of u (bit w)
  uint32_t d2row[MAXK];        // d2 bitmap row of u
  uint32_t key[MAXK][MAXK];    // key[w][u]: label key of S(m(w), L(u), l(w,u))
  uint32_t n2[MAXK][MAXK];     // n2[w][u]: counter row of directed query edge w -> u
};

struct DynArgs {
  SearchArgs s;
  const DynQuery* __restrict__ dq;   // indexed by query id
  const uint32_t* __restrict__ n2;   // SymBi n2 counters, rows of n2_stride entries
  uint32_t n2_stride;
};

template <int MK>
struct DynState {
  uint32_t mq[MK];             // data vertex of each query vertex (NONE: unmatched)
  uint32_t E[MK][MK];          // [depth][u]
  uint32_t A[MK][MK];          // [depth][u]: u_min | matched << 8 (u_min 0xFF: none)
  uint32_t ud[MK];             // vertex chosen at each depth
  uint32_t cur[MK], end[MK];   // its segment cursor
  uint32_t nc[MK], ci[MK];     // candidate buffer fill and read index
  uint32_t dirty[MK];          // the segment's list carries version words
  uint32_t jm[MK], jd[MK];     // join vertices (bit w) and their lists' dirty bits
  uint32_t jlo[MK][MK], jhi[MK][MK];
  uint32_t cand[MK][WAVE];
  uint32_t aborted, pad_;
  unsigned long long last_poll;
};

template <int MK, bool HASH>
__global__ __launch_bounds__(search_waves<MK>() * WAVE) void k_search_dyn(DynArgs da) {
  const SearchArgs& a = da.s;
  __shared__ DynState<MK> S_all[search_waves<MK>()];
  const uint32_t lane = lane_id();
  const uint64_t lt_mask = (1ull << lane) - 1ull;
  DynState<MK>& S = S_all[threadIdx.x / WAVE];
  if (lane == 0) {
    S.aborted = 0;
    S.last_poll = 0;
  }
  wave_sync();
  const uint32_t* __restrict__ adj = a.g.adj;
  const uint32_t* __restrict__ ats = a.g.ats;
  const uint32_t vbits = a.g.vbits, vmask = a.g.vmask;
  unsigned long long n_tasks = 0, n_cands = 0;

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
    return uniform(__shfl(v, 0)) != 0;
  };

  while (true) {
    uint32_t nch = uniform(__atomic_load_n(&a.ctrl->nchunks, __ATOMIC_RELAXED));
    nch = nch < a.chunk_cap ? nch : a.chunk_cap;
    uint32_t c = 0;
    if (lane == 0) c = atomicAdd(&a.ctrl->chunk_head, 1u);
    c = uniform(__shfl(c, 0));
    if (c >= nch) break;
    if (__atomic_load_n(&a.ctrl->abort, __ATOMIC_RELAXED)) break;
    const uint32_t sidx = uniform(a.chunk_owner[c]);
    if (sidx == NONE) continue;
    const Task* __restrict__ TK = a.tasks + sidx;
    const uint32_t t = uniform(TK->t), qo = uniform(TK->qo);
    const PlanRec* __restrict__ P = a.plans + qo;
    const uint32_t k = uniform(P->k);
    const uint32_t q = uniform(P->q);
    const uint32_t pos = t + 1 + a.pos_offset;
    if (k < 3 || k > (uint32_t)MK) continue;
    if (a.qabort && __atomic_load_n(&a.qabort[q], __ATOMIC_RELAXED)) continue;
    if (poll_abort()) {
      if (lane == 0) {
        atomicOr(&a.lim->ustatus[t], USTAT_TIME_LIMIT);
        atomicOr(&a.lim->qtl[q], 1u);
      }
      continue;
    }
    if (a.lflags & LIM_CAP) {
      uint32_t full = 0;
      if (lane == 0)
        full = __hip_atomic_load(&a.counts[(size_t)t * a.nq + q], __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT) >= a.lim->ucap;
      if (uniform(__shfl(full, 0))) continue;
    }
    ++n_tasks;
    const unsigned long long cands0 = n_cands;
    const DynQuery* __restrict__ DQ = da.dq + q;
    // this chunk's piece of the first dynamic level: the seed was cut into
    // chunking(hi - lo, chunk).x chunks with consecutive ids from base
    const uint32_t npiece = chunking(uniform(TK->hi) - uniform(TK->lo), uniform(P->chunk)).x;
    const uint32_t piece = c - uniform(TK->base);
    if (piece >= npiece) continue;   // not reached: ids [base, base + npiece) name this task

    // the seed: the DAG parent of the seed edge first
    const uint32_t ua = uniform(P->qv[0]), ub = uniform(P->qv[1]);
    const bool a_par = (uniform(DQ->parm[ub]) >> ua) & 1u;
    const uint32_t u1 = a_par ? ua : ub, u2 = a_par ? ub : ua;
    const uint32_t x1 = uniform(a_par ? TK->m[0] : TK->m[1]);
    const uint32_t x2 = uniform(a_par ? TK->m[1] : TK->m[0]);
    if (lane < (uint32_t)MK) S.mq[lane] = lane == u1 ? x1 : lane == u2 ? x2 : NONE;
    if (lane < k) {
      uint32_t E = NONE, umin = 0xFFu, matched = 0;
      if (lane != u1 && lane != u2) {
#pragma unroll
        for (int s = 0; s < 2; ++s) {
          const uint32_t uu = s == 0 ? u1 : u2, xx = s == 0 ? x1 : x2;
          if ((DQ->nbm[uu] >> lane) & 1u) {
            const uint32_t val = da.n2[(size_t)DQ->n2[uu][lane] * da.n2_stride + xx];
            if (val < E) { E = val; umin = uu; }
            ++matched;
          }
        }
      }
      S.E[2][lane] = E;
      S.A[2][lane] = umin | (matched << 8);
    }
    wave_sync();

    unsigned long long cnt = 0, flushed = 0, hsum = 0;
    bool stop = false, lost = false;
    const uint32_t L1 = k - 1;

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
        uint32_t capped = 0;
        if (lane == 0) {
          unsigned long long* cp = &a.counts[(size_t)t * a.nq + q];
          const unsigned long long tot_u =
              d ? atomicAdd(cp, d) + d
                : __hip_atomic_load(cp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          capped = tot_u >= a.lim->ucap;
        }
        if (!final_ && uniform(__shfl(capped, 0))) stop = true;
      }
      flushed = cnt;
      if (!final_ && a.qabort)
        stop = stop || uniform(__atomic_load_n(&a.qabort[q], __ATOMIC_RELAXED)) != 0;
    };

    // pick the next query vertex at depth d and set up its segment and joins;
    // false when nothing is left to match from here
    auto enter = [&](uint32_t d) CSM_AINL -> bool {
      uint64_t key = ~0ull;
      if (lane < k && S.mq[lane] == NONE) {
        const uint32_t E = S.E[d][lane];
        if (E != NONE) {
          const uint32_t iso = (S.A[d][lane] >> 8) == DQ->deg[lane] ? 1u : 0u;
          key = ((uint64_t)iso << 40) | ((uint64_t)E << 8) | lane;
        }
      }
      key = wave_min_u64(key);
      if (key == ~0ull) return false;
      const uint32_t u = uniform((uint32_t)key & 0xFFu);
      const uint32_t umin = uniform(S.A[d][u] & 0xFFu);
      uint32_t ry;
      const uint2 sg = seg_of(a.g, S.mq[umin], uniform(DQ->key[umin][u]), &ry);
      uint32_t lo = sg.x, hi = sg.y;
      if (d == 2) {
        const uint64_t L = hi - lo;
        hi = lo + (uint32_t)(L * (piece + 1) / npiece);
        lo = lo + (uint32_t)(L * piece / npiece);
      }
      // other matched neighbours of u: segments the candidates must be in
      bool j = lane < k && lane != umin && S.mq[lane] != NONE && ((DQ->nbm[u] >> lane) & 1u);
      bool jdirty = false;
      if (j) {
        uint32_t ry2;
        const uint2 s2 = seg_of(a.g, S.mq[lane], DQ->key[lane][u], &ry2);
        S.jlo[d][lane] = s2.x;
        S.jhi[d][lane] = s2.y;
        jdirty = (ry2 & DIRTY) != 0;
      }
      const uint32_t jm = (uint32_t)ballot(j), jd = (uint32_t)ballot(jdirty);
      if (lane == 0) {
        S.ud[d] = u;
        S.cur[d] = lo;
        S.end[d] = hi;
        S.nc[d] = 0;
        S.ci[d] = 0;
        S.dirty[d] = (ry & DIRTY) ? 1u : 0u;
        S.jm[d] = jm;
        S.jd[d] = jd;
      }
      wave_sync();
      return true;
    };

    uint32_t d = 2;
    bool alive = enter(2);
    uint32_t ticks = 0;
    while (alive) {
      if (++ticks >= 256) {
        ticks = 0;
        flush(false);
        if (stop) break;
      }
      if (poll_abort()) {
        lost = true;
        break;
      }
      const uint32_t ci = uniform(S.ci[d]), ncd = uniform(S.nc[d]);
      if (d < L1 && ci < ncd) {
        // descend with the next buffered candidate
        const uint32_t u = uniform(S.ud[d]);
        const uint32_t v = uniform(S.cand[d][ci]);
        if (lane == 0) {
          S.ci[d] = ci + 1;
          S.mq[u] = v;
        }
        wave_sync();
        if (lane < k) {
          uint32_t E = S.E[d][lane], A = S.A[d][lane];
          if (S.mq[lane] == NONE && ((DQ->nbm[u] >> lane) & 1u)) {
            const uint32_t val = da.n2[(size_t)DQ->n2[u][lane] * da.n2_stride + v];
            if (val < E) { E = val; A = (A & ~0xFFu) | u; }
            A += 1u << 8;
          }
          S.E[d + 1][lane] = E;
          S.A[d + 1][lane] = A;
        }
        wave_sync();
        ++d;
        if (!enter(d)) {
          --d;
          if (lane == 0) S.mq[S.ud[d]] = NONE;
          wave_sync();
        }
        continue;
      }
      const uint32_t cur = uniform(S.cur[d]), end = uniform(S.end[d]);
      if (cur < end) {
        // test the next 64 entries of the segment
        const uint32_t u = uniform(S.ud[d]);
        const uint32_t e = cur + lane;
        bool ok = e < end;
        uint32_t v = 0;
        if (ok) {
          v = adj[e] & vmask;
          if (S.dirty[d]) ok = ts_visible(ats[e], pos);
        }
        n_cands += popc64(ballot(ok));
        if (ok) ok = d2_bit(a.d2, DQ->d2row[u], a.d2_words, v);
        if (ok && !a.homo)
          for (uint32_t i = 0; i < k; ++i) ok = ok && S.mq[i] != v;
        uint32_t jm = uniform(S.jm[d]);
        const uint32_t jd = uniform(S.jd[d]);
        while (jm) {
          const uint32_t w = __builtin_ctz(jm);
          jm &= jm - 1u;
          if (!ok) continue;
          const uint32_t x = (DQ->key[w][u] << vbits) | v;
          const uint32_t slo = uniform(S.jlo[d][w]);
          ok = contains_u(adj, ats, slo, uniform(S.jhi[d][w]) - slo, x, (jd >> w) & 1u, pos);
        }
        const uint64_t m = ballot(ok);
        if (d == L1) {
          cnt += popc64(m);
          if (HASH && ok) {
            uint64_t h = 0;
            for (uint32_t i = 0; i < k; ++i) h = mix64(h + (uint64_t)(i == u ? v : S.mq[i]) + 1);
            hsum += h;
          }
          if (a.lflags & LIM_ENUM)
            emit_rows(a.lim, ok, t, q, k, [&](uint32_t i) { return i == u ? v : S.mq[i]; });
        } else if (ok) {
          S.cand[d][popc64(m & lt_mask)] = v;
        }
        if (lane == 0) {
          S.cur[d] = cur + WAVE;
          if (d < L1) {
            S.nc[d] = popc64(m);
            S.ci[d] = 0;
          }
        }
        wave_sync();
        continue;
      }
      // this level is done: back to the previous one
      if (d == 2) break;
      --d;
      if (lane == 0) S.mq[S.ud[d]] = NONE;
      wave_sync();
    }
    flush(true);
    if (HASH) {
      hsum = wave_sum(hsum);
      if (lane == 0 && hsum) atomicAdd(&a.hashes[(size_t)t * a.nq + q], hsum);
    }
    if (lane == 0 && lost) {
      atomicOr(&a.lim->ustatus[t], USTAT_TIME_LIMIT);
      atomicOr(&a.lim->qtl[q], 1u);
    }
    if (lane == 0 && (a.lflags & LIM_WORK) && n_cands > cands0)
      atomicAdd(&a.lim->uwork[t], n_cands - cands0);
  }
  if (lane == 0 && n_tasks) {
    a.ctrl->round_busy[a.round_slot] = 1u;
    atomicAdd(&a.ctrl->stat_tasks, n_tasks);
    atomicAdd(&a.ctrl->stat_cands, n_cands);
    atomicAdd(&a.ctrl->stat_bytes, 4ull * n_cands);
  }
}

}  // namespace csm
