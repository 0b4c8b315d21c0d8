// Host side of the MI355X incremental match engine and its C-ABI
// (include/csm_c.h).  One engine = one device-resident data graph + registered
// queries; all state lives in HBM and every launch goes to the engine's own
// HIP stream.
#include <hip/hip_runtime.h>

#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/csm_c.h"
#include "csm_kernels.hpp"
#include "csm_plan.hpp"
#include "csm_symbi.hpp"
#include "csm_dyn.hpp"
#include "csm_parse.hpp"

using namespace csm;

namespace {

enum Algo { ALGO_GRAPHFLOW = 0, ALGO_SYMBI = 1, ALGO_TURBOFLUX = 2, ALGO_NONE = 3 };

constexpr uint32_t DEFAULT_MAX_BATCH = 65534;
// --print-enum rows: update, query, then the data vertex of each query vertex
constexpr uint32_t EMB_W = 2 + MAXK;
constexpr uint64_t EMB_DEFAULT_ROWS = 1u << 20;
constexpr size_t TUNE_MIN_BATCH = 4096;   // smallest batch the autotuner scores
constexpr int BUILD_BS = 256;
constexpr int APPLY_BS = 256;

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

// Query size classes of the search settings: up to 8, 12, 16 and 32 vertices.
constexpr int NSIZE_CLASSES = 4;
inline int size_class(uint32_t k) { return k <= 8 ? 0 : k <= 12 ? 1 : k <= 16 ? 2 : 3; }

struct QueryInfo {
  QueryGraph g;                         // original labels
  std::vector<uint32_t> dlab;           // dense vertex labels
  std::vector<std::vector<uint32_t>> orders;
  uint32_t plan_base = 0;               // first PlanRec index
  Dag dag;
  uint32_t d2_base = 0;                 // first SymBi row
  bool sx_built = false;
  // TurboFlux's explicit-path counts per query vertex (its order rule's
  // ranking; computed for parallel_turboflux with reference_orders)
  std::vector<uint32_t> tf_paths;
  // n2 counter row of each directed query edge w -> u, at [w * k + u]
  // (NONE: not an edge)
  std::vector<uint32_t> n2row;
};

inline uint32_t bits_for(uint64_t n) {  // bits to hold values in [0, n)
  uint32_t b = 1;
  while ((1ull << b) < n) ++b;
  return b;
}

}  // namespace

struct csm_engine {
  int algo = ALGO_GRAPHFLOW;
  std::string algo_name;
  csm_opts opts{};
  int dev = 0;
  hipStream_t stream = nullptr;
  int num_cus = 0;
  uint32_t max_k = 0;       // largest registered query (selects the search LDS layout)
  std::string err;

  // labels
  std::vector<uint32_t> vl_map, el_map;  // sorted original labels; dense id = index
  uint32_t nlab = 0, nel = 0, vbits = 0, vmask = 0;
  uint32_t nv = 0, nv_cap = 0;
  std::vector<uint32_t> h_vlab;          // dense labels (NO_LABEL = absent)
  uint64_t num_edges = 0;

  // device graph
  uint2* d_vrec = nullptr;
  uint32_t* d_vlab = nullptr;
  uint32_t* d_adj[2] = {nullptr, nullptr};
  uint32_t* d_ats[2] = {nullptr, nullptr};
  int cur = 0;
  uint64_t pool_cap = 0;
  Ctrl* d_ctrl = nullptr;
  Ctrl* h_ctrl = nullptr;  // pinned
  uint32_t* d_dirty[2] = {nullptr, nullptr};
  int dcur = 0;            // d_dirty[dcur] = list compacted at the next batch start
  // key directories (segment offsets per key) of lists with >= dir_threshold entries
  uint32_t* d_dslot = nullptr;
  uint32_t* d_dir = nullptr;
  uint32_t dir_cap = 0, dstride = 0, dir_threshold = 0;
  uint32_t last_batch = 0;  // updates in the previous batch (bounds its compaction list)
  unsigned long long trace_life0 = 0, trace_waves0 = 0;  // phase-profile trace baselines

  // batch buffers
  uint32_t max_batch = DEFAULT_MAX_BATCH;
  UpdateRec* d_upd = nullptr;
  uint8_t *d_type = nullptr, *d_add = nullptr;
  uint32_t *d_v1 = nullptr, *d_v2 = nullptr, *d_lab = nullptr, *d_slab = nullptr;
  uint64_t *d_keys = nullptr, *d_keys2 = nullptr;
  uint32_t *d_vals = nullptr, *d_vals2 = nullptr, *d_runs = nullptr;
  void* d_sort_tmp = nullptr;
  size_t sort_tmp_bytes = 0;
  uint32_t* d_elmap = nullptr;
  Task* d_seeds = nullptr;
  uint32_t seed_cap = 0;
  uint32_t* d_chunk_owner = nullptr;
  // repack scratch, kept between repacks: list lengths, new offsets, scan space
  uint32_t *d_rp_lens = nullptr, *d_rp_offs = nullptr;
  uint8_t* d_rp_tmp = nullptr;
  size_t rp_tmp_bytes = 0;
  uint32_t chunk_cap = 0;
  Task* d_spill[2] = {nullptr, nullptr};          // ping-pong task lists of later rounds
  uint32_t* d_spill_owner[2] = {nullptr, nullptr};
  uint32_t spill_cap = 0, spill_chunk_cap = 0;
  // search settings per query size class (size_class): candidates per
  // chunk when a range is cut, and the spill budget; every plan carries its
  // class's pair (PlanRec::chunk, PlanRec::budget)
  uint32_t cls_chunk[NSIZE_CLASSES], cls_budget[NSIZE_CLASSES];
  // online autotuner of those settings per size class (the reference's
  // --auto-tuning): after one settling batch, the next large batch runs the
  // search of each class's queries once per candidate; the least search
  // time is kept for that class
  bool tune_active = false;
  uint32_t tune_idx = 0;
  std::vector<std::pair<uint32_t, uint32_t>> tune_cands;
  bool tune_mask = false;   // the tuning search skips the queries of other classes
  uint32_t max_rounds = 0;                        // stats: deepest round count seen
  unsigned long long* d_counts = nullptr;
  unsigned long long* d_hashes = nullptr;
  unsigned long long* d_qcount = nullptr;  // per-query running matches (budget)
  uint32_t* d_qabort = nullptr;
  std::vector<uint32_t> h_qabort;
  uint32_t* d_qtl = nullptr;             // per query: time limit dropped work
  std::vector<uint32_t> h_qtl;
  size_t q_cap = 0;
  uint8_t* d_safe = nullptr;
  unsigned long long* h_counts = nullptr;  // pinned
  unsigned long long* h_hashes = nullptr;
  uint8_t* h_safe = nullptr;
  UpdateRec* h_upd = nullptr;              // pinned staging

  // queries
  std::vector<std::unique_ptr<QueryInfo>> queries;
  std::vector<PlanRec> h_plans;
  PlanRec* d_plans = nullptr;
  size_t d_plans_cap = 0;
  std::vector<uint32_t> h_tab_off, h_tab_ent;
  uint32_t* d_tab_off = nullptr;
  uint32_t* d_tab_ent = nullptr;
  size_t d_tab_ent_cap = 0;
  bool preprocessed = false;
  // updates in the batch being searched (sets the spill budget's scale)
  uint32_t batch_n = UINT32_MAX;
  // parallel_symbi with reference orders: the next query vertex is chosen
  // during the search (k_search_dyn), from these per-query tables
  bool dyn = false;
  DynQuery* d_dynq = nullptr;
  size_t d_dynq_cap = 0;

  // SymBi index: rows of all queries; d_d2 aliases the d2 bitmap for the search
  uint32_t* d_d2 = nullptr;
  uint32_t d2_rows = 0, d2_words = 0;
  uint32_t sx_n1rows = 0, sx_n2rows = 0;
  uint32_t sx_cap_rows = 0, sx_cap_n1 = 0, sx_cap_n2 = 0;
  std::vector<SxRow> h_sxrows;
  SxRow* d_sxrows = nullptr;
  uint32_t *d_sx_d1 = nullptr, *d_sx_d2 = nullptr, *d_sx_np1 = nullptr, *d_sx_nc2 = nullptr;
  uint32_t *d_sx_n1 = nullptr, *d_sx_n2 = nullptr;
  std::vector<uint32_t> h_sxtab_off, h_sxtab_ent;
  uint32_t *d_sxtab_off = nullptr, *d_sxtab_ent = nullptr;
  uint8_t* d_sx_mark = nullptr;
  unsigned long long* d_sx_lists[4] = {nullptr, nullptr, nullptr, nullptr};
  uint32_t sx_list_cap = 0;
  uint32_t* d_sx_slots = nullptr;
  uint32_t sx_maxk = 0;
  // label histograms of the initial graph (SymBi root choice, symbi.cpp:113-175)
  std::vector<uint64_t> h_triple_freq, h_vfreq;

  // totals and stats
  uint64_t tot_init = 0, tot_pos = 0, tot_neg = 0;
  uint64_t stats[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  // HIP events around search launches on the engine stream (kernel time)
  // the start and end of each search launch of a group of rounds
  hipEvent_t ev_round[ROUND_GROUP_MAX][2] = {};
  uint64_t dev_bytes = 0;

  // time limit (csm_opts.time_limit_s per engine call): a word in pinned,
  // fine-grained host memory that the search polls (h_abort, mapped at
  // d_abort_map); the host sets it once the deadline has passed
  uint32_t* h_abort = nullptr;
  uint32_t* d_abort_map = nullptr;
  double deadline = 0;       // steady-clock seconds; 0 = no limit in force
  bool timed_out = false;
  int call_depth = 0;
  // per update of the last call: USTAT_* bits (d_ustatus per batch)
  uint32_t* d_ustatus = nullptr;
  std::vector<uint8_t> h_ustatus;
  // per update of the last call: candidate entries the search tested
  // (csm_opts.want_work)
  unsigned long long* d_uwork = nullptr;
  SearchLimits* d_limits = nullptr;      // the search's SearchLimits (device copy)
  unsigned long long* h_uwork_pin = nullptr;
  std::vector<uint64_t> h_uwork;
  // --print-enum: device rows of the current search, rows of the last call
  uint32_t* d_emb = nullptr;
  unsigned long long* d_emb_n = nullptr;
  uint64_t emb_cap = 0;
  std::vector<uint32_t> h_emb;           // EMB_W words per row
  uint64_t emb_found = 0;
  uint32_t* h_ustatus_pin = nullptr;

  // steady-state batches replayed as HIP graphs (see run_batch_graph); gen
  // counts device allocations, whose addresses a captured graph bakes in
  uint64_t gen = 0;
  struct GraphKey {
    uint32_t n, nq, dcur, cur, hash, nv, budgeted, capped;
    uint64_t gen;
    bool operator==(const GraphKey& o) const { return std::memcmp(this, &o, sizeof *this) == 0; }
  };
  struct GraphEntry {
    GraphKey key;
    hipGraphExec_t exec = nullptr;
    uint32_t rounds = 0;      // search rounds inside the graph
    bool bad = false;         // capture failed: this key runs eagerly
  };
  std::vector<GraphEntry> graphs;
  uint64_t graph_launches = 0;
  // pinned words for the allocation-free parse's read-backs
  uint64_t* h_parse = nullptr;
  // transfer stream (default priority, like the engine stream) and markers
  // for overlapping host-device copies with the engine's work
  hipStream_t xfer = nullptr;
  hipEvent_t marks[CSM_MARKS] = {};

  std::vector<DevBuf> allocs;
};

namespace {

int fail(csm_engine* e, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (e) e->err = buf;
  return code;
}

inline double steady_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// One engine call that may search.  The outermost call of a nesting starts
// the time limit's clock (csm_opts.time_limit_s; the reference's watchdog,
// utils/utils.h:19-36, polled as reach_time_limit in the searches,
// parallel_graphflow.cpp:288,386-387) and turns a successful return into
// CSM_E_TIME_LIMIT when the limit fired.
struct CallScope {
  csm_engine* e;
  bool outer;
  explicit CallScope(csm_engine* e_) : e(e_), outer(e_ && e_->call_depth++ == 0) {
    if (!outer) return;
    e->timed_out = false;
    e->deadline = e->opts.time_limit_s > 0 ? steady_s() + e->opts.time_limit_s : 0.0;
    if (e->h_abort) __atomic_store_n(e->h_abort, 0u, __ATOMIC_SEQ_CST);
  }
  int finish(int rc) {
    if (outer && rc == CSM_OK && e->timed_out) {
      e->err = "time limit reached; updates flagged in csm_update_status have partial counts";
      rc = CSM_E_TIME_LIMIT;
    }
    return rc;
  }
  ~CallScope() {
    if (!e) return;
    if (--e->call_depth == 0) {
      e->deadline = 0;
      if (e->h_abort) __atomic_store_n(e->h_abort, 0u, __ATOMIC_SEQ_CST);
    }
  }
};

// past the deadline: raise the device-visible abort word (once)
inline void check_deadline(csm_engine* e) {
  if (e->deadline > 0 && !e->timed_out && steady_s() >= e->deadline) {
    e->timed_out = true;
    __atomic_store_n(e->h_abort, 1u, __ATOMIC_SEQ_CST);
  }
}

#define HIP_TRY(e, call)                                                               \
  do {                                                                                 \
    hipError_t _st = (call);                                                           \
    if (_st != hipSuccess)                                                             \
      return fail((e), CSM_E_DEVICE, "%s failed: %s", #call, hipGetErrorString(_st));  \
  } while (0)

template <class T>
int dalloc(csm_engine* e, T** p, size_t count) {
  size_t bytes = std::max<size_t>(count * sizeof(T), 256);
  if (e->opts.device_budget && e->dev_bytes + bytes > e->opts.device_budget)
    return fail(e, CSM_E_ADMISSION, "device budget of %llu bytes exceeded (%llu + %zu)",
                (unsigned long long)e->opts.device_budget, (unsigned long long)e->dev_bytes, bytes);
  void* q = nullptr;
  hipError_t st = hipMalloc(&q, bytes);
  if (st != hipSuccess)
    return fail(e, CSM_E_DEVICE, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(st));
  e->allocs.push_back({q, bytes});
  e->dev_bytes += bytes;
  ++e->gen;
  *p = static_cast<T*>(q);
  return CSM_OK;
}

template <class T>
void dfree(csm_engine* e, T*& p) {
  if (!p) return;
  for (size_t i = 0; i < e->allocs.size(); ++i)
    if (e->allocs[i].p == p) {
      e->dev_bytes -= e->allocs[i].bytes;
      e->allocs.erase(e->allocs.begin() + i);
      break;
    }
  (void)hipFree(p);
  p = nullptr;
}

inline int grid_for(uint64_t n, int bs, int cap = 4096) {
  uint64_t g = (n + bs - 1) / bs;
  if (g < 1) g = 1;
  return (int)std::min<uint64_t>(g, (uint64_t)cap);
}

inline uint32_t dense_of(const std::vector<uint32_t>& map, uint32_t x) {
  auto it = std::lower_bound(map.begin(), map.end(), x);
  return (it != map.end() && *it == x) ? (uint32_t)(it - map.begin()) : NO_LABEL;
}

// SymBi and TurboFlux keep the candidate index (d1/d2 and its counters on
// SymBi's DAG or TurboFlux's spanning tree) and filter the search by d2
inline bool uses_index(const csm_engine* e) {
  return e->algo == ALGO_SYMBI || e->algo == ALGO_TURBOFLUX;
}

GraphView view(const csm_engine* e) {
  GraphView g;
  g.vrec = e->d_vrec;
  g.vlab = e->d_vlab;
  g.adj = e->d_adj[e->cur];
  g.ats = e->d_ats[e->cur];
  g.dslot = e->d_dslot;
  g.dir = e->d_dir;
  g.dstride = e->dstride;
  g.vbits = e->vbits;
  g.vmask = e->vmask;
  g.nel = e->nel;
  g.nv = e->nv;
  return g;
}

int parse_algo(const char* a, int* out) {
  std::string s(a ? a : "");
  if (s == "parallel_graphflow" || s == "graphflow") *out = ALGO_GRAPHFLOW;
  else if (s == "parallel_symbi" || s == "symbi") *out = ALGO_SYMBI;
  else if (s == "parallel_turboflux" || s == "turboflux") *out = ALGO_TURBOFLUX;
  else if (s == "none") *out = ALGO_NONE;  // the base plugin: updates only, no search
  else return CSM_E_UNKNOWN_ALGO;
  return CSM_OK;
}

// -------------------------------------------------------------- planning
int build_query(csm_engine* e, const csm_graph* q, std::unique_ptr<QueryInfo>& out) {
  if (!q || q->num_vertices == 0) return fail(e, CSM_E_ARG, "empty query graph");
  if (q->num_vertices > (uint32_t)MAXK)
    return fail(e, CSM_E_CAPACITY, "query has %u vertices; the search supports up to %d",
                q->num_vertices, MAXK);
  auto qi = std::make_unique<QueryInfo>();
  qi->g.init(q->num_vertices, q->vlabels);
  for (size_t i = 0; i < q->num_edges; ++i) qi->g.add_edge(q->src[i], q->dst[i], q->elabels[i]);
  if (!qi->g.connected()) return fail(e, CSM_E_ARG, "query graph is not connected");
  e->max_k = std::max(e->max_k, qi->g.k);
  out = std::move(qi);
  return CSM_OK;
}

// Dense label maps must cover the query's labels; data labels are fixed at
// create time, so a query label absent from the data can never match: it is
// mapped to a label no data vertex carries.
int turboflux_paths(csm_engine* e, QueryInfo& q, std::vector<uint32_t>& paths);

int plan_query(csm_engine* e, QueryInfo& qi, uint32_t qid) {
  const QueryGraph& g = qi.g;
  qi.dlab.resize(g.k);
  bool labels_known = true;
  for (uint32_t u = 0; u < g.k; ++u) {
    uint32_t d = dense_of(e->vl_map, g.vl[u]);
    qi.dlab[u] = d;
    labels_known = labels_known && d != NO_LABEL;
    for (uint32_t w : g.nb[u]) labels_known = labels_known && dense_of(e->el_map, g.elabel(u, w)) != NO_LABEL;
  }
  qi.tf_paths.clear();
  if (e->opts.reference_orders && e->algo == ALGO_TURBOFLUX && qi.sx_built && labels_known) {
    // TurboFlux's own rule, ranked by the path counts of its index
    int rc = turboflux_paths(e, qi, qi.tf_paths);
    if (rc) return rc;
    qi.orders = turboflux_orders(g, qi.dag, qi.tf_paths);
  } else if (!e->opts.reference_orders) {
    auto el_of = [&](uint32_t a, uint32_t b) { return dense_of(e->el_map, g.elabel(a, b)); };
    auto triple = [&](uint32_t a, uint32_t b) -> double {
      uint32_t l = el_of(a, b);
      if (qi.dlab[a] == NO_LABEL || qi.dlab[b] == NO_LABEL || l == NO_LABEL) return 0.0;
      return (double)e->h_triple_freq[(qi.dlab[a] * e->nlab + qi.dlab[b]) * e->nel + l];
    };
    auto vf = [&](uint32_t a) -> double {
      return qi.dlab[a] == NO_LABEL ? 0.0 : (double)e->h_vfreq[qi.dlab[a]];
    };
    // fanout(w, u): directed (L(w) -> L(u), l) entries per vertex of label L(w)
    auto fanout = [&](uint32_t w, uint32_t u) { double n = vf(w); return n > 0 ? triple(w, u) / n : 0.0; };
    // hit(w, u): chance that a candidate of u is adjacent to w's match
    auto hit = [&](uint32_t w, uint32_t u) { double n = vf(u); return n > 0 ? fanout(w, u) / n : 0.0; };
    qi.orders = costed_orders(g, fanout, hit);
  } else {
    qi.orders = graphflow_orders(g);
  }
  qi.plan_base = (uint32_t)e->h_plans.size();
  for (size_t oi = 0; oi < qi.orders.size(); ++oi) {
    const auto& o = qi.orders[oi];
    PlanRec P;
    std::memset(&P, 0, sizeof P);
    P.k = g.k;
    P.q = qid;
    P.symbi = uses_index(e) ? 1u : 0u;
    std::vector<uint32_t> posof(g.k, 0);
    for (uint32_t p = 0; p < g.k; ++p) posof[o[p]] = p;
    bool matchable = true;
    for (uint32_t p = 0; p < g.k; ++p) {
      uint32_t u = o[p];
      P.qv[p] = u;
      P.d2row[p] = qi.d2_base + u;
      uint32_t nb = 0;
      for (size_t j = 0; j < g.nb[u].size(); ++j) {
        uint32_t w = g.nb[u][j];
        if (posof[w] >= p) continue;
        uint32_t el = dense_of(e->el_map, g.el[u][j]);
        if (qi.dlab[u] == NO_LABEL || el == NO_LABEL) { matchable = false; el = 0; }
        uint32_t key = (qi.dlab[u] == NO_LABEL ? 0 : qi.dlab[u]) * e->nel + el;
        P.bw[p][nb++] = posof[w] | (key << 8);
      }
      P.nb[p] = nb;
    }
    if (!matchable) P.k = 0;  // never seeded (see the seed table below)
    e->h_plans.push_back(P);
  }
  return CSM_OK;
}

// Seed table: bucket (L(v1), L(v2), l) -> plans whose first query edge
// matches in that orientation (the per-edge checks of graphflow.cpp:300-352).
void build_seed_table(csm_engine* e) {
  const uint32_t nb = e->nlab * e->nlab * e->nel;
  std::vector<std::vector<uint32_t>> buckets(nb);
  for (auto& qp : e->queries) {
    if (e->algo == ALGO_NONE) break;
    QueryInfo& qi = *qp;
    for (size_t oi = 0; oi < qi.orders.size(); ++oi) {
      uint32_t qo = qi.plan_base + (uint32_t)oi;
      if (e->h_plans[qo].k == 0) continue;
      uint32_t ua = qi.orders[oi][0], ub = qi.orders[oi][1];
      uint32_t la = qi.dlab[ua], lb = qi.dlab[ub];
      uint32_t el = dense_of(e->el_map, qi.g.elabel(ua, ub));
      if (la == NO_LABEL || lb == NO_LABEL || el == NO_LABEL) continue;
      buckets[(la * e->nlab + lb) * e->nel + el].push_back(qo << 1);
      buckets[(lb * e->nlab + la) * e->nel + el].push_back((qo << 1) | 1u);
    }
  }
  e->h_tab_off.assign(nb + 1, 0);
  e->h_tab_ent.clear();
  for (uint32_t b = 0; b < nb; ++b) {
    e->h_tab_off[b] = (uint32_t)e->h_tab_ent.size();
    e->h_tab_ent.insert(e->h_tab_ent.end(), buckets[b].begin(), buckets[b].end());
  }
  e->h_tab_off[nb] = (uint32_t)e->h_tab_ent.size();
}

// Every plan takes the search settings of its query's size class.
void apply_plan_settings(csm_engine* e) {
  for (auto& qp : e->queries) {
    const int c = size_class(qp->g.k);
    for (size_t oi = 0; oi < qp->orders.size(); ++oi) {
      PlanRec& P = e->h_plans[qp->plan_base + oi];
      P.chunk = (uint16_t)std::min<uint32_t>(e->cls_chunk[c], 0xFFFFu);
      P.budget = (uint16_t)std::min<uint32_t>(e->cls_budget[c], 0xFFFFu);
    }
  }
}

// Per-query tables of the search-time order rule (k_search_dyn).
int upload_dyn(csm_engine* e) {
  std::vector<DynQuery> h(e->queries.size());
  for (size_t qi = 0; qi < e->queries.size(); ++qi) {
    const QueryInfo& q = *e->queries[qi];
    DynQuery& D = h[qi];
    std::memset(&D, 0, sizeof D);
    const uint32_t k = q.g.k;
    D.k = k;
    const bool have_rows = q.n2row.size() == (size_t)k * k && q.dag.parents.size() == k;
    for (uint32_t u = 0; u < k; ++u) {
      D.deg[u] = (uint32_t)q.g.nb[u].size();
      D.d2row[u] = q.d2_base + u;
      for (uint32_t w : q.g.nb[u]) {
        D.nbm[u] |= 1u << w;
        const uint32_t el = dense_of(e->el_map, q.g.elabel(w, u));
        const uint32_t lu = q.dlab[u];
        D.key[w][u] = (lu == NO_LABEL ? 0 : lu) * e->nel + (el == NO_LABEL ? 0 : el);
        D.n2[w][u] = have_rows ? q.n2row[(size_t)w * k + u] : 0u;
      }
      if (have_rows)
        for (uint32_t p : q.dag.parents[u]) D.parm[u] |= 1u << p;
    }
  }
  if (!e->d_dynq || h.size() > e->d_dynq_cap) {
    dfree(e, e->d_dynq);
    const size_t cap = std::max<size_t>(8, h.size() * 2);
    int rc = dalloc(e, &e->d_dynq, cap);
    if (rc) return rc;
    e->d_dynq_cap = cap;
  }
  if (!h.empty())
    HIP_TRY(e, hipMemcpyAsync(e->d_dynq, h.data(), h.size() * sizeof(DynQuery),
                              hipMemcpyHostToDevice, e->stream));
  return CSM_OK;
}

int upload_plans(csm_engine* e) {
  apply_plan_settings(e);
  if (e->dyn) {
    int rc = upload_dyn(e);
    if (rc) return rc;
  }
  if (e->h_plans.size() > e->d_plans_cap) {
    dfree(e, e->d_plans);
    size_t cap = std::max<size_t>(64, e->h_plans.size() * 2);
    int rc = dalloc(e, &e->d_plans, cap);
    if (rc) return rc;
    e->d_plans_cap = cap;
  }
  if (!e->h_plans.empty())
    HIP_TRY(e, hipMemcpyAsync(e->d_plans, e->h_plans.data(), e->h_plans.size() * sizeof(PlanRec),
                              hipMemcpyHostToDevice, e->stream));
  if (!e->d_tab_off) {
    int rc = dalloc(e, &e->d_tab_off, e->h_tab_off.size());
    if (rc) return rc;
  }
  if (e->h_tab_ent.size() > e->d_tab_ent_cap) {
    dfree(e, e->d_tab_ent);
    size_t cap = std::max<size_t>(256, e->h_tab_ent.size() * 2);
    int rc = dalloc(e, &e->d_tab_ent, cap);
    if (rc) return rc;
    e->d_tab_ent_cap = cap;
  }
  HIP_TRY(e, hipMemcpyAsync(e->d_tab_off, e->h_tab_off.data(), e->h_tab_off.size() * 4,
                            hipMemcpyHostToDevice, e->stream));
  if (!e->h_tab_ent.empty())
    HIP_TRY(e, hipMemcpyAsync(e->d_tab_ent, e->h_tab_ent.data(), e->h_tab_ent.size() * 4,
                              hipMemcpyHostToDevice, e->stream));
  HIP_TRY(e, hipStreamSynchronize(e->stream));
  return CSM_OK;
}

// --------------------------------------------------------------- sorting
int sort_keys(csm_engine* e, uint64_t* in, uint64_t* out, uint32_t* vin, uint32_t* vout,
              size_t n) {
  size_t need = 0;
  hipError_t st;
  if (vin)
    st = rocprim::radix_sort_pairs(nullptr, need, in, out, vin, vout, n, 0, 64, e->stream);
  else
    st = rocprim::radix_sort_keys(nullptr, need, in, out, n, 0, 64, e->stream);
  if (st != hipSuccess) return fail(e, CSM_E_DEVICE, "radix sort sizing failed");
  if (need > e->sort_tmp_bytes) {
    if (e->d_sort_tmp) {
      HIP_TRY(e, hipStreamSynchronize(e->stream));
      dfree(e, e->d_sort_tmp);
    }
    uint8_t* p = nullptr;
    int rc = dalloc(e, &p, need);
    if (rc) return rc;
    e->d_sort_tmp = p;
    e->sort_tmp_bytes = need;
  }
  if (vin)
    st = rocprim::radix_sort_pairs(e->d_sort_tmp, need, in, out, vin, vout, n, 0, 64, e->stream);
  else
    st = rocprim::radix_sort_keys(e->d_sort_tmp, need, in, out, n, 0, 64, e->stream);
  if (st != hipSuccess) return fail(e, CSM_E_DEVICE, "radix sort failed");
  return CSM_OK;
}

// ------------------------------------------------------------ graph build
int dir_init(csm_engine* e, uint64_t cap);

// The initial CSR from edge arrays in HBM (m edges; dense edge labels).
// Self-loops and repeated edges are dropped, the first copy of an edge kept
// (Graph::AddEdge, graph.cpp:58-93).  The arrays are freed here.
int build_graph_dev(csm_engine* e, uint32_t* d_src, uint32_t* d_dst, uint32_t* d_el, uint64_t m) {
  const uint64_t n2 = 2 * m;
  uint64_t *k1 = nullptr, *k2 = nullptr;
  uint32_t *v1 = nullptr, *v2 = nullptr, *vend = nullptr;
  int rc;
  if ((rc = dalloc(e, &k1, n2)) || (rc = dalloc(e, &k2, n2)) || (rc = dalloc(e, &v1, n2)) ||
      (rc = dalloc(e, &v2, n2)) || (rc = dalloc(e, &vend, e->nv_cap)))
    return rc;
  if (m) {
    k_build_keys<<<grid_for(m, BUILD_BS), BUILD_BS, 0, e->stream>>>(d_src, d_dst, m, k1, v1);
    if ((rc = sort_keys(e, k1, k2, v1, v2, n2))) return rc;
    k_build_pack<<<grid_for(n2, BUILD_BS), BUILD_BS, 0, e->stream>>>(k2, v2, n2, e->d_vlab, d_el,
                                                                    e->nel, e->vbits, k1);
    if ((rc = sort_keys(e, k1, k2, nullptr, nullptr, n2))) return rc;
  }
  HIP_TRY(e, hipMemsetAsync(e->d_vrec, 0, (size_t)e->nv_cap * sizeof(uint2), e->stream));
  HIP_TRY(e, hipMemsetAsync(vend, 0, (size_t)e->nv_cap * 4, e->stream));
  if (m) {
    k_build_lists<<<grid_for(n2, BUILD_BS), BUILD_BS, 0, e->stream>>>(
        k2, n2, e->d_adj[0], e->d_ats[0], e->d_vrec, vend);
    k_build_lens<<<grid_for(e->nv_cap, BUILD_BS), BUILD_BS, 0, e->stream>>>(e->d_vrec, vend,
                                                                          e->nv_cap);
  }
  HIP_TRY(e, hipStreamSynchronize(e->stream));
  // live entries = non-sentinel keys; count on the host from the list lengths
  std::vector<uint2> rec(e->nv_cap);
  HIP_TRY(e, hipMemcpy(rec.data(), e->d_vrec, rec.size() * sizeof(uint2), hipMemcpyDeviceToHost));
  uint64_t live = 0;
  for (auto& r : rec) live += r.y & LEN_MASK;
  e->num_edges = live / 2;
  std::memset(e->h_ctrl, 0, sizeof(Ctrl));
  e->h_ctrl->pool_top = (uint32_t)n2;
  e->h_ctrl->live = live;
  e->h_ctrl->err_at = 0xFFFFFFFFu;
  HIP_TRY(e, hipMemcpy(e->d_ctrl, e->h_ctrl, sizeof(Ctrl), hipMemcpyHostToDevice));
  dfree(e, d_src); dfree(e, d_dst); dfree(e, d_el);
  dfree(e, k1); dfree(e, k2); dfree(e, v1); dfree(e, v2); dfree(e, vend);
  // key directories of the long lists
  uint64_t nlong = 0;
  for (auto& r : rec) nlong += (r.y & LEN_MASK) >= e->dir_threshold;
  return dir_init(e, nlong + nlong / 4 + e->nv_cap / 64 + 1024);
}

// The initial CSR from host edge arrays.
int build_graph(csm_engine* e, const csm_graph* d) {
  const uint64_t m = d->num_edges;
  uint32_t *d_src = nullptr, *d_dst = nullptr, *d_el = nullptr;
  int rc;
  if ((rc = dalloc(e, &d_src, m)) || (rc = dalloc(e, &d_dst, m)) || (rc = dalloc(e, &d_el, m)))
    return rc;
  std::vector<uint32_t> del(m);
  for (uint64_t i = 0; i < m; ++i) {
    if (d->src[i] >= e->nv || d->dst[i] >= e->nv)
      return fail(e, CSM_E_ARG, "edge %llu references a vertex >= %u", (unsigned long long)i, e->nv);
    del[i] = dense_of(e->el_map, d->elabels[i]);
  }
  if (m) {
    HIP_TRY(e, hipMemcpyAsync(d_src, d->src, m * 4, hipMemcpyHostToDevice, e->stream));
    HIP_TRY(e, hipMemcpyAsync(d_dst, d->dst, m * 4, hipMemcpyHostToDevice, e->stream));
    HIP_TRY(e, hipMemcpyAsync(d_el, del.data(), m * 4, hipMemcpyHostToDevice, e->stream));
  }
  return build_graph_dev(e, d_src, d_dst, d_el, m);
}

// The initial CSR from graph-file records in HBM (n records): the records
// become the build's edge arrays on the device, and the label-triple
// histogram of the SymBi root choice is counted on the way.
int build_graph_records(csm_engine* e, const UpdateRec* d_rec, uint64_t n) {
  uint32_t *d_src = nullptr, *d_dst = nullptr, *d_el = nullptr;
  unsigned long long* hist = nullptr;
  const size_t nh = (size_t)e->nlab * e->nlab * e->nel;
  int rc;
  if ((rc = dalloc(e, &d_src, n)) || (rc = dalloc(e, &d_dst, n)) || (rc = dalloc(e, &d_el, n)) ||
      (rc = dalloc(e, &hist, nh)))
    return rc;
  HIP_TRY(e, hipMemsetAsync(hist, 0, nh * 8, e->stream));
  if (n)
    k_rec_edges<<<grid_for(n, 256), 256, 0, e->stream>>>(d_rec, n, e->d_vlab, e->d_elmap, e->nel,
                                                          e->nlab, d_src, d_dst, d_el, hist);
  e->h_triple_freq.assign(nh, 0);
  HIP_TRY(e, hipMemcpyAsync(e->h_triple_freq.data(), hist, nh * 8, hipMemcpyDeviceToHost,
                            e->stream));
  HIP_TRY(e, hipStreamSynchronize(e->stream));
  dfree(e, hist);
  return build_graph_dev(e, d_src, d_dst, d_el, n);
}

// Graph-file records in HBM: vertex count (largest id + 1), each vertex's
// label from its first 'v' record (NO_LABEL without one), the 'e' record
// count and the set of edge labels, all computed on the device; only the
// per-vertex labels and the label set come to the host.
int scan_records(csm_engine* e, const UpdateRec* d_rec, uint64_t n, uint32_t* nv,
                 std::vector<uint32_t>& raw_vl, uint64_t* ne, std::vector<uint32_t>& els) {
  int rc;
  uint32_t* cnt = nullptr;       // [0] vertex count, [2..3] edge count (u64)
  if ((rc = dalloc(e, &cnt, 4))) return rc;
  HIP_TRY(e, hipMemsetAsync(cnt, 0, 16, e->stream));
  if (n)
    k_rec_maxid<<<grid_for(n, 256), 256, 0, e->stream>>>(
        d_rec, n, cnt, reinterpret_cast<unsigned long long*>(cnt + 2));
  uint32_t h[4] = {0, 0, 0, 0};
  HIP_TRY(e, hipMemcpyAsync(h, cnt, 16, hipMemcpyDeviceToHost, e->stream));
  HIP_TRY(e, hipStreamSynchronize(e->stream));
  dfree(e, cnt);
  *nv = h[0];
  *ne = (uint64_t)h[2] | ((uint64_t)h[3] << 32);
  // labels from the first 'v' record of each vertex
  unsigned long long* first = nullptr;
  uint32_t* vl = nullptr;
  if ((rc = dalloc(e, &first, std::max<uint32_t>(*nv, 1))) ||
      (rc = dalloc(e, &vl, std::max<uint32_t>(*nv, 1))))
    return rc;
  HIP_TRY(e, hipMemsetAsync(first, 0xFF, (size_t)std::max<uint32_t>(*nv, 1) * 8, e->stream));
  HIP_TRY(e, hipMemsetAsync(vl, 0xFF, (size_t)std::max<uint32_t>(*nv, 1) * 4, e->stream));
  if (n) {
    k_rec_first_vertex<<<grid_for(n, 256), 256, 0, e->stream>>>(d_rec, n, first);
    k_rec_vlabels<<<grid_for(n, 256), 256, 0, e->stream>>>(d_rec, n, first, vl);
  }
  raw_vl.assign(*nv, NO_LABEL);
  if (*nv)
    HIP_TRY(e, hipMemcpyAsync(raw_vl.data(), vl, (size_t)*nv * 4, hipMemcpyDeviceToHost, e->stream));
  HIP_TRY(e, hipStreamSynchronize(e->stream));
  dfree(e, first);
  dfree(e, vl);
  // edge label set: sort the records' labels, keep the distinct ones
  els.clear();
  if (!*ne) return CSM_OK;
  uint32_t *lab = nullptr, *lab2 = nullptr, *uniq = nullptr;
  size_t* nuniq = nullptr;
  if ((rc = dalloc(e, &lab, n)) || (rc = dalloc(e, &lab2, n)) || (rc = dalloc(e, &uniq, n)) ||
      (rc = dalloc(e, &nuniq, 1)))
    return rc;
  k_rec_elabels<<<grid_for(n, 256), 256, 0, e->stream>>>(d_rec, n, lab);
  size_t need = 0, need2 = 0;
  HIP_TRY(e, rocprim::radix_sort_keys(nullptr, need, lab, lab2, (size_t)n, 0, 32, e->stream));
  HIP_TRY(e, rocprim::unique(nullptr, need2, lab2, uniq, nuniq, (size_t)n,
                             rocprim::equal_to<uint32_t>(), e->stream));
  uint8_t* tmp = nullptr;
  if ((rc = dalloc(e, &tmp, std::max(need, need2)))) return rc;
  HIP_TRY(e, rocprim::radix_sort_keys(tmp, need, lab, lab2, (size_t)n, 0, 32, e->stream));
  HIP_TRY(e, rocprim::unique(tmp, need2, lab2, uniq, nuniq, (size_t)n,
                             rocprim::equal_to<uint32_t>(), e->stream));
  size_t k = 0;
  HIP_TRY(e, hipMemcpyAsync(&k, nuniq, sizeof k, hipMemcpyDeviceToHost, e->stream));
  HIP_TRY(e, hipStreamSynchronize(e->stream));
  els.resize(k);
  if (k) HIP_TRY(e, hipMemcpy(els.data(), uniq, k * 4, hipMemcpyDeviceToHost));
  els.erase(std::remove(els.begin(), els.end(), NO_LABEL), els.end());
  dfree(e, lab); dfree(e, lab2); dfree(e, uniq); dfree(e, nuniq); dfree(e, tmp);
  return CSM_OK;
}

// Allocate the key directory pool for `cap` slots and build the directories
// of every list long enough to get one.
int dir_init(csm_engine* e, uint64_t cap) {
  int rc;
  e->dir_cap = (uint32_t)std::min<uint64_t>(cap, e->nv_cap);
  if ((rc = dalloc(e, &e->d_dslot, e->nv_cap)) ||
      (rc = dalloc(e, &e->d_dir, (size_t)e->dir_cap * e->dstride)))
    return rc;
  HIP_TRY(e, hipMemsetAsync(e->d_dslot, 0xFF, (size_t)e->nv_cap * 4, e->stream));
  k_dir_build<DIR_ALL><<<grid_for(e->nv_cap, 256 / WAVE, 8192), 256, 0, e->stream>>>(
      e->d_vrec, e->d_adj[e->cur], e->d_dslot, e->d_dir, e->dstride, e->vbits, e->dir_threshold,
      nullptr, nullptr, nullptr, e->nv_cap, e->d_ctrl, e->dir_cap);
  HIP_TRY(e, hipStreamSynchronize(e->stream));
  HIP_TRY(e, hipMemcpy(e->h_ctrl, e->d_ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost));
  return CSM_OK;
}

// Make room for the slots one batch of n updates can hand out: at most one per
// rewritten list, i.e. per merged vertex (2n) and per vertex compacted from the
// previous batch (2 per update).  Without room the extra lists keep using
// binary search, so a budget that forbids the growth is not an error.
int dir_reserve(csm_engine* e, uint32_t n) {
  const uint64_t top = e->h_ctrl->dir_top;
  const uint64_t need = top + 2ull * n + 2ull * e->last_batch + 1024;
  e->last_batch = n;
  if (need <= e->dir_cap || e->dir_cap >= e->nv_cap) return CSM_OK;
  const uint64_t ncap = std::min<uint64_t>(std::max<uint64_t>(2ull * e->dir_cap, need), e->nv_cap);
  const size_t bytes = (size_t)ncap * e->dstride * 4;
  if (e->opts.device_budget && e->dev_bytes + bytes > e->opts.device_budget) return CSM_OK;
  uint32_t* nd = nullptr;
  int rc;
  if ((rc = dalloc(e, &nd, (size_t)ncap * e->dstride))) return rc;
  HIP_TRY(e, hipMemcpyAsync(nd, e->d_dir, (size_t)std::min<uint64_t>(top, e->dir_cap) * e->dstride * 4,
                            hipMemcpyDeviceToDevice, e->stream));
  HIP_TRY(e, hipStreamSynchronize(e->stream));
  dfree(e, e->d_dir);
  e->d_dir = nd;
  e->dir_cap = (uint32_t)ncap;
  return CSM_OK;
}

// Repack every list into the other pool (drops garbage left by out-of-place
// rewrites).  Lists keep their versions and DIRTY flags.
int repack(csm_engine* e) {
  int other = 1 - e->cur;
  if (!e->d_adj[other]) {
    int rc;
    if ((rc = dalloc(e, &e->d_adj[other], e->pool_cap)) ||
        (rc = dalloc(e, &e->d_ats[other], e->pool_cap)))
      return rc;
  }
  int rc;
  if (!e->d_rp_lens) {
    if ((rc = dalloc(e, &e->d_rp_lens, e->nv_cap)) || (rc = dalloc(e, &e->d_rp_offs, e->nv_cap)))
      return rc;
    HIP_TRY(e, rocprim::exclusive_scan(nullptr, e->rp_tmp_bytes, e->d_rp_lens, e->d_rp_offs, 0u,
                                       e->nv_cap, rocprim::plus<uint32_t>(), e->stream));
    if ((rc = dalloc(e, &e->d_rp_tmp, e->rp_tmp_bytes))) return rc;
  }
  uint32_t *lens = e->d_rp_lens, *offs = e->d_rp_offs;
  k_repack_lens<<<grid_for(e->nv_cap, 256), 256, 0, e->stream>>>(e->d_vrec, lens, e->nv_cap);
  size_t need = e->rp_tmp_bytes;
  HIP_TRY(e, rocprim::exclusive_scan(e->d_rp_tmp, need, lens, offs, 0u, e->nv_cap,
                                     rocprim::plus<uint32_t>(), e->stream));
  k_repack_copy<256><<<grid_for(e->nv_cap, 16, 65536), 256, 0, e->stream>>>(
      e->d_vrec, offs, e->d_adj[e->cur], e->d_ats[e->cur], e->d_adj[other], e->d_ats[other],
      e->nv_cap);
  HIP_TRY(e, hipStreamSynchronize(e->stream));
  HIP_TRY(e, hipMemcpy(e->h_ctrl, e->d_ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost));
  e->h_ctrl->pool_top = (uint32_t)e->h_ctrl->live;
  HIP_TRY(e, hipMemcpy(e->d_ctrl, e->h_ctrl, sizeof(Ctrl), hipMemcpyHostToDevice));
  e->cur = other;
  ++e->stats[4];
  return CSM_OK;
}

int alloc_batch(csm_engine* e) {
  const uint32_t B = e->max_batch;
  int rc;
  if ((rc = dalloc(e, &e->d_upd, B)) || (rc = dalloc(e, &e->d_type, B)) ||
      (rc = dalloc(e, &e->d_add, B)) || (rc = dalloc(e, &e->d_v1, B)) ||
      (rc = dalloc(e, &e->d_v2, B)) || (rc = dalloc(e, &e->d_lab, B)) ||
      (rc = dalloc(e, &e->d_slab, B)) || (rc = dalloc(e, &e->d_keys, 2 * (size_t)B)) ||
      (rc = dalloc(e, &e->d_keys2, 2 * (size_t)B)) || (rc = dalloc(e, &e->d_vals, 2 * (size_t)B)) ||
      (rc = dalloc(e, &e->d_vals2, 2 * (size_t)B)) || (rc = dalloc(e, &e->d_runs, 2 * (size_t)B)) ||
      (rc = dalloc(e, &e->d_dirty[0], 2 * (size_t)B + 2)) ||
      (rc = dalloc(e, &e->d_dirty[1], 2 * (size_t)B + 2)) || (rc = dalloc(e, &e->d_safe, B)) ||
      (rc = dalloc(e, &e->d_ustatus, (size_t)B + 1)))
    return rc;
  HIP_TRY(e, hipMemset(e->d_ustatus, 0, ((size_t)B + 1) * 4));
  if (e->opts.want_work) {
    if ((rc = dalloc(e, &e->d_uwork, (size_t)B + 1))) return rc;
    HIP_TRY(e, hipMemset(e->d_uwork, 0, ((size_t)B + 1) * 8));
    HIP_TRY(e, hipHostMalloc((void**)&e->h_uwork_pin, ((size_t)B + 1) * 8));
  }
  e->seed_cap = std::max<uint32_t>(1u << 16, std::min<uint64_t>(16ull * B, 1u << 22));
  e->chunk_cap = 1u << 26;
  // spill buffers: each XCD writes into its own eighth, and a heavy subtree
  // keeps spilling into one region (2^22 tasks: 512k per region)
  e->spill_cap = 1u << 22;
  e->spill_chunk_cap = 1u << 25;
  if ((rc = dalloc(e, &e->d_seeds, e->seed_cap)) ||
      (rc = dalloc(e, &e->d_chunk_owner, e->chunk_cap)) ||
      (rc = dalloc(e, &e->d_spill[0], e->spill_cap)) ||
      (rc = dalloc(e, &e->d_spill[1], e->spill_cap)) ||
      (rc = dalloc(e, &e->d_spill_owner[0], e->spill_chunk_cap)) ||
      (rc = dalloc(e, &e->d_spill_owner[1], e->spill_chunk_cap)))
    return rc;
  HIP_TRY(e, hipHostMalloc((void**)&e->h_upd, (size_t)B * sizeof(UpdateRec)));
  HIP_TRY(e, hipHostMalloc((void**)&e->h_safe, B));
  HIP_TRY(e, hipHostMalloc((void**)&e->h_ustatus_pin, ((size_t)B + 1) * 4));
  // the abort word: coherent (fine-grained) pinned memory mapped into the
  // device address space, so that a running kernel sees the host's store
  HIP_TRY(e, hipHostMalloc((void**)&e->h_abort, 64, hipHostMallocMapped | hipHostMallocCoherent));
  *e->h_abort = 0;
  HIP_TRY(e, hipHostGetDevicePointer((void**)&e->d_abort_map, e->h_abort, 0));
  return CSM_OK;
}

// The search's SearchLimits: the addresses of the time-limit word, the
// status, work and per-query arrays, and the result cap (all fixed between
// reallocations, which call this again).
int upload_limits(csm_engine* e) {
  if (!e->d_limits) {
    int rc = dalloc(e, &e->d_limits, 1);
    if (rc) return rc;
  }
  SearchLimits l;
  l.habort = e->d_abort_map;
  l.ustatus = e->d_ustatus;
  l.qtl = e->d_qtl;
  l.uwork = e->d_uwork;
  l.ucap = e->opts.max_results;
  l.emb = e->d_emb;
  l.emb_n = e->d_emb_n;
  l.emb_cap = e->emb_cap;
  l.emb_w = EMB_W;
  HIP_TRY(e, hipMemcpy(e->d_limits, &l, sizeof l, hipMemcpyHostToDevice));
  return CSM_OK;
}

int ensure_results(csm_engine* e) {
  size_t need = (size_t)e->max_batch * std::max<size_t>(1, e->queries.size());
  static_assert(sizeof(unsigned long long) == 8, "u64");
  if (e->d_counts) {
    // capacity tracked via allocation size
    bool big = false;
    for (auto& a : e->allocs)
      if (a.p == e->d_counts && a.bytes >= need * 8) big = true;
    if (big && e->queries.size() <= e->q_cap) {
      e->h_qabort.assign(e->queries.size(), 0);
      e->h_qtl.assign(e->queries.size(), 0);
      return CSM_OK;
    }
    dfree(e, e->d_counts);
    dfree(e, e->d_hashes);
    if (e->h_counts) (void)hipHostFree(e->h_counts);
    if (e->h_hashes) (void)hipHostFree(e->h_hashes);
    e->h_counts = e->h_hashes = nullptr;
  }
  int rc;
  if ((rc = dalloc(e, &e->d_counts, need)) || (rc = dalloc(e, &e->d_hashes, need))) return rc;
  if (e->queries.size() > e->q_cap) {
    dfree(e, e->d_qcount);
    dfree(e, e->d_qabort);
    dfree(e, e->d_qtl);
    e->q_cap = std::max<size_t>(64, e->queries.size() * 2);
    if ((rc = dalloc(e, &e->d_qcount, e->q_cap)) || (rc = dalloc(e, &e->d_qabort, e->q_cap)) ||
        (rc = dalloc(e, &e->d_qtl, e->q_cap)))
      return rc;
    HIP_TRY(e, hipMemset(e->d_qtl, 0, e->q_cap * 4));
    HIP_TRY(e, hipMemset(e->d_qcount, 0, e->q_cap * 8));
    HIP_TRY(e, hipMemset(e->d_qabort, 0, e->q_cap * 4));
  }
  e->h_qabort.assign(e->queries.size(), 0);
  e->h_qtl.assign(e->queries.size(), 0);
  HIP_TRY(e, hipHostMalloc((void**)&e->h_counts, need * 8));
  HIP_TRY(e, hipHostMalloc((void**)&e->h_hashes, need * 8));
  if (e->opts.print_enum && !e->d_emb) {
    e->emb_cap = e->opts.enum_cap ? e->opts.enum_cap : EMB_DEFAULT_ROWS;
    if ((rc = dalloc(e, &e->d_emb, e->emb_cap * EMB_W)) || (rc = dalloc(e, &e->d_emb_n, 1))) return rc;
  }
  return upload_limits(e);
}

int read_ctrl(csm_engine* e) {
  HIP_TRY(e, hipMemcpyAsync(e->h_ctrl, e->d_ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, e->stream));
  if (e->deadline > 0) {
    // under a time limit the host watches the clock while the device works
    while (true) {
      hipError_t st = hipStreamQuery(e->stream);
      if (st == hipSuccess) break;
      if (st != hipErrorNotReady)
        return fail(e, CSM_E_DEVICE, "hipStreamQuery failed: %s", hipGetErrorString(st));
      check_deadline(e);
      std::this_thread::sleep_for(std::chrono::microseconds(50));
    }
    return CSM_OK;
  }
  HIP_TRY(e, hipStreamSynchronize(e->stream));
  return CSM_OK;
}

BatchView batch_view(const csm_engine* e, uint32_t n) {
  BatchView b;
  b.type = e->d_type;
  b.add = e->d_add;
  b.v1 = e->d_v1;
  b.v2 = e->d_v2;
  b.label = e->d_lab;
  b.n = n;
  return b;
}

template <int MK, bool H, bool S>
void launch_search_mk(csm_engine* e, const SearchArgs& sa) {
  static int per_cu = -1;
  if (per_cu < 0) {
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_search<MK, true, H, S>,
                                                     search_waves<MK>() * WAVE, 0) != hipSuccess ||
        per_cu < 1)
      per_cu = 2;
    if (getenv("CSM_TRACE_ROUNDS"))
      fprintf(stderr, "[csm] search kernel MK=%d: %d workgroups per CU, %d CUs\n", MK, per_cu,
              e->num_cus);
  }
  k_search<MK, true, H, S><<<e->num_cus * per_cu, search_waves<MK>() * WAVE, 0, e->stream>>>(sa);
}

// the smallest LDS state that holds the largest registered query
template <bool H, bool S>
void launch_search(csm_engine* e, const SearchArgs& sa) {
  if (e->max_k <= 8) launch_search_mk<8, H, S>(e, sa);
  else if (e->max_k <= 10) launch_search_mk<10, H, S>(e, sa);
  else if (e->max_k <= 12) launch_search_mk<12, H, S>(e, sa);
  else if (e->max_k <= 16) launch_search_mk<16, H, S>(e, sa);
  else if (e->max_k <= 24) launch_search_mk<24, H, S>(e, sa);
  else launch_search_mk<MAXK, H, S>(e, sa);
}

// One search launch in slot `slot` of a group of rounds, between that slot's
// two events.
template <int MK, bool H>
void launch_dyn_mk(csm_engine* e, const DynArgs& da) {
  static int per_cu = -1;
  if (per_cu < 0) {
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_search_dyn<MK, H>,
                                                     search_waves<MK>() * WAVE, 0) != hipSuccess ||
        per_cu < 1)
      per_cu = 1;
  }
  k_search_dyn<MK, H><<<e->num_cus * per_cu, search_waves<MK>() * WAVE, 0, e->stream>>>(da);
}

// the search-time order rule (parallel_symbi with reference orders)
template <bool H>
void launch_dyn(csm_engine* e, const SearchArgs& sa) {
  DynArgs da;
  da.s = sa;
  da.dq = e->d_dynq;
  da.n2 = e->d_sx_n2;
  da.n2_stride = e->nv_cap;
  if (e->max_k <= 8) launch_dyn_mk<8, H>(e, da);
  else if (e->max_k <= 16) launch_dyn_mk<16, H>(e, da);
  else launch_dyn_mk<MAXK, H>(e, da);
}

void launch_round(csm_engine* e, SearchArgs& sa, bool h, uint32_t slot) {
  const bool s = uses_index(e) && e->d_d2 != nullptr;
  sa.round_slot = slot;
  (void)hipEventRecord(e->ev_round[slot][0], e->stream);
  if (e->dyn && s) {
    if (h) launch_dyn<true>(e, sa); else launch_dyn<false>(e, sa);
  } else if (h) {
    if (s) launch_search<true, true>(e, sa); else launch_search<true, false>(e, sa);
  } else {
    if (s) launch_search<false, true>(e, sa); else launch_search<false, false>(e, sa);
  }
  (void)hipEventRecord(e->ev_round[slot][1], e->stream);
}

int read_ctrl(csm_engine* e);

// The search arguments of one batch (round buffers are set per round).
SearchArgs search_args(csm_engine* e, unsigned long long* counts, unsigned long long* hashes,
                       uint32_t pos_offset) {
  SearchArgs sa;
  sa.g = view(e);
  sa.plans = e->d_plans;
  sa.ctrl = e->d_ctrl;
  sa.round_slot = 0;
  sa.counts = counts;
  sa.hashes = hashes;
  sa.d2 = e->d_d2;
  sa.d2_words = e->d2_words;
  sa.nq = (uint32_t)e->queries.size();
  sa.homo = e->opts.homo;
  sa.pos_offset = pos_offset;
  sa.qcount = e->opts.match_budget ? e->d_qcount : nullptr;
  sa.qabort = (e->opts.match_budget || e->tune_mask) ? e->d_qabort : nullptr;
  sa.budget = e->opts.match_budget;
  sa.spill_cap = e->spill_cap;
  sa.spill_chunk_cap = e->spill_chunk_cap;
  sa.lflags = (e->deadline > 0 ? LIM_TIME : 0u) | (e->opts.max_results ? LIM_CAP : 0u) |
              (e->opts.want_work ? LIM_WORK : 0u) | (e->d_emb ? LIM_ENUM : 0u);
  sa.lim = e->d_limits;
  // spill budget for the batch's size: a quarter up to 2,048 updates, half
  // up to 8,192 (served batches: shorter rounds, lower latency; measured on
  // the config 2 workload)
  sa.budget_shift = e->batch_n <= 2048 ? 2u : e->batch_n <= 8192 ? 1u : 0u;
  return sa;
}

// Round r reads the tasks round r-1 spilled (round 0: the seeds) and spills
// into the other buffer of the ping-pong pair.
void round_buffers(csm_engine* e, SearchArgs& sa, uint32_t round) {
  if (round == 0) {
    sa.tasks = e->d_seeds;
    sa.chunk_owner = e->d_chunk_owner;
    sa.chunk_cap = e->chunk_cap;
  } else {
    const int in = (int)((round - 1) & 1u);
    sa.tasks = e->d_spill[in];
    sa.chunk_owner = e->d_spill_owner[in];
    sa.chunk_cap = e->spill_chunk_cap;
  }
  const int out = (int)(round & 1u);
  sa.spill = e->d_spill[out];
  sa.spill_owner = e->d_spill_owner[out];
  sa.xregions = round > 0 ? 1u : 0u;
}

// Search rounds from `first` on: a task that exceeds its iteration budget
// spills the rest of its DFS into the next round's task list, and rounds
// repeat until nothing spills.  The round's chunk counts are in place on the
// device (k_seeds for round 0, k_next_round after each round), so rounds are
// enqueued back to back, as many as the deepest batch so far needed (a round
// with nothing to do ends at once), with one read-back of Ctrl per group; a
// batch that needs more continues with another group.
int search(csm_engine* e, bool versioned, unsigned long long* counts, unsigned long long* hashes,
           uint32_t pos_offset = 0, uint32_t first = 0) {
  (void)versioned;  // every list carries version words; clean lists skip them
  SearchArgs sa = search_args(e, counts, hashes, pos_offset);
  const bool h = hashes != nullptr;
  static const bool trace = getenv("CSM_TRACE_ROUNDS") != nullptr;
  for (uint32_t round = first;;) {
    const uint32_t want = e->max_rounds > round ? e->max_rounds - round : 1u;
    const uint32_t group = trace ? 1u : std::min<uint32_t>(want, ROUND_GROUP_MAX);
    const uint32_t chunks_in = trace ? e->h_ctrl->nchunks : 0;
    check_deadline(e);   // after the limit, rounds only drain their chunks
    k_fill_u32<<<1, 64, 0, e->stream>>>(e->d_ctrl->round_busy, ROUND_GROUP_MAX, 0u);
    for (uint32_t i = 0; i < group; ++i) {
      round_buffers(e, sa, round + i);
      launch_round(e, sa, h, i);
      k_next_round<<<1, 64, 0, e->stream>>>(e->d_ctrl, e->spill_chunk_cap);
    }
    int rc = read_ctrl(e);
    if (rc) return rc;
    // the search time of the launches that ran a task
    uint32_t busy = 0;
    uint64_t ns_group = 0;
    for (uint32_t i = 0; i < group; ++i) {
      if (!e->h_ctrl->round_busy[i]) continue;
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, e->ev_round[i][0], e->ev_round[i][1]) == hipSuccess) {
        e->stats[5] += (uint64_t)(ms * 1e6);
        e->stats[6] += 1;
        ns_group += (uint64_t)(ms * 1e6);
      }
      busy = i + 1;
    }
    if (trace) {
      fprintf(stderr, "[csm] round %u: %s %u chunks, %.3f ms, spilled %u chunks\n", round,
              round ? "spilled" : "seed", chunks_in, ns_group / 1e6, e->h_ctrl->nchunks);
      unsigned long long& life0 = e->trace_life0;
      unsigned long long& waves0 = e->trace_waves0;
      const Ctrl* c = e->h_ctrl;
      if (c->prof[29] > waves0 && c->prof[31] > c->prof[30]) {
        const double span = (double)(c->prof[31] - c->prof[30]);
        const double nw = (double)(c->prof[29] - waves0);
        fprintf(stderr, "[csm]   %.0f waves, span %.3f ms, busy fraction %.3f\n", nw, span * 1e-5,
                (double)(c->prof[28] - life0) / (span * nw));
        life0 = c->prof[28];
        waves0 = c->prof[29];
      }
    }
    round += group;
    if (e->h_ctrl->nchunks == 0) {   // the group's last round spilled nothing
      e->max_rounds = std::max(e->max_rounds, round - group + busy);
      if (trace && e->h_ctrl->prof[8]) {
        static const char* names[8] = {"fetch", "enter", "prepare_last", "refill",
                                       "refill_k-2", "leaves_shared", "leaves_own", "spill"};
        for (int i = 0; i < 8; ++i)
          fprintf(stderr, "[csm] phase %-13s %14llu cycles %12llu calls\n", names[i],
                  e->h_ctrl->prof[i], e->h_ctrl->prof[8 + i]);
        for (int i = 16; i < 32; ++i)
          fprintf(stderr, "[csm] counter %2d %16llu\n", i - 16, e->h_ctrl->prof[i]);
      }
      return CSM_OK;
    }
    if (round > 100000) return fail(e, CSM_E_DEVICE, "search did not converge");
  }
}

void seeds(csm_engine* e, uint32_t n, unsigned long long* counts, unsigned long long* hashes,
           uint8_t* safe, bool versioned) {
  SeedArgs a;
  a.g = view(e);
  a.b = batch_view(e, n);
  a.seed_label = e->d_slab;
  a.tab_off = e->d_tab_off;
  a.tab_ent = e->d_tab_ent;
  a.plans = e->d_plans;
  a.d2 = uses_index(e) ? e->d_d2 : nullptr;
  a.d2_words = e->d2_words;
  a.nlab = e->nlab;
  a.nq = (uint32_t)e->queries.size();
  a.versioned = versioned;
  a.homo = e->opts.homo;
  a.want_hash = hashes != nullptr;
  a.seeds = e->d_seeds;
  a.seed_cap = e->seed_cap;
  a.chunk_owner = e->d_chunk_owner;
  a.chunk_cap = e->chunk_cap;
  a.ctrl = e->d_ctrl;
  a.counts = counts;
  a.hashes = hashes;
  a.safe = safe;
  a.ustatus = e->d_ustatus;
  a.lim = e->d_limits;
  a.enumerate = (hashes && e->d_emb) ? 1u : 0u;
  a.dyn = e->dyn ? 1u : 0u;
  k_seeds<<<grid_for((uint64_t)n * WAVE, 256, 8192), 256, 0, e->stream>>>(a);
}

int check_errors(csm_engine* e) {
  uint32_t err = e->h_ctrl->err;
  if (!err) return CSM_OK;
  if (err & ERR_DELETE_MISSING)
    return fail(e, CSM_E_DELETE_MISSING, "deletion error (update %u)", e->h_ctrl->err_at);
  if (err & ERR_VERTEX_UPDATE)
    return fail(e, CSM_E_UNSUPPORTED,
                "vertex record at %u in a device batch; vertex updates go through "
                "csm_apply_batch, csm_add_vertex or csm_remove_vertex", e->h_ctrl->err_at);
  if (err & ERR_BAD_UPDATE)
    return fail(e, CSM_E_ARG, "invalid update (unknown vertex or label, or self-loop) at %u",
                e->h_ctrl->err_at);
  if (err & (ERR_SEED_OVERFLOW | ERR_CHUNK_OVERFLOW))
    return fail(e, CSM_E_CAPACITY, "search task capacity exceeded; use smaller batches");
  if (err & ERR_POOL_OVERFLOW) return fail(e, CSM_E_CAPACITY, "adjacency pool exhausted");
  if (err & ERR_INDEX_OVERFLOW)
    return fail(e, CSM_E_CAPACITY, "SymBi index propagation exceeded its frontier capacity");
  return fail(e, CSM_E_DEVICE, "device error bits 0x%x", err);
}

int symbi_prepare_query(csm_engine* e, QueryInfo& q);
int symbi_batch_insert(csm_engine* e, uint32_t n);
int symbi_batch_delete(csm_engine* e, uint32_t n);
int symbi_add_vertex(csm_engine* e, uint32_t id);

// The batch pipeline; k_batch_begin has run and the updates are unpacked into
// d_type/d_add/d_v1/d_v2/d_lab.  With search == false the graph and the SymBi
// index are updated but no match is searched (vertex deletion).
int dir_reserve(csm_engine* e, uint32_t n);
int tune_batch(csm_engine* e, uint32_t n, unsigned long long* counts, unsigned long long* hashes,
               uint8_t* safe);

// Host decisions before a batch: room in the adjacency pool (repack) and in
// the key-directory pool.
int batch_prepare(csm_engine* e, uint32_t n) {
  int rc;
  // room for this batch's rewrites: at most every list twice plus the new entries
  uint64_t live = e->h_ctrl->live;
  if ((uint64_t)e->h_ctrl->pool_top + 2 * live + 2ull * n + 1024 > e->pool_cap) {
    if ((rc = repack(e))) return rc;
    if ((uint64_t)e->h_ctrl->pool_top + 2 * live + 2ull * n + 1024 > e->pool_cap)
      return fail(e, CSM_E_CAPACITY, "adjacency pool too small");
  }
  return dir_reserve(e, n);
}

// Zero a batch's result arrays with fill kernels (no memset, so that a
// captured graph replays them exactly).
void batch_clear(csm_engine* e, uint32_t n, unsigned long long* counts, unsigned long long* hashes) {
  const size_t nq = e->queries.size();
  const uint64_t w = 2ull * n * nq;   // u32 words of n x nq u64
  k_fill_u32<<<grid_for(w, 256), 256, 0, e->stream>>>(reinterpret_cast<uint32_t*>(counts), w, 0u);
  if (hashes)
    k_fill_u32<<<grid_for(w, 256), 256, 0, e->stream>>>(reinterpret_cast<uint32_t*>(hashes), w, 0u);
  k_fill_u32<<<grid_for(n, 256), 256, 0, e->stream>>>(e->d_ustatus, n, 0u);
  if (e->d_emb_n)
    k_fill_u32<<<1, 64, 0, e->stream>>>(reinterpret_cast<uint32_t*>(e->d_emb_n), 2, 0u);
  if (e->d_uwork)
    k_fill_u32<<<grid_for(2ull * n, 256), 256, 0, e->stream>>>(
        reinterpret_cast<uint32_t*>(e->d_uwork), 2ull * n, 0u);
  if (e->opts.match_budget) {
    k_fill_u32<<<1, 256, 0, e->stream>>>(reinterpret_cast<uint32_t*>(e->d_qcount), 2 * nq, 0u);
    k_fill_u32<<<1, 256, 0, e->stream>>>(e->d_qabort, nq, 0u);
  }
}

// The graph and index update of a batch whose records are unpacked.
int batch_update(csm_engine* e, uint32_t n) {
  int rc;
  const size_t nq = e->queries.size();
  k_fill_u32<<<1, 256, 0, e->stream>>>(e->d_qtl, nq, 0u);
  // 1. compact lists dirtied by the previous batch
  k_compact<APPLY_BS><<<2048, APPLY_BS, 0, e->stream>>>(
      e->d_vrec, e->d_adj[e->cur], e->d_ats[e->cur], e->d_dirty[e->dcur], &e->d_ctrl->ndirty_cur,
      e->d_ctrl, (uint32_t)e->pool_cap);
  // 2. merge insertions
  BatchView b = batch_view(e, n);
  k_batch_keys<<<grid_for(n, 256), 256, 0, e->stream>>>(b, e->d_vlab, e->nel, e->vbits, e->nv,
                                                        e->d_keys, e->d_vals, &e->d_ctrl->err,
                                                        &e->d_ctrl->err_at);
  if ((rc = sort_keys(e, e->d_keys, e->d_keys2, e->d_vals, e->d_vals2, 2 * (size_t)n))) return rc;
  k_runs<<<grid_for(2 * n, 256), 256, 0, e->stream>>>(e->d_keys2, 2 * n, e->d_runs,
                                                      &e->d_ctrl->nruns);
  k_merge<APPLY_BS><<<2048, APPLY_BS, 0, e->stream>>>(
      e->d_vrec, e->d_adj[e->cur], e->d_ats[e->cur], e->d_keys2, e->d_vals2, 2 * n, e->d_runs,
      e->d_ctrl, e->d_dirty[1 - e->dcur], (uint32_t)e->pool_cap);
  // key directories of the lists rewritten by compaction and merge
  k_dir_build<DIR_LIST><<<1024, 256, 0, e->stream>>>(
      e->d_vrec, e->d_adj[e->cur], e->d_dslot, e->d_dir, e->dstride, e->vbits, e->dir_threshold,
      e->d_dirty[e->dcur], &e->d_ctrl->ndirty_cur, nullptr, 0, e->d_ctrl, e->dir_cap);
  k_dir_build<DIR_RUNS><<<1024, 256, 0, e->stream>>>(
      e->d_vrec, e->d_adj[e->cur], e->d_dslot, e->d_dir, e->dstride, e->vbits, e->dir_threshold,
      e->d_runs, nullptr, e->d_keys2, 0, e->d_ctrl, e->dir_cap);
  // 3. deletions and duplicate insertions
  GraphView g = view(e);
  k_delete_mark<<<grid_for(n, 256), 256, 0, e->stream>>>(b, g, e->d_ats[e->cur], e->d_vrec,
                                                         e->d_slab, e->d_ctrl,
                                                         e->d_dirty[1 - e->dcur]);
  k_dedup<<<grid_for(n, 256), 256, 0, e->stream>>>(b, g, e->d_ats[e->cur]);
  // 4. SymBi index: insertions (before the search) ...
  if (uses_index(e) && e->d_d2)
    if ((rc = symbi_batch_insert(e, n))) return rc;
  return CSM_OK;
}

// After the search: the result cap, the index's deletions, the batch's end.
int batch_tail(csm_engine* e, uint32_t n, unsigned long long* counts, bool searched) {
  int rc;
  const size_t nq = e->queries.size();
  if (searched && e->opts.max_results)
    k_clamp_counts<<<grid_for((uint64_t)n * nq, 256), 256, 0, e->stream>>>(
        counts, n, (uint32_t)nq, e->opts.max_results, e->d_ustatus);
  // 6. ... and deletions (after it)
  if (uses_index(e) && e->d_d2)
    if ((rc = symbi_batch_delete(e, n))) return rc;
  // next batch compacts what this one dirtied
  k_batch_end<<<1, 64, 0, e->stream>>>(e->d_ctrl);
  e->dcur = 1 - e->dcur;
  ++e->stats[3];
  return CSM_OK;
}

// The batch pipeline; k_batch_begin has run and the updates are unpacked into
// d_type/d_add/d_v1/d_v2/d_lab.  With search_matches == false the graph and
// the SymBi index are updated but no match is searched (vertex deletion).
int run_batch(csm_engine* e, uint32_t n, unsigned long long* counts, unsigned long long* hashes,
              uint8_t* safe, bool search_matches = true) {
  int rc;
  if ((rc = batch_prepare(e, n))) return rc;
  batch_clear(e, n, counts, hashes);
  if ((rc = batch_update(e, n))) return rc;
  // 5. seeds + search
  if (search_matches) {
    if (e->tune_active && n >= TUNE_MIN_BATCH && e->tune_idx > 0) {
      if ((rc = tune_batch(e, n, counts, hashes, safe))) return rc;
    } else {
      seeds(e, n, counts, hashes, safe, true);
      if ((rc = search(e, true, counts, hashes))) return rc;
      if (e->tune_active && n >= TUNE_MIN_BATCH) ++e->tune_idx;   // the settling batch
    }
  }
  return batch_tail(e, n, counts, search_matches);
}

// ------------------------------------------------------- batches as graphs
// Small batches (up to GRAPH_MAX_BATCH updates: the plain per-update path,
// serving at low and moderate load), where launch and host round-trip
// latency weigh most, run as a HIP graph: k_batch_begin, unpack, the graph
// and index update, the seeds and a fixed number of search rounds (each
// followed by k_next_round, so that a round with nothing to do ends at
// once), replayed with one launch and one read-back.  The batch is padded
// with no-op records to a power-of-two bucket, so that batches of varying
// size share a few graphs.  The graph reads and writes the engine's own
// staging buffers (d_upd, d_counts, d_safe, d_hashes), so a caller's buffers
// are copied in and out around it, and is keyed by everything its launches
// bake in (bucket, ping-pong buffers, allocations).  Rounds still pending
// after it continue eagerly.  The first batch of a key runs eagerly (sizes
// and allocations settle), the second is captured.  Larger batches run
// launch by launch: there the overhead is well under 1 % of a batch, and
// the search kernel's time is measured per launch with HIP events.
constexpr size_t MAX_GRAPHS = 16;
constexpr uint32_t GRAPH_MAX_BATCH = 4096;
constexpr uint32_t GRAPH_MAX_ROUNDS = 6;

bool graphs_allowed(const csm_engine* e, uint32_t n) {
  static const bool off = getenv("CSM_NO_GRAPHS") != nullptr || getenv("CSM_TRACE_ROUNDS") != nullptr;
  return !off && !e->opts.no_graphs && n <= GRAPH_MAX_BATCH && e->deadline <= 0 &&
         e->algo != ALGO_NONE && !e->opts.print_enum && !e->dyn;
}

// the bucket a batch of n updates is padded to
uint32_t graph_bucket(const csm_engine* e, uint32_t n) {
  uint32_t b = 64;
  while (b < n) b <<= 1;
  return std::min(b, e->max_batch);
}

csm_engine::GraphKey graph_key(const csm_engine* e, uint32_t n, bool hash) {
  csm_engine::GraphKey k;
  std::memset(&k, 0, sizeof k);
  k.n = n;
  k.nq = (uint32_t)e->queries.size();
  k.dcur = (uint32_t)e->dcur;
  k.cur = (uint32_t)e->cur;
  k.hash = hash ? 1u : 0u;
  k.nv = e->nv;
  k.budgeted = e->opts.match_budget ? 1u : 0u;
  k.capped = e->opts.max_results ? 1u : 0u;
  k.gen = e->gen;
  return k;
}

void graph_release(csm_engine::GraphEntry& g) {
  if (g.exec) (void)hipGraphExecDestroy(g.exec);
  g.exec = nullptr;
}

// Enqueue (on a capturing stream) one batch's head: begin, unpack, clear,
// update, seeds, `rounds` search rounds between the search events.
int enqueue_batch_head(csm_engine* e, uint32_t n, bool hash, uint32_t rounds) {
  int rc;
  unsigned long long* counts = e->d_counts;
  unsigned long long* hashes = hash ? e->d_hashes : nullptr;
  k_batch_begin<<<1, 64, 0, e->stream>>>(e->d_ctrl);
  k_unpack<<<grid_for(n, 256), 256, 0, e->stream>>>(e->d_upd, n, e->d_type, e->d_add, e->d_v1,
                                                    e->d_v2, e->d_lab, e->d_elmap,
                                                    (uint32_t)e->el_map.size(), e->d_ctrl);
  batch_clear(e, n, counts, hashes);
  if ((rc = batch_update(e, n))) return rc;
  seeds(e, n, counts, hashes, e->d_safe, true);
  SearchArgs sa = search_args(e, counts, hashes, 0);
  const bool symbi = uses_index(e) && e->d_d2 != nullptr;
  for (uint32_t r = 0; r < rounds; ++r) {
    round_buffers(e, sa, r);
    sa.round_slot = r % ROUND_GROUP_MAX;
    if (hash) { if (symbi) launch_search<true, true>(e, sa); else launch_search<true, false>(e, sa); }
    else { if (symbi) launch_search<false, true>(e, sa); else launch_search<false, false>(e, sa); }
    k_next_round<<<1, 64, 0, e->stream>>>(e->d_ctrl, e->spill_chunk_cap);
  }
  return CSM_OK;
}

// Capture the head of a batch into an executable graph; false (and the entry
// marked bad) if the capture failed, allocated, or holds a memset node.
bool graph_capture(csm_engine* e, csm_engine::GraphEntry& g, uint32_t n, bool hash) {
  const uint64_t gen0 = e->gen;
  const uint32_t rounds = std::min<uint32_t>(std::max<uint32_t>(e->max_rounds, 2u), GRAPH_MAX_ROUNDS);
  if (hipStreamBeginCapture(e->stream, hipStreamCaptureModeRelaxed) != hipSuccess) {
    g.bad = true;
    return false;
  }
  const int rc = enqueue_batch_head(e, n, hash, rounds);
  hipGraph_t graph = nullptr;
  const hipError_t st = hipStreamEndCapture(e->stream, &graph);
  bool ok = rc == CSM_OK && st == hipSuccess && graph && e->gen == gen0;
  if (ok) {
    size_t nn = 0;
    ok = hipGraphGetNodes(graph, nullptr, &nn) == hipSuccess;
    std::vector<hipGraphNode_t> nodes(nn);
    if (ok && nn) ok = hipGraphGetNodes(graph, nodes.data(), &nn) == hipSuccess;
    for (size_t i = 0; ok && i < nn; ++i) {
      hipGraphNodeType t;
      ok = hipGraphNodeGetType(nodes[i], &t) == hipSuccess && t != hipGraphNodeTypeMemset;
    }
  }
  if (ok) ok = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0) == hipSuccess;
  if (graph) (void)hipGraphDestroy(graph);
  (void)hipGetLastError();
  if (!ok) {
    graph_release(g);
    g.bad = true;
    return false;
  }
  g.rounds = rounds;
  return true;
}

// A device batch through a graph.  Returns 1 if this batch should run
// eagerly instead (nothing has been enqueued then).
int run_batch_graph(csm_engine* e, const csm_update* d_updates, uint32_t n, uint64_t* d_counts,
                    uint8_t* d_safe, uint64_t* d_hash) {
  int rc;
  const uint32_t B = graph_bucket(e, n);
  if ((rc = batch_prepare(e, B))) return rc;
  const bool hash = d_hash != nullptr;
  const csm_engine::GraphKey key = graph_key(e, B, hash);
  csm_engine::GraphEntry* g = nullptr;
  for (auto& x : e->graphs)
    if (x.key == key) g = &x;
  if (!g) {
    // first batch of this key: eagerly (allocations settle); keep the key
    if (e->graphs.size() >= MAX_GRAPHS) {
      graph_release(e->graphs.front());
      e->graphs.erase(e->graphs.begin());
    }
    csm_engine::GraphEntry x;
    x.key = key;
    e->graphs.push_back(x);
    return 1;
  }
  if (g->bad) return 1;
  const size_t nq = e->queries.size();
  if (!g->exec && !graph_capture(e, *g, B, hash)) return 1;
  // the caller's records, then no-op records (type 0) up to the bucket
  if (reinterpret_cast<const void*>(d_updates) != e->d_upd)
    HIP_TRY(e, hipMemcpyAsync(e->d_upd, d_updates, (size_t)n * sizeof(UpdateRec),
                              hipMemcpyDeviceToDevice, e->stream));
  if (B > n) {
    const uint64_t w = (uint64_t)(B - n) * sizeof(UpdateRec) / 4;
    k_fill_u32<<<grid_for(w, 256), 256, 0, e->stream>>>(reinterpret_cast<uint32_t*>(e->d_upd + n),
                                                        w, 0u);
  }
  HIP_TRY(e, hipGraphLaunch(g->exec, e->stream));
  ++e->graph_launches;
  if ((rc = read_ctrl(e))) return rc;
  e->max_rounds = std::max(e->max_rounds, e->h_ctrl->rounds + 1);
  if (e->h_ctrl->nchunks > 0) {
    // more rounds than the graph holds: continue eagerly from round g->rounds
    if ((rc = search(e, true, e->d_counts, hash ? e->d_hashes : nullptr, 0, g->rounds))) return rc;
  }
  if ((rc = batch_tail(e, B, e->d_counts, true))) return rc;
  if (reinterpret_cast<void*>(d_counts) != e->d_counts)
    HIP_TRY(e, hipMemcpyAsync(d_counts, e->d_counts, (size_t)n * nq * 8, hipMemcpyDeviceToDevice,
                              e->stream));
  if (d_safe && d_safe != e->d_safe)
    HIP_TRY(e, hipMemcpyAsync(d_safe, e->d_safe, n, hipMemcpyDeviceToDevice, e->stream));
  if (hash && reinterpret_cast<void*>(d_hash) != e->d_hashes)
    HIP_TRY(e, hipMemcpyAsync(d_hash, e->d_hashes, (size_t)n * nq * 8, hipMemcpyDeviceToDevice,
                              e->stream));
  return CSM_OK;
}

void accumulate_stats(csm_engine* e) {
  e->stats[7] = e->max_rounds;
  e->stats[0] = e->h_ctrl->stat_tasks;
  e->stats[1] = e->h_ctrl->stat_cands;
  e->stats[2] = e->h_ctrl->stat_bytes;
}

// ------------------------------------------------------------- SymBi index
constexpr uint32_t SX_UNMATCHABLE = 0xFFFFFFFEu;
constexpr int SX_SLOTS = 256;

SxView sx_view(csm_engine* e) {
  SxView x;
  x.rows = e->d_sxrows;
  x.d1 = e->d_sx_d1;
  x.d2 = e->d_sx_d2;
  x.np1 = e->d_sx_np1;
  x.nc2 = e->d_sx_nc2;
  x.n1 = e->d_sx_n1;
  x.n2 = e->d_sx_n2;
  x.words = e->d2_words;
  x.nvc = e->nv_cap;
  return x;
}

// grow a row-major array of `rows_old` rows to `rows_new` rows of `per_row` words
int grow_rows(csm_engine* e, uint32_t*& p, uint64_t per_row, uint32_t rows_old, uint32_t rows_new) {
  uint32_t* q = nullptr;
  int rc = dalloc(e, &q, per_row * std::max<uint32_t>(rows_new, 1));
  if (rc) return rc;
  HIP_TRY(e, hipMemsetAsync(q, 0, per_row * std::max<uint32_t>(rows_new, 1) * 4, e->stream));
  if (p && rows_old)
    HIP_TRY(e, hipMemcpyAsync(q, p, per_row * rows_old * 4, hipMemcpyDeviceToDevice, e->stream));
  HIP_TRY(e, hipStreamSynchronize(e->stream));
  dfree(e, p);
  p = q;
  return CSM_OK;
}

int sx_reserve(csm_engine* e, uint32_t rows, uint32_t n1rows, uint32_t n2rows) {
  int rc;
  const uint64_t words = e->d2_words, nvc = e->nv_cap;
  if (rows > e->sx_cap_rows) {
    uint32_t cap = std::max<uint32_t>(rows, e->sx_cap_rows * 2);
    if ((rc = grow_rows(e, e->d_sx_d1, words, e->d2_rows, cap)) ||
        (rc = grow_rows(e, e->d_sx_d2, words, e->d2_rows, cap)) ||
        (rc = grow_rows(e, e->d_sx_np1, nvc, e->d2_rows, cap)) ||
        (rc = grow_rows(e, e->d_sx_nc2, nvc, e->d2_rows, cap)))
      return rc;
    e->sx_cap_rows = cap;
  }
  if (n1rows > e->sx_cap_n1) {
    uint32_t cap = std::max<uint32_t>(n1rows, e->sx_cap_n1 * 2);
    if ((rc = grow_rows(e, e->d_sx_n1, nvc, e->sx_n1rows, cap))) return rc;
    e->sx_cap_n1 = cap;
  }
  if (n2rows > e->sx_cap_n2) {
    uint32_t cap = std::max<uint32_t>(n2rows, e->sx_cap_n2 * 2);
    if ((rc = grow_rows(e, e->d_sx_n2, nvc, e->sx_n2rows, cap))) return rc;
    e->sx_cap_n2 = cap;
  }
  e->d_d2 = e->d_sx_d2;
  return CSM_OK;
}

// DAG-edge table: bucket (L(p), L(c), l) -> (parent row << 8) | child index
int sx_upload_tables(csm_engine* e) {
  const uint32_t nb = e->nlab * e->nlab * e->nel;
  std::vector<std::vector<uint32_t>> buckets(nb);
  for (uint32_t r = 0; r < e->h_sxrows.size(); ++r) {
    const SxRow& R = e->h_sxrows[r];
    if (R.label == SX_UNMATCHABLE) continue;
    for (uint32_t j = 0; j < R.nc; ++j) {
      const SxRow& C = e->h_sxrows[R.ch_row[j]];
      if (C.label == SX_UNMATCHABLE) continue;
      uint32_t l = R.ch_key[j] % e->nel;
      buckets[(R.label * e->nlab + C.label) * e->nel + l].push_back((r << 8) | j);
    }
  }
  e->h_sxtab_off.assign(nb + 1, 0);
  e->h_sxtab_ent.clear();
  for (uint32_t b = 0; b < nb; ++b) {
    e->h_sxtab_off[b] = (uint32_t)e->h_sxtab_ent.size();
    e->h_sxtab_ent.insert(e->h_sxtab_ent.end(), buckets[b].begin(), buckets[b].end());
  }
  e->h_sxtab_off[nb] = (uint32_t)e->h_sxtab_ent.size();
  dfree(e, e->d_sxtab_off);
  dfree(e, e->d_sxtab_ent);
  dfree(e, e->d_sxrows);
  int rc;
  if ((rc = dalloc(e, &e->d_sxtab_off, e->h_sxtab_off.size())) ||
      (rc = dalloc(e, &e->d_sxtab_ent, std::max<size_t>(1, e->h_sxtab_ent.size()))) ||
      (rc = dalloc(e, &e->d_sxrows, e->h_sxrows.size())))
    return rc;
  HIP_TRY(e, hipMemcpyAsync(e->d_sxtab_off, e->h_sxtab_off.data(), e->h_sxtab_off.size() * 4,
                            hipMemcpyHostToDevice, e->stream));
  if (!e->h_sxtab_ent.empty())
    HIP_TRY(e, hipMemcpyAsync(e->d_sxtab_ent, e->h_sxtab_ent.data(), e->h_sxtab_ent.size() * 4,
                              hipMemcpyHostToDevice, e->stream));
  HIP_TRY(e, hipMemcpyAsync(e->d_sxrows, e->h_sxrows.data(), e->h_sxrows.size() * sizeof(SxRow),
                            hipMemcpyHostToDevice, e->stream));
  HIP_TRY(e, hipStreamSynchronize(e->stream));
  return CSM_OK;
}

int symbi_prepare_query(csm_engine* e, QueryInfo& q) {
  const QueryGraph& g = q.g;
  const uint32_t k = g.k;
  std::vector<uint32_t> dl(k);
  for (uint32_t u = 0; u < k; ++u) dl[u] = dense_of(e->vl_map, g.vl[u]);
  auto el_of = [&](uint32_t a, uint32_t b) { return dense_of(e->el_map, g.elabel(a, b)); };
  auto efreq = [&](uint32_t a, uint32_t b) -> uint64_t {
    uint32_t l = el_of(a, b);
    if (dl[a] == NO_LABEL || dl[b] == NO_LABEL || l == NO_LABEL) return 0;
    return e->h_triple_freq[(dl[a] * e->nlab + dl[b]) * e->nel + l];
  };
  auto vfreq = [&](uint32_t a) -> uint64_t { return dl[a] == NO_LABEL ? 0 : e->h_vfreq[dl[a]]; };
  q.dag = e->algo == ALGO_TURBOFLUX ? turboflux_tree(g, efreq, vfreq) : symbi_dag(g, efreq, vfreq);
  // rows, n1 rows (one per DAG edge, at the child), n2 rows (one per directed edge)
  const uint32_t base = e->d2_rows;
  std::vector<std::vector<uint32_t>> n2id(k, std::vector<uint32_t>(k, NONE));
  uint32_t n2next = e->sx_n2rows;
  for (uint32_t u = 0; u < k; ++u)
    for (uint32_t w : g.nb[u]) n2id[u][w] = n2next++;
  q.n2row.assign((size_t)k * k, NONE);
  for (uint32_t u = 0; u < k; ++u)
    for (uint32_t w = 0; w < k; ++w) q.n2row[(size_t)u * k + w] = n2id[u][w];
  std::vector<std::vector<uint32_t>> n1id(k, std::vector<uint32_t>(k, NONE));
  uint32_t n1next = e->sx_n1rows;
  for (uint32_t u = 0; u < k; ++u)
    for (uint32_t p : q.dag.parents[u]) n1id[u][p] = n1next++;
  int rc = sx_reserve(e, base + k, n1next, n2next);
  if (rc) return rc;
  bool all_known = true;
  for (uint32_t u = 0; u < k; ++u) {
    SxRow R;
    std::memset(&R, 0, sizeof R);
    R.label = dl[u] == NO_LABEL ? SX_UNMATCHABLE : dl[u];
    all_known = all_known && dl[u] != NO_LABEL;
    R.np = (uint32_t)q.dag.parents[u].size();
    R.nc = (uint32_t)q.dag.children[u].size();
    for (uint32_t j = 0; j < R.np; ++j) {
      uint32_t p = q.dag.parents[u][j];
      uint32_t l = el_of(u, p);
      all_known = all_known && l != NO_LABEL;
      R.par_row[j] = base + p;
      R.par_key[j] = (dl[p] == NO_LABEL ? 0 : dl[p]) * e->nel + (l == NO_LABEL ? 0 : l);
      R.par_n1[j] = n1id[u][p];
      R.par_n2up[j] = n2id[u][p];
      R.par_n2dn[j] = n2id[p][u];
    }
    for (uint32_t j = 0; j < R.nc; ++j) {
      uint32_t c = q.dag.children[u][j];
      uint32_t l = el_of(u, c);
      R.ch_row[j] = base + c;
      R.ch_key[j] = (dl[c] == NO_LABEL ? 0 : dl[c]) * e->nel + (l == NO_LABEL ? 0 : l);
      R.ch_n1[j] = n1id[c][u];
      R.ch_n2dn[j] = n2id[u][c];
      R.ch_n2up[j] = n2id[c][u];
    }
    e->h_sxrows.push_back(R);
  }
  e->d2_rows = base + k;
  e->sx_n1rows = n1next;
  e->sx_n2rows = n2next;
  e->sx_maxk = std::max(e->sx_maxk, k);
  q.d2_base = base;
  if ((rc = sx_upload_tables(e))) return rc;
  // BuildDCS (symbi.cpp:246-377): top-down, bottom-up, then parent-side n2
  if (all_known) {
    GraphView gv = view(e);
    SxView x = sx_view(e);
    const int grid = grid_for(e->nv, 256);
    for (uint32_t i = 0; i < k; ++i)
      k_sx_build<<<grid, 256, 0, e->stream>>>(gv, x, base + q.dag.bfs[i], 0);
    for (uint32_t i = 0; i < k; ++i)
      k_sx_build<<<grid, 256, 0, e->stream>>>(gv, x, base + q.dag.bfs[k - 1 - i], 1);
    for (uint32_t i = 0; i < k; ++i) k_sx_build<<<grid, 256, 0, e->stream>>>(gv, x, base + i, 2);
    HIP_TRY(e, hipStreamSynchronize(e->stream));
  }
  q.sx_built = true;
  return CSM_OK;
}

// TurboFlux's explicit-path counts of a query whose spanning-tree index is
// built (k_tf_paths, one launch per tree edge in BFS order).
int turboflux_paths(csm_engine* e, QueryInfo& q, std::vector<uint32_t>& paths) {
  const uint32_t k = q.g.k, nv = e->nv;
  paths.assign(k, 0);
  uint32_t* calls = nullptr;
  unsigned long long* d_paths = nullptr;
  int rc;
  if ((rc = dalloc(e, &calls, (size_t)k * nv)) || (rc = dalloc(e, &d_paths, k))) {
    dfree(e, calls);
    return rc;
  }
  HIP_TRY(e, hipMemsetAsync(calls, 0, (size_t)k * nv * 4, e->stream));
  HIP_TRY(e, hipMemsetAsync(d_paths, 0, (size_t)k * 8, e->stream));
  const GraphView g = view(e);
  const SxView x = sx_view(e);
  const int grid = grid_for(nv, 256);
  const uint32_t root = q.dag.root;
  k_tf_root_calls<<<grid, 256, 0, e->stream>>>(e->d_vlab, q.dlab[root], calls + (size_t)root * nv, nv);
  for (uint32_t u : q.dag.bfs)
    for (uint32_t c : q.dag.children[u]) {
      const uint32_t key = q.dlab[c] * e->nel + dense_of(e->el_map, q.g.elabel(u, c));
      k_tf_paths<<<grid, 256, 0, e->stream>>>(g, x, key, q.d2_base + c, calls + (size_t)u * nv,
                                              calls + (size_t)c * nv, d_paths + c);
    }
  std::vector<unsigned long long> h(k);
  HIP_TRY(e, hipMemcpyAsync(h.data(), d_paths, k * 8, hipMemcpyDeviceToHost, e->stream));
  HIP_TRY(e, hipStreamSynchronize(e->stream));
  dfree(e, calls);
  dfree(e, d_paths);
  for (uint32_t u = 0; u < k; ++u) paths[u] = (uint32_t)h[u];   // 32-bit, as the reference's uint
  paths[root] = UINT32_MAX - 1;
  return CSM_OK;
}

int sx_buffers(csm_engine* e) {
  if (e->d_sx_mark) return CSM_OK;
  int rc;
  e->sx_list_cap = 1u << 24;
  if ((rc = dalloc(e, &e->d_sx_mark, e->max_batch)) || (rc = dalloc(e, &e->d_sx_slots, SX_SLOTS)))
    return rc;
  for (auto& l : e->d_sx_lists)
    if ((rc = dalloc(e, &l, e->sx_list_cap))) return rc;
  return CSM_OK;
}

SxList sx_list(csm_engine* e, int buf, int slot) {
  SxList l;
  l.buf = e->d_sx_lists[buf];
  l.count = e->d_sx_slots + slot;
  l.cap = e->sx_list_cap;
  return l;
}

// One phase: snapshot step, d1 propagation, then d2 propagation (insert: set,
// union view; delete: clear, end view).  Lists: 0 candidates, 1 frontier,
// 2 changed-d1 record, 3 spare.  Slots: a fresh counter per list use.
int sx_phase(csm_engine* e, uint32_t n, bool ins) {
  int rc;
  if ((rc = sx_buffers(e))) return rc;
  GraphView g = view(e);
  SxView x = sx_view(e);
  BatchView b = batch_view(e, n);
  uint32_t* err = &e->d_ctrl->err;
  k_sx_zero<<<1, 256, 0, e->stream>>>(e->d_sx_slots, SX_SLOTS);
  if (ins) k_sx_mark<<<grid_for(n, 256), 256, 0, e->stream>>>(b, g, e->d_slab, e->d_sx_mark);
  SxSeedArgs a;
  a.b = b;
  a.g = g;
  a.x = x;
  a.slab = e->d_slab;
  a.mark = e->d_sx_mark;
  a.tab_off = e->d_sxtab_off;
  a.tab_ent = e->d_sxtab_ent;
  a.nlab = e->nlab;
  a.want = ins ? 1u : 2u;
  a.err = err;
  int slot = 0;
  const int iters = (int)e->sx_maxk + 1;
  SxList none;
  none.buf = nullptr;
  none.count = e->d_sx_slots + SX_SLOTS - 1;
  none.cap = 0;
  const int pgrid = std::max(1, e->num_cus * 8);
  // d1
  SxList cand = sx_list(e, 0, slot++);
  SxList rec = sx_list(e, 2, slot++);
  a.stage = 0;
  a.out = cand;
  k_sx_seed<<<grid_for(n, 256), 256, 0, e->stream>>>(a);
  for (int it = 0; it < iters; ++it) {
    SxList front = sx_list(e, 1, slot++);
    k_sx_flags<<<grid_for(e->sx_list_cap, 256, 2048), 256, 0, e->stream>>>(cand, e->d_sx_d1, e->d2_words, ins, front, rec, err);
    cand = sx_list(e, 0, slot++);
    k_sx_propagate<<<pgrid, 256, 0, e->stream>>>(g, x, front, ins, !ins, 0, cand, err);
  }
  // d2
  SxList c2 = sx_list(e, 3, slot++);
  a.stage = 1;
  a.out = c2;
  k_sx_seed<<<grid_for(n, 256), 256, 0, e->stream>>>(a);
  k_sx_from_d1<<<grid_for(e->sx_list_cap, 256, 2048), 256, 0, e->stream>>>(rec, x, ins, c2, err);
  for (int it = 0; it < iters; ++it) {
    SxList front = sx_list(e, 1, slot++);
    k_sx_flags<<<grid_for(e->sx_list_cap, 256, 2048), 256, 0, e->stream>>>(c2, e->d_sx_d2, e->d2_words, ins, front, none, err);
    c2 = sx_list(e, it % 2 ? 3 : 0, slot++);
    k_sx_propagate<<<pgrid, 256, 0, e->stream>>>(g, x, front, ins, !ins, 1, c2, err);
  }
  k_sx_check<<<1, 64, 0, e->stream>>>(cand.count, err);
  k_sx_check<<<1, 64, 0, e->stream>>>(c2.count, err);
  if (slot >= SX_SLOTS - 1) return fail(e, CSM_E_CAPACITY, "SymBi propagation slots exhausted");
  return CSM_OK;
}

int symbi_batch_insert(csm_engine* e, uint32_t n) { return sx_phase(e, n, true); }
int symbi_batch_delete(csm_engine* e, uint32_t n) { return sx_phase(e, n, false); }

int symbi_add_vertex(csm_engine* e, uint32_t id) {
  if (!e->d_sxrows) return CSM_OK;
  k_sx_add_vertex<<<1, 256, 0, e->stream>>>(sx_view(e), e->d2_rows, id, e->h_vlab[id]);
  HIP_TRY(e, hipStreamSynchronize(e->stream));
  return CSM_OK;
}

// InitialMatching on the device: seeds are all directed data edges matching
// the first edge of order 0, searched from depth 2 (graphflow.cpp:268-283
// enumerates from single vertices; both count every embedding once).
// --print-enum: the rows of the search just finished to the host, their
// update indices shifted by t_offset, added to the rows already collected
// (append) or replacing them; at most emb_cap rows are kept per call, and
// emb_found counts every embedding.
int collect_embeddings(csm_engine* e, uint32_t t_offset, bool append) {
  if (!e->d_emb) return CSM_OK;
  unsigned long long n = 0;
  HIP_TRY(e, hipMemcpyAsync(&n, e->d_emb_n, 8, hipMemcpyDeviceToHost, e->stream));
  HIP_TRY(e, hipStreamSynchronize(e->stream));
  if (!append) {
    e->h_emb.clear();
    e->emb_found = 0;
  }
  e->emb_found += n;
  const uint64_t have = e->h_emb.size() / EMB_W;
  const uint64_t keep = std::min<uint64_t>(std::min<uint64_t>(n, e->emb_cap),
                                           e->emb_cap > have ? e->emb_cap - have : 0);
  if (keep == 0) return CSM_OK;
  const size_t old = e->h_emb.size();
  e->h_emb.resize(old + keep * EMB_W);
  HIP_TRY(e, hipMemcpy(e->h_emb.data() + old, e->d_emb, keep * EMB_W * 4, hipMemcpyDeviceToHost));
  for (uint64_t r = 0; r < keep; ++r) e->h_emb[old + r * EMB_W] += t_offset;
  return CSM_OK;
}

int initial_matching(csm_engine* e, QueryInfo& q, uint64_t* total) {
  const PlanRec& P = e->h_plans[q.plan_base];
  if (P.k == 0) { *total = 0; return CSM_OK; }
  e->batch_n = UINT32_MAX;   // the whole graph: full spill budget
  uint32_t ua = q.orders[0][0], ub = q.orders[0][1];
  uint32_t la = q.dlab[ua];
  uint32_t key01 = (q.dlab[ub]) * e->nel + dense_of(e->el_map, q.g.elabel(ua, ub));
  const size_t nq = e->queries.size();
  HIP_TRY(e, hipMemsetAsync(e->d_counts, 0, nq * 8, e->stream));
  k_fill_u32<<<1, 64, 0, e->stream>>>(e->d_ustatus, 1, 0u);
  k_fill_u32<<<1, 256, 0, e->stream>>>(e->d_qtl, nq, 0u);
  // --print-enum: the rows come from the hashing search
  unsigned long long* hashes = e->d_emb ? e->d_hashes : nullptr;
  if (e->d_emb) {
    k_fill_u32<<<1, 256, 0, e->stream>>>(reinterpret_cast<uint32_t*>(e->d_hashes), 2 * nq, 0u);
    k_fill_u32<<<1, 64, 0, e->stream>>>(reinterpret_cast<uint32_t*>(e->d_emb_n), 2, 0u);
  }
  uint32_t v0 = 0, step = e->nv;
  while (v0 < e->nv) {
    uint32_t v1 = (uint32_t)std::min<uint64_t>((uint64_t)v0 + step, e->nv);
    k_batch_begin<<<1, 64, 0, e->stream>>>(e->d_ctrl);
    GraphView g = view(e);
    k_initial_seeds<<<grid_for(v1 - v0, 256 / WAVE, 8192), 256, 0, e->stream>>>(
        g, e->d_plans, q.plan_base, la, key01, v0, v1, P.chunk, e->d_seeds,
        e->seed_cap, e->d_chunk_owner, e->chunk_cap, e->d_ctrl, e->d_d2, e->d2_words,
        (uses_index(e) && e->d_d2) ? 1u : 0u, e->d_counts, e->d_emb ? e->d_limits : nullptr);
    int rc = read_ctrl(e);
    if (rc) return rc;
    if (e->h_ctrl->err & (ERR_SEED_OVERFLOW | ERR_CHUNK_OVERFLOW)) {
      if (step == 1) return fail(e, CSM_E_CAPACITY, "initial matching: one vertex overflows the task buffers");
      step = std::max<uint32_t>(1, step / 4);
      continue;
    }
    // search with the end-of-batch view (position 0xFFFF)
    if ((rc = search(e, true, e->d_counts, hashes, 0xFFFEu))) return rc;
    v0 = v1;
  }
  if (int rc = collect_embeddings(e, 0, false)) return rc;
  if (e->opts.max_results)
    k_clamp_counts<<<1, 64, 0, e->stream>>>(e->d_counts, 1, (uint32_t)nq, e->opts.max_results,
                                            e->d_ustatus);
  unsigned long long c = 0;
  uint32_t st = 0;
  HIP_TRY(e, hipMemcpyAsync(&c, e->d_counts, 8, hipMemcpyDeviceToHost, e->stream));  // query 0
  HIP_TRY(e, hipMemcpyAsync(&st, e->d_ustatus, 4, hipMemcpyDeviceToHost, e->stream));
  HIP_TRY(e, hipStreamSynchronize(e->stream));
  e->h_ustatus.assign(1, (uint8_t)st);
  *total = c;
  return CSM_OK;
}

// Vertex deletion (Graph::RemoveVertex, graph.cpp:51-56, called by the
// plugins' RemoveVertex, e.g. parallel_graphflow.cpp:2311-2315 and
// symbi.cpp:992-1021): the vertex loses its label and its edges, and no match
// is counted, as in the reference.  The reference leaves the vertex's entries
// in its neighbours' lists (they no longer match its label); here every
// incident edge is deleted through the batch pipeline with the search off, so
// no list keeps an entry for it and the SymBi index stays the fix-point of the
// remaining graph.
int remove_vertex(csm_engine* e, uint32_t id) {
  if (id >= e->nv || e->h_vlab[id] == NO_LABEL) return CSM_OK;
  uint2 rec;
  HIP_TRY(e, hipMemcpy(&rec, e->d_vrec + id, sizeof(uint2), hipMemcpyDeviceToHost));
  const uint32_t len = rec.y & LEN_MASK;
  std::vector<uint32_t> ent(len), ver(len);
  if (len) {
    HIP_TRY(e, hipMemcpy(ent.data(), e->d_adj[e->cur] + rec.x, len * 4, hipMemcpyDeviceToHost));
    HIP_TRY(e, hipMemcpy(ver.data(), e->d_ats[e->cur] + rec.x, len * 4, hipMemcpyDeviceToHost));
  }
  std::vector<UpdateRec> dels;
  for (uint32_t i = 0; i < len; ++i) {
    if ((ver[i] & 0xFFFFu) != 0xFFFFu) continue;  // deleted in an earlier batch
    const uint32_t nbr = ent[i] & e->vmask, key = ent[i] >> e->vbits;
    dels.push_back(UpdateRec{'e', 0, 0, id, nbr, e->el_map[key % e->nel]});
  }
  for (size_t done = 0; done < dels.size();) {
    const uint32_t m = (uint32_t)std::min<size_t>(dels.size() - done, e->max_batch);
    std::memcpy(e->h_upd, dels.data() + done, m * sizeof(UpdateRec));
    HIP_TRY(e, hipMemcpyAsync(e->d_upd, e->h_upd, m * sizeof(UpdateRec), hipMemcpyHostToDevice,
                              e->stream));
    k_batch_begin<<<1, 64, 0, e->stream>>>(e->d_ctrl);
    k_unpack<<<grid_for(m, 256), 256, 0, e->stream>>>(e->d_upd, m, e->d_type, e->d_add, e->d_v1,
                                                      e->d_v2, e->d_lab, e->d_elmap,
                                                      (uint32_t)e->el_map.size(), e->d_ctrl);
    int rc = run_batch(e, m, e->d_counts, nullptr, e->d_safe, false);
    if (rc) return rc;
    if ((rc = read_ctrl(e))) return rc;
    if ((rc = check_errors(e))) return rc;
    done += m;
  }
  e->num_edges -= dels.size();
  e->h_vlab[id] = NO_LABEL;
  HIP_TRY(e, hipMemcpyAsync(e->d_vlab + id, &e->h_vlab[id], 4, hipMemcpyHostToDevice, e->stream));
  if (uses_index(e) && e->d_sxrows)
    k_sx_remove_vertex<<<1, 256, 0, e->stream>>>(sx_view(e), e->d2_rows, e->sx_n1rows,
                                                 e->sx_n2rows, id);
  HIP_TRY(e, hipStreamSynchronize(e->stream));
  return CSM_OK;
}

// Device parse of `len` bytes of text at d_text (see csm_parse.hpp) on
// stream st, scratch through alloc/release; deliver(d_records, count) copies
// the records out (it returns a CSM_* code).  Returns CSM_OK, CSM_E_ARG
// (malformed line: *err_line, 1-based), CSM_E_CAPACITY or CSM_E_DEVICE; msg
// gets the text.  *n = records found.
template <class Alloc, class Release, class Deliver>
int parse_text_on(hipStream_t st, int mode, const uint8_t* d_text, uint64_t len, size_t* n,
                  uint64_t* err_line, Alloc alloc, Release release, Deliver deliver,
                  std::string& msg) {
  *n = 0;
  if (err_line) *err_line = 0;
  if (mode != PARSE_GRAPH && mode != PARSE_STREAM) { msg = "unknown parse mode"; return CSM_E_ARG; }
  if (len == 0) return CSM_OK;
  if ((reinterpret_cast<uintptr_t>(d_text) & 15u) != 0) {
    msg = "text must be 16-byte aligned in device memory";
    return CSM_E_ARG;
  }
  const uint64_t ntiles = (len + PARSE_TILE - 1) / PARSE_TILE;
  if (ntiles > 0x7FFFFFFFull) { msg = "text too large"; return CSM_E_CAPACITY; }
  std::vector<void*> held;
  auto take = [&](size_t bytes) -> void* {
    void* p = alloc(std::max<size_t>(bytes, 256));
    if (p) held.push_back(p);
    return p;
  };
  auto done = [&](int rc) {
    (void)hipStreamSynchronize(st);
    for (void* p : held) release(p);
    return rc;
  };
#define PARSE_CHECK(x)                                                        \
  do {                                                                        \
    hipError_t e_ = (x);                                                      \
    if (e_ != hipSuccess) { msg = hipGetErrorString(e_); return done(CSM_E_DEVICE); } \
  } while (0)
  auto* tcount = static_cast<uint32_t*>(take(ntiles * 4));
  auto* toff = static_cast<uint32_t*>(take(ntiles * 4));
  if (!tcount || !toff) { msg = "device allocation failed"; return done(CSM_E_DEVICE); }
  k_parse_count<<<(uint32_t)ntiles, PARSE_BS, 0, st>>>(d_text, len, tcount);
  size_t need = 0;
  PARSE_CHECK(rocprim::exclusive_scan(nullptr, need, tcount, toff, 0u, (size_t)ntiles,
                                      rocprim::plus<uint32_t>(), st));
  void* tmp = take(need);
  if (!tmp) { msg = "device allocation failed"; return done(CSM_E_DEVICE); }
  PARSE_CHECK(rocprim::exclusive_scan(tmp, need, tcount, toff, 0u, (size_t)ntiles,
                                      rocprim::plus<uint32_t>(), st));
  uint32_t last_off = 0, last_cnt = 0;
  uint8_t last_byte = 0;
  PARSE_CHECK(hipMemcpyAsync(&last_off, toff + ntiles - 1, 4, hipMemcpyDeviceToHost, st));
  PARSE_CHECK(hipMemcpyAsync(&last_cnt, tcount + ntiles - 1, 4, hipMemcpyDeviceToHost, st));
  PARSE_CHECK(hipMemcpyAsync(&last_byte, d_text + len - 1, 1, hipMemcpyDeviceToHost, st));
  PARSE_CHECK(hipStreamSynchronize(st));
  const uint64_t n_nl = (uint64_t)last_off + last_cnt;
  const uint64_t n_lines = n_nl + (last_byte != '\n' ? 1 : 0);
  auto* lend = static_cast<uint64_t*>(take(std::max<uint64_t>(n_nl, 1) * 8));
  auto* rec = static_cast<UpdateRec*>(take(n_lines * sizeof(UpdateRec)));
  auto* rec2 = static_cast<UpdateRec*>(take(n_lines * sizeof(UpdateRec)));
  auto* keep = static_cast<uint8_t*>(take(n_lines));
  auto* flags = static_cast<uint64_t*>(take(16));   // [0] error bit, [1] first bad line
  auto* count = static_cast<uint64_t*>(take(8));
  if (!lend || !rec || !rec2 || !keep || !flags || !count) {
    msg = "device allocation failed";
    return done(CSM_E_DEVICE);
  }
  const uint64_t init[2] = {0, ~0ull};
  PARSE_CHECK(hipMemcpyAsync(flags, init, 16, hipMemcpyHostToDevice, st));
  k_parse_lines<<<(uint32_t)ntiles, PARSE_BS, 0, st>>>(d_text, len, toff, lend);
  k_parse_records<<<grid_for(n_lines, 256, 65536), 256, 0, st>>>(
      d_text, len, lend, n_nl, n_lines, mode, rec, keep, reinterpret_cast<uint32_t*>(flags),
      reinterpret_cast<unsigned long long*>(flags + 1));
  size_t need2 = 0;
  PARSE_CHECK(rocprim::select(nullptr, need2, rec, keep, rec2, count, (size_t)n_lines, st));
  void* tmp2 = take(need2);
  if (!tmp2) { msg = "device allocation failed"; return done(CSM_E_DEVICE); }
  PARSE_CHECK(rocprim::select(tmp2, need2, rec, keep, rec2, count, (size_t)n_lines, st));
  uint64_t h_flags[2] = {0, 0}, h_count = 0;
  PARSE_CHECK(hipMemcpyAsync(h_flags, flags, 16, hipMemcpyDeviceToHost, st));
  PARSE_CHECK(hipMemcpyAsync(&h_count, count, 8, hipMemcpyDeviceToHost, st));
  PARSE_CHECK(hipStreamSynchronize(st));
  if (h_flags[0] & 1u) {
    if (err_line) *err_line = h_flags[1] + 1;
    char b[96];
    snprintf(b, sizeof b, "malformed line %llu", (unsigned long long)(h_flags[1] + 1));
    msg = b;
    return done(CSM_E_ARG);
  }
  *n = h_count;
  const int rc = deliver(rec2, (size_t)h_count);
  if (rc == CSM_E_CAPACITY) msg = "output capacity too small";
  else if (rc) msg = "copying the records out failed";
#undef PARSE_CHECK
  return done(rc);
}

// ---------------------------------------------- allocation-free text parse
// csm_parse_text_device_static parses text already in HBM with no device
// allocation: its working space is this file-scope array, PARSE_SCRATCH_BYTES
// per GPU (each device holds its own copy until the process ends).  Host code
// must never pass the array's name to a kernel or copy: its device address on
// each GPU comes from hipGetSymbolAddress at engine creation (scratch_init)
// and is kept per device in g_scratch.  Every call on one GPU shares that
// GPU's array, so calls that overlap on one GPU race; callers serialise them.
constexpr size_t PARSE_SCRATCH_BYTES = 12u << 20;
constexpr uint64_t PARSE_STATIC_MAX_TEXT = 8u << 20;   // bytes of text per call
constexpr uint64_t PARSE_STATIC_MAX_LINES = 1u << 18;  // lines per call
static __device__ __align__(256) uint8_t scratch[PARSE_SCRATCH_BYTES];
constexpr int SCRATCH_MAX_DEVICES = 64;
static uint8_t* g_scratch[SCRATCH_MAX_DEVICES] = {};
static std::mutex g_scratch_mu;

// with `dev` current: look up this GPU's copy of the scratch array (once)
int scratch_init(int dev) {
  if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) return CSM_E_CAPACITY;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[dev]) return CSM_OK;
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return CSM_E_DEVICE;
  g_scratch[dev] = static_cast<uint8_t*>(p);
  return CSM_OK;
}

__global__ void k_parse_flags_init(unsigned long long* flags) {
  if (threadIdx.x == 0) {
    flags[0] = 0;
    flags[1] = ~0ull;
  }
}

int parse_static(csm_engine* e, int mode, const uint8_t* d_text, uint64_t len,
                 csm_update* d_out, size_t cap, size_t* n, uint64_t* err_line) {
  *n = 0;
  if (err_line) *err_line = 0;
  if (mode != PARSE_GRAPH && mode != PARSE_STREAM) return fail(e, CSM_E_ARG, "unknown parse mode");
  if (len > PARSE_STATIC_MAX_TEXT)
    return fail(e, CSM_E_CAPACITY, "text of %llu bytes exceeds the %llu-byte limit of the "
                "allocation-free parse", (unsigned long long)len,
                (unsigned long long)PARSE_STATIC_MAX_TEXT);
  if (len == 0) return CSM_OK;
  if ((reinterpret_cast<uintptr_t>(d_text) & 15u) != 0)
    return fail(e, CSM_E_ARG, "text must be 16-byte aligned in device memory");
  uint8_t* base = e->dev < SCRATCH_MAX_DEVICES ? g_scratch[e->dev] : nullptr;
  if (!base) return fail(e, CSM_E_DEVICE, "scratch of the allocation-free parse is not set up");
  hipStream_t st = e->stream;
  // bump allocation inside the scratch array, 256-byte aligned
  size_t used = 0;
  auto take = [&](size_t bytes) -> uint8_t* {
    size_t at = (used + 255) & ~size_t(255);
    if (at + bytes > PARSE_SCRATCH_BYTES) return nullptr;
    used = at + bytes;
    return base + at;
  };
  const uint64_t ntiles = (len + PARSE_TILE - 1) / PARSE_TILE;
  auto* tcount = reinterpret_cast<uint32_t*>(take(ntiles * 4));
  auto* toff = reinterpret_cast<uint32_t*>(take(ntiles * 4));
  auto* flags = reinterpret_cast<unsigned long long*>(take(16));
  auto* count = reinterpret_cast<uint64_t*>(take(8));
  size_t need = 0;
  if (rocprim::exclusive_scan(nullptr, need, tcount, toff, 0u, (size_t)ntiles,
                              rocprim::plus<uint32_t>(), st) != hipSuccess)
    return fail(e, CSM_E_DEVICE, "scan size query failed");
  void* tmp = take(need);
  if (!tcount || !toff || !flags || !count || !tmp)
    return fail(e, CSM_E_CAPACITY, "allocation-free parse: scratch too small");
  k_parse_count<<<(uint32_t)ntiles, PARSE_BS, 0, st>>>(d_text, len, tcount);
  HIP_TRY(e, rocprim::exclusive_scan(tmp, need, tcount, toff, 0u, (size_t)ntiles,
                                     rocprim::plus<uint32_t>(), st));
  uint64_t* hp = e->h_parse;
  hp[0] = hp[1] = hp[2] = 0;
  HIP_TRY(e, hipMemcpyAsync(&hp[0], toff + ntiles - 1, 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(e, hipMemcpyAsync(&hp[1], tcount + ntiles - 1, 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(e, hipMemcpyAsync(&hp[2], d_text + len - 1, 1, hipMemcpyDeviceToHost, st));
  HIP_TRY(e, hipStreamSynchronize(st));
  const uint64_t n_nl = (hp[0] & 0xFFFFFFFFull) + (hp[1] & 0xFFFFFFFFull);
  const uint64_t n_lines = n_nl + ((hp[2] & 0xFFu) != '\n' ? 1 : 0);
  if (n_lines > PARSE_STATIC_MAX_LINES)
    return fail(e, CSM_E_CAPACITY, "%llu lines exceed the %llu-line limit of the "
                "allocation-free parse", (unsigned long long)n_lines,
                (unsigned long long)PARSE_STATIC_MAX_LINES);
  auto* lend = reinterpret_cast<uint64_t*>(take(std::max<uint64_t>(n_nl, 1) * 8));
  auto* rec = reinterpret_cast<UpdateRec*>(take(n_lines * sizeof(UpdateRec)));
  auto* rec2 = reinterpret_cast<UpdateRec*>(take(n_lines * sizeof(UpdateRec)));
  auto* keep = take(n_lines);
  size_t need2 = 0;
  if (rocprim::select(nullptr, need2, rec, keep, rec2, count, (size_t)n_lines, st) != hipSuccess)
    return fail(e, CSM_E_DEVICE, "select size query failed");
  void* tmp2 = take(need2);
  if (!lend || !rec || !rec2 || !keep || !tmp2)
    return fail(e, CSM_E_CAPACITY, "allocation-free parse: scratch too small");
  // flags start at (no error, no bad line) on every call: the array keeps
  // the last call's contents
  k_parse_flags_init<<<1, 64, 0, st>>>(flags);
  k_parse_lines<<<(uint32_t)ntiles, PARSE_BS, 0, st>>>(d_text, len, toff, lend);
  k_parse_records<<<grid_for(n_lines, 256, 65536), 256, 0, st>>>(
      d_text, len, lend, n_nl, n_lines, mode, rec, keep, reinterpret_cast<uint32_t*>(flags),
      flags + 1);
  HIP_TRY(e, rocprim::select(tmp2, need2, rec, keep, rec2, count, (size_t)n_lines, st));
  HIP_TRY(e, hipMemcpyAsync(&hp[3], flags, 16, hipMemcpyDeviceToHost, st));
  HIP_TRY(e, hipMemcpyAsync(&hp[5], count, 8, hipMemcpyDeviceToHost, st));
  HIP_TRY(e, hipStreamSynchronize(st));
  if (hp[3] & 1u) {
    if (err_line) *err_line = hp[4] + 1;
    return fail(e, CSM_E_ARG, "malformed line %llu", (unsigned long long)(hp[4] + 1));
  }
  *n = hp[5];
  if (hp[5] > cap) return fail(e, CSM_E_CAPACITY, "output capacity too small");
  if (hp[5]) {
    HIP_TRY(e, hipMemcpyAsync(d_out, rec2, hp[5] * sizeof(UpdateRec), hipMemcpyDeviceToDevice, st));
    HIP_TRY(e, hipStreamSynchronize(st));
  }
  return CSM_OK;
}

// Autotuning on one batch whose graph update has run.  For each size class
// with registered queries, the search of that class's queries (the others'
// chunks are skipped) runs once per candidate setting: the same work each
// time, so the least search time picks the class's setting for the engine's
// lifetime.  Then the whole search runs with the chosen settings (with a
// single class the last candidate's search has counted everything already:
// the counts do not depend on the settings).
int tune_batch(csm_engine* e, uint32_t n, unsigned long long* counts, unsigned long long* hashes,
               uint8_t* safe) {
  int rc;
  std::vector<int> classes;
  for (auto& qp : e->queries) {
    const int c = size_class(qp->g.k);
    if (std::find(classes.begin(), classes.end(), c) == classes.end()) classes.push_back(c);
  }
  std::sort(classes.begin(), classes.end());
  const size_t nq = e->queries.size();
  // skip flags of the other classes' queries (the per-query flags a match
  // budget uses; with a budget set, the classes are tuned together)
  const bool mask = classes.size() > 1 && !e->opts.match_budget;
  static const bool trace = getenv("CSM_TRACE_ROUNDS") != nullptr;
  bool first = true;
  for (size_t ci = 0; ci < (mask ? classes.size() : 1); ++ci) {
    if (mask) {
      std::vector<uint32_t> skip(nq);
      for (size_t q = 0; q < nq; ++q) skip[q] = size_class(e->queries[q]->g.k) != classes[ci];
      HIP_TRY(e, hipMemcpyAsync(e->d_qabort, skip.data(), nq * 4, hipMemcpyHostToDevice, e->stream));
      HIP_TRY(e, hipStreamSynchronize(e->stream));
      e->tune_mask = true;
    }
    size_t best = 0;
    double best_ns = 0;
    if (trace) fprintf(stderr, "[csm] autotune, queries of up to %s vertices:",
                       mask ? (classes[ci] == 0 ? "8" : classes[ci] == 1 ? "12" : classes[ci] == 2 ? "16" : "32")
                            : "32");
    for (size_t i = 0; i < e->tune_cands.size(); ++i) {
      if (!first) {
        k_search_reset<<<1, 64, 0, e->stream>>>(e->d_ctrl);
        batch_clear(e, n, counts, hashes);
      }
      first = false;
      for (int c : classes)
        if (!mask || c == classes[ci]) {
          e->cls_chunk[c] = e->tune_cands[i].first;
          e->cls_budget[c] = e->tune_cands[i].second;
        }
      if ((rc = upload_plans(e))) return rc;
      const uint64_t ns0 = e->stats[5];
      seeds(e, n, counts, hashes, safe, true);
      if ((rc = search(e, true, counts, hashes))) return rc;
      const double ns = (double)(e->stats[5] - ns0);
      if (i == 0 || ns < best_ns) { best = i; best_ns = ns; }
      if (trace) fprintf(stderr, " (%u,%u)=%.2fms", e->tune_cands[i].first, e->tune_cands[i].second,
                         ns * 1e-6);
    }
    for (int c : classes)
      if (!mask || c == classes[ci]) {
        e->cls_chunk[c] = e->tune_cands[best].first;
        e->cls_budget[c] = e->tune_cands[best].second;
      }
    if (trace) fprintf(stderr, " -> split %u, budget %u\n", e->tune_cands[best].first,
                       e->tune_cands[best].second);
  }
  e->tune_active = false;
  if ((rc = upload_plans(e))) return rc;
  if (mask) {
    e->tune_mask = false;
    k_fill_u32<<<1, 256, 0, e->stream>>>(e->d_qabort, nq, 0u);
    k_search_reset<<<1, 64, 0, e->stream>>>(e->d_ctrl);
    batch_clear(e, n, counts, hashes);
    seeds(e, n, counts, hashes, safe, true);
    if ((rc = search(e, true, counts, hashes))) return rc;
  }
  return CSM_OK;
}

}  // namespace

// ================================================================== C-ABI
extern "C" {

// Where an engine's data graph comes from: host arrays, or graph-file
// records in HBM.
struct DataSrc {
  const csm_graph* host = nullptr;
  const UpdateRec* d_rec = nullptr;
  size_t n_rec = 0;
};

static csm_engine* create_engine(const char* algo, const csm_graph* query, const DataSrc& data,
                                 const csm_opts* opts, int* err);

csm_engine* csm_create(const char* algo, const csm_graph* query, const csm_graph* data,
                       const csm_opts* opts, int* err) {
  if (!data) { if (err) *err = CSM_E_ARG; return nullptr; }
  DataSrc src;
  src.host = data;
  return create_engine(algo, query, src, opts, err);
}

csm_engine* csm_create_from_records(const char* algo, const csm_graph* query,
                                    const csm_update* d_records, size_t n, const csm_opts* opts,
                                    int* err) {
  if (n && !d_records) { if (err) *err = CSM_E_ARG; return nullptr; }
  DataSrc src;
  src.d_rec = reinterpret_cast<const UpdateRec*>(d_records);
  src.n_rec = n;
  return create_engine(algo, query, src, opts, err);
}

static csm_engine* create_engine(const char* algo, const csm_graph* query, const DataSrc& data,
                                 const csm_opts* opts, int* err) {
  auto set = [&](int c) { if (err) *err = c; };
  int a;
  if (parse_algo(algo, &a)) { set(CSM_E_UNKNOWN_ALGO); return nullptr; }
  if (!query) { set(CSM_E_ARG); return nullptr; }
  std::unique_ptr<csm_engine> e(new csm_engine());
  e->algo = a;
  e->algo_name = algo;
  if (opts) e->opts = *opts;
  if (e->opts.max_batch == 0 || e->opts.max_batch > DEFAULT_MAX_BATCH) e->opts.max_batch = DEFAULT_MAX_BATCH;
  e->dyn = a == ALGO_SYMBI && e->opts.reference_orders != 0;
  // the search-time order has no split or spill settings to tune
  e->tune_active = e->opts.auto_tuning != 0 && e->opts.split_chunk == 0 &&
                   e->opts.spill_budget == 0 && !e->dyn;
  if (e->tune_active)
    e->tune_cands = {{64, 1024}, {64, 2048}, {64, 512}, {64, 384}, {64, 320}, {64, 256}, {32, 320}};
  for (int c = 0; c < NSIZE_CLASSES; ++c) {
    e->cls_chunk[c] = e->opts.split_chunk ? e->opts.split_chunk : 64;
    e->cls_budget[c] = e->opts.spill_budget ? e->opts.spill_budget : 1024;
  }
  e->max_batch = e->opts.max_batch;
  e->dev = (int)e->opts.device;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= e->dev) { set(CSM_E_DEVICE); return nullptr; }
  if (hipSetDevice(e->dev) != hipSuccess) { set(CSM_E_DEVICE); return nullptr; }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, e->dev) != hipSuccess) { set(CSM_E_DEVICE); return nullptr; }
  e->num_cus = prop.multiProcessorCount;
  if (e->opts.device_budget == 0) e->opts.device_budget = (uint64_t)(prop.totalGlobalMem * 0.9);
  if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) { set(CSM_E_DEVICE); return nullptr; }
  for (auto& pair : e->ev_round)
    for (auto& ev : pair)
      if (hipEventCreate(&ev) != hipSuccess) { set(CSM_E_DEVICE); return nullptr; }
  if (hipStreamCreateWithFlags(&e->xfer, hipStreamNonBlocking) != hipSuccess) { set(CSM_E_DEVICE); return nullptr; }
  for (auto& ev : e->marks)
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) { set(CSM_E_DEVICE); return nullptr; }


  // the data graph's vertex labels, edge count and edge label set
  std::vector<uint32_t> raw_vl, els;
  uint32_t nv = 0;
  uint64_t ne = 0;
  if (data.host) {
    nv = data.host->num_vertices;
    ne = data.host->num_edges;
    raw_vl.assign(data.host->vlabels, data.host->vlabels + nv);
    els.assign(data.host->elabels, data.host->elabels + ne);
  } else {
    const int rc = scan_records(e.get(), data.d_rec, data.n_rec, &nv, raw_vl, &ne, els);
    if (rc) { set(rc); return nullptr; }
  }
  // label maps from data and query labels
  std::vector<uint32_t> vls(raw_vl);
  els.insert(els.end(), query->elabels, query->elabels + query->num_edges);
  std::sort(vls.begin(), vls.end());
  vls.erase(std::unique(vls.begin(), vls.end()), vls.end());
  vls.erase(std::remove(vls.begin(), vls.end(), NO_LABEL), vls.end());
  std::sort(els.begin(), els.end());
  els.erase(std::unique(els.begin(), els.end()), els.end());
  if (els.empty()) els.push_back(0);
  if (vls.empty()) vls.push_back(0);
  e->vl_map = vls;
  e->el_map = els;
  e->nlab = (uint32_t)vls.size();
  e->nel = (uint32_t)els.size();
  e->nv = nv;
  uint64_t keys = (uint64_t)e->nlab * e->nel;
  uint32_t kbits = bits_for(keys);
  uint64_t want_cap = (uint64_t)e->nv + e->nv / 8 + 1024;
  uint32_t vb = bits_for(want_cap);
  if (vb + kbits > 32) vb = bits_for(std::max<uint64_t>(e->nv, 2));
  if (vb + kbits > 32) {
    set(CSM_E_CAPACITY);
    return nullptr;
  }
  e->vbits = vb;
  e->vmask = vb >= 32 ? 0xFFFFFFFFu : ((1u << vb) - 1u);
  // a directory holds nkeys + 1 offsets; lists shorter than an eighth of that
  // (or than 8) keep binary search
  e->dstride = (uint32_t)keys + 1;
  e->dir_threshold = std::max<uint32_t>(8u, e->dstride / 8);
  e->nv_cap = (uint32_t)std::min<uint64_t>(want_cap, 1ull << vb);
  if (e->nv_cap < e->nv) e->nv_cap = e->nv;
  e->d2_words = (e->nv_cap + 31) / 32;
  e->h_vlab.assign(e->nv_cap, NO_LABEL);
  for (uint32_t v = 0; v < e->nv; ++v) e->h_vlab[v] = dense_of(e->vl_map, raw_vl[v]);

  // label histograms of the initial graph for the SymBi root choice (the
  // triple histogram of device records is counted while they are built)
  e->h_triple_freq.assign((size_t)e->nlab * e->nlab * e->nel, 0);
  e->h_vfreq.assign(e->nlab, 0);
  for (uint32_t v = 0; v < e->nv; ++v)
    if (e->h_vlab[v] != NO_LABEL) ++e->h_vfreq[e->h_vlab[v]];
  if (data.host) {
    const csm_graph* d = data.host;
    for (size_t i = 0; i < d->num_edges; ++i) {
      uint32_t a = d->src[i], b = d->dst[i];
      if (a >= e->nv || b >= e->nv || a == b) continue;
      uint32_t la = e->h_vlab[a], lb = e->h_vlab[b], l = dense_of(e->el_map, d->elabels[i]);
      if (la == NO_LABEL || lb == NO_LABEL || l == NO_LABEL) continue;
      ++e->h_triple_freq[(la * e->nlab + lb) * e->nel + l];
      ++e->h_triple_freq[(lb * e->nlab + la) * e->nel + l];
    }
  }

  int rc = CSM_OK;
  // 8x the initial directed entries: a batch may need room for twice the
  // live entries (every list rewritten by compaction and merge), so a pool
  // of 4x left room for about one batch of hub rewrites between repacks
  // (records build from every record, a non-edge one as a dropped self-loop)
  const uint64_t build_edges = data.host ? ne : data.n_rec;
  e->pool_cap = std::max<uint64_t>(8 * 2 * build_edges + 8ull * e->max_batch, 1u << 20);
  if (e->pool_cap > 0xFFFFFFF0ull) e->pool_cap = 0xFFFFFFF0ull;
  if ((rc = dalloc(e.get(), &e->d_vrec, e->nv_cap)) || (rc = dalloc(e.get(), &e->d_vlab, e->nv_cap)) ||
      (rc = dalloc(e.get(), &e->d_adj[0], e->pool_cap)) ||
      (rc = dalloc(e.get(), &e->d_ats[0], e->pool_cap)) || (rc = dalloc(e.get(), &e->d_ctrl, 1)) ||
      (rc = dalloc(e.get(), &e->d_elmap, e->el_map.size()))) {
    set(rc);
    return nullptr;
  }
  if (hipHostMalloc((void**)&e->h_ctrl, sizeof(Ctrl)) != hipSuccess) { set(CSM_E_DEVICE); return nullptr; }
  if (hipHostMalloc((void**)&e->h_parse, 8 * sizeof(uint64_t)) != hipSuccess) { set(CSM_E_DEVICE); return nullptr; }
  // the allocation-free parse's scratch on this device, and room for error
  // texts so that setting one allocates nothing
  if ((rc = scratch_init(e->dev))) { set(rc); return nullptr; }
  e->err.reserve(512);
  if (hipMemcpy(e->d_vlab, e->h_vlab.data(), (size_t)e->nv_cap * 4, hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(e->d_elmap, e->el_map.data(), e->el_map.size() * 4, hipMemcpyHostToDevice) !=
          hipSuccess) {
    set(CSM_E_DEVICE);
    return nullptr;
  }
  rc = data.host ? build_graph(e.get(), data.host)
                 : build_graph_records(e.get(), data.d_rec, data.n_rec);
  if (rc || (rc = alloc_batch(e.get()))) { set(rc); return nullptr; }
  std::unique_ptr<QueryInfo> qi;
  if ((rc = build_query(e.get(), query, qi))) { set(rc); return nullptr; }
  e->queries.push_back(std::move(qi));
  if ((rc = ensure_results(e.get()))) { set(rc); return nullptr; }
  set(CSM_OK);
  return e.release();
}

int csm_register_query(csm_engine* e, const csm_graph* query, uint32_t* qid) {
  if (!e) return CSM_E_ARG;
  std::unique_ptr<QueryInfo> qi;
  int rc = build_query(e, query, qi);
  if (rc) return rc;
  uint32_t id = (uint32_t)e->queries.size();
  e->queries.push_back(std::move(qi));
  if (qid) *qid = id;
  if ((rc = ensure_results(e))) return rc;
  if (e->preprocessed) {
    QueryInfo& q = *e->queries.back();
    if (uses_index(e)) {
      q.d2_base = e->d2_rows;
      if ((rc = symbi_prepare_query(e, q))) return rc;
    }
    if ((rc = plan_query(e, q, id))) return rc;
    build_seed_table(e);
    if ((rc = upload_plans(e))) return rc;
  }
  return CSM_OK;
}

uint32_t csm_num_queries(const csm_engine* e) { return e ? (uint32_t)e->queries.size() : 0; }

int csm_preprocess(csm_engine* e) {
  if (!e) return CSM_E_ARG;
  if (e->preprocessed) return CSM_OK;
  int rc;
  HIP_TRY(e, hipSetDevice(e->dev));
  e->h_plans.clear();
  for (size_t i = 0; i < e->queries.size(); ++i) {
    QueryInfo& q = *e->queries[i];
    if (uses_index(e)) {
      q.d2_base = e->d2_rows;
      if ((rc = symbi_prepare_query(e, q))) return rc;
    }
    if ((rc = plan_query(e, q, (uint32_t)i))) return rc;
  }
  build_seed_table(e);
  if ((rc = upload_plans(e))) return rc;
  e->preprocessed = true;
  if (e->opts.print_prep) {
    QueryInfo& q = *e->queries[0];
    printf("matching order: \n-vertex(backward neighbors)-\n");
    for (size_t i = 0; i < q.orders.size(); ++i) {
      printf("#%zu: ", i);
      for (uint32_t p = 0; p < q.g.k; ++p) printf("%u%s", q.orders[i][p], p + 1 < q.g.k ? "-" : "");
      printf("\n");
    }
    printf("\n");
  }
  return CSM_OK;
}

static int initial_impl(csm_engine* e, uint64_t* num_results);

int csm_initial(csm_engine* e, uint64_t* num_results) {
  CallScope cs(e);
  return cs.finish(initial_impl(e, num_results));
}

static int initial_impl(csm_engine* e, uint64_t* num_results) {
  if (!e || !num_results) return CSM_E_ARG;
  int rc = csm_preprocess(e);
  if (rc) return rc;
  HIP_TRY(e, hipSetDevice(e->dev));
  QueryInfo& q = *e->queries[0];
  uint64_t total = 0;
  if (e->algo == ALGO_NONE) {  // matching::InitialMatching does nothing
    *num_results = 0;
    return CSM_OK;
  }
  if (q.g.ne == 0) {
    e->h_emb.clear();
    for (uint32_t v = 0; v < e->nv; ++v) {
      const bool hit = e->h_vlab[v] != NO_LABEL && e->h_vlab[v] == q.dlab[0];
      total += hit;
      if (hit && e->d_emb && e->h_emb.size() / EMB_W < e->emb_cap) {
        e->h_emb.resize(e->h_emb.size() + EMB_W, NONE);
        uint32_t* row = e->h_emb.data() + e->h_emb.size() - EMB_W;
        row[0] = 0;
        row[1] = 0;
        row[2] = v;
      }
    }
    e->emb_found = e->d_emb ? total : 0;
    *num_results = total;
    e->tot_init = total;
    return CSM_OK;
  }
  rc = initial_matching(e, q, &total);
  if (rc) return rc;
  *num_results = total;
  e->tot_init = total;
  return CSM_OK;
}

int csm_classify(csm_engine* e, uint32_t v1, uint32_t v2, uint32_t label, int* safe) {
  if (!e || !safe) return CSM_E_ARG;
  if (!e->preprocessed) { int rc = csm_preprocess(e); if (rc) return rc; }
  if (v1 >= e->nv || v2 >= e->nv) return fail(e, CSM_E_ARG, "vertex out of range");
  if (e->algo == ALGO_NONE) { *safe = 0; return CSM_OK; }  // matching::Classify (matching.cpp:66-68)
  uint32_t l1 = e->h_vlab[v1], l2 = e->h_vlab[v2], el = dense_of(e->el_map, label);
  if (l1 == NO_LABEL || l2 == NO_LABEL || el == NO_LABEL) { *safe = 1; return CSM_OK; }
  uint32_t bucket = (l1 * e->nlab + l2) * e->nel + el;
  *safe = e->h_tab_off[bucket] == e->h_tab_off[bucket + 1];
  return CSM_OK;
}

// Plain per-update path: the batch pipeline run on one update, one launch
// sequence and one host round trip per update.
static int single_update(csm_engine* e, bool add, uint32_t v1, uint32_t v2, uint32_t label,
                         uint64_t* out) {
  if (!e) return CSM_E_ARG;
  csm_update u;
  std::memset(&u, 0, sizeof u);
  u.type = 'e';
  u.is_add = add ? 1 : 0;
  u.id1 = v1;
  u.id2 = v2;
  u.label = label;
  if (v1 >= e->nv || v2 >= e->nv || e->h_vlab[v1] == NO_LABEL || e->h_vlab[v2] == NO_LABEL)
    return fail(e, add ? CSM_E_ARG : CSM_E_DELETE_MISSING, "vertex out of range");
  std::vector<uint64_t> pos(e->queries.size()), neg(e->queries.size());
  int rc = csm_apply_batch(e, &u, 1, pos.data(), neg.data(), nullptr, nullptr);
  if (rc) return rc;
  uint64_t tot = 0;
  for (size_t q = 0; q < pos.size(); ++q) tot += add ? pos[q] : neg[q];
  if (out) *out = tot;
  return CSM_OK;
}

int csm_add_edge(csm_engine* e, uint32_t v1, uint32_t v2, uint32_t label, uint64_t* pos) {
  return single_update(e, true, v1, v2, label, pos);
}

int csm_remove_edge(csm_engine* e, uint32_t v1, uint32_t v2, uint64_t* neg) {
  return single_update(e, false, v1, v2, 0, neg);
}

int csm_add_vertex(csm_engine* e, uint32_t id, uint32_t label) {
  if (!e) return CSM_E_ARG;
  if (id >= e->nv_cap) return fail(e, CSM_E_CAPACITY, "vertex id %u beyond capacity %u", id, e->nv_cap);
  uint32_t d = dense_of(e->vl_map, label);
  if (d == NO_LABEL) return fail(e, CSM_E_CAPACITY, "vertex label %u unknown to the engine", label);
  // Graph::AddVertex (graph.cpp:24-44): only a vertex without a label gets one
  if (e->h_vlab[id] != NO_LABEL) return CSM_OK;
  e->h_vlab[id] = d;
  if (id >= e->nv) e->nv = id + 1;
  HIP_TRY(e, hipMemcpyAsync(e->d_vlab + id, &e->h_vlab[id], 4, hipMemcpyHostToDevice, e->stream));
  HIP_TRY(e, hipStreamSynchronize(e->stream));
  if (uses_index(e) && e->d_d2) return symbi_add_vertex(e, id);
  return CSM_OK;
}

int csm_remove_vertex(csm_engine* e, uint32_t id) {
  if (!e) return CSM_E_ARG;
  if (!e->preprocessed) { int rc = csm_preprocess(e); if (rc) return rc; }
  HIP_TRY(e, hipSetDevice(e->dev));
  return remove_vertex(e, id);
}

// emb_offset / emb_append: embedding rows (--print-enum) of this batch are
// numbered from emb_offset and added to the rows already collected by the
// same host call (its earlier slices), or replace the last call's rows.
static int apply_batch_device_impl(csm_engine* e, const csm_update* d_updates, size_t n,
                                   uint64_t* d_counts, uint8_t* d_safe, uint64_t* d_hash,
                                   uint32_t emb_offset = 0, bool emb_append = false) {
  if (!e || (n && (!d_updates || !d_counts))) return CSM_E_ARG;
  if (n > e->max_batch) return fail(e, CSM_E_CAPACITY, "batch of %zu > max_batch %u", n, e->max_batch);
  if (!e->preprocessed) { int rc = csm_preprocess(e); if (rc) return rc; }
  if (d_hash && !e->opts.want_hash) d_hash = nullptr;
  // --print-enum: the rows come from the hashing search
  if (!d_hash && e->d_emb) d_hash = reinterpret_cast<uint64_t*>(e->d_hashes);
  e->h_ustatus.assign(n, 0);
  e->h_uwork.assign(n, 0);
  e->batch_n = (uint32_t)n;
  if (!emb_append) {
    e->h_emb.clear();
    e->emb_found = 0;
  }
  if (n == 0) return CSM_OK;
  HIP_TRY(e, hipSetDevice(e->dev));
  int rc = 1;
  if (graphs_allowed(e, (uint32_t)n)) rc = run_batch_graph(e, d_updates, (uint32_t)n, d_counts, d_safe, d_hash);
  if (rc == 1) {
    k_batch_begin<<<1, 64, 0, e->stream>>>(e->d_ctrl);
    k_unpack<<<grid_for(n, 256), 256, 0, e->stream>>>(
        reinterpret_cast<const UpdateRec*>(d_updates), (uint32_t)n, e->d_type, e->d_add, e->d_v1,
        e->d_v2, e->d_lab, e->d_elmap, (uint32_t)e->el_map.size(), e->d_ctrl);
    rc = run_batch(e, (uint32_t)n, reinterpret_cast<unsigned long long*>(d_counts),
                   reinterpret_cast<unsigned long long*>(d_hash), d_safe);
  }
  if (rc) return rc;
  if (e->opts.match_budget)
    HIP_TRY(e, hipMemcpyAsync(e->h_qabort.data(), e->d_qabort, e->queries.size() * 4,
                              hipMemcpyDeviceToHost, e->stream));
  HIP_TRY(e, hipMemcpyAsync(e->h_ustatus_pin, e->d_ustatus, n * 4, hipMemcpyDeviceToHost,
                            e->stream));
  HIP_TRY(e, hipMemcpyAsync(e->h_qtl.data(), e->d_qtl, e->queries.size() * 4,
                            hipMemcpyDeviceToHost, e->stream));
  if (e->d_uwork)
    HIP_TRY(e, hipMemcpyAsync(e->h_uwork_pin, e->d_uwork, n * 8, hipMemcpyDeviceToHost,
                              e->stream));
  if ((rc = read_ctrl(e))) return rc;
  for (size_t i = 0; i < n; ++i) e->h_ustatus[i] = (uint8_t)e->h_ustatus_pin[i];
  e->h_uwork.assign(e->h_uwork_pin ? e->h_uwork_pin : nullptr,
                    e->h_uwork_pin ? e->h_uwork_pin + n : nullptr);
  accumulate_stats(e);
  if ((rc = collect_embeddings(e, emb_offset, true))) return rc;
  return check_errors(e);
}

int csm_apply_batch_device(csm_engine* e, const csm_update* d_updates, size_t n,
                           uint64_t* d_counts, uint8_t* d_safe, uint64_t* d_hash) {
  CallScope cs(e);
  return cs.finish(apply_batch_device_impl(e, d_updates, n, d_counts, d_safe, d_hash));
}

int csm_update_status(const csm_engine* e, uint8_t* status, size_t n) {
  if (!e || (n && !status)) return CSM_E_ARG;
  for (size_t i = 0; i < n; ++i) status[i] = i < e->h_ustatus.size() ? e->h_ustatus[i] : 0;
  return CSM_OK;
}

int csm_update_work(const csm_engine* e, uint64_t* work, size_t n) {
  if (!e || (n && !work)) return CSM_E_ARG;
  if (!e->opts.want_work) return fail(const_cast<csm_engine*>(e), CSM_E_UNSUPPORTED,
                                      "per-update work needs csm_opts.want_work");
  for (size_t i = 0; i < n; ++i) work[i] = i < e->h_uwork.size() ? e->h_uwork[i] : 0;
  return CSM_OK;
}

int csm_embeddings(const csm_engine* e, uint32_t* rows, size_t cap_rows, size_t* n_rows,
                   uint64_t* n_found, uint32_t* width) {
  if (!e || (cap_rows && !rows)) return CSM_E_ARG;
  if (!e->opts.print_enum)
    return fail(const_cast<csm_engine*>(e), CSM_E_UNSUPPORTED,
                "embedding rows need csm_opts.print_enum");
  const uint32_t w = 2 + e->max_k;
  const size_t have = e->h_emb.size() / EMB_W;
  const size_t n = rows ? std::min(have, cap_rows) : have;
  for (size_t r = 0; r < n && rows; ++r) {
    const uint32_t* src = e->h_emb.data() + r * EMB_W;
    uint32_t* dst = rows + r * w;
    dst[0] = src[0];
    dst[1] = src[1];
    const uint32_t k = src[1] < e->queries.size() ? e->queries[src[1]]->g.k : 0u;
    for (uint32_t i = 0; i < e->max_k; ++i) dst[2 + i] = i < k ? src[2 + i] : NONE;
  }
  if (n_rows) *n_rows = n;
  if (n_found) *n_found = e->emb_found;
  if (width) *width = w;
  return CSM_OK;
}

int csm_set_time_limit(csm_engine* e, double seconds) {
  if (!e || !(seconds >= 0)) return CSM_E_ARG;
  e->opts.time_limit_s = seconds;
  return CSM_OK;
}

int csm_query_status(const csm_engine* e, uint8_t* aborted, size_t n) {
  if (!e || !aborted) return CSM_E_ARG;
  for (size_t i = 0; i < n && i < e->h_qabort.size(); ++i)
    aborted[i] = (e->h_qabort[i] ? 1 : 0) | (i < e->h_qtl.size() && e->h_qtl[i] ? 2 : 0);
  return CSM_OK;
}

int csm_sync(csm_engine* e) {
  if (!e) return CSM_E_ARG;
  HIP_TRY(e, hipStreamSynchronize(e->stream));
  return CSM_OK;
}

static int apply_batch_impl(csm_engine* e, const csm_update* updates, size_t n, uint64_t* pos,
                            uint64_t* neg, uint8_t* safe, uint64_t* emb_hash,
                            std::vector<uint8_t>& status) {
  if (!e || (n && !updates)) return CSM_E_ARG;
  if (!e->preprocessed) { int rc = csm_preprocess(e); if (rc) return rc; }
  HIP_TRY(e, hipSetDevice(e->dev));
  const size_t nq = e->queries.size();
  size_t done = 0;
  status.assign(n, 0);
  e->h_emb.clear();       // the slices below add their embedding rows
  e->emb_found = 0;
  while (done < n) {
    // a vertex deletion is applied on its own, in stream order
    if (updates[done].type == 'v' && !updates[done].is_add) {
      int rc = remove_vertex(e, updates[done].id1);
      if (rc) return rc;
      for (size_t q = 0; q < nq; ++q) {
        size_t o = done * nq + q;
        if (pos) pos[o] = 0;
        if (neg) neg[o] = 0;
        if (emb_hash) emb_hash[o] = 0;
      }
      if (safe) safe[done] = 0;
      ++done;
      continue;
    }
    size_t m = std::min<size_t>(n - done, e->max_batch);
    for (size_t i = 1; i < m; ++i)
      if (updates[done + i].type == 'v' && !updates[done + i].is_add) { m = i; break; }
    // vertex insertions take effect before the slice's edges; vertex records
    // travel to the device as no-ops so that positions are kept
    for (size_t i = 0; i < m; ++i) {
      const csm_update& u = updates[done + i];
      std::memcpy(&e->h_upd[i], &u, sizeof(UpdateRec));
      if (u.type == 'v') {
        int rc = csm_add_vertex(e, u.id1, u.label);
        if (rc) return rc;
        e->h_upd[i].type = 0;
      }
    }
    HIP_TRY(e, hipMemcpyAsync(e->d_upd, e->h_upd, m * sizeof(UpdateRec), hipMemcpyHostToDevice,
                              e->stream));
    bool hash = emb_hash && e->opts.want_hash;
    int rc = apply_batch_device_impl(e, reinterpret_cast<const csm_update*>(e->d_upd), m,
                                     reinterpret_cast<uint64_t*>(e->d_counts), e->d_safe,
                                     hash ? reinterpret_cast<uint64_t*>(e->d_hashes) : nullptr,
                                     (uint32_t)done, true);
    if (rc) return rc;
    for (size_t i = 0; i < m && i < e->h_ustatus.size(); ++i) status[done + i] = e->h_ustatus[i];
    HIP_TRY(e, hipMemcpyAsync(e->h_counts, e->d_counts, m * nq * 8, hipMemcpyDeviceToHost, e->stream));
    if (hash)
      HIP_TRY(e, hipMemcpyAsync(e->h_hashes, e->d_hashes, m * nq * 8, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(e, hipMemcpyAsync(e->h_safe, e->d_safe, m, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(e, hipStreamSynchronize(e->stream));
    for (size_t i = 0; i < m; ++i) {
      const csm_update& u = updates[done + i];
      bool is_edge = u.type == 'e';
      for (size_t q = 0; q < nq; ++q) {
        uint64_t c = is_edge ? e->h_counts[i * nq + q] : 0;
        size_t o = (done + i) * nq + q;
        if (pos) pos[o] = (is_edge && u.is_add) ? c : 0;
        if (neg) neg[o] = (is_edge && !u.is_add) ? c : 0;
        if (emb_hash) emb_hash[o] = hash ? e->h_hashes[i * nq + q] : 0;
        if (u.is_add) e->tot_pos += (is_edge ? c : 0); else e->tot_neg += (is_edge ? c : 0);
      }
      if (safe) safe[done + i] = is_edge ? e->h_safe[i] : 1;
      if (is_edge) e->num_edges += u.is_add ? 1 : -1;
    }
    done += m;
  }
  return CSM_OK;
}

int csm_apply_batch(csm_engine* e, const csm_update* updates, size_t n, uint64_t* pos,
                    uint64_t* neg, uint8_t* safe, uint64_t* emb_hash) {
  CallScope cs(e);
  std::vector<uint8_t> status;
  int rc = apply_batch_impl(e, updates, n, pos, neg, safe, emb_hash, status);
  if (e) e->h_ustatus.swap(status);
  return cs.finish(rc);
}

int csm_totals(const csm_engine* e, uint64_t* initial, uint64_t* positive, uint64_t* negative) {
  if (!e) return CSM_E_ARG;
  if (initial) *initial = e->tot_init;
  if (positive) *positive = e->tot_pos;
  if (negative) *negative = e->tot_neg;
  return CSM_OK;
}

int csm_memory(const csm_engine* e, size_t* num_edges, size_t* num_vertices) {
  if (!e) return CSM_E_ARG;
  if (num_edges) *num_edges = (size_t)e->num_edges;
  if (num_vertices) *num_vertices = e->nv;
  return CSM_OK;
}

int csm_device_bytes(const csm_engine* e, uint64_t* bytes) {
  if (!e || !bytes) return CSM_E_ARG;
  *bytes = e->dev_bytes;
  return CSM_OK;
}

int csm_stats(const csm_engine* e, uint64_t* out, size_t n_out) {
  if (!e || !out) return CSM_E_ARG;
  for (size_t i = 0; i < n_out && i < 8; ++i) out[i] = e->stats[i];
  // [8], [9]: the settings of the first registered query's size class
  const int c0 = e->queries.empty() ? 0 : size_class(e->queries[0]->g.k);
  if (n_out > 8) out[8] = e->cls_chunk[c0];
  if (n_out > 9) out[9] = e->cls_budget[c0];
  if (n_out > 10) out[10] = e->tune_active ? 1 : 0;
  if (n_out > 11) out[11] = e->graph_launches;
  for (int c = 0; c < NSIZE_CLASSES; ++c) {
    if (n_out > 12 + 2 * (size_t)c) out[12 + 2 * c] = e->cls_chunk[c];
    if (n_out > 13 + 2 * (size_t)c) out[13 + 2 * c] = e->cls_budget[c];
  }
  return CSM_OK;
}

int csm_query_orders(csm_engine* e, uint32_t qid, uint32_t* out, size_t n, uint32_t* paths,
                     size_t* n_orders) {
  if (!e || !n_orders) return CSM_E_ARG;
  if (qid >= e->queries.size()) return fail(e, CSM_E_ARG, "no query %u", qid);
  if (!e->preprocessed) { int rc = csm_preprocess(e); if (rc) return rc; }
  const QueryInfo& q = *e->queries[qid];
  const uint32_t k = q.g.k;
  *n_orders = q.orders.size();
  if (out) {
    if (n < q.orders.size() * k) return fail(e, CSM_E_ARG, "order buffer too small");
    for (size_t i = 0; i < q.orders.size(); ++i)
      for (uint32_t j = 0; j < k; ++j) out[i * k + j] = q.orders[i][j];
  }
  if (paths)
    for (uint32_t u = 0; u < k; ++u) paths[u] = q.tf_paths.empty() ? 0u : q.tf_paths[u];
  return CSM_OK;
}

int csm_index_dump(csm_engine* e, uint32_t qid, uint32_t what, uint32_t* out, size_t n) {
  if (!e || !out) return CSM_E_ARG;
  if (!uses_index(e) || qid >= e->queries.size() || !e->preprocessed)
    return fail(e, CSM_E_ARG, "no candidate index for query %u", qid);
  HIP_TRY(e, hipSetDevice(e->dev));
  HIP_TRY(e, hipStreamSynchronize(e->stream));
  QueryInfo& q = *e->queries[qid];
  const uint32_t k = q.g.k, nv = e->nv, base = q.d2_base;
  std::vector<uint32_t> rows;  // device rows to copy, in canonical order
  uint32_t* arr = nullptr;
  bool bits = false;
  switch (what) {
    case 0:
      if (n < 1) return CSM_E_ARG;
      out[0] = q.dag.root;
      return CSM_OK;
    case 1: case 2:
      bits = true;
      arr = what == 1 ? e->d_sx_d1 : e->d_sx_d2;
      for (uint32_t u = 0; u < k; ++u) rows.push_back(base + u);
      break;
    case 3: case 4:
      arr = what == 3 ? e->d_sx_np1 : e->d_sx_nc2;
      for (uint32_t u = 0; u < k; ++u) rows.push_back(base + u);
      break;
    case 5:
      arr = e->d_sx_n1;
      for (uint32_t u = 0; u < k; ++u) {
        const SxRow& R = e->h_sxrows[base + u];
        for (uint32_t j = 0; j < R.np; ++j) rows.push_back(R.par_n1[j]);
      }
      break;
    case 6:
      arr = e->d_sx_n2;
      for (uint32_t u = 0; u < k; ++u) {
        const SxRow& R = e->h_sxrows[base + u];
        for (uint32_t w : q.g.nb[u]) {
          uint32_t row = NONE;
          for (uint32_t j = 0; j < R.np; ++j)
            if (R.par_row[j] == base + w) row = R.par_n2up[j];
          for (uint32_t j = 0; j < R.nc; ++j)
            if (R.ch_row[j] == base + w) row = R.ch_n2dn[j];
          rows.push_back(row);
        }
      }
      break;
    default:
      return CSM_E_ARG;
  }
  if (n < rows.size() * (size_t)nv) return fail(e, CSM_E_ARG, "dump buffer too small");
  for (size_t i = 0; i < rows.size(); ++i) {
    uint32_t* dst = out + i * nv;
    if (rows[i] == NONE) {
      // a query edge outside TurboFlux's tree keeps no counters (zero rows)
      std::fill(dst, dst + nv, 0u);
    } else if (bits) {
      std::vector<uint32_t> w(e->d2_words);
      HIP_TRY(e, hipMemcpy(w.data(), arr + (size_t)rows[i] * e->d2_words, e->d2_words * 4,
                           hipMemcpyDeviceToHost));
      for (uint32_t v = 0; v < nv; ++v) dst[v] = (w[v >> 5] >> (v & 31)) & 1u;
    } else {
      HIP_TRY(e, hipMemcpy(dst, arr + (size_t)rows[i] * e->nv_cap, (size_t)nv * 4,
                           hipMemcpyDeviceToHost));
    }
  }
  return CSM_OK;
}

int csm_device_alloc(csm_engine* e, size_t bytes, void** ptr) {
  if (!e || !ptr) return CSM_E_ARG;
  HIP_TRY(e, hipSetDevice(e->dev));
  uint8_t* p = nullptr;
  const uint64_t gen = e->gen;   // a caller's buffer does not move the engine's own
  int rc = dalloc(e, &p, bytes);
  e->gen = gen;
  if (rc) return rc;
  *ptr = p;
  return CSM_OK;
}

int csm_device_free(csm_engine* e, void* ptr) {
  if (!e) return CSM_E_ARG;
  HIP_TRY(e, hipSetDevice(e->dev));
  HIP_TRY(e, hipStreamSynchronize(e->stream));
  uint8_t* p = static_cast<uint8_t*>(ptr);
  dfree(e, p);
  return CSM_OK;
}

int csm_host_alloc(csm_engine* e, size_t bytes, void** ptr) {
  if (!e || !ptr) return CSM_E_ARG;
  HIP_TRY(e, hipSetDevice(e->dev));
  HIP_TRY(e, hipHostMalloc(ptr, std::max<size_t>(bytes, 64)));
  return CSM_OK;
}

int csm_host_free(csm_engine* e, void* ptr) {
  if (!e) return CSM_E_ARG;
  if (ptr) HIP_TRY(e, hipHostFree(ptr));
  return CSM_OK;
}

int csm_copy_to_device(csm_engine* e, void* dst, const void* src, size_t bytes) {
  if (!e || (bytes && (!dst || !src))) return CSM_E_ARG;
  HIP_TRY(e, hipSetDevice(e->dev));
  HIP_TRY(e, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, e->stream));
  HIP_TRY(e, hipStreamSynchronize(e->stream));
  return CSM_OK;
}

int csm_copy_to_host(csm_engine* e, void* dst, const void* src, size_t bytes) {
  if (!e || (bytes && (!dst || !src))) return CSM_E_ARG;
  HIP_TRY(e, hipSetDevice(e->dev));
  HIP_TRY(e, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, e->stream));
  HIP_TRY(e, hipStreamSynchronize(e->stream));
  return CSM_OK;
}

int csm_parse_text_device(csm_engine* e, int mode, const char* d_text, size_t len,
                          csm_update* d_out, size_t cap, size_t* n, uint64_t* err_line) {
  if (!e || !n || (len && !d_text)) return CSM_E_ARG;
  HIP_TRY(e, hipSetDevice(e->dev));
  std::string msg;
  int rc = parse_text_on(
      e->stream, mode, reinterpret_cast<const uint8_t*>(d_text), len, n, err_line,
      [&](size_t b) -> void* { uint8_t* p = nullptr; return dalloc(e, &p, b) ? nullptr : p; },
      [&](void* p) { uint8_t* q = static_cast<uint8_t*>(p); dfree(e, q); },
      [&](const UpdateRec* recs, size_t cnt) -> int {
        if (cnt > cap) return CSM_E_CAPACITY;
        if (cnt && hipMemcpyAsync(d_out, recs, cnt * sizeof(UpdateRec), hipMemcpyDeviceToDevice,
                                  e->stream) != hipSuccess)
          return CSM_E_DEVICE;
        return CSM_OK;
      },
      msg);
  if (rc) return fail(e, rc, "%s", msg.c_str());
  return CSM_OK;
}

int csm_parse_text_device_static(csm_engine* e, int mode, const char* d_text, size_t len,
                                 csm_update* d_out, size_t cap, size_t* n, uint64_t* err_line) {
  if (!e || !n || (len && !d_text)) return CSM_E_ARG;
  HIP_TRY(e, hipSetDevice(e->dev));
  return parse_static(e, mode, reinterpret_cast<const uint8_t*>(d_text), len, d_out, cap, n,
                      err_line);
}

int csm_parse_text_to_device(int device, int mode, const char* text, size_t len,
                             csm_update** d_records, size_t* n, uint64_t* err_line) {
  if (!n || !d_records || (len && !text)) return CSM_E_ARG;
  *n = 0;
  *d_records = nullptr;
  if (hipSetDevice(device) != hipSuccess) return CSM_E_DEVICE;
  hipStream_t st = nullptr;
  if (hipStreamCreate(&st) != hipSuccess) return CSM_E_DEVICE;
  auto alloc = [](size_t b) -> void* { void* p = nullptr; return hipMalloc(&p, b) == hipSuccess ? p : nullptr; };
  auto release = [](void* p) { (void)hipFree(p); };
  int rc = CSM_OK;
  uint8_t* d_text = static_cast<uint8_t*>(len ? alloc(len) : nullptr);
  if (len && !d_text) rc = CSM_E_DEVICE;
  if (!rc && len && hipMemcpyAsync(d_text, text, len, hipMemcpyHostToDevice, st) != hipSuccess)
    rc = CSM_E_DEVICE;
  std::string msg;
  if (!rc)
    rc = parse_text_on(st, mode, d_text, len, n, err_line, alloc, release,
                       [&](const UpdateRec* recs, size_t cnt) -> int {
                         void* p = alloc(std::max<size_t>(cnt, 1) * sizeof(UpdateRec));
                         if (!p) return CSM_E_DEVICE;
                         if (cnt && hipMemcpyAsync(p, recs, cnt * sizeof(UpdateRec),
                                                   hipMemcpyDeviceToDevice, st) != hipSuccess) {
                           release(p);
                           return CSM_E_DEVICE;
                         }
                         *d_records = static_cast<csm_update*>(p);
                         return CSM_OK;
                       },
                       msg);
  if (d_text) release(d_text);
  (void)hipStreamSynchronize(st);
  (void)hipStreamDestroy(st);
  if (rc && *d_records) {
    release(*d_records);
    *d_records = nullptr;
  }
  return rc;
}

void csm_free_records(int device, csm_update* d_records) {
  if (!d_records || hipSetDevice(device) != hipSuccess) return;
  (void)hipFree(d_records);
}

int csm_parse_text(int device, int mode, const char* text, size_t len, csm_update* out,
                   size_t cap, size_t* n, uint64_t* err_line) {
  if (!n || (len && !text)) return CSM_E_ARG;
  *n = 0;
  if (hipSetDevice(device) != hipSuccess) return CSM_E_DEVICE;
  hipStream_t st = nullptr;
  if (hipStreamCreate(&st) != hipSuccess) return CSM_E_DEVICE;
  auto alloc = [](size_t b) -> void* { void* p = nullptr; return hipMalloc(&p, b) == hipSuccess ? p : nullptr; };
  auto release = [](void* p) { (void)hipFree(p); };
  int rc = CSM_OK;
  uint8_t* d_text = static_cast<uint8_t*>(len ? alloc(len) : nullptr);
  if (len && !d_text) rc = CSM_E_DEVICE;
  if (!rc && len && hipMemcpyAsync(d_text, text, len, hipMemcpyHostToDevice, st) != hipSuccess)
    rc = CSM_E_DEVICE;
  std::string msg;
  if (!rc)
    rc = parse_text_on(st, mode, d_text, len, n, err_line, alloc, release,
                       [&](const UpdateRec* recs, size_t cnt) -> int {
                         if (cnt > cap) return CSM_E_CAPACITY;
                         if (cnt && out &&
                             (hipMemcpyAsync(out, recs, cnt * sizeof(UpdateRec),
                                             hipMemcpyDeviceToHost, st) != hipSuccess ||
                              hipStreamSynchronize(st) != hipSuccess))
                           return CSM_E_DEVICE;
                         return CSM_OK;
                       },
                       msg);
  if (d_text) release(d_text);
  (void)hipStreamDestroy(st);
  return rc;
}

int csm_copy_device(csm_engine* e, void* dst, const void* src, size_t bytes) {
  if (!e || (bytes && (!dst || !src))) return CSM_E_ARG;
  HIP_TRY(e, hipSetDevice(e->dev));
  HIP_TRY(e, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, e->stream));
  HIP_TRY(e, hipStreamSynchronize(e->stream));
  return CSM_OK;
}

void* csm_stream(csm_engine* e) { return e ? (void*)e->stream : nullptr; }

// The transfer-stream entry points touch only the transfer stream and the
// marker events (never the engine's error text), so one host thread may call
// them while another runs a batch on the engine.
static hipStream_t stream_of(csm_engine* e, int which) {
  return which == CSM_STREAM_ENGINE ? e->stream : e->xfer;
}

int csm_copy_async(csm_engine* e, void* dst, const void* src, size_t bytes) {
  if (!e || (bytes && (!dst || !src))) return CSM_E_ARG;
  if (hipSetDevice(e->dev) != hipSuccess ||
      hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, e->xfer) != hipSuccess)
    return CSM_E_DEVICE;
  return CSM_OK;
}

int csm_mark(csm_engine* e, int stream, uint32_t slot) {
  if (!e || slot >= (uint32_t)CSM_MARKS || (stream != CSM_STREAM_ENGINE && stream != CSM_STREAM_TRANSFER))
    return CSM_E_ARG;
  if (hipSetDevice(e->dev) != hipSuccess ||
      hipEventRecord(e->marks[slot], stream_of(e, stream)) != hipSuccess)
    return CSM_E_DEVICE;
  return CSM_OK;
}

int csm_wait_mark(csm_engine* e, int stream, uint32_t slot) {
  if (!e || slot >= (uint32_t)CSM_MARKS || (stream != CSM_STREAM_ENGINE && stream != CSM_STREAM_TRANSFER))
    return CSM_E_ARG;
  if (hipSetDevice(e->dev) != hipSuccess ||
      hipStreamWaitEvent(stream_of(e, stream), e->marks[slot], 0) != hipSuccess)
    return CSM_E_DEVICE;
  return CSM_OK;
}

int csm_sync_mark(csm_engine* e, uint32_t slot) {
  if (!e || slot >= (uint32_t)CSM_MARKS) return CSM_E_ARG;
  if (hipEventSynchronize(e->marks[slot]) != hipSuccess) return CSM_E_DEVICE;
  return CSM_OK;
}

int csm_mem_info(csm_engine* e, uint64_t* free_bytes, uint64_t* total_bytes) {
  if (!e) return CSM_E_ARG;
  size_t f = 0, t = 0;
  if (hipSetDevice(e->dev) != hipSuccess || hipMemGetInfo(&f, &t) != hipSuccess) return CSM_E_DEVICE;
  if (free_bytes) *free_bytes = f;
  if (total_bytes) *total_bytes = t;
  return CSM_OK;
}

const char* csm_last_error(const csm_engine* e) { return e ? e->err.c_str() : "null engine"; }

void csm_destroy(csm_engine* e) {
  if (!e) return;
  (void)hipSetDevice(e->dev);
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  for (auto& g : e->graphs) graph_release(g);
  e->graphs.clear();
  for (auto& a : e->allocs) (void)hipFree(a.p);
  e->allocs.clear();
  if (e->h_ctrl) (void)hipHostFree(e->h_ctrl);
  if (e->h_counts) (void)hipHostFree(e->h_counts);
  if (e->h_hashes) (void)hipHostFree(e->h_hashes);
  if (e->h_safe) (void)hipHostFree(e->h_safe);
  if (e->h_upd) (void)hipHostFree(e->h_upd);
  if (e->h_ustatus_pin) (void)hipHostFree(e->h_ustatus_pin);
  if (e->h_uwork_pin) (void)hipHostFree(e->h_uwork_pin);
  if (e->h_abort) (void)hipHostFree(e->h_abort);
  if (e->h_parse) (void)hipHostFree(e->h_parse);
  for (auto& pair : e->ev_round)
    for (auto& ev : pair)
      if (ev) (void)hipEventDestroy(ev);
  if (e->xfer) (void)hipStreamSynchronize(e->xfer);
  for (auto& ev : e->marks)
    if (ev) (void)hipEventDestroy(ev);
  if (e->xfer) (void)hipStreamDestroy(e->xfer);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
}

}  // extern "C"
