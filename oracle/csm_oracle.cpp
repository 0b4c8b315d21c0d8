// CPU oracle for the incremental match search — TEST INFRASTRUCTURE ONLY.
//
// This file restates, on the CPU, the reference's per-update semantics so the
// HIP engine can be checked against it.  Only tests, the smoke check and the
// benchmark's CPU-baseline leg load it; the engine never does.
//
// What it restates (all paths relative to the reference repository):
//   Graph store and mutation  ParaCOSM/CSM/graph/graph.cpp:24-116
//   GraphFlow orders          ParaCOSM/CSM/matching/GraphFlow/graphflow.cpp:36-160
//   GraphFlow search          graphflow.cpp:181-265, AddEdge :289-355,
//                             RemoveEdge :357-416, InitialMatching :268-283
//   Parallel GraphFlow shape  ParaCOSM/CSM/matching/Parallel_GraphFlow/
//                             parallel_graphflow.cpp:780-992 (serial depth-2
//                             expansion, depth-3 task vector, omp parallel for),
//                             :1030-1100 and :1362-1448 (thread-local DFS).  The
//                             edge-label bug at :878 is NOT reproduced: the
//                             correct label of the chosen backward neighbour is
//                             used, as in the sequential plugin.
//   SymBi DAG/DCS/search      ParaCOSM/CSM/matching/SymBi/symbi.cpp:113-965
//   Classify (label filter)   parallel_graphflow.cpp:1990-2016
//
// Parity pinning (see tests/test_oracle.py): brute-force enumeration on tiny
// graphs (pos = |Emb(G+e)| - |Emb(G)|), closed forms (triangles, 2-paths,
// stars), the telescoping identity against InitialMatching, GraphFlow == SymBi
// per update, and the SymBi index fix-point (incremental == from scratch).
// SymBi's search-time order rule (symbi.cpp:474-506) and its per-update work
// (orc_symbi_work) are pinned by hand-worked cases in
// tests/test_oracle_order_rule.py.
//
// The DCS lists of SymBi are not stored: DCS[e(u->u')][v] is exactly the set of
// neighbours w of v with label L(u') reached over an edge labelled l(u,u')
// (symbi.cpp:404-529 builds it that way), so it is read from the graph.  To
// reproduce the reference's per-query-edge insertion/erasure order inside one
// AddEdge/RemoveEdge, the updated edge is hidden from or shown to DCS scans
// according to which (query edge, endpoint) pairs have been processed so far.

#include <algorithm>
#include <array>
#include <chrono>
#include <climits>
#include <cstdint>
#include <cstring>
#include <deque>
#include <string>
#include <unordered_map>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr uint32_t NOT_EXIST = UINT32_MAX;

inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

// Order-independent embedding-set hash: sum over embeddings of a fold of the
// mapping in query-vertex order.
inline uint64_t emb_hash(const uint32_t* m, uint32_t k) {
  uint64_t h = 0;
  for (uint32_t i = 0; i < k; ++i) h = mix64(h + (uint64_t)m[i] + 1);
  return h;
}

// ------------------------------------------------------------------ graph
// ParaCOSM/CSM/graph/graph.cpp:24-116: sorted neighbour vectors with a
// parallel edge-label vector per vertex.
struct Graph {
  std::vector<uint32_t> vl;
  std::vector<std::vector<uint32_t>> nb, el;
  size_t ne = 0;

  void add_vertex(uint32_t id, uint32_t label) {  // graph.cpp:24-44
    if (id >= vl.size()) {
      vl.resize(id + 1, NOT_EXIST);
      vl[id] = label;
      nb.resize(id + 1);
      el.resize(id + 1);
    } else if (vl[id] == NOT_EXIST) {
      vl[id] = label;
    }
  }
  bool add_edge(uint32_t a, uint32_t b, uint32_t l) {  // graph.cpp:58-93
    auto& na = nb[a];
    auto it = std::lower_bound(na.begin(), na.end(), b);
    if (it != na.end() && *it == b) return false;
    size_t d = it - na.begin();
    na.insert(it, b);
    el[a].insert(el[a].begin() + d, l);
    auto& nbb = nb[b];
    it = std::lower_bound(nbb.begin(), nbb.end(), a);
    d = it - nbb.begin();
    nbb.insert(it, a);
    el[b].insert(el[b].begin() + d, l);
    ++ne;
    return true;
  }
  int remove_edge(uint32_t a, uint32_t b) {  // graph.cpp:95-116
    auto& na = nb[a];
    auto it = std::lower_bound(na.begin(), na.end(), b);
    if (it == na.end() || *it != b) return -1;
    size_t d = it - na.begin();
    na.erase(it);
    el[a].erase(el[a].begin() + d);
    auto& nbb = nb[b];
    it = std::lower_bound(nbb.begin(), nbb.end(), a);
    if (it == nbb.end() || *it != a) return -1;
    d = it - nbb.begin();
    nbb.erase(it);
    el[b].erase(el[b].begin() + d);
    --ne;
    return 0;
  }
  uint32_t nv() const { return (uint32_t)vl.size(); }
  uint32_t label(uint32_t v) const { return vl[v]; }
  uint32_t deg(uint32_t v) const { return (uint32_t)nb[v].size(); }
  // label of edge (a,b), NOT_EXIST if absent
  uint32_t elabel(uint32_t a, uint32_t b) const {
    const auto& n = nb[a];
    auto it = std::lower_bound(n.begin(), n.end(), b);
    if (it == n.end() || *it != b) return NOT_EXIST;
    return el[a][it - n.begin()];
  }
};

struct Opts {
  bool homo = false;
  bool want_hash = false;
  // per-update result cap (0: none): the search of an update stops once its
  // count reaches the cap, and the count reported is min(exact, cap); the
  // reference checks max_num_results_ after each candidate
  // (parallel_graphflow.cpp:386) or seed (graphflow.cpp:328, 349) and keeps
  // the count it had reached
  uint64_t max_results = 0;
  // time limit of one update's search (steady-clock seconds; 0: none): the
  // reference's reach_time_limit polls (graphflow.cpp:329, 350); once it
  // passes, the search unwinds and the update is reported unfinished
  double deadline = 0;
};

inline double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// ------------------------------------------------------------- GraphFlow
// ParaCOSM/CSM/matching/GraphFlow/graphflow.cpp
struct GraphFlow {
  const Graph& q;
  Graph& d;
  Opts o;
  uint32_t k;
  std::vector<std::vector<uint32_t>> ovs, ocsr, ooff;  // order_vs_/csrs_/offs_
  std::vector<uint8_t> visited;
  mutable uint64_t polls = 0;
  mutable bool timed_out = false;

  GraphFlow(const Graph& q_, Graph& d_, Opts o_) : q(q_), d(d_), o(o_), k(q_.nv()) {
    uint32_t ne = (uint32_t)q.ne;
    ovs.assign(ne, std::vector<uint32_t>(k, 0));
    ocsr.assign(ne, std::vector<uint32_t>(ne + 1, 0));
    ooff.assign(ne, std::vector<uint32_t>(k, 0));
    visited.assign(d.nv(), 0);
    gen_orders();
  }

  // greedy next vertex: most already-ordered neighbours, tie -> higher degree
  // (graphflow.cpp:55-80, :116-141)
  uint32_t pick_next(const std::vector<uint8_t>& vis) const {
    uint32_t best = NOT_EXIST, best_adj = 0;
    for (uint32_t j = 0; j < k; ++j) {
      if (vis[j]) continue;
      uint32_t adj = 0;
      for (uint32_t w : q.nb[j]) adj += vis[w] ? 1 : 0;
      if (!adj) continue;
      if (best == NOT_EXIST || (adj == best_adj && q.deg(j) > q.deg(best)) || adj > best_adj) {
        best_adj = adj;
        best = j;
      }
    }
    return best;
  }

  void gen_orders() {  // graphflow.cpp:36-160
    uint32_t ne = (uint32_t)q.ne;
    if (ne == 0) return;
    std::vector<uint8_t> vis(k, 0);
    uint32_t maxd = 0;
    for (uint32_t i = 0; i < k; ++i)
      if (q.deg(i) > maxd) { maxd = q.deg(i); ovs[0][0] = i; }
    vis[ovs[0][0]] = 1;
    for (uint32_t i = 1; i < k; ++i) {
      uint32_t u = pick_next(vis);
      ovs[0][i] = u;
      vis[u] = 1;
      ooff[0][i] = ooff[0][i - 1];
      for (uint32_t w : q.nb[u])
        if (vis[w]) ocsr[0][ooff[0][i]++] = w;
    }
    for (uint32_t i = 1; i < ne; ++i) {
      std::vector<uint8_t> v2(k, 0);
      auto it = std::lower_bound(ooff[0].begin(), ooff[0].end(), i + 1);
      ovs[i][0] = ovs[0][it - ooff[0].begin()];
      ovs[i][1] = ocsr[0][i];
      ocsr[i][0] = ovs[i][0];
      v2[ovs[i][0]] = 1;
      v2[ovs[i][1]] = 1;
      if (k > 2) ooff[i][2] = 1;
      ooff[i][1] = 1;
      for (uint32_t j = 2; j < k; ++j) {
        uint32_t u = pick_next(v2);
        ovs[i][j] = u;
        v2[u] = 1;
        ooff[i][j] = ooff[i][j - 1];
        for (uint32_t w : q.nb[u])
          if (v2[w]) ocsr[i][ooff[i][j]++] = w;
      }
    }
  }

  // graphflow.cpp:181-265 (visited via the m array when visited is null)
  void find(uint32_t oi, uint32_t depth, std::vector<uint32_t>& m, uint64_t& cnt,
            uint64_t& h, std::vector<uint8_t>& vis) const {
    if (o.deadline > 0) {
      if (timed_out) return;
      if ((++polls & 4095) == 0 && now_s() >= o.deadline) { timed_out = true; return; }
    }
    uint32_t u = ovs[oi][depth];
    uint32_t u_min = NOT_EXIST, u_min_label = NOT_EXIST, u_min_size = UINT32_MAX;
    const auto& qn = q.nb[u];
    const auto& ql = q.el[u];
    for (size_t i = 0; i < qn.size(); ++i) {
      if (m[qn[i]] == NOT_EXIST) continue;
      uint32_t s = d.deg(m[qn[i]]);
      if (s < u_min_size) { u_min_size = s; u_min = qn[i]; u_min_label = ql[i]; }
    }
    const auto& cn = d.nb[m[u_min]];
    const auto& cl = d.el[m[u_min]];
    uint32_t ulab = q.label(u);
    for (size_t i = 0; i < cn.size(); ++i) {
      uint32_t v = cn[i];
      if (d.label(v) != ulab || cl[i] != u_min_label) continue;
      bool ok = true;
      for (size_t j = 0; j < qn.size() && ok; ++j) {
        uint32_t uo = qn[j];
        if (m[uo] == NOT_EXIST || uo == u_min) continue;
        if (d.elabel(m[uo], v) != ql[j]) ok = false;
      }
      if (!ok) continue;
      if (!o.homo && vis[v]) continue;
      m[u] = v;
      vis[v] = 1;
      if (depth == k - 1) {
        ++cnt;
        if (o.want_hash) h += emb_hash(m.data(), k);
      } else {
        find(oi, depth + 1, m, cnt, h, vis);
      }
      vis[v] = 0;
      m[u] = NOT_EXIST;
      if (o.max_results && cnt >= o.max_results) return;
      if (timed_out) return;
    }
  }

  // seeds of one update: (order, m[u1], m[u2]) for each matching orientation
  // (graphflow.cpp:300-352)
  void seeds(uint32_t v1, uint32_t v2, uint32_t l,
             std::vector<std::array<uint32_t, 3>>& out) const {
    for (uint32_t i = 0; i < (uint32_t)q.ne; ++i) {
      uint32_t u1 = ovs[i][0], u2 = ovs[i][1];
      uint32_t ql = q.elabel(u1, u2);
      if (q.label(u1) == d.label(v1) && q.label(u2) == d.label(v2) && ql == l)
        out.push_back({i, v1, v2});
      if (q.label(u1) == d.label(v2) && q.label(u2) == d.label(v1) && ql == l)
        out.push_back({i, v2, v1});
    }
  }

  void search_update(uint32_t v1, uint32_t v2, uint32_t l, uint64_t& cnt, uint64_t& h) {
    std::vector<std::array<uint32_t, 3>> sd;
    seeds(v1, v2, l, sd);
    std::vector<uint32_t> m(k, NOT_EXIST);
    for (auto& s : sd) {
      uint32_t u1 = ovs[s[0]][0], u2 = ovs[s[0]][1];
      m[u1] = s[1];
      m[u2] = s[2];
      visited[s[1]] = visited[s[2]] = 1;
      if (k == 2) {
        ++cnt;
        if (o.want_hash) h += emb_hash(m.data(), k);
      } else {
        find(s[0], 2, m, cnt, h, visited);
      }
      visited[s[1]] = visited[s[2]] = 0;
      m[u1] = m[u2] = NOT_EXIST;
      if (o.max_results && cnt >= o.max_results) break;
      if (timed_out) break;
    }
    if (o.max_results && cnt > o.max_results) cnt = o.max_results;
  }

  int add_edge(uint32_t v1, uint32_t v2, uint32_t l, uint64_t& cnt, uint64_t& h) {
    d.add_edge(v1, v2, l);  // graphflow.cpp:296
    search_update(v1, v2, l, cnt, h);
    return 0;
  }
  int remove_edge(uint32_t v1, uint32_t v2, uint64_t& cnt, uint64_t& h) {
    uint32_t l = d.elabel(v1, v2);  // graphflow.cpp:361
    if (l == NOT_EXIST) return -2;
    search_update(v1, v2, l, cnt, h);
    return d.remove_edge(v1, v2) == 0 ? 0 : -2;
  }
  uint64_t initial(uint64_t& h) {  // graphflow.cpp:268-283
    uint64_t cnt = 0;
    if (k == 0) return 0;
    std::vector<uint32_t> m(k, NOT_EXIST);
    uint32_t u0 = ovs.empty() ? 0 : ovs[0][0];
    for (uint32_t v = 0; v < d.nv(); ++v) {
      if (d.label(v) == NOT_EXIST || d.label(v) != q.label(u0)) continue;
      m[u0] = v;
      visited[v] = 1;
      if (k == 1) { ++cnt; if (o.want_hash) h += emb_hash(m.data(), k); }
      else find(0, 1, m, cnt, h, visited);
      visited[v] = 0;
      m[u0] = NOT_EXIST;
    }
    return cnt;
  }

  // ---- the reference's parallel shape (parallel_graphflow.cpp:780-992)
  // depth-2 expanded serially, one task per (depth-2 match, depth-3 neighbour
  // index of the chosen backward neighbour), then an OpenMP loop over tasks
  // each running a thread-local DFS (:1030-1100, :1362-1448).
  struct Task { uint32_t v2match; uint32_t i3; };

  void find_local(uint32_t oi, uint32_t depth, uint32_t* m, uint64_t& cnt) const {
    uint32_t u = ovs[oi][depth];
    uint32_t u_min = NOT_EXIST, u_min_label = NOT_EXIST, u_min_size = UINT32_MAX;
    const auto& qn = q.nb[u];
    const auto& ql = q.el[u];
    for (size_t i = 0; i < qn.size(); ++i) {
      if (m[qn[i]] == NOT_EXIST) continue;
      uint32_t s = d.deg(m[qn[i]]);
      if (s < u_min_size) { u_min_size = s; u_min = qn[i]; u_min_label = ql[i]; }
    }
    const auto& cn = d.nb[m[u_min]];
    const auto& cl = d.el[m[u_min]];
    for (size_t i = 0; i < cn.size(); ++i)
      try_candidate(oi, depth, u, u_min, u_min_label, cn[i], cl[i], m, cnt);
  }

  inline void try_candidate(uint32_t oi, uint32_t depth, uint32_t u, uint32_t u_min,
                            uint32_t u_min_label, uint32_t v, uint32_t vel, uint32_t* m,
                            uint64_t& cnt) const {
    if (d.label(v) != q.label(u) || vel != u_min_label) return;
    const auto& qn = q.nb[u];
    const auto& ql = q.el[u];
    for (size_t j = 0; j < qn.size(); ++j) {
      uint32_t uo = qn[j];
      if (m[uo] == NOT_EXIST || uo == u_min) continue;
      if (d.elabel(m[uo], v) != ql[j]) return;
    }
    if (!o.homo)
      for (uint32_t j = 0; j < k; ++j)
        if (m[j] == v) return;
    if (depth == k - 1) { ++cnt; return; }
    m[u] = v;
    find_local(oi, depth + 1, m, cnt);
    m[u] = NOT_EXIST;
  }

  uint64_t parallel_seed(uint32_t oi, std::vector<uint32_t>& m, int threads) const {
    uint64_t total = 0;
    if (k == 2) return 1;
    uint32_t u = ovs[oi][2];
    uint32_t u_min = NOT_EXIST, u_min_label = NOT_EXIST, u_min_size = UINT32_MAX;
    for (size_t i = 0; i < q.nb[u].size(); ++i) {
      uint32_t uo = q.nb[u][i];
      if (m[uo] == NOT_EXIST) continue;
      uint32_t s = d.deg(m[uo]);
      if (s < u_min_size) { u_min_size = s; u_min = uo; u_min_label = q.el[u][i]; }
    }
    if (k == 3) {
      const auto& cn = d.nb[m[u_min]];
      for (size_t i = 0; i < cn.size(); ++i)
        try_candidate(oi, 2, u, u_min, u_min_label, cn[i], d.el[m[u_min]][i], m.data(), total);
      return total;
    }
    // serial depth-2 expansion -> tasks for depth 3
    std::vector<Task> tasks;
    std::vector<uint32_t> task_umin, task_umin_label;
    uint32_t u3 = ovs[oi][3];
    const auto& cn = d.nb[m[u_min]];
    const auto& cl = d.el[m[u_min]];
    std::vector<uint32_t> v2s;
    std::vector<std::array<uint32_t, 2>> umin3;  // (u_min3, label) per depth-2 match
    for (size_t i = 0; i < cn.size(); ++i) {
      uint32_t v = cn[i];
      if (d.label(v) != q.label(u) || cl[i] != u_min_label) continue;
      bool ok = true;
      for (size_t j = 0; j < q.nb[u].size() && ok; ++j) {
        uint32_t uo = q.nb[u][j];
        if (m[uo] == NOT_EXIST || uo == u_min) continue;
        if (d.elabel(m[uo], v) != q.el[u][j]) ok = false;
      }
      if (!ok) continue;
      if (!o.homo) {
        bool used = false;
        for (uint32_t j = 0; j < k; ++j) used |= (m[j] == v);
        if (used) continue;
      }
      m[u] = v;
      uint32_t um = NOT_EXIST, uml = NOT_EXIST, ums = UINT32_MAX;
      for (size_t j = 0; j < q.nb[u3].size(); ++j) {
        uint32_t uo = q.nb[u3][j];
        if (m[uo] == NOT_EXIST) continue;
        uint32_t s = d.deg(m[uo]);
        if (s < ums) { ums = s; um = uo; uml = q.el[u3][j]; }
      }
      uint32_t idx = (uint32_t)v2s.size();
      v2s.push_back(v);
      umin3.push_back({um, uml});
      for (uint32_t i3 = 0; i3 < d.deg(m[um]); ++i3) tasks.push_back({idx, i3});
      m[u] = NOT_EXIST;
    }
    int nt = threads;
    if ((size_t)nt > tasks.size()) nt = tasks.empty() ? 1 : (int)tasks.size();
#pragma omp parallel num_threads(nt) reduction(+ : total)
    {
      std::vector<uint32_t> ml(m);
#pragma omp for schedule(dynamic, 64)
      for (size_t t = 0; t < tasks.size(); ++t) {
        const Task& tk = tasks[t];
        ml[u] = v2s[tk.v2match];
        uint32_t um = umin3[tk.v2match][0], uml = umin3[tk.v2match][1];
        uint32_t base = ml[um];
        try_candidate(oi, 3, u3, um, uml, d.nb[base][tk.i3], d.el[base][tk.i3], ml.data(), total);
        ml[u] = NOT_EXIST;
      }
    }
    return total;
  }

  uint64_t parallel_update(uint32_t v1, uint32_t v2, uint32_t l, int threads) const {
    std::vector<std::array<uint32_t, 3>> sd;
    seeds(v1, v2, l, sd);
    std::vector<uint32_t> m(k, NOT_EXIST);
    uint64_t cnt = 0;
    for (auto& s : sd) {
      uint32_t u1 = ovs[s[0]][0], u2 = ovs[s[0]][1];
      m[u1] = s[1];
      m[u2] = s[2];
      cnt += parallel_seed(s[0], m, threads);
      m[u1] = m[u2] = NOT_EXIST;
    }
    return cnt;
  }
};

// ------------------------------------------------------------------ SymBi
// ParaCOSM/CSM/matching/SymBi/symbi.cpp, with dense per-(query edge, vertex)
// counters in place of the hash maps.
struct SymBi {
  const Graph& q;
  Graph& d;
  Opts o;
  uint32_t k, nde;  // query vertices, directed query edges
  std::vector<std::vector<uint32_t>> eidx;  // eidx_[u][u'] (symbi.cpp:36-47)
  struct Node { std::vector<uint32_t> fw, fwl, bw, bwl, nbrs; };
  std::vector<Node> tn;
  uint32_t root = 0;
  std::vector<uint32_t> ser;  // serialized_tree_
  std::vector<std::vector<uint8_t>> d1, d2;
  std::vector<std::vector<uint32_t>> n1, n2, np1, nc2;
  std::deque<std::pair<uint32_t, uint32_t>> Q1, Q2;
  std::vector<uint8_t> visited;
  // search work: DCS entries the searches walked (one per candidate of the
  // vertex the ExtendableVertex rule picked, at every node of the search tree)
  uint64_t work = 0;
  // updated edge and the (directed query edge, from-vertex) pairs whose DCS
  // lists currently hold it (insert) / no longer hold it (delete)
  uint32_t pe_a = NOT_EXIST, pe_b = NOT_EXIST;
  bool pe_insert = false;
  std::vector<std::pair<uint32_t, uint32_t>> pe_done;

  // tree = true: TurboFlux's spanning tree in place of SymBi's DAG (the index
  // is the same construction over it).  dag_src: the graph whose label
  // frequencies choose the root and the tree (default: d).
  SymBi(const Graph& q_, Graph& d_, Opts o_, bool tree = false, const Graph* dag_src = nullptr)
      : q(q_), d(d_), o(o_), k(q_.nv()) {
    eidx.assign(k, std::vector<uint32_t>(k, NOT_EXIST));
    uint32_t pos = 0;
    for (uint32_t i = 0; i < k; ++i)
      for (uint32_t w : q.nb[i]) eidx[i][w] = pos++;
    nde = pos;
    tn.assign(k, Node());
    ser.assign(k, 0);
    visited.assign(d.nv(), 0);
    build_dag(tree, dag_src ? *dag_src : d);
    alloc();
    build_dcs(d1, d2, n1, n2, np1, nc2);
  }

  void alloc() {
    uint32_t V = d.nv();
    d1.assign(k, std::vector<uint8_t>(V, 0));
    d2.assign(k, std::vector<uint8_t>(V, 0));
    np1.assign(k, std::vector<uint32_t>(V, 0));
    nc2.assign(k, std::vector<uint32_t>(V, 0));
    n1.assign(nde, std::vector<uint32_t>(V, 0));
    n2.assign(nde, std::vector<uint32_t>(V, 0));
  }

  // symbi.cpp:113-244; with `tree`, parallel_turboflux.cpp:123-320
  void build_dag(bool tree, const Graph& g) {
    std::vector<std::vector<uint64_t>> freq(k, std::vector<uint64_t>(k, 0));
    uint64_t min_freq = UINT64_MAX;
    uint32_t mv0 = 0, mv1 = 0;
    for (uint32_t u = 0; u < k; ++u) {
      for (size_t i = 0; i < q.nb[u].size(); ++i) {
        uint32_t uo = q.nb[u][i], uol = q.el[u][i];
        if (u > uo) { freq[u][uo] = freq[uo][u]; continue; }
        for (uint32_t v = 0; v < g.nv(); ++v) {
          if (g.label(v) == NOT_EXIST) continue;
          if (g.label(v) != q.label(u)) continue;
          for (size_t j = 0; j < g.nb[v].size(); ++j)
            if (g.label(g.nb[v][j]) == q.label(uo) && g.el[v][j] == uol) ++freq[u][uo];
        }
        if (freq[u][uo] < min_freq) { min_freq = freq[u][uo]; mv0 = u; mv1 = uo; }
      }
    }
    uint64_t f0 = 0, f1 = 0;
    for (uint32_t v = 0; v < g.nv(); ++v) {
      if (g.label(v) == NOT_EXIST) continue;
      if (g.label(v) == q.label(mv0)) ++f0;
      if (g.label(v) == q.label(mv1)) ++f1;
    }
    if (f0 > f1) root = mv1;
    else if (f0 < f1) root = mv0;
    else if (q.deg(mv0) > q.deg(mv1)) root = mv0;
    else root = mv1;
    if (tree) { build_tree(freq); return; }
    std::vector<uint8_t> vis(k, 0), on(k, 0);
    std::deque<uint32_t> bfs{root};
    on[root] = 1;
    uint32_t p = 0;
    ser[p++] = root;
    while (!bfs.empty()) {
      uint32_t u = bfs.front();
      bfs.pop_front();
      vis[u] = 1;
      for (size_t j = 0; j < q.nb[u].size(); ++j) {
        uint32_t uo = q.nb[u][j], uol = q.el[u][j];
        if (!on[uo]) { bfs.push_back(uo); on[uo] = 1; ser[p++] = uo; }
        if (!vis[uo]) { tn[u].fw.push_back(uo); tn[u].fwl.push_back(uol); }
        else { tn[u].bw.push_back(uo); tn[u].bwl.push_back(uol); }
        tn[u].nbrs.push_back(uo);
      }
    }
  }

  // TurboFlux's spanning tree (parallel_turboflux.cpp:239-320): level by
  // level from the root, repeatedly take the least frequent edge from the
  // current level to an unvisited vertex (the first one on ties), drop every
  // queued edge into that vertex, and queue its edges for the next level.
  // The serialized tree is the BFS order over tree children.
  void build_tree(const std::vector<std::vector<uint64_t>>& freq) {
    struct E { uint32_t a, b; uint64_t f; };
    std::vector<uint8_t> vis(k, 0);
    vis[root] = 1;
    std::vector<E> level;
    for (uint32_t w : q.nb[root]) level.push_back({root, w, freq[root][w]});
    while (!level.empty()) {
      std::vector<E> next;
      while (!level.empty()) {
        size_t best = 0;
        for (size_t i = 1; i < level.size(); ++i)
          if (level[i].f < level[best].f) best = i;
        const E sel = level[best];
        std::vector<E> keep;
        for (const E& x : level)
          if (x.b != sel.b) keep.push_back(x);
        level.swap(keep);
        if (vis[sel.b]) continue;
        vis[sel.b] = 1;
        const uint32_t l = q.elabel(sel.a, sel.b);
        tn[sel.a].fw.push_back(sel.b);
        tn[sel.a].fwl.push_back(l);
        tn[sel.b].bw.push_back(sel.a);
        tn[sel.b].bwl.push_back(l);
        for (uint32_t w : q.nb[sel.b])
          if (!vis[w]) next.push_back({sel.b, w, freq[sel.b][w]});
      }
      level.swap(next);
    }
    for (uint32_t u = 0; u < k; ++u) tn[u].nbrs = q.nb[u];
    std::deque<uint32_t> fifo{root};
    uint32_t p = 0;
    while (!fifo.empty()) {
      uint32_t u = fifo.front();
      fifo.pop_front();
      ser[p++] = u;
      for (uint32_t c : tn[u].fw) fifo.push_back(c);
    }
  }

  // visit DCS[e(u->u')][v] = neighbours w of v with L(w)=L(u'), edge label l(u,u')
  template <class F>
  void dcs_for(uint32_t u, uint32_t up, uint32_t v, F&& f) const {
    uint32_t e = eidx[u][up];
    uint32_t lab = q.label(up), el = q.elabel(u, up);
    const auto& n = d.nb[v];
    const auto& l = d.el[v];
    for (size_t i = 0; i < n.size(); ++i) {
      uint32_t w = n[i];
      if (d.label(w) != lab || l[i] != el) continue;
      if ((v == pe_a && w == pe_b) || (v == pe_b && w == pe_a)) {
        bool done = false;
        for (auto& pr : pe_done) done |= (pr.first == e && pr.second == v);
        if (pe_insert ? !done : done) continue;
      }
      f(w);
    }
  }
  bool dcs_has(uint32_t u, uint32_t up, uint32_t v, uint32_t w) const {
    return d.elabel(v, w) == q.elabel(u, up) && d.label(w) == q.label(up);
  }

  void build_dcs(std::vector<std::vector<uint8_t>>& D1, std::vector<std::vector<uint8_t>>& D2,
                 std::vector<std::vector<uint32_t>>& N1, std::vector<std::vector<uint32_t>>& N2,
                 std::vector<std::vector<uint32_t>>& NP1,
                 std::vector<std::vector<uint32_t>>& NC2) const {  // symbi.cpp:246-377
    uint32_t V = d.nv();
    for (uint32_t i = 0; i < k; ++i) {  // top-down
      uint32_t u = ser[i];
      for (uint32_t v = 0; v < V; ++v) {
        if (d.label(v) != q.label(u)) continue;
        for (size_t b = 0; b < tn[u].bw.size(); ++b) {
          uint32_t up = tn[u].bw[b];
          uint32_t e = eidx[u][up];
          dcs_for(u, up, v, [&](uint32_t w) { if (D1[up][w]) N1[e][v] += 1; });
          if (N1[e][v]) NP1[u][v] += 1;
        }
        if (NP1[u][v] == tn[u].bw.size()) D1[u][v] = 1;
      }
    }
    for (uint32_t i = 0; i < k; ++i) {  // bottom-up
      uint32_t u = ser[k - 1 - i];
      for (uint32_t v = 0; v < V; ++v) {
        if (d.label(v) != q.label(u)) continue;
        for (size_t c = 0; c < tn[u].fw.size(); ++c) {
          uint32_t uc = tn[u].fw[c];
          uint32_t e = eidx[u][uc];
          dcs_for(u, uc, v, [&](uint32_t w) { if (D2[uc][w]) N2[e][v] += 1; });
          if (N2[e][v]) NC2[u][v] += 1;
        }
        if (NC2[u][v] == tn[u].fw.size() && D1[u][v] == 1) D2[u][v] = 1;
      }
    }
    for (uint32_t i = 0; i < k; ++i) {  // top-down n2 fix-up
      uint32_t u = ser[i];
      for (size_t b = 0; b < tn[u].bw.size(); ++b) {
        uint32_t up = tn[u].bw[b];
        uint32_t e = eidx[u][up];
        for (uint32_t v = 0; v < V; ++v) {
          if (d.label(v) != q.label(u)) continue;
          dcs_for(u, up, v, [&](uint32_t w) { if (D2[up][w]) N2[e][v] += 1; });
        }
      }
    }
  }

  // symbi.cpp:404-466
  void ins_td(uint32_t u, uint32_t uc, uint32_t v, uint32_t vc) {
    uint32_t e = eidx[uc][u];
    if (n1[e][vc] == 0) {
      np1[uc][vc] += 1;
      if (np1[uc][vc] == tn[uc].bw.size()) {
        d1[uc][vc] = 1;
        Q1.emplace_back(uc, vc);
        if (nc2[uc][vc] == tn[uc].fw.size()) {
          d2[uc][vc] = 1;
          Q2.emplace_back(uc, vc);
        }
      }
    }
    n1[e][vc] += 1;
  }
  void ins_bu(uint32_t u, uint32_t up, uint32_t v, uint32_t vp) {
    uint32_t e = eidx[up][u];
    if (n2[e][vp] == 0) {
      nc2[up][vp] += 1;
      if (d1[up][vp] && nc2[up][vp] == tn[up].fw.size()) {
        d2[up][vp] = 1;
        Q2.emplace_back(up, vp);
      }
    }
    n2[e][vp] += 1;
  }
  void del_td(uint32_t u, uint32_t uc, uint32_t v, uint32_t vc) {
    uint32_t e = eidx[uc][u];
    n1[e][vc] -= 1;
    if (n1[e][vc] == 0) {
      if (d1[uc][vc] == 1) {
        if (d2[uc][vc] == 1) { Q2.emplace_back(uc, vc); d2[uc][vc] = 0; }
        Q1.emplace_back(uc, vc);
        d1[uc][vc] = 0;
      }
      np1[uc][vc] -= 1;
    }
  }
  void del_bu(uint32_t u, uint32_t up, uint32_t v, uint32_t vp) {
    uint32_t e = eidx[up][u];
    n2[e][vp] -= 1;
    if (n2[e][vp] == 0) {
      if (d2[up][vp] == 1) { Q2.emplace_back(up, vp); d2[up][vp] = 0; }
      nc2[up][vp] -= 1;
    }
  }

  bool is_bw(uint32_t u, uint32_t x) const {
    return std::find(tn[u].bw.begin(), tn[u].bw.end(), x) != tn[u].bw.end();
  }

  // index maintenance of symbi.cpp:683-747 (insert) and :875-943 (delete)
  void index_update(uint32_t v1, uint32_t v2, uint32_t l, bool insert) {
    pe_a = v1; pe_b = v2; pe_insert = insert;
    pe_done.clear();
    for (uint32_t a = 0; a < k; ++a) {
      if (d.label(v1) != q.label(a)) continue;
      for (uint32_t b = 0; b < k; ++b) {
        if (d.label(v2) != q.label(b)) continue;
        if (q.elabel(a, b) != l) continue;
        uint32_t u1 = a, u2 = b, x1 = v1, x2 = v2;
        if (is_bw(u1, u2)) { std::swap(u1, u2); std::swap(x1, x2); }
        if (!is_bw(u2, u1)) continue;
        pe_done.push_back({eidx[u1][u2], x1});
        pe_done.push_back({eidx[u2][u1], x2});
        bool p_d1 = d1[u1][x1], p_d2 = d2[u1][x1], c_d2 = d2[u2][x2];
        if (insert) {
          if (p_d1) ins_td(u1, u2, x1, x2);
          if (c_d2) ins_bu(u2, u1, x2, x1);
          if (p_d2) n2[eidx[u2][u1]][x2] += 1;
          while (!Q1.empty()) {
            auto [uq, vq] = Q1.front();
            Q1.pop_front();
            for (uint32_t uc : tn[uq].fw)
              dcs_for(uq, uc, vq, [&](uint32_t w) { ins_td(uq, uc, vq, w); });
          }
          while (!Q2.empty()) {
            auto [uq, vq] = Q2.front();
            Q2.pop_front();
            for (uint32_t up : tn[uq].bw)
              dcs_for(uq, up, vq, [&](uint32_t w) { ins_bu(uq, up, vq, w); });
            for (uint32_t uc : tn[uq].fw)
              dcs_for(uq, uc, vq, [&](uint32_t w) { n2[eidx[uc][uq]][w] += 1; });
          }
        } else {
          if (c_d2) del_bu(u2, u1, x2, x1);
          if (p_d2) n2[eidx[u2][u1]][x2] -= 1;
          if (p_d1) del_td(u1, u2, x1, x2);
          while (!Q1.empty()) {
            auto [uq, vq] = Q1.front();
            Q1.pop_front();
            for (uint32_t uc : tn[uq].fw)
              dcs_for(uq, uc, vq, [&](uint32_t w) { del_td(uq, uc, vq, w); });
          }
          while (!Q2.empty()) {
            auto [uq, vq] = Q2.front();
            Q2.pop_front();
            for (uint32_t up : tn[uq].bw)
              dcs_for(uq, up, vq, [&](uint32_t w) { del_bu(uq, up, vq, w); });
            for (uint32_t uc : tn[uq].fw)
              dcs_for(uq, uc, vq, [&](uint32_t w) { n2[eidx[uc][uq]][w] -= 1; });
          }
        }
      }
    }
    pe_a = pe_b = NOT_EXIST;
    pe_done.clear();
  }

  struct Ext { uint32_t E = NOT_EXIST, matched = 0, u_min = NOT_EXIST; };

  void find(uint32_t depth, std::vector<uint32_t>& m, std::vector<Ext>& ext, uint64_t& cnt,
            uint64_t& h) {  // symbi.cpp:471-587
    uint32_t niso = NOT_EXIST, iso = NOT_EXIST, nisoE = NOT_EXIST, isoE = NOT_EXIST;
    for (uint32_t i = 0; i < k; ++i) {
      if (m[i] != NOT_EXIST) continue;
      if (ext[i].matched == q.nb[i].size()) {
        if (ext[i].E < isoE) { isoE = ext[i].E; iso = i; }
      } else {
        if (ext[i].E < nisoE) { nisoE = ext[i].E; niso = i; }
      }
    }
    uint32_t u = (nisoE == NOT_EXIST) ? iso : niso;
    uint32_t u_min = ext[u].u_min;
    ext[u] = Ext();
    std::vector<uint32_t> cands;
    dcs_for(u_min, u, m[u_min], [&](uint32_t w) { cands.push_back(w); });
    work += cands.size();
    for (uint32_t v : cands) {
      if (d2[u][v] == 0) continue;
      bool ok = true;
      for (uint32_t uo : tn[u].nbrs) {
        if (m[uo] == NOT_EXIST || uo == u_min) continue;
        if (!dcs_has(uo, u, m[uo], v)) { ok = false; break; }
      }
      if (!ok) continue;
      if (!o.homo && visited[v]) continue;
      m[u] = v;
      visited[v] = 1;
      std::vector<Ext> t(ext);
      for (uint32_t uo : tn[u].nbrs) {
        if (m[uo] != NOT_EXIST) continue;
        uint32_t val = n2[eidx[u][uo]][m[u]];
        if (val < t[uo].E) { t[uo].E = val; t[uo].u_min = u; }
        t[uo].matched += 1;
      }
      if (depth == k - 1) {
        ++cnt;
        if (o.want_hash) h += emb_hash(m.data(), k);
      } else {
        find(depth + 1, m, t, cnt, h);
      }
      visited[v] = 0;
      m[u] = NOT_EXIST;
    }
  }

  // symbi.cpp:749-813 / :817-873: seeds on DAG-oriented (u1->u2) with d2 on both
  void search(uint32_t v1, uint32_t v2, uint32_t l, bool require_dcs, uint64_t& cnt,
              uint64_t& h) {
    std::vector<uint32_t> m(k, NOT_EXIST);
    for (uint32_t a = 0; a < k; ++a) {
      if (d.label(v1) != q.label(a)) continue;
      for (uint32_t b = 0; b < k; ++b) {
        if (d.label(v2) != q.label(b)) continue;
        if (require_dcs) {
          if (!dcs_has(a, b, v1, v2)) continue;
        } else if (q.elabel(a, b) != l) {
          continue;
        }
        uint32_t u1 = a, u2 = b, x1 = v1, x2 = v2;
        if (is_bw(u1, u2)) { std::swap(u1, u2); std::swap(x1, x2); }
        if (!(is_bw(u2, u1) && d2[u1][x1] == 1 && d2[u2][x2] == 1)) continue;
        m[u1] = x1;
        m[u2] = x2;
        visited[x1] = visited[x2] = 1;
        std::vector<Ext> ext(k);
        for (uint32_t uu : {u1, u2}) {
          for (uint32_t uo : tn[uu].nbrs) {
            if (m[uo] != NOT_EXIST) continue;
            uint32_t val = n2[eidx[uu][uo]][m[uu]];
            if (val < ext[uo].E) { ext[uo].E = val; ext[uo].u_min = uu; }
            ext[uo].matched += 1;
          }
        }
        if (k == 2) {
          ++cnt;
          if (o.want_hash) h += emb_hash(m.data(), k);
        } else {
          find(2, m, ext, cnt, h);
        }
        visited[x1] = visited[x2] = 0;
        m[u1] = m[u2] = NOT_EXIST;
      }
    }
  }

  int add_edge(uint32_t v1, uint32_t v2, uint32_t l, uint64_t& cnt, uint64_t& h) {
    d.add_edge(v1, v2, l);
    index_update(v1, v2, l, true);
    search(v1, v2, l, false, cnt, h);
    return 0;
  }
  int remove_edge(uint32_t v1, uint32_t v2, uint64_t& cnt, uint64_t& h) {
    uint32_t l = d.elabel(v1, v2);
    if (l == NOT_EXIST) return -2;
    search(v1, v2, l, true, cnt, h);
    index_update(v1, v2, l, false);
    return d.remove_edge(v1, v2) == 0 ? 0 : -2;
  }
  uint64_t initial(uint64_t& h) {  // symbi.cpp:379-402
    uint64_t cnt = 0;
    if (k == 0) return 0;
    std::vector<uint32_t> m(k, NOT_EXIST);
    if (q.nb[root].empty()) {
      for (uint32_t v = 0; v < d.nv(); ++v)
        if (d.label(v) == q.label(root)) { ++cnt; if (o.want_hash) { m[root] = v; h += emb_hash(m.data(), k); } }
      return cnt;
    }
    for (uint32_t v = 0; v < d.nv(); ++v) {
      if (d.label(v) != q.label(root) || !d2[root][v]) continue;
      m[root] = v;
      visited[v] = 1;
      std::vector<Ext> ext(k);
      for (uint32_t uo : tn[root].nbrs) {
        uint32_t val = n2[eidx[root][uo]][v];
        if (val < ext[uo].E) { ext[uo].E = val; ext[uo].u_min = root; }
        ext[uo].matched += 1;
      }
      if (k == 1) ++cnt;
      else find(1, m, ext, cnt, h);
      visited[v] = 0;
      m[root] = NOT_EXIST;
    }
    return cnt;
  }

  // index fix-point check: incremental state == from-scratch BuildDCS
  bool index_consistent() const {
    uint32_t V = d.nv();
    std::vector<std::vector<uint8_t>> D1(k, std::vector<uint8_t>(V, 0)), D2 = D1;
    std::vector<std::vector<uint32_t>> NP1(k, std::vector<uint32_t>(V, 0)), NC2 = NP1;
    std::vector<std::vector<uint32_t>> N1(nde, std::vector<uint32_t>(V, 0)), N2 = N1;
    build_dcs(D1, D2, N1, N2, NP1, NC2);
    for (uint32_t u = 0; u < k; ++u)
      for (uint32_t v = 0; v < V; ++v) {
        if (d.label(v) != q.label(u)) continue;
        if (D1[u][v] != d1[u][v] || D2[u][v] != d2[u][v]) return false;
        if (NP1[u][v] != np1[u][v] || NC2[u][v] != nc2[u][v]) return false;
        for (uint32_t w : q.nb[u]) {
          uint32_t e = eidx[u][w];
          if (N2[e][v] != n2[e][v]) return false;
          if (is_bw(u, w) && N1[e][v] != n1[e][v]) return false;
        }
      }
    return true;
  }
};

// Same result as calling Graph::AddEdge for every edge in order
// (graph.cpp:217-249): sorted lists, the first copy of a duplicate pair wins.
Graph make_graph(uint32_t nv, const uint32_t* vl, size_t ne, const uint32_t* s, const uint32_t* t,
                 const uint32_t* l) {
  Graph g;
  for (uint32_t i = 0; i < nv; ++i) g.add_vertex(i, vl[i]);
  std::vector<uint64_t> cnt(nv + 1, 0);
  for (size_t i = 0; i < ne; ++i) {
    if (s[i] == t[i]) continue;
    ++cnt[s[i]];
    ++cnt[t[i]];
  }
  std::vector<std::vector<std::pair<uint32_t, size_t>>> tmp(nv);
  for (uint32_t v = 0; v < nv; ++v) tmp[v].reserve(cnt[v]);
  for (size_t i = 0; i < ne; ++i) {
    if (s[i] == t[i]) continue;
    tmp[s[i]].push_back({t[i], i});
    tmp[t[i]].push_back({s[i], i});
  }
  size_t m2 = 0;
  for (uint32_t v = 0; v < nv; ++v) {
    auto& a = tmp[v];
    std::sort(a.begin(), a.end());
    auto& nb = g.nb[v];
    auto& el = g.el[v];
    nb.reserve(a.size());
    el.reserve(a.size());
    for (size_t j = 0; j < a.size(); ++j) {
      if (j > 0 && a[j].first == a[j - 1].first) continue;
      nb.push_back(a[j].first);
      el.push_back(l[a[j].second]);
    }
    m2 += nb.size();
    std::vector<std::pair<uint32_t, size_t>>().swap(a);
  }
  g.ne = m2 / 2;
  return g;
}

// Brute force: all injective (or not, with homo) label- and edge-label-
// preserving maps of the query into the data graph.  Tiny inputs only.
struct Brute {
  const Graph& q;
  const Graph& d;
  bool homo, want_hash;
  std::vector<uint32_t> order, m;
  uint64_t cnt = 0, h = 0;
  const std::vector<std::vector<uint8_t>>* d2 = nullptr;  // optional candidate filter
  void rec(uint32_t i) {
    uint32_t k = q.nv();
    if (i == k) { ++cnt; if (want_hash) h += emb_hash(m.data(), k); return; }
    uint32_t u = order[i];
    for (uint32_t v = 0; v < d.nv(); ++v) {
      if (d.label(v) != q.label(u)) continue;
      if (d2 && !(*d2)[u][v]) continue;
      bool ok = true;
      if (!homo)
        for (uint32_t j = 0; j < i && ok; ++j) ok = m[order[j]] != v;
      for (size_t j = 0; j < q.nb[u].size() && ok; ++j) {
        uint32_t w = q.nb[u][j];
        if (m[w] == NOT_EXIST) continue;
        ok = d.elabel(m[w], v) == q.el[u][j];
      }
      if (!ok) continue;
      m[u] = v;
      rec(i + 1);
      m[u] = NOT_EXIST;
    }
  }
};

}  // namespace

extern "C" {

struct orc_graph { Graph g; };

orc_graph* orc_graph_create(uint32_t nv, const uint32_t* vl, size_t ne, const uint32_t* s,
                            const uint32_t* t, const uint32_t* l) {
  orc_graph* o = new orc_graph;
  o->g = make_graph(nv, vl, ne, s, t, l);
  return o;
}
void orc_graph_destroy(orc_graph* g) { delete g; }
uint64_t orc_graph_num_edges(const orc_graph* g) { return g->g.ne; }

// Label-triple filter of parallel_graphflow.cpp:1990-2016: 1 = safe.
int orc_classify(const orc_graph* q, uint32_t l1, uint32_t l2, uint32_t el) {
  const Graph& Q = q->g;
  for (uint32_t a = 0; a < Q.nv(); ++a)
    for (size_t j = 0; j < Q.nb[a].size(); ++j) {
      uint32_t b = Q.nb[a][j];
      if (Q.el[a][j] == el && Q.label(a) == l1 && Q.label(b) == l2) return 0;
    }
  return 1;
}

uint64_t orc_brute_count_filtered(const orc_graph* d, const orc_graph* q, int homo,
                                  uint64_t* hash, const std::vector<std::vector<uint8_t>>* d2);

uint64_t orc_brute_count(const orc_graph* d, const orc_graph* q, int homo, uint64_t* hash) {
  return orc_brute_count_filtered(d, q, homo, hash, nullptr);
}

// Embeddings whose every vertex image passes the d2 flag of the index with
// the given shape (tree != 0: TurboFlux's tree, else SymBi's DAG) built on d.
// d2 is a sound filter iff this equals orc_brute_count.
uint64_t orc_d2_filtered_count(const orc_graph* d, const orc_graph* q, int tree) {
  Graph g = d->g;
  SymBi sb(q->g, g, Opts(), tree != 0);
  return orc_brute_count_filtered(d, q, 0, nullptr, &sb.d2);
}

uint64_t orc_brute_count_filtered(const orc_graph* d, const orc_graph* q, int homo,
                                  uint64_t* hash, const std::vector<std::vector<uint8_t>>* d2) {
  Brute b{q->g, d->g, homo != 0, hash != nullptr, {}, {}};
  b.d2 = d2;
  uint32_t k = q->g.nv();
  b.m.assign(k, NOT_EXIST);
  // BFS-ish order: each next vertex adjacent to an earlier one when possible
  std::vector<uint8_t> in(k, 0);
  for (uint32_t s = 0; s < k; ++s) {
    if (in[s]) continue;
    std::deque<uint32_t> dq{s};
    in[s] = 1;
    while (!dq.empty()) {
      uint32_t u = dq.front();
      dq.pop_front();
      b.order.push_back(u);
      for (uint32_t w : q->g.nb[u])
        if (!in[w]) { in[w] = 1; dq.push_back(w); }
    }
  }
  b.rec(0);
  if (hash) *hash = b.h;
  return b.cnt;
}

// Run an update stream one update at a time through the named algorithm.
//   algo: "graphflow" | "symbi" (sequential semantics) |
//         "parallel_graphflow" (reference parallel shape, `threads` threads)
//   add[i]: 1 insert, 0 delete; lab[i] is used for inserts only (deletions
//   look the label up in the graph, as RemoveEdge does).
//   pos/neg: per-update positive/negative counts; hash (nullable): per-update
//   embedding-set hash; initial (nullable): InitialMatching count first.
//   check_index: SymBi only, verify the index fix-point after every update.
//   seconds (nullable): wall time of the update loop only.
// Returns 0, -1 unknown algo, -2 deletion of an absent edge (index of the
// failing update in *fail_at), -3 index inconsistency (at *fail_at).
int orc_run(const char* algo, const orc_graph* data, const orc_graph* query, size_t n,
            const uint8_t* add, const uint32_t* v1, const uint32_t* v2, const uint32_t* lab,
            uint64_t* pos, uint64_t* neg, uint64_t* hash, int threads, int homo,
            uint64_t* initial, int check_index, double* seconds, size_t* fail_at,
            uint64_t max_results) {
  Graph d = data->g;
  const Graph& q = query->g;
  Opts o;
  o.homo = homo != 0;
  o.want_hash = hash != nullptr;
  o.max_results = max_results;
  std::string a(algo);
  if (max_results && a != "graphflow") return -4;   // the cap is restated for GraphFlow only
  auto t0 = std::chrono::steady_clock::now();
  auto finish = [&](int rc) {
    if (seconds)
      *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return rc;
  };
  if (a == "graphflow" || a == "parallel_graphflow") {
    GraphFlow gf(q, d, o);
    if (initial) { uint64_t h = 0; *initial = gf.initial(h); }
    t0 = std::chrono::steady_clock::now();
    bool par = a == "parallel_graphflow";
    for (size_t i = 0; i < n; ++i) {
      uint64_t c = 0, h = 0;
      int rc = 0;
      if (par) {
        if (add[i]) {
          d.add_edge(v1[i], v2[i], lab[i]);
          c = gf.parallel_update(v1[i], v2[i], lab[i], threads);
        } else {
          uint32_t l = d.elabel(v1[i], v2[i]);
          if (l == NOT_EXIST) rc = -2;
          else {
            c = gf.parallel_update(v1[i], v2[i], l, threads);
            d.remove_edge(v1[i], v2[i]);
          }
        }
      } else {
        rc = add[i] ? gf.add_edge(v1[i], v2[i], lab[i], c, h) : gf.remove_edge(v1[i], v2[i], c, h);
      }
      if (rc) { if (fail_at) *fail_at = i; return finish(rc); }
      if (pos) pos[i] = add[i] ? c : 0;
      if (neg) neg[i] = add[i] ? 0 : c;
      if (hash) hash[i] = h;
    }
    return finish(0);
  }
  if (a == "symbi_index" || a == "turboflux_index") {
    // index maintenance only (no search; counts stay 0), checked against a
    // from-scratch build after every update when check_index is set
    SymBi sb(q, d, o, a == "turboflux_index");
    t0 = std::chrono::steady_clock::now();
    for (size_t i = 0; i < n; ++i) {
      if (add[i]) {
        d.add_edge(v1[i], v2[i], lab[i]);
        sb.index_update(v1[i], v2[i], lab[i], true);
      } else {
        uint32_t l = d.elabel(v1[i], v2[i]);
        if (l == NOT_EXIST) { if (fail_at) *fail_at = i; return finish(-2); }
        sb.index_update(v1[i], v2[i], l, false);
        d.remove_edge(v1[i], v2[i]);
      }
      if (check_index && !sb.index_consistent()) { if (fail_at) *fail_at = i; return finish(-3); }
      if (pos) pos[i] = 0;
      if (neg) neg[i] = 0;
      if (hash) hash[i] = 0;
    }
    return finish(0);
  }
  if (a == "symbi") {
    SymBi sb(q, d, o);
    if (initial) { uint64_t h = 0; *initial = sb.initial(h); }
    t0 = std::chrono::steady_clock::now();
    for (size_t i = 0; i < n; ++i) {
      uint64_t c = 0, h = 0;
      int rc = add[i] ? sb.add_edge(v1[i], v2[i], lab[i], c, h) : sb.remove_edge(v1[i], v2[i], c, h);
      if (rc) { if (fail_at) *fail_at = i; return finish(rc); }
      if (check_index && !sb.index_consistent()) { if (fail_at) *fail_at = i; return finish(-3); }
      if (pos) pos[i] = add[i] ? c : 0;
      if (neg) neg[i] = add[i] ? 0 : c;
      if (hash) hash[i] = h;
    }
    return finish(0);
  }
  return -1;
}

// Per-update search work of sequential SymBi (symbi.cpp:471-587, the
// ExtendableVertex order rule of :474-506): the DCS entries its search walks
// for each update (the entries of DCS[u_min -> u][m(u_min)] at every node of
// the search tree, valid or not), and the per-update match counts.
// Returns 0, -2 deletion of an absent edge (at *fail_at).
int orc_symbi_work(const orc_graph* data, const orc_graph* query, size_t n, const uint8_t* add,
                   const uint32_t* v1, const uint32_t* v2, const uint32_t* lab, int homo,
                   uint64_t* work, uint64_t* count, size_t* fail_at) {
  Graph d = data->g;
  const Graph& q = query->g;
  Opts o;
  o.homo = homo != 0;
  SymBi sb(q, d, o);
  for (size_t i = 0; i < n; ++i) {
    uint64_t c = 0, h = 0;
    sb.work = 0;
    int rc = add[i] ? sb.add_edge(v1[i], v2[i], lab[i], c, h) : sb.remove_edge(v1[i], v2[i], c, h);
    if (rc) { if (fail_at) *fail_at = i; return rc; }
    work[i] = sb.work;
    count[i] = c;
  }
  return 0;
}

// Sequential GraphFlow over a stream with a per-update result cap and time
// limits: each update's search gets at most `update_s` seconds and the whole
// run `total_s` (0: none); an update whose search ran out of time has
// unfinished[i] = 1 (its count is a lower bound), and once the total is spent
// the remaining updates are applied without a search (unfinished).  The graph
// is updated either way.  Returns 0, or -2 at a deletion of an absent edge.
int orc_graphflow_limited(const orc_graph* data, const orc_graph* query, size_t n,
                          const uint8_t* add, const uint32_t* v1, const uint32_t* v2,
                          const uint32_t* lab, uint64_t* count, uint8_t* unfinished,
                          uint64_t max_results, double update_s, double total_s) {
  Graph d = data->g;
  Opts o;
  o.max_results = max_results;
  GraphFlow gf(query->g, d, o);
  const double t_end = total_s > 0 ? now_s() + total_s : 0.0;
  for (size_t i = 0; i < n; ++i) {
    uint64_t c = 0, h = 0;
    const double t = now_s();
    bool skip = t_end > 0 && t >= t_end;
    double dl = update_s > 0 ? t + update_s : 0.0;
    if (t_end > 0 && (dl == 0.0 || dl > t_end)) dl = t_end;
    gf.o.deadline = dl;
    gf.timed_out = false;
    uint32_t l = lab[i];
    if (!add[i]) {
      l = d.elabel(v1[i], v2[i]);
      if (l == NOT_EXIST) return -2;
    }
    if (add[i]) d.add_edge(v1[i], v2[i], l);
    if (!skip) gf.search_update(v1[i], v2[i], l, c, h);
    if (!add[i]) d.remove_edge(v1[i], v2[i]);
    count[i] = c;
    unfinished[i] = (skip || gf.timed_out) ? 1 : 0;
  }
  return 0;
}

// SymBi index built from scratch on `data` (symbi.cpp:246-377), dumped in the
// layout of csm_index_dump: what = 0 root, 1 d1, 2 d2, 3 np1, 4 nc2 ([k][nv]),
// 5 n1 (per vertex u, per DAG parent in BFS order), 6 n2 (per u, per
// neighbour in id order).  Returns the number of words written, or -1.
long orc_index_dump(const orc_graph* dag_src, const orc_graph* data, const orc_graph* query,
                    int tree, uint32_t what, uint32_t* out, size_t n);

long orc_symbi_dump(const orc_graph* data, const orc_graph* query, uint32_t what, uint32_t* out,
                    size_t n) {
  return orc_index_dump(nullptr, data, query, 0, what, out, n);
}

// The same for either index shape: tree != 0 builds TurboFlux's spanning
// tree (parallel_turboflux.cpp:123-469) instead of SymBi's DAG; the root and
// the tree/DAG come from dag_src's label frequencies (nullptr: data's).
long orc_index_dump(const orc_graph* dag_src, const orc_graph* data, const orc_graph* query,
                    int tree, uint32_t what, uint32_t* out, size_t n) {
  Graph d = data->g;
  SymBi sb(query->g, d, Opts(), tree != 0, dag_src ? &dag_src->g : nullptr);
  const uint32_t k = query->g.nv(), V = d.nv();
  size_t w = 0;
  auto put_row = [&](const auto& row) {
    for (uint32_t v = 0; v < V; ++v) {
      if (w >= n) return false;
      out[w++] = (uint32_t)row[v];
    }
    return true;
  };
  if (what == 0) { if (n < 1) return -1; out[0] = sb.root; return 1; }
  for (uint32_t u = 0; u < k; ++u) {
    bool ok = true;
    if (what == 1) ok = put_row(sb.d1[u]);
    else if (what == 2) ok = put_row(sb.d2[u]);
    else if (what == 3) ok = put_row(sb.np1[u]);
    else if (what == 4) ok = put_row(sb.nc2[u]);
    else if (what == 5) { for (uint32_t p : sb.tn[u].bw) ok = ok && put_row(sb.n1[sb.eidx[u][p]]); }
    else if (what == 6) { for (uint32_t x : query->g.nb[u]) ok = ok && put_row(sb.n2[sb.eidx[u][x]]); }
    else return -1;
    if (!ok) return -1;
  }
  return (long)w;
}

// Parents of every query vertex in the index's DAG (tree != 0: TurboFlux's
// spanning tree) built on `data`: out[u*k + j] = j-th parent of u, NOT_EXIST
// past the last.  Returns the root.
uint32_t orc_dag_parents(const orc_graph* data, const orc_graph* query, int tree, uint32_t* out) {
  Graph d = data->g;
  SymBi sb(query->g, d, Opts(), tree != 0);
  const uint32_t k = query->g.nv();
  for (uint32_t u = 0; u < k; ++u)
    for (uint32_t j = 0; j < k; ++j) out[u * k + j] = j < sb.tn[u].bw.size() ? sb.tn[u].bw[j] : NOT_EXIST;
  return sb.root;
}

// TurboFlux's matching orders (parallel_turboflux.cpp:469-836), over the
// spanning-tree index of the initial graph.
//   CountDownwards (:469-491): from every data vertex carrying the root's
//   label (the keys of DCS_[root -> first child], :536-541), walk the DCS
//   lists down the tree, counting per query vertex the d2-valid children
//   reached; a (child, parent data vertex) pair is walked once and its count
//   reused afterwards (num_dp).  Counts are 32-bit and wrap, as `uint` does;
//   the root's count is UINT_MAX - 1.
//   initial_order (:544-575): repeatedly take the leaf of the remaining tree
//   with the fewest paths (ties: the smallest id; the reference takes the
//   first in an unordered_set's iteration order) and fill the order from the
//   back; a parent becomes a leaf once its children are all taken.
//   Per query edge (u < u' in id order, :577-700): a tree edge is oriented
//   child first; the order is the two endpoints, the tree path from the
//   second up to the root, then the rest in initial_order.
// out[i*k + j]: order i (query edges in that enumeration); paths[u]: the
// path counts.  Returns the number of orders.
struct TurboFluxOrders {
  const SymBi& sb;
  std::vector<uint32_t> paths;
  std::vector<std::unordered_map<uint32_t, uint32_t>> dp;
  explicit TurboFluxOrders(const SymBi& s) : sb(s), paths(s.k, 0), dp(s.k) {}
  void count_downwards(uint32_t u, uint32_t v) {
    for (uint32_t uc : sb.tn[u].fw) {
      auto it = dp[uc].find(v);
      if (it == dp[uc].end()) {
        sb.dcs_for(u, uc, v, [&](uint32_t vc) {
          if (sb.d2[uc][vc]) {
            ++paths[uc];
            ++dp[uc][v];
            count_downwards(uc, vc);
          }
        });
      } else {
        paths[uc] += it->second;
      }
    }
  }
};

int orc_turboflux_orders(const orc_graph* data, const orc_graph* query, uint32_t* out,
                         uint32_t* paths_out) {
  Graph d = data->g;
  const Graph& q = query->g;
  SymBi sb(q, d, Opts(), true);
  const uint32_t k = q.nv(), root = sb.root;
  TurboFluxOrders t(sb);
  for (uint32_t v = 0; v < d.nv(); ++v)
    if (d.label(v) == q.label(root)) t.count_downwards(root, v);
  t.paths[root] = UINT_MAX - 1;
  if (paths_out)
    for (uint32_t u = 0; u < k; ++u) paths_out[u] = t.paths[u];
  std::vector<uint8_t> leaf(k, 0), taken(k, 0);
  for (uint32_t u = 0; u < k; ++u) leaf[u] = sb.tn[u].fw.empty();
  std::vector<uint32_t> initial(k);
  for (uint32_t i = 0; i < k; ++i) {
    uint32_t best = NOT_EXIST;
    for (uint32_t u = 0; u < k; ++u)
      if (leaf[u] && (best == NOT_EXIST || t.paths[u] < t.paths[best])) best = u;
    leaf[best] = 0;
    taken[best] = 1;
    initial[k - 1 - i] = best;
    if (best != root) {
      const uint32_t par = sb.tn[best].bw.front();
      bool all = true;
      for (uint32_t c : sb.tn[par].fw) all = all && taken[c];
      if (all) leaf[par] = 1;
    }
  }
  int n = 0;
  for (uint32_t a = 0; a < k; ++a)
    for (uint32_t b : q.nb[a]) {
      if (a > b) continue;
      uint32_t u = a, uo = b;
      if (!sb.tn[uo].bw.empty() && sb.tn[uo].bw.front() == u) std::swap(u, uo);
      std::vector<uint8_t> vis(k, 0);
      std::vector<uint32_t> ord{u, uo};
      vis[u] = vis[uo] = 1;
      for (uint32_t cur = uo; cur != root;) {
        cur = sb.tn[cur].bw.front();
        vis[cur] = 1;
        ord.push_back(cur);
      }
      for (uint32_t x : initial)
        if (!vis[x]) { vis[x] = 1; ord.push_back(x); }
      for (uint32_t j = 0; j < k; ++j) out[(size_t)n * k + j] = ord[j];
      ++n;
    }
  return n;
}

// Matching orders of the GraphFlow plugin (graphflow.cpp:36-160), for
// checking the engine's host-side planner: out[i*k + j] = order_vs_[i][j].
int orc_graphflow_orders(const orc_graph* data, const orc_graph* query, uint32_t* out) {
  Graph d = data->g;
  GraphFlow gf(query->g, d, Opts());
  uint32_t k = query->g.nv();
  for (size_t i = 0; i < gf.ovs.size(); ++i)
    for (uint32_t j = 0; j < k; ++j) out[i * k + j] = gf.ovs[i][j];
  return (int)gf.ovs.size();
}

}  // extern "C"
