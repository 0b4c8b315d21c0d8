// Host-side query planning: matching orders (GraphFlow) and the BFS DAG
// (SymBi).  Plans stay on the host; the device receives flat PlanRec tables.
#pragma once
#include <stdint.h>

#include <algorithm>
#include <deque>
#include <vector>

namespace csm {

// A small labelled query graph with sorted adjacency (like the reference
// Graph, ParaCOSM/CSM/graph/graph.h:63-100, restricted to what planning needs).
struct QueryGraph {
  uint32_t k = 0;
  size_t ne = 0;
  std::vector<uint32_t> vl;                    // original labels
  std::vector<std::vector<uint32_t>> nb, el;   // sorted neighbours, edge labels

  void init(uint32_t n, const uint32_t* labels) {
    k = n;
    vl.assign(labels, labels + n);
    nb.assign(n, {});
    el.assign(n, {});
    ne = 0;
  }
  bool add_edge(uint32_t a, uint32_t b, uint32_t l) {
    if (a >= k || b >= k || a == b) return false;
    auto it = std::lower_bound(nb[a].begin(), nb[a].end(), b);
    if (it != nb[a].end() && *it == b) return false;
    el[a].insert(el[a].begin() + (it - nb[a].begin()), l);
    nb[a].insert(it, b);
    it = std::lower_bound(nb[b].begin(), nb[b].end(), a);
    el[b].insert(el[b].begin() + (it - nb[b].begin()), l);
    nb[b].insert(it, a);
    ++ne;
    return true;
  }
  uint32_t deg(uint32_t u) const { return (uint32_t)nb[u].size(); }
  uint32_t elabel(uint32_t a, uint32_t b) const {
    auto it = std::lower_bound(nb[a].begin(), nb[a].end(), b);
    if (it == nb[a].end() || *it != b) return UINT32_MAX;
    return el[a][it - nb[a].begin()];
  }
  bool connected() const {
    if (k == 0) return true;
    std::vector<uint8_t> seen(k, 0);
    std::deque<uint32_t> q{0};
    seen[0] = 1;
    uint32_t n = 1;
    while (!q.empty()) {
      uint32_t u = q.front();
      q.pop_front();
      for (uint32_t w : nb[u])
        if (!seen[w]) { seen[w] = 1; ++n; q.push_back(w); }
    }
    return n == k;
  }
};

// GraphFlow matching orders, one per query edge.  Same rule as the
// reference (ParaCOSM/CSM/matching/GraphFlow/graphflow.cpp:36-160): order 0
// starts at the highest-degree vertex; order i (i >= 1) starts with the i-th
// backward edge of order 0; each next vertex has the most already-ordered
// neighbours, ties going to the higher degree, then the lower id.
inline std::vector<std::vector<uint32_t>> graphflow_orders(const QueryGraph& q) {
  const uint32_t k = q.k;
  std::vector<std::vector<uint32_t>> orders;
  if (q.ne == 0) return orders;
  auto next_vertex = [&](const std::vector<uint8_t>& placed) {
    uint32_t best = UINT32_MAX, best_adj = 0;
    for (uint32_t j = 0; j < k; ++j) {
      if (placed[j]) continue;
      uint32_t adj = 0;
      for (uint32_t w : q.nb[j]) adj += placed[w];
      if (adj == 0) continue;
      bool better = best == UINT32_MAX || adj > best_adj ||
                    (adj == best_adj && q.deg(j) > q.deg(best));
      if (better) { best = j; best_adj = adj; }
    }
    return best;
  };
  // order 0 and its backward-edge list (the query edges, enumerated)
  std::vector<uint32_t> o0(k);
  std::vector<std::pair<uint32_t, uint32_t>> edges;  // (vertex, backward neighbour)
  {
    std::vector<uint8_t> placed(k, 0);
    uint32_t first = 0, maxd = 0;
    for (uint32_t i = 0; i < k; ++i)
      if (q.deg(i) > maxd) { maxd = q.deg(i); first = i; }
    o0[0] = first;
    placed[first] = 1;
    for (uint32_t p = 1; p < k; ++p) {
      uint32_t u = next_vertex(placed);
      o0[p] = u;
      placed[u] = 1;
      for (uint32_t w : q.nb[u])
        if (placed[w] && w != u) edges.push_back({u, w});
    }
  }
  orders.push_back(o0);
  for (size_t i = 1; i < q.ne; ++i) {
    std::vector<uint32_t> o(k);
    std::vector<uint8_t> placed(k, 0);
    o[0] = edges[i].first;
    o[1] = edges[i].second;
    placed[o[0]] = placed[o[1]] = 1;
    for (uint32_t p = 2; p < k; ++p) {
      uint32_t u = next_vertex(placed);
      o[p] = u;
      placed[u] = 1;
    }
    orders.push_back(o);
  }
  return orders;
}

// The same orders with positions 2..k-1 re-chosen by estimated cost (the
// first two positions, which make every query edge the seed of exactly one
// order, are kept).  `fanout(w, u)` estimates how many candidates for u one
// data vertex matched to w yields (neighbours with u's label over the query
// edge's label); `hit(w, u)` the probability that a given candidate of u is
// adjacent to w's match.  The next vertex minimizes fanout from its best
// placed neighbour times the hit rates of its other placed neighbours; ties
// go to more placed neighbours, then the higher degree, then the lower id.
// Any order enumerates the same embeddings: only the search effort changes.
template <class Fanout, class Hit>
std::vector<std::vector<uint32_t>> costed_orders(const QueryGraph& q, Fanout fanout, Hit hit) {
  std::vector<std::vector<uint32_t>> orders = graphflow_orders(q);
  const uint32_t k = q.k;
  for (auto& o : orders) {
    if (k <= 3) continue;
    std::vector<uint8_t> placed(k, 0);
    placed[o[0]] = placed[o[1]] = 1;
    for (uint32_t p = 2; p < k; ++p) {
      uint32_t best = UINT32_MAX, best_adj = 0;
      double best_cost = 0.0;
      for (uint32_t u = 0; u < k; ++u) {
        if (placed[u]) continue;
        uint32_t adj = 0, wbest = UINT32_MAX;
        double f = 1e300;
        for (uint32_t w : q.nb[u]) {
          if (!placed[w]) continue;
          ++adj;
          const double fw = fanout(w, u);
          if (fw < f) { f = fw; wbest = w; }
        }
        if (adj == 0) continue;
        // selectivity of the placed neighbours other than the cheapest one
        double h = 1.0;
        for (uint32_t w : q.nb[u])
          if (placed[w] && w != wbest) h *= hit(w, u);
        const double cost = f * h;
        bool better = best == UINT32_MAX || cost < best_cost ||
                      (cost == best_cost && (adj > best_adj ||
                                             (adj == best_adj && q.deg(u) > q.deg(best))));
        if (better) { best = u; best_adj = adj; best_cost = cost; }
      }
      o[p] = best;
      placed[best] = 1;
    }
  }
  return orders;
}

// SymBi's DAG: BFS from the root chosen by edge frequency (the rarest query
// edge's less frequent endpoint; ParaCOSM/CSM/matching/SymBi/symbi.cpp:113-244).
// `edge_freq(u, u2)` counts directed data edges whose labels match (u -> u2).
struct Dag {
  uint32_t root = 0;
  std::vector<uint32_t> bfs;                       // serialized tree
  std::vector<std::vector<uint32_t>> parents, children;
};

// The index root shared by SymBi and TurboFlux: the endpoint of the rarest
// query edge (first in (u < w) order on ties) with the rarer vertex label,
// ties to the higher query degree.
template <class EdgeFreq, class VertexFreq>
uint32_t index_root(const QueryGraph& q, EdgeFreq edge_freq, VertexFreq vertex_freq) {
  uint64_t min_f = UINT64_MAX;
  uint32_t a0 = 0, a1 = 0;
  for (uint32_t u = 0; u < q.k; ++u)
    for (uint32_t w : q.nb[u]) {
      if (u > w) continue;
      uint64_t f = edge_freq(u, w);
      if (f < min_f) { min_f = f; a0 = u; a1 = w; }
    }
  uint64_t f0 = vertex_freq(a0), f1 = vertex_freq(a1);
  if (f0 > f1) return a1;
  if (f0 < f1) return a0;
  return q.deg(a0) > q.deg(a1) ? a0 : a1;
}

template <class EdgeFreq, class VertexFreq>
Dag symbi_dag(const QueryGraph& q, EdgeFreq edge_freq, VertexFreq vertex_freq) {
  Dag d;
  const uint32_t k = q.k;
  d.root = index_root(q, edge_freq, vertex_freq);
  d.parents.assign(k, {});
  d.children.assign(k, {});
  std::vector<uint8_t> done(k, 0), queued(k, 0);
  std::deque<uint32_t> fifo{d.root};
  queued[d.root] = 1;
  d.bfs.push_back(d.root);
  while (!fifo.empty()) {
    uint32_t u = fifo.front();
    fifo.pop_front();
    done[u] = 1;
    for (uint32_t w : q.nb[u]) {
      if (!queued[w]) { queued[w] = 1; fifo.push_back(w); d.bfs.push_back(w); }
      if (!done[w]) d.children[u].push_back(w);
      else d.parents[u].push_back(w);
    }
  }
  return d;
}

// TurboFlux's spanning tree (ParaCOSM/CSM/matching/Parallel_TurboFlux/
// parallel_turboflux.cpp:123-320): from the same root, level by level, take
// the rarest queued edge into an unvisited vertex (the earliest queued on
// ties), drop the other queued edges into that vertex, and queue the new
// vertex's edges to unvisited vertices for the next level.  Edge frequency is
// that of the (lower id, higher id) orientation.  `bfs` is the BFS order over
// tree children.
template <class EdgeFreq, class VertexFreq>
Dag turboflux_tree(const QueryGraph& q, EdgeFreq edge_freq, VertexFreq vertex_freq) {
  Dag d;
  const uint32_t k = q.k;
  d.root = index_root(q, edge_freq, vertex_freq);
  d.parents.assign(k, {});
  d.children.assign(k, {});
  auto freq = [&](uint32_t a, uint32_t b) { return a < b ? edge_freq(a, b) : edge_freq(b, a); };
  struct Cand { uint32_t from, to; uint64_t f; };
  std::vector<uint8_t> seen(k, 0);
  seen[d.root] = 1;
  std::vector<Cand> level, next;
  for (uint32_t w : q.nb[d.root]) level.push_back({d.root, w, freq(d.root, w)});
  while (!level.empty()) {
    next.clear();
    while (!level.empty()) {
      size_t best = 0;
      for (size_t i = 1; i < level.size(); ++i)
        if (level[i].f < level[best].f) best = i;
      const Cand c = level[best];
      size_t o = 0;
      for (size_t i = 0; i < level.size(); ++i)
        if (level[i].to != c.to) level[o++] = level[i];
      level.resize(o);
      if (seen[c.to]) continue;
      seen[c.to] = 1;
      d.children[c.from].push_back(c.to);
      d.parents[c.to].push_back(c.from);
      for (uint32_t w : q.nb[c.to])
        if (!seen[w]) next.push_back({c.to, w, freq(c.to, w)});
    }
    level.swap(next);
  }
  std::deque<uint32_t> fifo{d.root};
  while (!fifo.empty()) {
    uint32_t u = fifo.front();
    fifo.pop_front();
    d.bfs.push_back(u);
    for (uint32_t c : d.children[u]) fifo.push_back(c);
  }
  return d;
}

// TurboFlux's matching orders (ParaCOSM/CSM/matching/Parallel_TurboFlux/
// parallel_turboflux.cpp:533-700), one per query edge, from its spanning tree
// and its explicit-path counts per query vertex (`paths`, the root's being
// UINT32_MAX - 1):
//   * initial order: repeatedly take the leaf of the remaining tree with the
//     fewest paths (ties: the smallest id -- the reference takes the first in
//     an unordered_set's iteration order) and fill the order from the back; a
//     parent becomes a leaf once its children are all taken;
//   * per query edge (u < u' in id order; a tree edge oriented child first):
//     the two endpoints, the tree path from the second up to the root, then
//     the rest in the initial order.
inline std::vector<std::vector<uint32_t>> turboflux_orders(const QueryGraph& q, const Dag& t,
                                                           const std::vector<uint32_t>& paths) {
  const uint32_t k = q.k;
  auto parent = [&](uint32_t u) { return t.parents[u].empty() ? UINT32_MAX : t.parents[u][0]; };
  std::vector<uint8_t> leaf(k, 0), taken(k, 0);
  for (uint32_t u = 0; u < k; ++u) leaf[u] = t.children[u].empty();
  std::vector<uint32_t> initial(k);
  for (uint32_t i = 0; i < k; ++i) {
    uint32_t best = UINT32_MAX;
    for (uint32_t u = 0; u < k; ++u)
      if (leaf[u] && (best == UINT32_MAX || paths[u] < paths[best])) best = u;
    leaf[best] = 0;
    taken[best] = 1;
    initial[k - 1 - i] = best;
    if (best != t.root) {
      const uint32_t par = parent(best);
      bool all = true;
      for (uint32_t c : t.children[par]) all = all && taken[c];
      if (all) leaf[par] = 1;
    }
  }
  std::vector<std::vector<uint32_t>> orders;
  for (uint32_t a = 0; a < k; ++a)
    for (uint32_t b : q.nb[a]) {
      if (a > b) continue;
      uint32_t u = a, uo = b;
      if (parent(uo) == u) std::swap(u, uo);
      std::vector<uint8_t> vis(k, 0);
      std::vector<uint32_t> ord{u, uo};
      vis[u] = vis[uo] = 1;
      for (uint32_t cur = uo; cur != t.root;) {
        cur = parent(cur);
        vis[cur] = 1;
        ord.push_back(cur);
      }
      for (uint32_t x : initial)
        if (!vis[x]) { vis[x] = 1; ord.push_back(x); }
      orders.push_back(ord);
    }
  return orders;
}

}  // namespace csm
