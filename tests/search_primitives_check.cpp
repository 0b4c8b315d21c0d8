// Host check of the search primitives (paracosm_amd/csrc/csm_search.hpp)
// against exhaustive searches on random ascending arrays with repeated
// values and version windows.  Prints "cases N bad B"; exits 1 on any bad.
#define CSM_SEARCH_HOST
#include "../paracosm_amd/csrc/csm_search.hpp"

#include <algorithm>
#include <cstdio>
#include <random>
#include <vector>

int main() {
  std::mt19937 rng(20241203);
  long cases = 0, bad = 0;
  for (int it = 0; it < 40000; ++it) {
    const uint32_t n = rng() % 300, off = rng() % 5;
    std::vector<uint32_t> v(n), t(n);
    for (auto& x : v) x = rng() % (n + 7);
    std::sort(v.begin(), v.end());
    for (auto& x : t) {
      const uint32_t a = rng() % 8, b = a + rng() % 8;
      x = (a << 16) | b;
    }
    // guards around the segment: nothing outside [off, off + n) may decide
    std::vector<uint32_t> adj(n + off + 4, 0u), ats(n + off + 4, 0u);
    for (uint32_t i = 0; i < off; ++i) adj[i] = 0xFFFFFFFFu;
    for (uint32_t i = 0; i < n; ++i) {
      adj[off + i] = v[i];
      ats[off + i] = t[i];
    }
    for (uint32_t i = off + n; i < adj.size(); ++i) adj[i] = 0u;
    for (int q = 0; q < 8; ++q) {
      const uint32_t x = rng() % (n + 9), pos = rng() % 10;
      const bool ver = rng() & 1u;
      const uint32_t lb = off + (uint32_t)(std::lower_bound(v.begin(), v.end(), x) - v.begin());
      bool in = false;
      for (uint32_t i = 0; i < n; ++i)
        if (v[i] == x && (!ver || csm::ts_visible(t[i], pos))) in = true;
      bad += csm::lower_bound(adj.data(), off, off + n, x) != lb;
      bad += csm::contains(adj.data(), ats.data(), off, off + n, x, ver, pos) != in;
      bad += csm::contains_u(adj.data(), ats.data(), off, n, x, ver, pos) != in;
      bad += csm::lds_contains(v.data(), n, x) != std::binary_search(v.begin(), v.end(), x);
      cases += 4;
    }
  }
  // run ends of sorted insertion keys (vertex << 32 | entry, ~0 padding)
  // against a scan key by key from the run start
  for (int it = 0; it < 20000; ++it) {
    const uint32_t n = rng() % 400, pad = rng() % 4, nv = 1 + rng() % 40;
    const uint32_t top = (it & 1) ? 0xFFFFFFFEu : nv;  // also vertices next to ~0's
    std::vector<uint64_t> keys(n);
    for (auto& k : keys) {
      const uint32_t v = top - rng() % nv;
      k = ((uint64_t)v << 32) | (rng() & 0xFFFFFFFFu);
    }
    std::sort(keys.begin(), keys.end());
    for (uint32_t i = 0; i < pad; ++i) keys.push_back(~0ull);
    const uint32_t m = (uint32_t)keys.size();
    for (uint32_t bs = 0; bs < m; ++bs) {
      if (keys[bs] == ~0ull || (bs > 0 && (keys[bs - 1] >> 32) == (keys[bs] >> 32))) continue;
      const uint32_t v = (uint32_t)(keys[bs] >> 32);
      uint32_t e = bs + 1;
      while (e < m && keys[e] != ~0ull && (uint32_t)(keys[e] >> 32) == v) ++e;
      bad += csm::run_end(keys.data(), bs, m, v) != e;
      ++cases;
    }
  }
  std::printf("cases %ld bad %ld\n", cases, bad);
  return bad != 0;
}
