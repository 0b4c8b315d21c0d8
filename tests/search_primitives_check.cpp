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
  std::printf("cases %ld bad %ld\n", cases, bad);
  return bad != 0;
}
