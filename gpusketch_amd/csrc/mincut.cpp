// Exact global minimum cut of an undirected multigraph (host code).
//
// Replaces the VieCut call of MCSketchAlg::calc_minimum_cut (src/mc_sketch_alg.cpp:966-1003),
// which is an inexact heuristic; the min-cut answers of this library are exact.
//
// Method (Nagamochi–Ono–Ibaraki contraction): keep an upper bound `best` (initially the minimum
// weighted degree). Each pass computes a maximum-adjacency (MA) ordering; when vertex x is
// visited, every unvisited neighbour y gets r(y) += w(x,y) and q = r(y) is a lower bound on the
// x-y connectivity, so x and y may be contracted whenever q >= best without losing any cut
// smaller than `best`. The last two vertices of the ordering are also contracted, after the
// cut of the phase (the degree of the last vertex) has been offered to `best` -- the
// Stoer–Wagner step -- which guarantees progress. The contracted graph's minimum degree is
// offered to `best` after every pass. When one vertex is left, `best` is the minimum cut.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <queue>
#include <utility>
#include <vector>

#include "../../include/gpusketch.h"

namespace {

struct Graph {
  uint32_t n = 0;
  std::vector<uint64_t> off;                       // CSR offsets
  std::vector<std::pair<uint32_t, uint64_t>> adj;  // (neighbour, weight), no self loops
};

struct Dsu {
  std::vector<uint32_t> p;
  explicit Dsu(uint32_t n) : p(n) { std::iota(p.begin(), p.end(), 0u); }
  uint32_t find(uint32_t x) {
    while (p[x] != x) { p[x] = p[p[x]]; x = p[x]; }
    return x;
  }
  bool unite(uint32_t a, uint32_t b) {
    a = find(a); b = find(b);
    if (a == b) return false;
    if (a > b) std::swap(a, b);
    p[b] = a;
    return true;
  }
};

// build a CSR graph from (a, b, w) triples, merging parallel edges and dropping self loops
Graph build(uint32_t n, std::vector<std::pair<uint64_t, uint64_t>>& keyed) {
  // keyed: (min<<32|max, weight)
  std::sort(keyed.begin(), keyed.end());
  std::vector<std::pair<uint64_t, uint64_t>> merged;
  merged.reserve(keyed.size());
  for (auto& kv : keyed) {
    if (!merged.empty() && merged.back().first == kv.first) merged.back().second += kv.second;
    else merged.push_back(kv);
  }
  Graph g;
  g.n = n;
  std::vector<uint64_t> deg(n + 1, 0);
  for (auto& kv : merged) {
    deg[kv.first >> 32]++;
    deg[(uint32_t)kv.first]++;
  }
  g.off.assign(n + 1, 0);
  for (uint32_t v = 0; v < n; v++) g.off[v + 1] = g.off[v] + deg[v];
  g.adj.resize(g.off[n]);
  std::vector<uint64_t> pos(g.off.begin(), g.off.end() - 1);
  for (auto& kv : merged) {
    uint32_t a = (uint32_t)(kv.first >> 32), b = (uint32_t)kv.first;
    g.adj[pos[a]++] = {b, kv.second};
    g.adj[pos[b]++] = {a, kv.second};
  }
  return g;
}

}  // namespace

extern "C" int gs_min_cut(uint32_t n, const uint32_t* pairs, uint64_t m, uint64_t* value,
                          uint8_t* side_out) {
  if (!value || (m && !pairs)) return GS_ERR_INVALID;
  if (n < 2) {
    *value = UINT64_MAX;
    if (side_out && n) std::memset(side_out, 0, n);
    return GS_OK;
  }
  for (uint64_t e = 0; e < 2 * m; e++)
    if (pairs[e] >= n) return GS_ERR_INVALID;

  // disconnected input: a component is a cut of value 0
  {
    Dsu d(n);
    for (uint64_t e = 0; e < m; e++) d.unite(pairs[2 * e], pairs[2 * e + 1]);
    uint32_t r0 = d.find(0);
    for (uint32_t v = 1; v < n; v++) {
      if (d.find(v) != r0) {
        *value = 0;
        if (side_out)
          for (uint32_t u = 0; u < n; u++) side_out[u] = d.find(u) == r0;
        return GS_OK;
      }
    }
  }

  std::vector<std::pair<uint64_t, uint64_t>> keyed;
  keyed.reserve(m);
  for (uint64_t e = 0; e < m; e++) {
    uint32_t a = pairs[2 * e], b = pairs[2 * e + 1];
    if (a == b) continue;
    if (a > b) std::swap(a, b);
    keyed.push_back({((uint64_t)a << 32) | b, 1});
  }
  Graph g = build(n, keyed);
  std::vector<uint32_t> label(n);  // original vertex -> current super vertex
  std::iota(label.begin(), label.end(), 0u);

  uint64_t best = UINT64_MAX;
  std::vector<uint8_t> best_side(n, 0);
  auto offer_degrees = [&](const Graph& cg) {
    uint64_t dmin = UINT64_MAX;
    uint32_t arg = 0;
    for (uint32_t x = 0; x < cg.n; x++) {
      uint64_t d = 0;
      for (uint64_t i = cg.off[x]; i < cg.off[x + 1]; i++) d += cg.adj[i].second;
      if (d < dmin) { dmin = d; arg = x; }
    }
    if (dmin < best) {
      best = dmin;
      for (uint32_t u = 0; u < n; u++) best_side[u] = label[u] == arg;
    }
  };
  offer_degrees(g);

  std::vector<uint64_t> r;
  std::vector<uint8_t> visited;
  while (g.n > 1 && best > 0) {
    const uint32_t cn = g.n;
    r.assign(cn, 0);
    visited.assign(cn, 0);
    Dsu dsu(cn);
    std::priority_queue<std::pair<uint64_t, uint32_t>> pq;
    pq.push({0, 0});
    uint32_t prev = UINT32_MAX, last = UINT32_MAX;
    uint32_t nvis = 0;
    while (!pq.empty()) {
      auto [rv, x] = pq.top();
      pq.pop();
      if (visited[x] || rv != r[x]) continue;
      visited[x] = 1;
      nvis++;
      prev = last;
      last = x;
      for (uint64_t i = g.off[x]; i < g.off[x + 1]; i++) {
        uint32_t y = g.adj[i].first;
        if (visited[y]) continue;
        r[y] += g.adj[i].second;
        if (r[y] >= best) dsu.unite(x, y);
        pq.push({r[y], y});
      }
    }
    (void)nvis;  // the graph stays connected under contraction, so every vertex is visited
    // cut of the phase: r(last) is the degree of `last`, already offered via offer_degrees;
    // contracting the last two vertices is the Stoer–Wagner step
    if (prev != UINT32_MAX) dsu.unite(prev, last);

    // contract
    std::vector<uint32_t> newid(cn, UINT32_MAX);
    uint32_t nn = 0;
    for (uint32_t x = 0; x < cn; x++) {
      uint32_t rt = dsu.find(x);
      if (newid[rt] == UINT32_MAX) newid[rt] = nn++;
    }
    for (uint32_t u = 0; u < n; u++) label[u] = newid[dsu.find(label[u])];
    std::vector<std::pair<uint64_t, uint64_t>> kk;
    kk.reserve(g.adj.size() / 2);
    for (uint32_t x = 0; x < cn; x++) {
      uint32_t a = newid[dsu.find(x)];
      for (uint64_t i = g.off[x]; i < g.off[x + 1]; i++) {
        uint32_t y = g.adj[i].first;
        if (y < x) continue;  // each undirected edge once
        uint32_t b = newid[dsu.find(y)];
        if (a == b) continue;
        uint32_t lo = std::min(a, b), hi = std::max(a, b);
        kk.push_back({((uint64_t)lo << 32) | hi, g.adj[i].second});
      }
    }
    g = build(nn, kk);
    if (g.n > 1) offer_degrees(g);
  }
  *value = best;
  if (side_out) std::memcpy(side_out, best_side.data(), n);
  return GS_OK;
}
