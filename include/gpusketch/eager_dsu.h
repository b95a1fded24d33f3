// The host-side connectivity cache of the reference's CCSketchAlg / MCSketchAlg
// (src/cc_sketch_alg.cpp:105-130): while every update only adds connectivity, a union-find over a
// spanning forest answers connectivity queries without sampling the sketches. An update that
// toggles a forest edge invalidates the cache until the next sketch query rebuilds it.
#pragma once
#include <atomic>
#include <memory>
#include <mutex>
#include <unordered_set>
#include <vector>

#include "dsu.h"
#include "return_types.h"
#include "types.h"

class EagerDSU {
 public:
  EagerDSU(node_id_t n, bool valid)
      : n(n), dsu(n), dsu_valid(valid), shared_valid(valid), forest(n), forest_mtx(new std::mutex[n]) {}

  // called for every stream update before it is buffered (thread-safe)
  void pre_insert(const GraphUpdate& upd) {
    if (!dsu_valid) return;
    const node_id_t a = std::min(upd.edge.src, upd.edge.dst);
    const node_id_t b = std::max(upd.edge.src, upd.edge.dst);
    std::lock_guard<std::mutex> lk(forest_mtx[a]);
    if (dsu.merge(a, b).merged) {
      forest[a].insert(b);  // new connectivity: a forest edge
    } else if (forest[a].count(b)) {
      dsu_valid = false;  // a forest edge is toggled away: the cached answer may be stale
      shared_valid = false;
    }
  }
  // replace the cache by a spanning forest from a sketch query
  void adopt(const std::vector<Edge>& edges) {
    dsu.reset();
    for (auto& s : forest) s.clear();
    for (const Edge& e : edges) {
      const node_id_t a = std::min(e.src, e.dst), b = std::max(e.src, e.dst);
      dsu.merge(a, b);
      forest[a].insert(b);
    }
    dsu_valid = true;
    shared_valid = true;
  }
  void invalidate() {
    dsu_valid = false;
    shared_valid = false;
  }
  bool valid() const { return shared_valid.load(); }
  ConnectedComponents components() { return ConnectedComponents(n, dsu); }
  SpanningForest spanning_forest() { return SpanningForest(n, forest.data()); }

 private:
  node_id_t n;
  DisjointSetUnion_MT<node_id_t> dsu;
  bool dsu_valid;
  std::atomic<bool> shared_valid;
  std::vector<std::unordered_set<node_id_t>> forest;
  std::unique_ptr<std::mutex[]> forest_mtx;
};
