// EdgeStore (src/edge_store.cpp:1-303 of the reference; semantics in edge_store.h).
#include "gpusketch/edge_store.h"

#include <algorithm>

#include "gpusketch/bucket.h"
#include "gpusketch/util.h"

EdgeStore::EdgeStore(size_t seed, node_id_t n, size_t sketch_bytes, size_t num_subgraphs,
                     size_t start_subgraph)
    : seed(seed),
      num_vertices(n),
      num_subgraphs(num_subgraphs),
      cur_subgraph(start_subgraph),
      true_min_subgraph(start_subgraph),
      adjlist(n),
      vertex_contracted(n, 1),
      max_edges(sketch_bytes / store_edge_bytes),
      adj_mutex(new std::mutex[n]) {
  needs_contraction = n;  // no contraction running: nothing to claim
}

EdgeStore::~EdgeStore() = default;

// Tags with the deepest subgraph index num_subgraphs - 1 as the cap, the cap the caller's routing
// uses once sketch subgraphs exist (the reference caps these first inserts at num_subgraphs,
// src/edge_store.cpp:54 against src/mc_gpu_sketch_alg.cpp:160, so an edge at the deepest level
// inserted before the first contraction and deleted after it would be stored as two entries
// that never cancel).
TaggedUpdateBatch EdgeStore::insert_adj_edges(node_id_t src, const std::vector<node_id_t>& dsts) {
  std::vector<SubgraphTaggedUpdate> tagged(dsts.size());
  for (size_t i = 0; i < dsts.size(); i++) {
    const vec_t idx = concat_pairing_fn(src, dsts[i]);
    tagged[i] = {(node_id_t)Bucket_Boruvka::get_index_depth(idx, (long)seed, num_subgraphs - 1),
                 dsts[i]};
  }
  return insert_adj_edges(src, 0, tagged.data(), tagged.size());
}

TaggedUpdateBatch EdgeStore::insert_adj_edges(node_id_t src, node_id_t caller_first_es,
                                              SubgraphTaggedUpdate* data, size_t size) {
  std::vector<SubgraphTaggedUpdate> ret;
  size_t first_es = cur_subgraph;
  if (size == 0) return {src, (node_id_t)(first_es - 1), (node_id_t)first_es, ret};
  std::sort(data, data + size);
  // an entry that occurs twice in one batch (insert and delete) cancels itself
  size_t w = 0;
  for (size_t r = 0; r < size;) {
    size_t e = r;
    while (e < size && data[e] == data[r]) e++;
    if ((e - r) & 1) data[w++] = data[r];
    r = e;
  }
  const size_t kept = w;
  {
    std::lock_guard<std::mutex> lk(adj_mutex[src]);
    first_es = cur_subgraph;
    if (true_min_subgraph < first_es && !vertex_contracted[src]) ret = vertex_contract(src);
    // the caller routed with an older first stored subgraph: it has not applied these updates
    // to the subgraphs that have become sketches since, so hand them all back
    if (caller_first_es < first_es) ret.insert(ret.end(), data, data + kept);

    // symmetric difference of the sorted buffer and the entries at or above first_es
    const SubgraphTaggedUpdate* in = data;
    const SubgraphTaggedUpdate* in_end = data + kept;
    while (in < in_end && in->subgraph < first_es) ++in;
    auto& buf = adjlist[src];
    const size_t before = buf.size();
    std::vector<SubgraphTaggedUpdate> merged;
    merged.reserve(before + (size_t)(in_end - in));
    size_t b = 0;
    while (b < before && in < in_end) {
      if (buf[b] < *in) merged.push_back(buf[b++]);
      else if (*in < buf[b]) merged.push_back(*in++);
      else { ++b; ++in; }  // present twice: the edge is toggled away
    }
    merged.insert(merged.end(), buf.begin() + b, buf.end());
    merged.insert(merged.end(), in, in_end);
    buf.swap(merged);
    num_edges += buf.size() - before;  // modular arithmetic handles a shrink
  }
  if (ret.empty() && true_min_subgraph < first_es) return vertex_advance_subgraph((node_id_t)first_es);
  check_if_too_big();
  return {src, (node_id_t)(first_es - 1), (node_id_t)first_es, ret};
}

std::vector<Edge> EdgeStore::get_edges() {
  std::vector<Edge> out;
  out.reserve(num_edges);
  for (node_id_t src = 0; src < num_vertices; src++)
    for (const auto& e : adjlist[src]) out.push_back({src, e.dst});
  return out;
}

std::vector<SubgraphTaggedUpdate> EdgeStore::vertex_contract(node_id_t src) {
  std::vector<SubgraphTaggedUpdate> ret;
  if (vertex_contracted[src]) return ret;
  vertex_contracted[src] = 1;
  auto& buf = adjlist[src];
  if (buf.empty()) return ret;
  ret = buf;  // every entry belongs to the newly sketched subgraph
  const size_t before = buf.size();
  buf.erase(std::remove_if(buf.begin(), buf.end(),
                           [&](const SubgraphTaggedUpdate& e) { return e.subgraph < cur_subgraph; }),
            buf.end());
  num_edges -= before - buf.size();
  return ret;
}

// Hands back the next uncontracted vertex's entries during a contraction. The contraction ends
// (true_min_subgraph catches up) only when every vertex has been handled, not merely claimed: a
// thread that claimed a vertex and waits for its lock still belongs to this contraction, so the
// next one cannot start under it (it would reset the vertex's flag and move the first stored
// subgraph past entries that are then handed to the older subgraph only). The epoch is read
// after the claim, which pins it until this vertex is handled.
TaggedUpdateBatch EdgeStore::vertex_advance_subgraph(node_id_t) {
  for (;;) {
    // every vertex claimed: no claim past them either (callers may poll while the last ones are
    // handled, and the counter must not wrap around)
    node_id_t src = needs_contraction.load();
    if (src < num_vertices) src = needs_contraction.fetch_add(1);
    if (src >= num_vertices) {
      const node_id_t first = (node_id_t)cur_subgraph.load();
      return {0, first - 1, first, {}};
    }
    const node_id_t first = (node_id_t)cur_subgraph.load();
    TaggedUpdateBatch out{src, first - 1, first, {}};
    {
      std::lock_guard<std::mutex> lk(adj_mutex[src]);
      if (!adjlist[src].empty() && !vertex_contracted[src]) out.dsts_data = vertex_contract(src);
      vertex_contracted[src] = 1;
    }
    if (vertices_handled.fetch_add(1) + 1 == num_vertices) {
      std::lock_guard<std::mutex> lk(contract_lock);
      ++true_min_subgraph;  // every vertex has been handed back once
    }
    if (!out.dsts_data.empty()) return out;
  }
}

void EdgeStore::check_if_too_big() {
  if (num_edges < max_edges) return;
  std::lock_guard<std::mutex> lk(contract_lock);
  if (true_min_subgraph < cur_subgraph) return;  // a contraction is already running
  // no vertex can be claimed until the first stored subgraph has moved (needs_contraction is
  // still past the last vertex), so every claim of this contraction sees its epoch
  std::fill(vertex_contracted.begin(), vertex_contracted.end(), 0);
  vertices_handled = 0;
  ++cur_subgraph;
  needs_contraction = 0;
}
