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
      adj_mutex(new std::mutex[n]) {}

EdgeStore::~EdgeStore() = default;

TaggedUpdateBatch EdgeStore::insert_adj_edges(node_id_t src, const std::vector<node_id_t>& dsts) {
  std::vector<SubgraphTaggedUpdate> tagged(dsts.size());
  for (size_t i = 0; i < dsts.size(); i++) {
    const vec_t idx = concat_pairing_fn(src, dsts[i]);
    tagged[i] = {(node_id_t)Bucket_Boruvka::get_index_depth(idx, (long)seed, num_subgraphs),
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

TaggedUpdateBatch EdgeStore::vertex_advance_subgraph(node_id_t cur_first_es) {
  const node_id_t lo = cur_first_es - 1;
  for (;;) {
    const node_id_t src = needs_contraction.fetch_add(1);
    if (src >= num_vertices) {
      if (src == num_vertices) {
        std::lock_guard<std::mutex> lk(contract_lock);
        ++true_min_subgraph;  // every vertex has been handed back once
      }
      return {0, lo, cur_first_es, {}};
    }
    std::lock_guard<std::mutex> lk(adj_mutex[src]);
    if (!adjlist[src].empty() && !vertex_contracted[src])
      return {src, lo, cur_first_es, vertex_contract(src)};
    vertex_contracted[src] = 1;
  }
}

void EdgeStore::check_if_too_big() {
  if (num_edges < max_edges) return;
  std::lock_guard<std::mutex> lk(contract_lock);
  if (true_min_subgraph < cur_subgraph) return;  // a contraction is already running
  std::fill(vertex_contracted.begin(), vertex_contracted.end(), 0);
  needs_contraction = 0;
  ++cur_subgraph;
}
