// Device checks of the C++ API: CCGPUSketchAlg, SKGPUSketchAlg and MCGPUSketchAlg driven by
// GraphSketchDriver over binary streams, against host Sketch::update (itself checked bit-exact
// against the oracle on the host) and exact union-find / minimum cut. Usage:
//   facade_gpu <multiples_graph_1024_stream.data> <work_dir>
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <random>
#include <set>

#include "gpusketch/cc_gpu_sketch_alg.h"
#include "gpusketch/graph_sketch_driver.h"
#include "gpusketch/mc_gpu_sketch_alg.h"
#include "gpusketch/sk_gpu_sketch_alg.h"
#include "gpusketch/sketch_server.h"

static int failures = 0;
#define CHECK(cond)                                                           \
  do {                                                                        \
    if (!(cond)) {                                                            \
      std::printf("  check failed: %s (%s:%d)\n", #cond, __FILE__, __LINE__); \
      std::fflush(stdout);                                                    \
      failures++;                                                             \
      return;                                                                 \
    }                                                                         \
  } while (0)

static std::string fixture, work_dir;
static const uint64_t kSeed = 0xA11CE5EEDULL;

static void run(const char* name, void (*fn)()) {
  std::printf("RUN %s\n", name);
  std::fflush(stdout);
  const int before = failures;
  try {
    fn();
  } catch (const std::exception& e) {
    std::printf("  exception: %s\n", e.what());
    failures++;
  }
  std::printf("%s %s\n", failures == before ? "PASS" : "FAIL", name);
  std::fflush(stdout);
}

// every update of a stream file, in order
static std::vector<GraphStreamUpdate> read_all(const std::string& path, node_id_t* n) {
  BinaryFileStream s(path);
  *n = s.vertices();
  std::vector<GraphStreamUpdate> out, buf(4096);
  for (;;) {
    size_t got = s.get_update_buffer(buf.data(), buf.size());
    if (got == 1 && buf[0].type == BREAKPOINT) break;
    out.insert(out.end(), buf.begin(), buf.begin() + got);
  }
  return out;
}

static void write_stream(const std::string& path, node_id_t n, const std::vector<Edge>& edges) {
  std::vector<GraphStreamUpdate> ups(edges.size());
  for (size_t i = 0; i < edges.size(); i++) ups[i] = {edges[i], INSERT};
  BinaryFileStream w(path, false);
  w.write_header(n, ups.size());
  w.write_updates(ups.data(), ups.size());
}

// host sketches of the toggled edge multiset, optionally only edges of subgraph depth >= g
static std::vector<Bucket> host_sketches(node_id_t n, size_t samples, uint64_t seed,
                                         const std::vector<Edge>& edges, size_t min_depth = 0,
                                         size_t depth_cap = 0) {
  const vec_t len = Sketch::calc_vector_length(n);
  Sketch proto(len, seed, samples);
  std::vector<Bucket> all((size_t)n * proto.get_buckets());
  std::vector<std::unique_ptr<Sketch>> sk;
  for (node_id_t v = 0; v < n; v++) sk.emplace_back(new Sketch(len, seed, v, all.data(), samples));
  for (const Edge& e : edges) {
    const vec_t idx = concat_pairing_fn(e.src, e.dst);
    if (min_depth && Bucket_Boruvka::get_index_depth(idx, (long)seed, depth_cap) < min_depth)
      continue;
    sk[e.src]->update(idx);
    sk[e.dst]->update(idx);
  }
  return all;
}

// net edge set (toggle semantics) of a stream
static std::set<std::pair<node_id_t, node_id_t>> net_edges(const std::vector<Edge>& edges) {
  std::set<std::pair<node_id_t, node_id_t>> live;
  for (const Edge& e : edges) {
    auto key = std::make_pair(std::min(e.src, e.dst), std::max(e.src, e.dst));
    if (!live.erase(key)) live.insert(key);
  }
  return live;
}

static std::vector<node_id_t> dsu_labels(node_id_t n,
                                         const std::set<std::pair<node_id_t, node_id_t>>& g) {
  DisjointSetUnion<node_id_t> d(n, 1);
  for (auto& e : g) d.merge(e.first, e.second);
  std::vector<node_id_t> lab(n);
  for (node_id_t v = 0; v < n; v++) lab[v] = d.find_root(v);
  return lab;
}

static bool same_partition(const ConnectedComponents& cc, const std::vector<node_id_t>& lab) {
  std::map<node_id_t, node_id_t> a2b, b2a;
  for (node_id_t v = 0; v < lab.size(); v++) {
    const node_id_t a = cc.labels()[v], b = lab[v];
    if (a2b.count(a) && a2b[a] != b) return false;
    if (b2a.count(b) && b2a[b] != a) return false;
    a2b[a] = b;
    b2a[b] = a;
  }
  return true;
}

static std::vector<Bucket> export_all(DeviceSketchStore& st) {
  const gs_geometry& g = st.geometry();
  std::vector<Bucket> out((size_t)g.num_vertices * g.num_buckets);
  st.export_buckets(0, g.num_vertices, out.data());
  return out;
}

static bool equal_buckets(const std::vector<Bucket>& a, const std::vector<Bucket>& b) {
  return a.size() == b.size() && std::memcmp(a.data(), b.data(), a.size() * sizeof(Bucket)) == 0;
}

static void test_cc_fixture_through_driver() {
  node_id_t n = 0;
  auto ups = read_all(fixture, &n);
  std::vector<Edge> edges;
  for (auto& u : ups) edges.push_back(u.edge);
  BinaryFileStream stream(fixture);
  CCGPUSketchAlg alg(n, stream.edges(), 3, make_sketch_params(n, kSeed));
  auto cfg = DriverConfiguration().worker_threads(3);
  GraphSketchDriver<CCGPUSketchAlg> drv(&alg, &stream, cfg, 2);
  drv.process_stream_until(END_OF_STREAM);
  drv.prep_query(KSPANNINGFORESTS);  // never cached: flushes the gutters
  alg.flush_buffers();
  const size_t S = Sketch::calc_cc_samples(n, 1);
  CHECK(equal_buckets(export_all(alg.device_store()), host_sketches(n, S, kSeed, edges)));
  auto net = net_edges(edges);
  ConnectedComponents cc = alg.connected_components();
  CHECK(cc.size() == 78);
  CHECK(same_partition(cc, dsu_labels(n, net)));
  CHECK(alg.point_query(0, 0));
  // checkpoint and resume: the restored algorithm has no cached answer and queries the device
  const std::string ckpt = work_dir + "/facade_cc.bin";
  alg.write_binary(ckpt);
  std::unique_ptr<CCGPUSketchAlg> back(CCGPUSketchAlg::construct_from_serialized_data(ckpt, 2));
  CHECK(!back->has_cached_query(CONNECTIVITY));
  CHECK(equal_buckets(export_all(back->device_store()), export_all(alg.device_store())));
  ConnectedComponents cc2 = back->connected_components();
  CHECK(cc2.size() == 78 && same_partition(cc2, dsu_labels(n, net)));
  CHECK(back->has_cached_query(CONNECTIVITY));
  // k disjoint forests: edge-disjoint, edges of the graph, the sketches restored afterwards
  const auto before = export_all(back->device_store());
  auto sfs = back->calc_disjoint_spanning_forests(3);
  std::set<std::pair<node_id_t, node_id_t>> seen;
  for (auto& sf : sfs)
    for (const Edge& e : sf.get_edges()) {
      auto key = std::make_pair(std::min(e.src, e.dst), std::max(e.src, e.dst));
      CHECK(net.count(key));
      CHECK(seen.insert(key).second);
    }
  CHECK(sfs[0].get_edges().size() == n - 78);
  CHECK(equal_buckets(export_all(back->device_store()), before));
  std::remove(ckpt.c_str());
}

static void test_eager_dsu_matches_device() {
  const node_id_t n = 3000;
  std::mt19937_64 rng(23);
  std::vector<Edge> edges;
  std::set<std::pair<node_id_t, node_id_t>> have;
  while (edges.size() < 2500) {  // insert-only: the host union-find stays valid
    node_id_t u = rng() % n, v = rng() % n;
    if (u == v || !have.insert({std::min(u, v), std::max(u, v)}).second) continue;
    edges.push_back({u, v});
  }
  const std::string path = work_dir + "/facade_eager.data";
  write_stream(path, n, edges);
  BinaryFileStream stream(path);
  CCGPUSketchAlg alg(n, edges.size(), 2, make_sketch_params(n, kSeed + 1));
  GraphSketchDriver<CCGPUSketchAlg> drv(&alg, &stream, DriverConfiguration().worker_threads(2));
  drv.process_stream_until(END_OF_STREAM);
  CHECK(alg.has_cached_query(CONNECTIVITY));
  drv.prep_query(CONNECTIVITY);  // cached: returns without flushing
  ConnectedComponents eager = alg.connected_components();
  auto lab = dsu_labels(n, have);
  CHECK(same_partition(eager, lab));
  drv.prep_query(KSPANNINGFORESTS);
  alg.flush_buffers();
  const std::string ckpt = work_dir + "/facade_eager.bin";
  alg.write_binary(ckpt);
  std::unique_ptr<CCGPUSketchAlg> dev(CCGPUSketchAlg::construct_from_serialized_data(ckpt));
  ConnectedComponents cc = dev->connected_components();
  CHECK(same_partition(cc, lab) && cc.size() == eager.size());
  std::remove(path.c_str());
  std::remove(ckpt.c_str());
}

static void test_sk_paths_match_cc() {
  node_id_t n = 0;
  auto ups = read_all(fixture, &n);
  std::vector<Edge> edges;
  for (auto& u : ups) edges.push_back(u.edge);
  const size_t S = Sketch::calc_cc_samples(n, 1);
  const auto expect = host_sketches(n, S, kSeed + 2, edges);
  for (int single = 0; single < 2; single++) {
    BinaryFileStream stream(fixture);
    SKGPUSketchAlg sk(true, n, stream.edges(), 2, make_sketch_params(n, kSeed + 2));
    GraphSketchDriver<SKGPUSketchAlg> drv(&sk, &stream, DriverConfiguration().worker_threads(2));
    drv.process_stream_until(END_OF_STREAM);
    drv.prep_query(KSPANNINGFORESTS);
    if (single) sk.launch_gpu_kernel();
    else sk.launch_gpu_update();
    CHECK(sk.get_batch_count() > 0);
    CHECK(equal_buckets(export_all(sk.device_store()), expect));
    CHECK(sk.connected_components().size() == 78);
  }
  bool threw = false;
  try {
    SKGPUSketchAlg cpu(false, n, 10, 1, make_sketch_params(n, kSeed));
  } catch (const std::invalid_argument&) {
    threw = true;
  }
  CHECK(threw);
}

static void test_mc_planted_cut() {
  const node_id_t n = 1024;
  const uint32_t cut = 3;
  const int k = 4;
  uint64_t count = 0;
  gs_throw_if(gs_gen_planted_cut(n, 0.2, cut, 0.1, 77, nullptr, &count));
  std::vector<uint32_t> pairs(2 * count);
  gs_throw_if(gs_gen_planted_cut(n, 0.2, cut, 0.1, 77, pairs.data(), &count));
  std::vector<Edge> edges(count);
  for (uint64_t i = 0; i < count; i++) edges[i] = {pairs[2 * i], pairs[2 * i + 1]};
  auto net = net_edges(edges);
  std::vector<Edge> net_list;
  for (auto& e : net) net_list.push_back({e.first, e.second});
  const std::string path = work_dir + "/facade_mc.data";
  write_stream(path, n, edges);
  const int num_graphs = 1 + 2 * 10;
  SketchParams p = make_sketch_params(n, kSeed + 3, 1.0);
  // a store budget far below the graph forces contractions into sketch subgraphs
  const size_t budget = 4000 * sizeof(SubgraphTaggedUpdate);
  MCGPUSketchAlg mc(n, 2, 1, p, num_graphs, num_graphs, k, budget, 0);
  {
    BinaryFileStream stream(path);
    GraphSketchDriver<MCGPUSketchAlg> drv(&mc, &stream, DriverConfiguration().worker_threads(2));
    drv.process_stream_until(END_OF_STREAM);
    drv.prep_query(KSPANNINGFORESTS);
  }
  mc.apply_flush_updates();
  const int sg = mc.get_num_sketch_graphs();
  CHECK(sg >= 1);
  // every sketch subgraph g holds exactly the edges of depth >= g
  for (int g = 0; g < sg; g++)
    CHECK(equal_buckets(export_all(mc.subgraph_store(g)),
                        host_sketches(n, p.num_samples, p.seed, edges, (size_t)g,
                                      (size_t)num_graphs - 1)));
  // the EdgeStore holds exactly the net edges of depth >= first stored subgraph (both directions)
  size_t stored = 0;
  for (auto& e : net)
    if (Bucket_Boruvka::get_index_depth(concat_pairing_fn(e.first, e.second), (long)p.seed,
                                        num_graphs - 1) >= (size_t)sg)
      stored += 2;
  CHECK(mc.get_num_adjlist_edges() == stored);
  MinCut exact = mc.calc_minimum_cut(net_list);
  CHECK(exact.value == cut);
  MinCutAnswer ans = mc_minimum_cut(mc, num_graphs, k);
  CHECK(ans.graph_id == 0 && ans.value == cut);
  CHECK(ans.cut.left_vertices.size() + ans.cut.right_vertices.size() == n);
  // forests of subgraph 0 are edge-disjoint edges of the graph; the sketches are restored
  const auto before = export_all(mc.subgraph_store(0));
  auto sfs = mc.calc_disjoint_spanning_forests(0, k);
  std::set<std::pair<node_id_t, node_id_t>> seen;
  for (auto& sf : sfs)
    for (const Edge& e : sf.get_edges()) {
      auto key = std::make_pair(std::min(e.src, e.dst), std::max(e.src, e.dst));
      CHECK(net.count(key) && seen.insert(key).second);
    }
  CHECK(equal_buckets(export_all(mc.subgraph_store(0)), before));
  ConnectedComponents cc = mc.connected_components();
  CHECK(same_partition(cc, dsu_labels(n, net)));
  std::remove(path.c_str());
}

// admission by sketch memory: a tenant that does not fit waits until another is removed, then
// its queued requests are answered; answers equal those of a plain server
static void test_server_admission() {
  TenantSpec a{3000, 11, 1}, b{5000, 12, 1}, c{4000, 13, 1};
  gs_geometry ga{}, gb{}, gc{};
  gs_throw_if(gs_geometry_for(a.num_vertices, 1, 1, &ga));
  gs_throw_if(gs_geometry_for(b.num_vertices, 1, 1, &gb));
  gs_throw_if(gs_geometry_for(c.num_vertices, 1, 1, &gc));
  ServerConfig cfg;
  cfg.memory_budget_bytes = ga.device_bytes + gb.device_bytes + gc.device_bytes / 2;
  cfg.keep_labels = true;
  std::mt19937_64 rng(77);
  auto batch = [&](uint32_t n, size_t m) {
    std::vector<uint32_t> p(2 * m);
    for (size_t i = 0; i < m; i++) {
      p[2 * i] = rng() % n;
      do p[2 * i + 1] = rng() % n; while (p[2 * i + 1] == p[2 * i]);
    }
    return p;
  };
  auto pc = batch(c.num_vertices, 3000), pa = batch(a.num_vertices, 2000);
  std::vector<std::vector<RequestResult>> answers;
  for (int batching = 0; batching < 2; batching++) {
    cfg.batching = batching != 0;
    SketchServer srv(cfg);
    uint32_t ta = srv.add_tenant(a), tb = srv.add_tenant(b), tc = srv.add_tenant(c);
    srv.wait_admissions();
    CHECK(srv.is_admitted(ta) && srv.is_admitted(tb) && !srv.is_admitted(tc));
    srv.submit_update(ta, pa.data(), pa.size() / 2);
    srv.submit_update(tc, pc.data(), pc.size() / 2);  // queued until tc is admitted
    srv.submit_cc_query(ta);
    srv.submit_cc_query(tc);
    srv.remove_tenant(tb);  // frees b's memory: c gets in
    srv.drain();
    CHECK(srv.is_admitted(tc) && srv.error().empty());
    auto res = srv.results();
    std::sort(res.begin(), res.end(),
              [](const RequestResult& x, const RequestResult& y) { return x.id < y.id; });
    CHECK(res.size() == 4);
    answers.push_back(res);
    // the CC answers are the exact partitions of the net graphs
    std::vector<Edge> ea, ec;
    for (size_t i = 0; i < pa.size(); i += 2) ea.push_back({pa[i], pa[i + 1]});
    for (size_t i = 0; i < pc.size(); i += 2) ec.push_back({pc[i], pc[i + 1]});
    ConnectedComponents cca(a.num_vertices, res[2].labels.data());
    ConnectedComponents ccc(c.num_vertices, res[3].labels.data());
    CHECK(same_partition(cca, dsu_labels(a.num_vertices, net_edges(ea))));
    CHECK(same_partition(ccc, dsu_labels(c.num_vertices, net_edges(ec))));
  }
  for (size_t i = 0; i < 4; i++)
    CHECK(answers[0][i].num_cc == answers[1][i].num_cc &&
          answers[0][i].partition_hash == answers[1][i].partition_hash);
}

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: facade_gpu <fixture.data> <work_dir>\n");
    return 2;
  }
  fixture = argv[1];
  work_dir = argv[2];
  run("cc_fixture_through_driver", test_cc_fixture_through_driver);
  run("eager_dsu_matches_device", test_eager_dsu_matches_device);
  run("sk_paths_match_cc", test_sk_paths_match_cc);
  run("mc_planted_cut", test_mc_planted_cut);
  run("server_admission", test_server_admission);
  std::printf("%d failure(s)\n", failures);
  return failures ? 1 : 0;
}
