// Host-only checks of the C++ API (no device calls): Sketch, pairing, union-find, streams,
// gutters + driver, EdgeStore. Usage:
//   facade_cpu all <work_dir>                       run every check, print PASS/FAIL lines
//   facade_cpu sketch_dump <pairs.bin> <n> <seed> <out.bin>
//       apply every pair {u,v} with Sketch::update to the sketches of u and v and write the
//       n * num_buckets reference-ordered buckets (checked against the oracle by the caller)
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <random>
#include <set>
#include <sstream>
#include <thread>

#include "gpusketch/dsu.h"
#include "gpusketch/edge_store.h"
#include "gpusketch/graph_sketch_driver.h"
#include "gpusketch/graph_verifier.h"
#include "gpusketch/sketch.h"

static int failures = 0;
#define CHECK(cond)                                                         \
  do {                                                                      \
    if (!(cond)) {                                                          \
      std::printf("  check failed: %s (%s:%d)\n", #cond, __FILE__, __LINE__); \
      failures++;                                                           \
      return;                                                               \
    }                                                                       \
  } while (0)

static void run(const char* name, void (*fn)()) {
  const int before = failures;
  fn();
  std::printf("%s %s\n", failures == before ? "PASS" : "FAIL", name);
}

static std::string work_dir = ".";

static void test_sketch_toggle_and_merge() {
  const node_id_t n = 1000;
  const vec_t len = Sketch::calc_vector_length(n);
  const size_t S = Sketch::calc_cc_samples(n, 1);
  std::mt19937_64 rng(7);
  Sketch a(len, 42, S), b(len, 42, S), ab(len, 42, S), zero(len, 42, S);
  std::set<vec_t> A, B;
  for (int i = 0; i < 300; i++) A.insert(concat_pairing_fn(rng() % n, rng() % n) | 1);
  for (int i = 0; i < 300; i++) B.insert(concat_pairing_fn(rng() % n, rng() % n) | 1);
  for (vec_t x : A) a.update(x);
  for (vec_t x : B) b.update(x);
  for (vec_t x : A)
    if (!B.count(x)) ab.update(x);
  for (vec_t x : B)
    if (!A.count(x)) ab.update(x);
  Sketch m(a);
  m.merge(b);
  CHECK(m == ab);  // linearity: sk(A) ^ sk(B) = sk(A xor B)
  Sketch t(a);
  for (vec_t x : A) t.update(x);
  CHECK(t == zero);  // every index toggled twice
  Sketch r(a);
  r.merge_raw_bucket_buffer(a.get_readonly_bucket_ptr());
  CHECK(r == zero);
}

static void test_sketch_sampling() {
  const node_id_t n = 4096;
  const vec_t len = Sketch::calc_vector_length(n);
  const size_t S = Sketch::calc_cc_samples(n, 1);
  Sketch s(len, 99, S);
  CHECK(s.sample().result == ZERO);
  s.reset_sample_state();
  const vec_t idx = concat_pairing_fn(3, 77);
  s.update(idx);
  SketchSample r = s.sample();
  CHECK(r.result == GOOD && r.idx == idx);  // a single index is always recovered
  s.zero_contents();
  // many indices: every GOOD sample is one of them; samples run out after S
  std::mt19937_64 rng(3);
  std::set<vec_t> in;
  for (int i = 0; i < 500; i++) in.insert(concat_pairing_fn(rng() % n, rng() % n) | 1);
  for (vec_t x : in) s.update(x);
  size_t good = 0;
  for (size_t i = 0; i < S; i++) {
    SketchSample q = s.sample();
    if (q.result == GOOD) {
      good++;
      CHECK(in.count(q.idx));
    }
    CHECK(q.result != ZERO);
  }
  CHECK(good > S / 2);
  bool thrown = false;
  try {
    s.sample();
  } catch (const OutOfSamplesException& e) {
    thrown = std::string(e.what()).find("cannot be sampled more times") != std::string::npos;
  }
  CHECK(thrown);
  s.reset_sample_state();
  ExhaustiveSketchSample ex = s.exhaustive_sample();
  CHECK(ex.result != ZERO && (ex.result == GOOD) == !ex.idxs.empty());
  for (vec_t x : ex.idxs) CHECK(in.count(x));
}

static void test_sketch_range_merge_and_serialize() {
  const node_id_t n = 512;
  const vec_t len = Sketch::calc_vector_length(n);
  const size_t S = Sketch::calc_cc_samples(n, 1);
  Sketch a(len, 5, S), b(len, 5, S);
  for (node_id_t v = 1; v < 200; v++) b.update(concat_pairing_fn(0, v));
  a.range_merge(b, 2, 3);
  const size_t B = a.get_bkt_per_col();
  const size_t cps = a.get_columns() / a.get_num_samples();  // columns per sample
  const Bucket* pa = a.get_readonly_bucket_ptr();
  const Bucket* pb = b.get_readonly_bucket_ptr();
  for (size_t i = 0; i < a.get_buckets() - 1; i++) {
    const bool inside = i >= 2 * cps * B && i < 5 * cps * B;
    CHECK((pa[i].alpha == (inside ? pb[i].alpha : 0)) &&
          (pa[i].gamma == (inside ? pb[i].gamma : 0)));
  }
  CHECK(pa[a.get_buckets() - 1].alpha == pb[b.get_buckets() - 1].alpha);
  std::stringstream ss;
  b.serialize(ss);
  Sketch c(len, 5, ss, S);
  CHECK(c == b);
  CHECK(b.bucket_array_bytes() == b.get_buckets() * 16);
}

static void test_pairing() {
  std::mt19937_64 rng(11);
  for (int i = 0; i < 10000; i++) {
    node_id_t u = (node_id_t)rng(), v = (node_id_t)rng();
    Edge e = inv_concat_pairing_fn(concat_pairing_fn(u, v));
    CHECK(e.src == std::min(u, v) && e.dst == std::max(u, v));
    node_id_t a = (node_id_t)(rng() % 100000), b = (node_id_t)(rng() % 100000);
    if (a == b) continue;
    Edge f = inv_nondir_non_self_edge_pairing_fn(nondirectional_non_self_edge_pairing_fn(a, b));
    CHECK(f.src == std::min(a, b) && f.dst == std::max(a, b));
  }
  size_t total = 0;
  for (size_t i = 0; i < 7; i++) {
    auto p = get_ith_partition(100, i, 7);
    CHECK(p.first == total);
    total = p.second;
  }
  CHECK(total == 100);
}

static void test_dsu() {
  const node_id_t n = 5000;
  std::mt19937_64 rng(5);
  std::vector<std::pair<node_id_t, node_id_t>> ops;
  for (int i = 0; i < 4000; i++) ops.push_back({(node_id_t)(rng() % n), (node_id_t)(rng() % n)});
  DisjointSetUnion<node_id_t> s(n, 1);
  for (auto& o : ops) s.merge(o.first, o.second);
  DisjointSetUnion_MT<node_id_t> m(n, 2);
  std::vector<std::thread> th;
  for (int t = 0; t < 4; t++)
    th.emplace_back([&, t] {
      for (size_t i = t; i < ops.size(); i += 4) m.merge(ops[i].first, ops[i].second);
    });
  for (auto& x : th) x.join();
  for (node_id_t u = 0; u < n; u++)
    for (node_id_t v : {(node_id_t)0, (node_id_t)(u / 2), (node_id_t)(n - 1 - u)})
      CHECK((s.find_root(u) == s.find_root(v)) == (m.find_root(u) == m.find_root(v)));
  auto r = s.merge(0, 0);
  CHECK(!r.merged);
  s.reset();
  CHECK(s.find_root(17) == 17);
}

static void test_binary_stream() {
  const std::string path = work_dir + "/facade_stream.data";
  std::mt19937_64 rng(13);
  std::vector<GraphStreamUpdate> ups(10007);
  for (auto& u : ups) {
    u.edge = {(node_id_t)(rng() % 300), (node_id_t)(rng() % 300)};
    u.type = (rng() & 1) ? DELETE : INSERT;
  }
  {
    BinaryFileStream w(path, false);
    w.write_header(300, ups.size());
    w.write_updates(ups.data(), 5000);
    w.write_updates(ups.data() + 5000, ups.size() - 5000);
  }
  BinaryFileStream r(path);
  CHECK(r.vertices() == 300 && r.edges() == ups.size());
  CHECK(r.set_break_point(4000));
  std::vector<GraphStreamUpdate> buf(333);
  size_t pos = 0;
  for (;;) {
    size_t got = r.get_update_buffer(buf.data(), buf.size());
    if (got == 1 && buf[0].type == BREAKPOINT) break;
    for (size_t i = 0; i < got; i++, pos++)
      CHECK(buf[i].edge == ups[pos].edge && buf[i].type == ups[pos].type);
  }
  CHECK(pos == 4000);
  CHECK(!r.set_break_point(100));
  CHECK(r.set_break_point(END_OF_STREAM));
  // the rest with concurrent readers: every record exactly once
  std::atomic<uint64_t> xsum{0}, cnt{0};
  std::vector<std::thread> th;
  for (int t = 0; t < 4; t++)
    th.emplace_back([&] {
      std::vector<GraphStreamUpdate> b(97);
      for (;;) {
        size_t got = r.get_update_buffer(b.data(), b.size());
        if (got == 1 && b[0].type == BREAKPOINT) return;
        for (size_t i = 0; i < got; i++) {
          xsum += (uint64_t)b[i].edge.src * 1000003ULL + b[i].edge.dst + (b[i].type == DELETE);
          cnt++;
        }
      }
    });
  for (auto& x : th) x.join();
  uint64_t expect = 0;
  for (size_t i = 4000; i < ups.size(); i++)
    expect += (uint64_t)ups[i].edge.src * 1000003ULL + ups[i].edge.dst + (ups[i].type == DELETE);
  CHECK(cnt == ups.size() - 4000 && xsum == expect);
  std::remove(path.c_str());
}

// Alg stand-in that records what the driver delivers
struct RecordingAlg {
  node_id_t n;
  std::atomic<uint64_t> pre{0}, delivered{0}, checksum{0};
  std::vector<std::atomic<uint64_t>> per_src;
  explicit RecordingAlg(node_id_t n) : n(n), per_src(n) {}
  void allocate_worker_memory(size_t) {}
  size_t get_desired_updates_per_batch() { return 37; }
  node_id_t get_num_vertices() { return n; }
  void pre_insert(GraphUpdate, int) { pre++; }
  void apply_update_batch(int, node_id_t src, const std::vector<node_id_t>& dsts) {
    if (dsts.size() > 37) checksum += 1ULL << 60;  // larger than the desired batch size
    for (node_id_t d : dsts) checksum += (uint64_t)src * 7919ULL + d;
    per_src[src] += dsts.size();
    delivered += dsts.size();
  }
  bool has_cached_query(int) { return false; }
  void print_configuration() {}
};

static void test_driver_delivers_every_update() {
  const std::string path = work_dir + "/facade_driver.data";
  const node_id_t n = 700;
  std::mt19937_64 rng(17);
  std::vector<GraphStreamUpdate> ups(50000);
  uint64_t expect = 0;
  std::vector<uint64_t> deg(n, 0);
  for (auto& u : ups) {
    u.edge = {(node_id_t)(rng() % n), (node_id_t)(rng() % n)};
    u.type = INSERT;
    expect += (uint64_t)u.edge.src * 7919ULL + u.edge.dst + (uint64_t)u.edge.dst * 7919ULL + u.edge.src;
    deg[u.edge.src]++;
    deg[u.edge.dst]++;
  }
  {
    BinaryFileStream w(path, false);
    w.write_header(n, ups.size());
    w.write_updates(ups.data(), ups.size());
  }
  BinaryFileStream s(path);
  RecordingAlg alg(n);
  auto cfg = DriverConfiguration().worker_threads(4);
  cfg.gutter_conf().queue_factor(2).wq_batch_per_elm(3);
  GraphSketchDriver<RecordingAlg> drv(&alg, &s, cfg, 2);
  drv.process_stream_until(20000);
  drv.prep_query(0);
  CHECK(alg.delivered == 40000);
  drv.process_stream_until(END_OF_STREAM);
  drv.prep_query(0);
  CHECK(alg.pre == ups.size());
  CHECK(alg.delivered == 2 * ups.size() && drv.get_total_updates() == 2 * ups.size());
  CHECK(alg.checksum == expect);
  for (node_id_t v = 0; v < n; v++) CHECK(alg.per_src[v] == deg[v]);
  std::remove(path.c_str());
}

static void test_edge_store() {
  const node_id_t n = 64;
  const size_t subgraphs = 8;
  // budget of 40 stored entries forces contractions
  EdgeStore es(123, n, 40 * sizeof(SubgraphTaggedUpdate), subgraphs, 0);
  std::map<std::pair<node_id_t, node_id_t>, int> live;
  std::vector<SubgraphTaggedUpdate> returned_all;
  std::mt19937_64 rng(19);
  for (int i = 0; i < 400; i++) {
    node_id_t u = rng() % n, v = rng() % n;
    if (u == v) continue;
    live[{u, v}] ^= 1;
    TaggedUpdateBatch r = es.insert_adj_edges(u, std::vector<node_id_t>{v});
    while (es.contract_in_progress()) {
      TaggedUpdateBatch m = es.vertex_advance_subgraph((node_id_t)es.get_first_store_subgraph());
      (void)m;
    }
    (void)r;
  }
  CHECK(es.get_first_store_subgraph() > 0);
  // what remains stored is exactly the live entries whose tag reaches the first stored subgraph
  std::set<std::pair<node_id_t, node_id_t>> expect;
  for (auto& kv : live) {
    if (!kv.second) continue;
    const vec_t idx = concat_pairing_fn(kv.first.first, kv.first.second);
    const size_t tag = Bucket_Boruvka::get_index_depth(idx, 123, subgraphs - 1);
    if (tag >= es.get_first_store_subgraph()) expect.insert(kv.first);
  }
  std::set<std::pair<node_id_t, node_id_t>> got;
  for (const Edge& e : es.get_edges()) got.insert({e.src, e.dst});
  CHECK(got == expect);
  CHECK(es.get_num_edges() == expect.size());
}

// The min-cut routing of MCGPUSketchAlg::apply_update_batch / complete_update_batch with a
// recording sink in place of the device subgraphs, driven by GraphSketchDriver with several
// workers: every subgraph must receive exactly the edges of depth >= its index, and the
// EdgeStore must keep exactly the rest.
struct RoutingAlg {
  node_id_t n;
  size_t nsub;
  uint64_t seed;
  EdgeStore es;
  std::mutex sink_mtx;
  std::vector<std::map<std::pair<node_id_t, node_id_t>, int>> sink;  // per subgraph, toggles
  RoutingAlg(node_id_t n, size_t nsub, uint64_t seed, size_t budget)
      : n(n), nsub(nsub), seed(seed), es(seed, n, budget, nsub, 0), sink(nsub) {}
  void allocate_worker_memory(size_t) {}
  size_t get_desired_updates_per_batch() { return 23; }
  node_id_t get_num_vertices() { return n; }
  void pre_insert(GraphUpdate, int) {}
  bool has_cached_query(int) { return false; }
  void print_configuration() {}
  void complete(const TaggedUpdateBatch& u) {
    std::lock_guard<std::mutex> lk(sink_mtx);
    for (const auto& d : u.dsts_data) {
      const size_t top = std::min<size_t>(d.subgraph, u.first_es_subgraph - 1);
      for (size_t g = u.min_subgraph; g <= top; g++) sink[g][{u.src, d.dst}] ^= 1;
    }
  }
  void apply_update_batch(int, node_id_t src, const std::vector<node_id_t>& dsts) {
    const node_id_t first_es = (node_id_t)es.get_first_store_subgraph();
    if (first_es == 0) {
      TaggedUpdateBatch more = es.insert_adj_edges(src, dsts);
      if (!more.dsts_data.empty()) complete(more);
      return;
    }
    std::vector<SubgraphTaggedUpdate> st, sk;
    for (node_id_t d : dsts) {
      const node_id_t t =
          (node_id_t)Bucket_Boruvka::get_index_depth(concat_pairing_fn(src, d), (long)seed, nsub - 1);
      if (t >= first_es) st.push_back({t, d});
      sk.push_back({t, d});
    }
    TaggedUpdateBatch more = es.insert_adj_edges(src, first_es, st.data(), st.size());
    complete({src, 0, first_es, sk});
    if (!more.dsts_data.empty()) complete(more);
  }
};

// workers driver threads, a store budget of budget_edges entries, updates random toggles
static void mc_routing_case(uint64_t stream_seed, int workers, size_t budget_edges, int updates) {
  const node_id_t n = 400;
  const size_t nsub = 12;
  const std::string path = work_dir + "/facade_routing_" + std::to_string(stream_seed) + ".data";
  std::mt19937_64 rng(stream_seed);
  std::vector<GraphStreamUpdate> ups;
  std::map<std::pair<node_id_t, node_id_t>, int> live;
  for (int i = 0; i < updates; i++) {
    node_id_t u = rng() % n, v = rng() % n;
    if (u == v) continue;
    ups.push_back({{u, v}, INSERT});
    live[{std::min(u, v), std::max(u, v)}] ^= 1;
  }
  {
    BinaryFileStream w(path, false);
    w.write_header(n, ups.size());
    w.write_updates(ups.data(), ups.size());
  }
  RoutingAlg alg(n, nsub, 31, budget_edges * sizeof(SubgraphTaggedUpdate));
  {
    BinaryFileStream s(path);
    GraphSketchDriver<RoutingAlg> drv(&alg, &s, DriverConfiguration().worker_threads(workers), 2);
    drv.process_stream_until(END_OF_STREAM);
    drv.prep_query(1);
  }
  while (alg.es.contract_in_progress()) {
    TaggedUpdateBatch m = alg.es.vertex_advance_subgraph((node_id_t)alg.es.get_first_store_subgraph());
    if (!m.dsts_data.empty()) alg.complete(m);
  }
  const size_t first = alg.es.get_first_store_subgraph();
  CHECK(first >= 2);
  size_t stored = 0;
  for (size_t g = 0; g < first; g++) {
    std::set<std::pair<node_id_t, node_id_t>> want, got;
    for (auto& kv : live) {
      if (!kv.second) continue;
      const size_t t = Bucket_Boruvka::get_index_depth(
          concat_pairing_fn(kv.first.first, kv.first.second), 31, nsub - 1);
      if (t >= g) {
        want.insert(kv.first);
        want.insert({kv.first.second, kv.first.first});
      }
    }
    for (auto& kv : alg.sink[g])
      if (kv.second) got.insert(kv.first);
    CHECK(got == want);
  }
  for (auto& kv : live)
    if (kv.second && Bucket_Boruvka::get_index_depth(
                         concat_pairing_fn(kv.first.first, kv.first.second), 31, nsub - 1) >= first)
      stored += 2;
  if (alg.es.get_num_edges() != stored) {
    std::printf("  stored %zu expected %zu (first %zu)\n", (size_t)alg.es.get_num_edges(), stored,
                first);
    std::map<std::pair<node_id_t, node_id_t>, int> cnt;
    for (const Edge& e : alg.es.get_edges()) cnt[{e.src, e.dst}]++;
    int shown = 0;
    for (auto& kv : cnt) {
      auto key = std::make_pair(std::min(kv.first.first, kv.first.second),
                                std::max(kv.first.first, kv.first.second));
      const size_t t = Bucket_Boruvka::get_index_depth(concat_pairing_fn(key.first, key.second), 31, nsub - 1);
      const bool want = live[key] && t >= first;
      if ((kv.second != 1 || !want) && shown++ < 10)
        std::printf("  extra %u->%u x%d tag %zu live %d\n", kv.first.first, kv.first.second, kv.second, t, live[key]);
    }
  }
  CHECK(alg.es.get_num_edges() == stored);
  CHECK(alg.es.get_edges().size() == stored);
  std::remove(path.c_str());
}

static void test_mc_routing_with_driver() { mc_routing_case(29, 3, 3000, 30000); }

// Contractions back to back under many workers: a budget barely above what one contraction
// leaves makes the next contraction start as soon as the previous one ends, so a vertex claimed
// by one contraction and handled late would meet the next one's state.
static void test_mc_routing_contractions_back_to_back() {
  for (uint64_t seed = 1; seed <= 12; seed++) mc_routing_case(1000 + seed, 8, 300, 20000);
}

static int sketch_dump(const char* pairs_path, node_id_t n, uint64_t seed, const char* out_path) {
  std::ifstream in(pairs_path, std::ios::binary);
  std::vector<uint32_t> p((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
  (void)p;
  in.clear();
  in.seekg(0, std::ios::end);
  const size_t bytes = (size_t)in.tellg();
  in.seekg(0);
  std::vector<uint32_t> pairs(bytes / 4);
  in.read(reinterpret_cast<char*>(pairs.data()), (std::streamsize)bytes);
  const vec_t len = Sketch::calc_vector_length(n);
  const size_t S = Sketch::calc_cc_samples(n, 1);
  Sketch proto(len, seed, S);
  const size_t nb = proto.get_buckets();
  std::vector<Bucket> all((size_t)n * nb);
  std::vector<std::unique_ptr<Sketch>> sk;
  for (node_id_t v = 0; v < n; v++) sk.emplace_back(new Sketch(len, seed, v, all.data(), S));
  for (size_t i = 0; i + 1 < pairs.size(); i += 2) {
    const node_id_t u = pairs[i], v = pairs[i + 1];
    const vec_t idx = concat_pairing_fn(u, v);
    sk[u]->update(idx);
    sk[v]->update(idx);
  }
  std::ofstream out(out_path, std::ios::binary);
  out.write(reinterpret_cast<const char*>(all.data()), (std::streamsize)(all.size() * sizeof(Bucket)));
  return out ? 0 : 1;
}

// GraphVerifier (the reference's test/util/graph_verifier.cpp checks) and AsciiFileStream: the
// cumulative text format, toggles, combine, and each check accepting a right answer and throwing
// on a wrong one
static void test_graph_verifier() {
  const std::string path = work_dir + "/verifier_graph.txt";
  {
    std::ofstream f(path);
    // two triangles {0,1,2}, {3,4,5}, a path 6-7, vertex 8 alone; the edge 2-3 listed twice
    f << "9 10\n0 1\n1 2\n2 0\n3 4\n4 5\n5 3\n6 7\n2 3\n3 2\n8 8\n";
  }
  GraphVerifier g(9);
  for (Edge e : std::vector<Edge>{{0, 1}, {1, 2}, {2, 0}, {3, 4}, {4, 5}, {5, 3}, {6, 7}}) g.edge_update(e);
  GraphVerifier from_file(9, path);  // (8, 8) is a self loop: it toggles nothing but itself
  from_file.edge_update({8, 8});
  CHECK(from_file == g);
  CHECK(g.get_num_kruskal_ccs() == 4);
  // the text stream reads the same file as updates
  {
    AsciiFileStream as(path, false);
    CHECK(as.vertices() == 9 && as.edges() == 10);
    GraphStreamUpdate buf[16];
    size_t got = as.get_update_buffer(buf, 16);
    CHECK(got == 10 && buf[7].edge.src == 2 && buf[7].edge.dst == 3 && buf[7].type == INSERT);
    CHECK(as.get_update_buffer(buf, 16) == 1 && buf[0].type == BREAKPOINT);
  }
  // connectivity: the right partition passes, a merged or split one throws
  const node_id_t right[9] = {0, 0, 0, 3, 3, 3, 6, 6, 8};
  g.verify_connected_components(ConnectedComponents(9, right));
  const node_id_t merged[9] = {0, 0, 0, 0, 0, 0, 6, 6, 8};
  const node_id_t split[9] = {0, 0, 2, 3, 3, 3, 6, 6, 8};
  bool t1 = false, t2 = false;
  try { g.verify_connected_components(ConnectedComponents(9, merged)); } catch (const IncorrectCCException&) { t1 = true; }
  try { g.verify_connected_components(ConnectedComponents(9, split)); } catch (const IncorrectCCException&) { t2 = true; }
  CHECK(t1 && t2);
  // edges
  g.verify_edge({1, 0});
  bool t3 = false;
  try { g.verify_edge({0, 3}); } catch (const BadEdgeException&) { t3 = true; }
  CHECK(t3);
  // forests: two disjoint spanning forests of the triangles' graph pass (the second spans the
  // graph minus the first); a cycle, a non-edge, a repeated edge and a non-spanning forest throw
  auto sf = [](std::vector<Edge> e) { return SpanningForest(9, std::move(e)); };
  g.verify_spanning_forests({sf({{0, 1}, {1, 2}, {3, 4}, {4, 5}, {6, 7}}), sf({{2, 0}, {5, 3}})});
  CHECK(g.num_edges() == 7);  // left unchanged
  bool c = false, ne = false, rep = false, span = false;
  try { g.verify_spanning_forests({sf({{0, 1}, {1, 2}, {2, 0}})}); } catch (const IncorrectForestException&) { c = true; }
  try { g.verify_spanning_forests({sf({{0, 3}})}); } catch (const BadEdgeException&) { ne = true; }
  try {  // the second forest reuses an edge the first one took
    g.verify_spanning_forests({sf({{0, 1}, {1, 2}, {3, 4}, {4, 5}, {6, 7}}), sf({{0, 1}})});
  } catch (const BadEdgeException&) { rep = true; }
  try { g.verify_spanning_forests({sf({{0, 1}, {3, 4}, {4, 5}, {6, 7}})}); } catch (const IncorrectForestException&) { span = true; }
  CHECK(c && ne && rep && span);
  CHECK(g.num_edges() == 7);
  // combine XORs the edge sets (two streams' graphs)
  GraphVerifier h(9);
  h.edge_update({6, 7});
  h.edge_update({7, 8});
  g.combine(h);
  CHECK(g.num_edges() == 7 && g.get_num_kruskal_ccs() == 4);
  auto m = g.extract_adj_matrix();
  CHECK(m[7][1] && !m[6][1] && m[0][1] && m[0][2]);
  CHECK(g != from_file);
}

int main(int argc, char** argv) {
  if (argc >= 2 && std::strcmp(argv[1], "sketch_dump") == 0 && argc == 6)
    return sketch_dump(argv[2], (node_id_t)std::stoul(argv[3]), std::stoull(argv[4]), argv[5]);
  if (argc >= 2 && std::strcmp(argv[1], "all") == 0) {
    if (argc >= 3) work_dir = argv[2];
    run("sketch_toggle_and_merge", test_sketch_toggle_and_merge);
    run("sketch_sampling", test_sketch_sampling);
    run("sketch_range_merge_and_serialize", test_sketch_range_merge_and_serialize);
    run("pairing", test_pairing);
    run("dsu", test_dsu);
    run("binary_stream", test_binary_stream);
    run("driver_delivers_every_update", test_driver_delivers_every_update);
    run("edge_store", test_edge_store);
    run("mc_routing_with_driver", test_mc_routing_with_driver);
    run("mc_routing_contractions_back_to_back", test_mc_routing_contractions_back_to_back);
    run("graph_verifier", test_graph_verifier);
    std::printf("%d failure(s)\n", failures);
    return failures ? 1 : 0;
  }
  std::fprintf(stderr, "usage: facade_cpu all <dir> | sketch_dump <pairs> <n> <seed> <out>\n");
  return 2;
}
