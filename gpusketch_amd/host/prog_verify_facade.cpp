// The C++ API as the reference's VERIFY_SAMPLES_F builds use it: this program is compiled with
// -DVERIFY_SAMPLES_F, so GraphSketchDriver records the exact graph of every stream it reads and
// hands it to the algorithm (set_verifier), whose queries then check their answers against it.
// The checks restate the reference's test/cc_alg_test.cpp (the multiples_graph_1024 fixture read
// as text, check_verifier against the cumulative file, 78 components; random streams) through the
// base classes the reference's code holds (CCSketchAlg*, MCSketchAlg*), and make sure a wrong
// verifier is noticed. Usage: verify_facade <multiples_graph_1024.txt> <work_dir>
#ifndef VERIFY_SAMPLES_F
#error "build with -DVERIFY_SAMPLES_F"
#endif
#include <cstdio>
#include <memory>
#include <random>

#include "gpusketch/cc_gpu_sketch_alg.h"
#include "gpusketch/graph_sketch_driver.h"
#include "gpusketch/mc_gpu_sketch_alg.h"
#include "gpusketch/sk_gpu_sketch_alg.h"

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

// test/cc_alg_test.cpp:41-57 on CCSketchAlg itself: the driver's verifier must equal the
// cumulative file's graph, and the answer (checked by the verifier inside the query) has 78
// components
static void small_graph_connectivity() {
  AsciiFileStream stream{fixture, false};
  const node_id_t n = stream.vertices();
  CCSketchAlg cc_alg{n, 0x5EED};
  GraphSketchDriver<CCSketchAlg> driver(&cc_alg, &stream, DriverConfiguration().worker_threads(2));
  driver.process_stream_until(END_OF_STREAM);
  driver.prep_query(CONNECTIVITY);
  driver.check_verifier(GraphVerifier(1024, fixture));
  CHECK(cc_alg.connected_components().size() == 78);
  // the spanning forest and k disjoint forests are verified inside the queries as well; they
  // read the sketches, so the gutters are flushed first (the cached connectivity answer let
  // prep_query(CONNECTIVITY) skip that), as the reference's tools do (tools/min_cut.cu:180)
  driver.prep_query(KSPANNINGFORESTS);
  CHECK(cc_alg.calc_spanning_forest().get_edges().size() == 1024 - 78);
  auto sfs = cc_alg.calc_disjoint_spanning_forests(3);
  CHECK(sfs.size() == 3);
  CHECK(cc_alg.connected_components().size() == 78);  // the forests were put back
}

// the GPU algorithms through base-class pointers, as code written for CCSketchAlg holds them
static void gpu_algorithms_through_the_base() {
  for (int which = 0; which < 2; which++) {
    AsciiFileStream stream{fixture, false};
    const node_id_t n = stream.vertices();
    std::unique_ptr<CCSketchAlg> alg;
    SketchParams p = make_sketch_params(n, 0xB0B + which);
    if (which == 0) {
      auto* g = new CCGPUSketchAlg(n, stream.edges(), 2, p);
      alg.reset(g);
      GraphSketchDriver<CCGPUSketchAlg> driver(g, &stream, DriverConfiguration().worker_threads(2));
      driver.process_stream_until(END_OF_STREAM);
      driver.prep_query(CONNECTIVITY);
      g->flush_buffers();
      driver.check_verifier(GraphVerifier(1024, fixture));
    } else {
      auto* s = new SKGPUSketchAlg(true, n, stream.edges(), 2, p);
      alg.reset(s);
      GraphSketchDriver<SKGPUSketchAlg> driver(s, &stream, DriverConfiguration().worker_threads(2));
      driver.process_stream_until(END_OF_STREAM);
      driver.prep_query(CONNECTIVITY);
      s->launch_gpu_kernel();
      driver.check_verifier(GraphVerifier(1024, fixture));
    }
    CCSketchAlg* base = alg.get();
    CHECK(base->connected_components().size() == 78);
    CHECK(base->point_query(2, 4) && !base->point_query(0, 1));
  }
}

// test/cc_alg_test.cpp:59-78 (random streams): a binary stream of random insert/delete toggles,
// its answer verified inside the query, and the recorded graph equal to the stream's
static void random_streams_verified() {
  std::mt19937_64 rng(77);
  for (int trial = 0; trial < 4; trial++) {
    const node_id_t n = 200 + 300 * trial;
    std::vector<GraphStreamUpdate> ups;
    GraphVerifier expect(n);
    for (int i = 0; i < 6 * (int)n; i++) {
      node_id_t a = (node_id_t)(rng() % n), b = (node_id_t)(rng() % n);
      if (a == b) continue;
      ups.push_back({{a, b}, INSERT});
      expect.edge_update({a, b});
      if (rng() % 10 == 0) {  // a later deletion of the same edge
        ups.push_back({{b, a}, DELETE});
        expect.edge_update({b, a});
      }
    }
    const std::string path = work_dir + "/verify_random_" + std::to_string(trial) + ".data";
    {
      BinaryFileStream w(path, false);
      w.write_header(n, ups.size());
      w.write_updates(ups.data(), ups.size());
    }
    BinaryFileStream stream(path);
    CCSketchAlg alg{n, 1000u + (size_t)trial};
    GraphSketchDriver<CCSketchAlg> driver(&alg, &stream, DriverConfiguration().worker_threads(3), 2);
    driver.process_stream_until(END_OF_STREAM);
    driver.prep_query(CONNECTIVITY);
    driver.check_verifier(expect);
    CHECK(alg.connected_components().size() == expect.get_num_kruskal_ccs());
  }
}

// a verifier that disagrees with the sketches must make the query throw
static void wrong_verifier_is_noticed() {
  AsciiFileStream stream{fixture, false};
  const node_id_t n = stream.vertices();
  CCSketchAlg alg{n, 0x5EED};
  GraphSketchDriver<CCSketchAlg> driver(&alg, &stream, DriverConfiguration().worker_threads(2));
  driver.process_stream_until(END_OF_STREAM);
  driver.prep_query(CONNECTIVITY);
  auto wrong = std::make_unique<GraphVerifier>(1024, fixture);
  wrong->edge_update({0, 1});  // joins two of the 78 components
  alg.set_verifier(std::move(wrong));
  bool threw = false;
  try {
    alg.connected_components();
  } catch (const IncorrectCCException&) {
    threw = true;
  }
  CHECK(threw);
  bool threw_bad = false;
  try {
    driver.check_verifier(GraphVerifier(1024));
  } catch (const DriverException&) {
    threw_bad = true;
  }
  CHECK(threw_bad);
}

// MCGPUSketchAlg through MCSketchAlg*: k disjoint spanning forests of sketch graph 0 are
// verified (forest i spans the graph minus forests 0..i-1) and connectivity too
static void mc_through_the_base() {
  AsciiFileStream stream{fixture, false};
  const node_id_t n = stream.vertices();
  SketchParams p = make_sketch_params(n, 0xC0C0, 2.0);
  const int num_subgraphs = 1 + 2 * 10;
  auto* mc = new MCGPUSketchAlg(n, 2, 1, p, num_subgraphs, num_subgraphs, 4, size_t(1) << 30, 1,
                                CCAlgConfiguration().sketches_factor(2.0));
  std::unique_ptr<MCSketchAlg> base(mc);
  GraphSketchDriver<MCGPUSketchAlg> driver(mc, &stream, DriverConfiguration().worker_threads(2));
  driver.process_stream_until(END_OF_STREAM);
  driver.prep_query(KSPANNINGFORESTS);
  mc->apply_flush_updates();
  driver.check_verifier(GraphVerifier(1024, fixture));
  CHECK(mc->get_num_sketch_graphs() >= 1);
  auto sfs = base->calc_disjoint_spanning_forests(0, 4);
  CHECK(sfs.size() == 4);
  CHECK(base->connected_components().size() == 78);

  // MCSketchAlg constructed directly, sketch graph 0 created by the caller
  AsciiFileStream s2{fixture, false};
  MCSketchAlg direct(n, 0xC0C1, 1);
  direct.create_sketch_graph(0, make_sketch_params(n, 0xC0C1));
  GraphSketchDriver<MCSketchAlg> d2(&direct, &s2, DriverConfiguration().worker_threads(2));
  d2.process_stream_until(END_OF_STREAM);
  d2.prep_query(KSPANNINGFORESTS);
  CHECK(direct.calc_disjoint_spanning_forests(0, 2).size() == 2);
  CHECK(direct.connected_components().size() == 78);
}

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s <multiples_graph_1024.txt> <work_dir>\n", argv[0]);
    return 2;
  }
  fixture = argv[1];
  work_dir = argv[2];
  run("small_graph_connectivity", small_graph_connectivity);
  run("gpu_algorithms_through_the_base", gpu_algorithms_through_the_base);
  run("random_streams_verified", random_streams_verified);
  run("wrong_verifier_is_noticed", wrong_verifier_is_noticed);
  run("mc_through_the_base", mc_through_the_base);
  std::printf("%d failure(s)\n", failures);
  return failures ? 1 : 0;
}
