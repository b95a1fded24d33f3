/*
 * libgpusketch -- C ABI of the MI355X (gfx950) l0-sampling sketch engine.
 *
 * Plain pointers and sizes only. Every function returns a gs_status (0 = GS_OK) and never exits
 * the process; gs_last_error() returns a message for the calling thread's last failure.
 * Host pointers are caller-owned and may be reused as soon as a call returns (the library copies
 * them into its own pinned staging first). Functions named *_device take device pointers.
 *
 * Each declaration names the reference interface it replaces (paths relative to the reference
 * repository).
 */
#ifndef GPUSKETCH_H
#define GPUSKETCH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gs_store gs_store;
typedef struct gs_stream gs_stream;

/* One bucket, byte-identical to the reference Bucket {vec_t alpha; vec_hash_t gamma}
 * (include/bucket.h:7-12) with vec_hash_t 64-bit. */
typedef struct gs_bucket {
  uint64_t alpha;
  uint64_t gamma;
} gs_bucket;

enum gs_status {
  GS_OK = 0,
  GS_ERR_INVALID = 1,         /* bad argument or state */
  GS_ERR_OOM = 2,             /* device or pinned allocation failed */
  GS_ERR_HIP = 3,             /* HIP runtime error (message has the call and code) */
  GS_ERR_OUT_OF_SAMPLES = 4,  /* Sketch::sample past num_samples (include/sketch.h:219-230) */
  GS_ERR_ADMISSION = 5,       /* serving: request exceeds the sketch-memory budget */
  GS_ERR_UNSUPPORTED = 6
};

/* Sketch geometry; mirrors SketchParams (include/cuda_kernel.cuh:23-37) plus the derived sizes
 * of include/sketch.h:63-77,198-216. */
typedef struct gs_geometry {
  uint32_t num_vertices;
  uint32_t bkt_per_col;     /* depths per column, B */
  uint32_t num_samples;     /* S */
  uint32_t cols_per_sample; /* columns scanned per sample: 1 by default, 7 under L0_SAMPLING */
  uint32_t num_columns;     /* C = S * cols_per_sample */
  uint32_t num_buckets;     /* C * B + 1, the reference per-vertex bucket count */
  uint64_t seed;
  uint64_t vector_len;      /* n(n-1)/2 */
  uint64_t device_bytes;    /* HBM held by the store */
  uint32_t flags;           /* GS_STORE_* flags of the store (or of the sizing request) */
} gs_geometry;

/* gs_store_create flags.
 * GS_STORE_CUBESKETCH: the L0_SAMPLING update rule (src/sketch.cpp:66-83): an update of depth d in
 * a column toggles every depth 0..d of it (default: depth d only, src/sketch.cpp:85-99). */
#define GS_STORE_CUBESKETCH 1u

/* Thread-local message of the last failing call ("" if none). */
const char* gs_last_error(void);

/* Build provenance (no counterpart in the reference): writes the SHA-256 (64 hex digits and a
 * NUL, so cap >= 65) that the build recipe took over the library's sources and compile flags, so
 * a caller can check that the loaded library was built from the sources beside it. */
int gs_source_digest(char* out, size_t cap);

/* Sizing: Sketch::calc_vector_length / calc_cc_samples / calc_bkt_per_col
 * (include/sketch.h:63-77,198-206) for n vertices and a CCAlgConfiguration sketches_factor. */
int gs_geometry_for(uint32_t n, double sketches_factor, uint32_t cols_per_sample,
                    gs_geometry* out);
/* The same for a sketch variant: flags GS_STORE_CUBESKETCH takes the L0_SAMPLING num_samples_div
 * (include/sketch.h:208-211, log2(3) - 1) instead of the default (:213-215). */
int gs_geometry_for_sketch(uint32_t n, double sketches_factor, uint32_t cols_per_sample,
                           uint32_t flags, gs_geometry* out);

/* Device XXH3 known-answer check: out[i] = XXH3_64bits_withSeed(&keys[i], 8, seeds[i]) computed
 * by a device kernel (replaces CUDA_XXH64, include/cuda_xxhash64.cuh:587-591). Host arrays. */
int gs_hash_known_answer(const uint64_t* keys, const uint64_t* seeds, uint64_t count,
                         uint64_t* out);

/* Create a zeroed sketch store of n vertex sketches resident in HBM, all with the same seed
 * (the reference gives every vertex the same seed, src/cc_sketch_alg.cpp:11-22). Replaces the
 * cudaMallocManaged sketch buffer + SketchParams set-up (tools/cuda_process_stream.cu:97-116,
 * include/cc_gpu_sketch_alg.h:36-97). flags: GS_STORE_* (0 = the default update rule).
 * num_samples * cols_per_sample columns, at most 4096. */
int gs_store_create(uint32_t n, uint64_t seed, uint32_t num_samples, uint32_t cols_per_sample,
                    uint32_t flags, gs_store** out);
/* Applies what its streams hold pending, then frees the store and its streams. Refused
 * (GS_ERR_INVALID, nothing freed) while the store is a member of a group: destroy the group
 * first. */
int gs_store_destroy(gs_store* st);
int gs_store_geometry(const gs_store* st, gs_geometry* out);
/* Sketch::zero_contents for every vertex (src/sketch.cpp:102-108). */
int gs_store_zero(gs_store* st);

/* A submission stream (one HIP stream + staging + scratch). Mirrors CudaStream<Alg>
 * (include/cuda_stream.h:51-97): one per worker thread id. Streams of one store may be used
 * from different host threads concurrently; a single stream must not. */
int gs_stream_create(gs_store* st, gs_stream** out);
int gs_stream_destroy(gs_stream* s);

/* Pre-grouped vertex batches: batch b toggles concat(src[b], dsts[offsets[b] + i]) for
 * i < sizes[b] in the sketch of src[b]. The reference contract of CudaKernel::sketchUpdate /
 * sketchUpdate_UVM_kernel (include/cuda_kernel.cuh:46, src/cuda_kernel.cu:152-257) as driven by
 * CudaStream::process_batch (include/cuda_stream.h:99-213). Asynchronous: returns after the
 * host arrays are staged. */
int gs_submit_vertex_batches(gs_stream* s, const uint32_t* src, const uint64_t* sizes,
                             const uint64_t* offsets, uint32_t nbatches, const uint32_t* dsts);

/* Subgraph-filtered vertex batches for the min-cut path: only destinations whose edge depth
 * get_index_depth(concat(src, dst), seed, depth_cap) is >= min_depth are applied, the routing of
 * MCGPUSketchAlg::apply_update_batch / complete_update_batch (src/mc_gpu_sketch_alg.cpp:80-178)
 * for the sketch subgraph min_depth. min_depth = 0 applies everything. */
int gs_submit_vertex_batches_filtered(gs_stream* s, const uint32_t* src, const uint64_t* sizes,
                                      const uint64_t* offsets, uint32_t nbatches,
                                      const uint32_t* dsts, uint32_t min_depth,
                                      uint32_t depth_cap);

/* Raw undirected updates: pairs[2i], pairs[2i+1] toggles edge {u,v} in both endpoints'
 * sketches, as GraphSketchDriver inserts {u,v} and {v,u} (include/graph_sketch_driver.h:171-172)
 * and CCSketchAlg::update applies them (src/cc_sketch_alg.cpp:154-160). Grouping by vertex
 * happens on the device. Host memory (pinned or pageable); asynchronous after staging. For every
 * edge-pair submission (host or device memory): self loops are no-ops, and pairs naming a vertex
 * >= n are skipped and make the stream's next gs_flush return GS_ERR_INVALID with their count;
 * the other pairs of the submission are applied. */
int gs_submit_edge_pairs(gs_stream* s, const uint32_t* pairs, uint64_t nedges);
/* Same, input already resident in device memory. It must stay valid and unchanged until the
 * stream's next gs_flush (or gs_sync): long batches (2^18 edges and more) are read in place by
 * the launch that applies them, shorter ones are copied on the stream. The library never writes
 * it; pairs naming a vertex >= n are skipped and reported by that flush. */
int gs_submit_edge_pairs_device(gs_stream* s, const uint32_t* d_pairs, uint64_t nedges);
/* Subgraph-filtered variant for the min-cut path: only edges whose subgraph depth
 * get_index_depth(idx, seed, depth_cap) (src/mc_gpu_sketch_alg.cpp:160) is >= min_depth are
 * applied. Host memory. */
int gs_submit_edge_pairs_filtered(gs_stream* s, const uint32_t* pairs, uint64_t nedges,
                                  uint32_t min_depth, uint32_t depth_cap);

/* Wait until all work submitted on the stream is finished (CudaStream::flush_buffers,
 * include/cuda_stream.h:215-276). */
int gs_flush(gs_stream* s);
/* Launch the updates the stream holds pending (a partly filled batch) without waiting for them:
 * the half of gs_flush that the reference's flush_buffers performs before its wait. A later
 * gs_flush, gs_sync or query of the store waits for them; dropped pairs are reported by the next
 * gs_flush. For callers that overlap their own host work with the device's (a serving loop). */
int gs_flush_async(gs_stream* s);
/* Wait for every stream of the store (CCGPUSketchAlg::flush_buffers,
 * src/cc_gpu_sketch_alg.cpp:21-31). */
int gs_sync(gs_store* st);

/* Copy `count` vertex sketches starting at v0 to host memory in reference bucket order
 * [column * bkt_per_col + depth], deterministic bucket last (src/sketch.cpp:8-37); count *
 * num_buckets buckets. The byte image equals Sketch::serialize (src/sketch.cpp:204-206). */
int gs_export_buckets(gs_store* st, uint32_t v0, uint32_t count, gs_bucket* out);
/* Overwrite `count` vertex sketches from reference-ordered buckets (the istream constructor,
 * src/sketch.cpp:39-51, used by CCSketchAlg::construct_from_serialized_data :57-91). */
int gs_import_buckets(gs_store* st, uint32_t v0, uint32_t count, const gs_bucket* in);
/* Device-order snapshot of a whole store (its buckets as the device holds them, and whether its
 * vertex sketches XOR to zero) for spilling it to host memory: one DMA each way, straight from or
 * into pinned memory (gs_host_alloc). gs_store_snapshot waits for the store's pending updates;
 * gs_store_restore refuses a snapshot of another geometry or seed (GS_ERR_INVALID) and replaces
 * the store's contents. The format is this library's own, not the reference order
 * (gs_export_buckets writes that). No counterpart in the reference; the serving loop uses it to
 * evict and restore tenants. */
int gs_store_snapshot_bytes(gs_store* st, uint64_t* bytes);
int gs_store_snapshot(gs_store* st, void* host_dst);
int gs_store_restore(gs_store* st, const void* host_src);

/* Connected components by Borůvka emulation on the device (CCSketchAlg::boruvka_emulation +
 * connected_components, src/cc_sketch_alg.cpp:491-584). label_out[v] = smallest vertex id of
 * v's component (n entries). forest_out (may be NULL) receives up to forest_cap spanning-forest
 * edges as (min, max) pairs; *forest_len is the forest size. Returns GS_ERR_OUT_OF_SAMPLES when
 * the sketches run out of columns (OutOfSamplesException). */
int gs_query_cc(gs_store* st, uint32_t* label_out, uint32_t* num_cc, uint32_t* rounds,
                uint32_t* forest_out, uint64_t forest_cap, uint64_t* forest_len);

/* Test hook: Sketch::sample (src/sketch.cpp:110-132) of every vertex sketch at sample index
 * `sample_idx`, computed by the device Borůvka's round-0 samplers. result[v] = 0 GOOD (idx_out[v]
 * = the sampled edge index), 1 ZERO, 2 FAIL; n entries each. GS_ERR_OUT_OF_SAMPLES when
 * sample_idx >= num_samples (OutOfSamplesException, include/sketch.h:219-230). */
int gs_sample_vertices(gs_store* st, uint32_t sample_idx, int32_t* result, uint64_t* idx_out);

/* k edge-disjoint spanning forests (MCSketchAlg::calc_disjoint_spanning_forests,
 * src/mc_sketch_alg.cpp:889-920): forest i samples column i in round 0, its edges are toggled
 * out before forest i+1, and all are toggled back in before return. edges_out holds
 * sum(counts) (min,max) pairs, forest after forest; capacity edge_cap pairs. */
int gs_k_spanning_forests(gs_store* st, uint32_t k, uint32_t* edges_out, uint64_t edge_cap,
                          uint64_t* counts);

/* Exact global minimum cut of an undirected multigraph given as an edge list (replaces the
 * VieCut call of MCSketchAlg::calc_minimum_cut, src/mc_sketch_alg.cpp:966-1003). side_out[v]
 * = 1 for the vertices of one side. Host-only computation. */
int gs_min_cut(uint32_t n, const uint32_t* pairs, uint64_t m, uint64_t* value,
               uint8_t* side_out);

/* Kernel timing of the update path, for measurement. When enabled, the aggregation kernel's
 * launches on this stream are bracketed by HIP events; gs_stream_timing returns the summed
 * milliseconds, the launch count and the directed updates those launches processed. */
int gs_stream_set_profiling(gs_stream* s, int enable);
int gs_stream_timing(gs_stream* s, double* agg_ms, uint64_t* launches, uint64_t* updates);
/* The same, plus the summed milliseconds of the grouping kernels (count, offsets, scatter, sort)
 * that preceded the profiled edge-pair launches. Like gs_stream_timing it returns and resets the
 * sums. No counterpart in the reference (its kernel time is cudaEvent-timed in
 * SKGPUSketchAlg::launch_gpu_kernel, src/sk_gpu_sketch_alg.cpp:156-179). */
int gs_stream_kernel_times(gs_stream* s, double* grouping_ms, double* agg_ms, uint64_t* launches,
                           uint64_t* updates);
/* HBM work of the profiled aggregation launches, summed and reset: device buckets read-modify-
 * written by flushes (non-zero delta slots and deterministic buckets) and (vertex, bucket)
 * updates beyond the LDS window XORed straight into HBM. No counterpart in the reference (its
 * benchmarks report throughput only); used for the bytes of the roofline. */
int gs_stream_traffic(gs_stream* s, uint64_t* flushed_buckets, uint64_t* direct_updates);

/* ---- Tuning ----------------------------------------------------------------------------------
 * Launch parameters of a stream's update path (0 = default): lds_depths = depths of every column
 * kept in the aggregation kernel's LDS delta (1..16, default 9 to 12 by width; at most 12 are used
 * for sketches of up to 64 columns), item_updates = destinations per work
 * item (64..1024, default 512; 1024 for wide sketches), bucket_bits = vertex-id bits resolved
 * inside one grouping bucket (default log2(n) - 8, at most 13), launch_edges = submitted edge
 * pairs coalesced into one launch of the grouping and aggregation kernels (2^12..2^28; default
 * 256 per vertex of the stream's vertex space, within [2^16, 2^27]; scratch ≈ 40 B per edge of
 * it; a flush or a query applies a partial launch). The reference fixes its launch shape at
 * compile time
 * (num_device_threads = 1024, num_batch_per_buffer = 540, include/cc_gpu_sketch_alg.h:16-23). */
typedef struct gs_tuning {
  uint32_t lds_depths;
  uint32_t item_updates;
  uint32_t bucket_bits;
  uint32_t launch_edges;
  float kernel_ms; /* autotune: the winner's time for the sample batch */
} gs_tuning;
int gs_stream_set_tuning(gs_stream* s, const gs_tuning* t);
int gs_stream_get_tuning(gs_stream* s, gs_tuning* t);
/* Time every candidate on a device-resident sample batch (each applied twice, which leaves the
 * sketches unchanged) and keep the fastest on the stream. */
int gs_stream_autotune(gs_stream* s, const uint32_t* d_pairs, uint64_t nedges, gs_tuning* best);

/* ---- Batched streams: groups of stores ------------------------------------------------------
 * Many independent graphs (stores of any sizes, seeds and geometries) updated by one launch
 * sequence of the grouping and aggregation kernels and queried by one batched Borůvka. Member i's
 * vertex v is global vertex base_i + v of one virtual vertex space (at most 2^26 vertices). The
 * group owns a stream registered as a writer of every member; destroy the group before its
 * members; calls on one group are not concurrent. The reference holds one graph per algorithm
 * object (include/cc_gpu_sketch_alg.h:8-110); the per-member results equal what that graph's own
 * store would give (each member's sketches bit-exact, partitions identical). */
typedef struct gs_group gs_group;
int gs_group_create(gs_store* const* stores, uint32_t count, gs_group** out);
/* Applies the updates the group's stream holds pending, then frees the group (its members stay). */
int gs_group_destroy(gs_group* g);
int gs_group_info(const gs_group* g, uint32_t member, uint32_t* base, uint32_t* total);
/* The group's stream, for gs_flush / gs_stream_set_profiling / gs_stream_timing (not for
 * gs_stream_destroy). */
int gs_group_stream(gs_group* g, gs_stream** out);
/* Host edge pairs of every member (member-local ids; NULL with nedges[i] = 0 skips a member),
 * ingested as one stream of the group's vertex space. Every id is checked first: an id past its
 * member's end fails the call (GS_ERR_INVALID) before anything is submitted. */
int gs_group_submit_pairs(gs_group* g, const uint32_t* const* pairs, const uint64_t* nedges);
/* Host edge pairs of one member in its own ids, taken as they are: the DMA reads them straight
 * from the caller's memory when it is pinned (gs_host_alloc), and the device adds the member's
 * base and checks the range while appending them. A pair with an id >= the member's vertex count
 * is skipped and counted like an out-of-range pair (the group stream's next gs_flush returns
 * GS_ERR_INVALID); the other pairs are applied. Returns once the pairs have been read, so the
 * buffer may be reused. No host pass over the pairs (a serving loop's payloads go to the device
 * as they arrived). No counterpart in the reference. */
int gs_group_submit_member_pairs(gs_group* g, uint32_t member, const uint32_t* pairs,
                                 uint64_t nedges);
/* gs_group_submit_member_pairs without the wait for the DMA: pinned pairs (gs_host_alloc) are
 * read by a DMA that may still run after the call returns, so the caller keeps them unchanged
 * until gs_stream_wait_copies on the group's stream (or its gs_flush / gs_sync) has returned;
 * other memory is staged before the call returns. A serving loop enqueues a step's payloads back
 * to back this way. No counterpart in the reference. */
int gs_group_submit_member_pairs_async(gs_group* g, uint32_t member, const uint32_t* pairs,
                                       uint64_t nedges);
/* gs_group_submit_member_pairs_async for each i < count in order (batch i: members[i],
 * pairs[i], nedges[i]), with the per-batch work on the host cut to the DMA: the device rebase and
 * range check run as one launch per up to 64 batches. Every argument is checked before anything
 * is enqueued (GS_ERR_INVALID and nothing applied). A serving loop hands a step's payloads over
 * in one call. No counterpart in the reference. */
int gs_group_submit_member_batches_async(gs_group* g, uint32_t count, const uint32_t* members,
                                         const uint32_t* const* pairs, const uint64_t* nedges);
/* Waits until every host batch submitted on the stream so far has been read by its DMA (the
 * buffers of gs_group_submit_member_pairs_async may then be reused). No counterpart in the
 * reference. */
int gs_stream_wait_copies(gs_stream* s);
/* Device-resident pairs already in global ids (base_i added); must stay valid until gs_flush.
 * A pair whose endpoints lie in two different members is skipped and counted like an
 * out-of-range pair (the group stream's next gs_flush returns GS_ERR_INVALID). */
int gs_group_submit_pairs_device(gs_group* g, const uint32_t* d_pairs, uint64_t nedges);
/* Allocates now what the group's first submission and first query would allocate (the stream's
 * fill buffers and grouping scratch at its current launch size, the pinned staging, the batched
 * query's scratch and captured round), so that a serving loop allocates nothing after it starts.
 * Call after gs_stream_set_tuning on the group's stream. No counterpart in the reference. */
int gs_group_prepare(gs_group* g);
/* Connected components of the members selected by member_mask (NULL: all) in one batched
 * Borůvka whose rounds >= 1 replay a captured HIP graph; its kernels walk the selected members'
 * vertices only. labels: total entries, member-local smallest vertex of each component for the
 * selected members (the entries of the others are left unchanged); num_cc / rounds / status: one
 * entry per member (n / 0 / GS_OK for members not selected), status GS_OK or
 * GS_ERR_OUT_OF_SAMPLES (the call then also returns GS_ERR_OUT_OF_SAMPLES). */
int gs_group_query_cc(gs_group* g, const uint8_t* member_mask, uint32_t* labels,
                      uint32_t* num_cc, uint32_t* rounds, int32_t* status);
/* Launches the group's pending updates and marks the point its next gs_group_query_cc waits
 * for: that query then runs on the group's query stream after the updates submitted before the
 * fence (and after the queried members' other writers), while group updates submitted after the
 * fence may run beside it. The caller must not submit updates of the queried members between
 * the fence and the query. Without a fence the query waits for every writer of every member.
 * No counterpart in the reference; the serving loop overlaps a step's queries with the updates of
 * the tenants it does not query. */
int gs_group_fence(gs_group* g);

/* ---- Allocation-free update ----------------------------------------------------------------
 * No counterpart in the reference: an entry point for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * It applies device-resident edge pairs exactly as gs_submit_edge_pairs_device followed by
 * gs_flush would (the same grouping and aggregation kernels), but takes every scratch buffer
 * from a fixed device array of GS_NOALLOC_SCRATCH_BYTES that each GPU holds from the library's
 * first use of it until the process ends, and calls no hipMalloc/hipFree. At most
 * GS_NOALLOC_MAX_EDGES pairs per call (GS_ERR_INVALID beyond, or when the store's vertex count
 * would need more scratch); it runs on the store's internal stream and returns when the updates
 * are applied. Pairs naming a vertex >= n are skipped and reported (GS_ERR_INVALID) by the same
 * call. The store's streams must have nothing pending or in flight (gs_sync first), and calls
 * that overlap on one GPU race: they share that GPU's scratch array. */
#define GS_NOALLOC_MAX_EDGES (1ull << 20)
#define GS_NOALLOC_SCRATCH_BYTES (100ull << 20)
int gs_apply_edge_pairs_device_noalloc(gs_store* st, const uint32_t* d_pairs, uint64_t nedges);

/* Pinned host memory for staging buffers (replaces the cudaMallocHost calls of CudaStream and
 * SKGPUSketchAlg, include/cuda_stream.h:51-97, include/sk_gpu_sketch_alg.h:72). */
int gs_host_alloc(size_t bytes, void** hptr);
int gs_host_free(void* hptr);

/* Device memory for callers that keep inputs resident (bench, batched and serving modes). */
int gs_device_alloc(size_t bytes, void** dptr);
int gs_device_free(void* dptr);
int gs_memcpy_h2d(void* dptr, const void* src, size_t bytes);
int gs_memcpy_d2h(void* dst, const void* dptr, size_t bytes);
int gs_device_synchronize(void);
/* Device-to-device copy by a plain streaming kernel (16-byte loads and stores); also the
 * known-size reference kernel that calibrates HBM byte counters. */
int gs_device_copy(void* dst, const void* src, size_t bytes);

/* Synthetic seeded streams (the reference's tools/stream_generators write similar binary
 * streams). Each returns undirected updates as (u, v) pairs with toggle semantics: every
 * deletion removes a previously inserted, still live edge; no self loops; no duplicate live
 * edges. Call with pairs == NULL to get the count first. */
/* Erdős–Rényi G(n, p) inserts in seeded random order, then `del_frac` of them re-deleted at
 * random later positions. */
int gs_gen_erdos_renyi(uint32_t n, double p, double del_frac, uint64_t seed, uint32_t* pairs,
                       uint64_t* count);
/* R-MAT(a, b, c, 1-a-b-c) with n = 2^scale: `count` updates, about `del_frac` of which delete an
 * earlier inserted edge. R-MAT samples are drawn in order; self loops and edges that were drawn
 * before are dropped (every edge is inserted at most once), and each insertion is deleted with
 * probability del_frac / (1 - del_frac) behind a uniformly chosen later insertion. Generated on
 * all host cores (GPUSKETCH_GEN_THREADS overrides the count); the stream depends only on the
 * arguments. */
int gs_gen_rmat(uint32_t scale, double a, double b, double c, uint64_t count, double del_frac,
                uint64_t seed, uint32_t* pairs);
/* Two ER(p) halves of n/2 vertices joined by exactly `cut` crossing edges, shuffled, plus
 * churn_frac * m insert/delete churn pairs of random non-edges. */
int gs_gen_planted_cut(uint32_t n, double p, uint32_t cut, double churn_frac, uint64_t seed,
                       uint32_t* pairs, uint64_t* count);

#ifdef __cplusplus
}
#endif

#endif /* GPUSKETCH_H */
