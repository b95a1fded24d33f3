// Internal definitions shared by the HIP translation units of libgpusketch.
//
// Device layout of a sketch store (gfx950-first, see the store section below):
//   rows : per vertex, bkt_per_col x num_columns buckets of {alpha, gamma}: depths 0..11 as depth
//          rows of num_columns buckets (the low depths that almost every update touches are
//          contiguous and a wave's 64 lanes flush consecutive 16-byte buckets), then each
//          column's deeper depths as one run (bucket_slot); the per-vertex stride is padded to a
//          multiple of 8 buckets (128 bytes) so every vertex starts on an L2 line.
//   det  : the deterministic bucket of every vertex, in a separate dense array.
// Reference numbering (src/sketch.cpp:8-37) is bucket [column*bkt_per_col + depth] with the
// deterministic bucket last; export/import convert.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gpusketch.h"
#include "../../include/gpusketch/xxh3_len8.h"

namespace gs {

// ------------------------------------------------------------------------------------------
// error reporting
// ------------------------------------------------------------------------------------------
void set_error(const char* fmt, ...);
int fail(int code, const char* fmt, ...);

#define GS_HIP(call)                                                                         \
  do {                                                                                       \
    hipError_t _e = (call);                                                                  \
    if (_e != hipSuccess)                                                                    \
      return ::gs::fail(GS_ERR_HIP, "%s:%d %s: %s", __FILE__, __LINE__, #call,               \
                        hipGetErrorString(_e));                                              \
  } while (0)

#define GS_CHECK(cond, code, ...)                                                            \
  do {                                                                                       \
    if (!(cond)) return ::gs::fail((code), __VA_ARGS__);                                     \
  } while (0)

// ------------------------------------------------------------------------------------------
// store geometry and device view
// ------------------------------------------------------------------------------------------
struct Bucket16 {
  uint64_t alpha, gamma;
};

struct StoreView {
  Bucket16* rows;       // n * vstride buckets
  Bucket16* det;        // n buckets
  uint32_t n;           // vertices
  uint32_t bpc;         // depths per column (B)
  uint32_t cols;        // columns (C)
  uint32_t vstride;     // buckets per vertex in rows (>= B*C, multiple of 8)
  uint64_t seed;
};

#if defined(__HIP__)
// XOR of a 32-bit value over the 64 lanes with DPP (quad swaps, half-row and row mirrors, then
// row broadcasts); the total ends in lane 63 and is returned as a uniform value.
__device__ __forceinline__ uint32_t wave_xor_u32(uint32_t v) {
  v ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);   // quad_perm 1,0,3,2
  v ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false);   // quad_perm 2,3,0,1
  v ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, false);  // row_half_mirror
  v ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x140, 0xF, 0xF, false);  // row_mirror
  v ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);  // row_bcast15
  v ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);  // row_bcast31
  return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

__device__ __forceinline__ uint64_t wave_xor_reduce(uint64_t x) {
  return ((uint64_t)wave_xor_u32((uint32_t)(x >> 32)) << 32) | wave_xor_u32((uint32_t)x);
}

#endif

// Bucket (depth d, column c) inside a vertex's row of B x C buckets. Depths below kRowDepths
// are depth-major rows of C buckets (the LDS window of the aggregation kernel, where almost every
// update lands: a flush writes consecutive buckets); the deeper depths of each column follow as
// one run of B - kRowDepths buckets per column (reached by about 2^-12 of the column updates, and
// read by a query about 3 lines at a time instead of one 128-byte line per depth). A query reads
// one column per vertex and round: kRowDepths + 3..4 lines instead of B.
constexpr uint32_t kRowDepths = 12;
__host__ __device__ __forceinline__ uint32_t bucket_slot(uint32_t d, uint32_t c, uint32_t cols,
                                                         uint32_t bpc) {
  return d < kRowDepths ? d * cols + c : kRowDepths * cols + c * (bpc - kRowDepths) + (d - kRowDepths);
}

// work item of the aggregation kernel: `len` destinations of vertex `vertex` starting at
// `start` in the destination array; bit 31 of `len` = the vertex is owned by this item alone;
// `graph` = member index when a group of stores is ingested together (0 otherwise).
struct WorkItem {
  uint32_t vertex;
  uint32_t len;
  uint32_t start;
  uint32_t graph;
};
constexpr uint32_t kItemExclusive = 0x80000000u;

struct Store;
struct Group;

// one member of a group of stores that share a virtual vertex space: global vertex
// base + v is vertex v of this store
struct GraphDesc {
  Bucket16* rows;
  Bucket16* det;
  uint32_t n, bpc, cols, vstride;
  uint64_t seed;
  uint32_t base;
  uint32_t samples;  // num_samples: Boruvka rounds available
};

// vertex -> bucket mapping of the edge-pair grouping: hashed = (v * mul) mod 2^nbits (a
// bijection), bucket = hashed >> local_bits, local id = low local_bits bits
struct PartParams {
  uint32_t nbits, local_bits, nbuckets, mask, mul, inv;
};

// The pairs of one launch: up to kMaxSegs segments (the fill buffer, then device batches read in
// place), each covered by whole K1/K3 workgroups of 4096 edges; blk0[i] is segment i's first
// workgroup, blk0[nseg] the total.
constexpr int kMaxSegs = 48;  // a 2^25-edge launch of 2^20-edge device batches takes 32
struct PairSegs {
  const uint2* ptr[kMaxSegs];
  uint64_t n[kMaxSegs];
  uint32_t blk0[kMaxSegs + 1];
  uint32_t nseg;
};

// per-stream scratch for the edge-pair pipeline, grown on demand
// Work items are grouped into length classes (class = floor(log2(len)), len <= 1023) so that
// the aggregation kernel can hand out the longest items first.
constexpr int kClasses = 11;
constexpr int kClassBase = 4;  // counters[kClassBase + k]: items of class k in the current batch
struct ItemClasses {
  uint64_t off[kClasses];  // start of each class's region in the item array
};

// most edges coalesced per launch of the grouping and aggregation kernels by default (the
// default is 256 per vertex of the stream's vertex space, within [2^16, kLaunchEdges])
constexpr uint64_t kLaunchEdges = 1ULL << 27;

struct PairScratch {
  uint64_t cap_edges = 0;        // capacity in undirected edges
  uint2* d_entries = nullptr;    // directed entries (vertex, other) grouped by bucket (2 x cap)
  uint32_t* d_sorted = nullptr;  // destinations grouped by vertex (2 x cap)
  uint64_t cap_items = 0;
  WorkItem* d_items = nullptr;
  uint64_t cap_hist = 0;          // words of d_hist
  uint32_t* d_hist = nullptr;     // [bucket][partition block]: entries, then their offsets
  uint32_t cap_buckets = 0;
  uint32_t* d_totals = nullptr;   // [bucket]: entries of the launch, then [nbuckets + bucket]:
                                  // start of the bucket's region
  ItemClasses cls{};
  uint32_t* d_counters = nullptr;  // [0..2] unused, [3]=dropped out-of-range entries since the
                                   // last flush, [kClassBase + k]=items of class k
};
constexpr int kCounterWords = 16;

struct Stream {
  Store* store = nullptr;  // the store this stream updates, or
  Group* group = nullptr;  // the group of stores it updates together
  hipStream_t stream = nullptr;
  bool own_stream = true;  // false: `stream` belongs to someone else (a store's query stream)
  hipStream_t copy_stream = nullptr;
  // Edge-pair launches are pipelined: the appends into the fill buffer and the grouping kernels
  // run on gstream, the aggregation kernel on `stream` (after an event of its grouping), so the
  // grouping of launch k+1 runs while launch k aggregates. The two launches use two grouping
  // scratch sets: ws, and ws_alt, allocated the first time a launch finds ws still read by an
  // aggregation. grouped[i] / aggregated[i]: set i's grouping done / aggregation done.
  hipStream_t gstream = nullptr;
  PairScratch ws;
  PairScratch ws_alt;
  int ws_next = 0;  // set of the next launch (0: ws, 1: ws_alt)
  hipEvent_t grouped[2] = {nullptr, nullptr};
  hipEvent_t aggregated[2] = {nullptr, nullptr};
  bool agg_used[2] = {false, false};
  // host staging (pinned) and matching device buffers for vertex-batch submissions
  uint32_t* h_stage = nullptr;
  size_t h_stage_bytes = 0;
  uint32_t* d_stage = nullptr;
  size_t d_stage_bytes = 0;
  hipEvent_t stage_free = nullptr;  // recorded when the kernels that read the staging are done
  bool stage_pending = false;
  // Coalescing: submitted edge pairs are appended (range-checked and filtered) to the fill
  // buffer d_pend[pend_buf], or, for long unfiltered device batches, kept as segments read in
  // place; the grouping and aggregation kernels run once pend_cap edges are pending, or at a
  // flush. Host batches are copied on copy_stream into the fill buffer while the kernels read
  // the other one; pend_free[b] is recorded after the kernel that read buffer b. `mu`
  // serializes submissions and drains of one stream.
  std::mutex mu;
  uint32_t* d_pend[2] = {nullptr, nullptr};
  uint64_t pend_cap = 0;  // edges per launch (and per buffer)
  uint64_t pend_n = 0;    // edges pending: the fill buffer's plus the segments'
  uint64_t fill_n = 0;    // edges in the fill buffer
  std::vector<std::pair<const uint32_t*, uint64_t>> zsegs;  // device batches read in place
  int pend_buf = 0;
  hipEvent_t pend_free[2] = {nullptr, nullptr};
  bool pend_used[2] = {false, false};
  hipEvent_t copied = nullptr;  // recorded on copy_stream after a host batch has landed
  bool copies_pending = false;  // host batches in the fill buffer the compute stream has not
                                // waited for yet
  // tuning (0 = default): depths kept in the aggregation kernel's LDS delta, destinations per
  // work item, bits of the vertex id resolved inside one grouping bucket
  uint32_t tune_dl = 0, tune_chunk = 0, tune_bucket_bits = 0, tune_launch = 0;
  PartParams tuned_part{};
  // optional kernel timing of the aggregation kernel
  bool profile = false;
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pairs;
  // per aggregation launch fed by the edge-pair grouping: the events recorded before and after
  // the grouping kernels (on gstream); nullptrs for vertex-batch launches
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_group;
  double grp_ms = 0;
  double agg_ms = 0;
  uint64_t agg_launches = 0;
  uint64_t agg_updates = 0;
  unsigned long long* d_traffic = nullptr;  // [0] flushed buckets, [1] direct HBM updates
};

struct Store {
  StoreView view{};
  PartParams part{};
  uint32_t num_samples = 0;
  uint32_t cols_per_sample = 1;
  uint32_t flags = 0;
  int device = 0;
  std::mutex mu;
  std::vector<Stream*> streams;
  // Borůvka scratch, allocated on first query
  uint32_t* d_parent = nullptr;      // union-find being merged this round
  uint32_t* d_comp = nullptr;        // supernode root of every vertex at round start
  uint32_t* d_vflags = nullptr;      // per-root query state (ZERO / merged / done)
  Bucket16* d_acc = nullptr;         // per-root XOR of one column + deterministic bucket (the
                                     // query kernels use it as an alpha and a gamma plane)
  Bucket16* d_cpart = nullptr;       // per-wave partial column sums outside the giant supernode
  Bucket16* d_bpart = nullptr;       // per-wave partial column sums of the big supernodes
  // false once the store took updates that need not be symmetric (vertex batches, imports): the
  // XOR of all its vertex sketches is then not known to be zero
  std::atomic<bool> symmetric{true};
  uint32_t* d_flags = nullptr;       // [0]=modified, [1]=out of samples, [2]=forest count,
                                     // [3]=giant supernode of the round
  uint32_t* d_forest = nullptr;      // forest edges (min,max) pairs, capacity k * n
  uint64_t forest_cap = 0;           // pairs
  uint32_t* h_flags = nullptr;       // pinned mirror of d_flags
  uint32_t* h_labels = nullptr;      // pinned staging of the query's labels (n words)
  uint32_t last_rounds = 0;          // rounds the store's last Boruvka pass took
  hipStream_t qstream = nullptr;
  Stream* qs = nullptr;              // internal stream for forest toggles (not in `streams`)
  size_t bytes() const {
    return (size_t)view.n * view.vstride * sizeof(Bucket16) + (size_t)view.n * sizeof(Bucket16);
  }
};

// A group of stores ingested (and queried) together in one virtual vertex space.
struct Group {
  std::vector<Store*> members;
  std::vector<uint32_t> bases;  // host copy; bases[i] = sum of the members' n before i
  uint32_t n_total = 0;
  uint32_t max_cols = 0, min_bpc = 0, max_bpc = 0;
  PartParams part{};
  GraphDesc* d_descs = nullptr;
  uint32_t* d_bases = nullptr;
  Stream* s = nullptr;          // the group's stream (registered as a writer of every member)
  gs_stream* handle = nullptr;   // owns `s`
  uint32_t* d_vgraph = nullptr;  // member of every global vertex
  // batched Boruvka scratch (virtual vertex space), allocated on the first query
  uint32_t* d_parent = nullptr;
  uint32_t* d_comp = nullptr;
  uint32_t* d_vflags = nullptr;
  uint32_t* d_gstate = nullptr;  // [0,G) modified this round, [G,2G) status, [2G,3G) rounds,
                                 // [3G] round counter, [3G+1] any member still active,
                                 // [3G+2] vertices of the queried members, then per queried
                                 // member (base, first position in d_vmap), and the total
  uint32_t* d_vmap = nullptr;    // the queried members' global vertices, member after member
  uint32_t* h_gstate = nullptr;  // pinned mirror
  Bucket16* d_acc = nullptr;
  uint64_t acc_stride = 0;       // buckets per vertex in d_acc (max_bpc + 1)
  uint32_t* d_giant = nullptr;   // per member, each round: the supernode summed from the rest
  uint32_t* d_gsym = nullptr;    // per member, each query: its vertex sketches XOR to zero
                                 // (inside d_gstate)
  hipGraphExec_t round_exec = nullptr;  // one captured Boruvka round over every member
  hipStream_t qstream = nullptr;  // the batched queries' stream
  hipEvent_t ev_fence = nullptr;  // gs_group_fence: the updates the next query waits for
  bool fenced = false;
};

// vertex space and partition of whatever a stream updates
inline const PartParams& stream_part(const Stream* s) {
  if (s->tune_bucket_bits) return s->tuned_part;
  return s->group ? s->group->part : s->store->part;
}
inline uint32_t stream_vertices(const Stream* s) {
  return s->group ? s->group->n_total : s->store->view.n;
}
// a stream of a wide store (more than 64 columns, or the CubeSketch rule): the lane-per-column
// update kernel, which hashes every column of every update and so gains most from long items
inline bool stream_wide(const Stream* s) {
  return !s->group && ((s->store->flags & GS_STORE_CUBESKETCH) || s->store->view.cols > 64);
}

// pinned (+ device, unless `device` is false) staging of a stream (api.hip): ensure_stage waits
// until the previous staged submission has been consumed; release_stage marks the staging busy
// until the stream gets there
int ensure_stage(Stream* s, size_t bytes, bool device = true);
int release_stage(Stream* s);
// HIP objects of a stream (its own HIP stream unless `stream` is already set, the copy stream,
// the staging and fill-buffer events) and their release
int init_stream_objects(Stream* s);
void destroy_stream_resources(Stream& s);
// every pending submission of every writer of a store applied and finished, queries included
int sync_store(Store* st);

// kernels/launchers implemented in ingest.hip
int launch_vertex_batches(Stream* s, const uint32_t* d_class_counts, const WorkItem* d_items,
                          uint32_t nitems, const uint32_t* d_dsts, uint64_t updates,
                          uint32_t min_depth, uint32_t depth_cap);
// edge-pair submissions: append to the stream's fill buffer (processing it whenever it is full);
// drain_stream processes whatever is pending; ingest_pairs_device appends and drains at once
int submit_pairs_device(Stream* s, const uint32_t* d_pairs, uint64_t nedges, uint32_t min_depth,
                        uint32_t depth_cap);
// host pairs of one group member in its own ids (n > 0): rebased by `base` and range-checked on
// the device as they are appended
struct MemberIds {
  uint32_t base = 0, n = 0;
};
// wait_read: return only once the caller's memory has been read; else pinned memory is left to
// the DMA (the caller keeps it unchanged until wait_copies) and other memory is staged before
// the call returns
int submit_pairs_host(Stream* s, const uint32_t* pairs, uint64_t nedges, uint32_t min_depth,
                      uint32_t depth_cap, const MemberIds& mid = MemberIds{},
                      bool wait_read = true);
// submit_pairs_host(s, pairs[i], nedges[i], 0, 0, mids[i], false) for each i in order, with one
// rebase launch per up to kAppendSegs batches
constexpr uint32_t kAppendSegs = 64;
int submit_member_batches_host(Stream* s, uint32_t count, const MemberIds* mids,
                               const uint32_t* const* pairs, const uint64_t* nedges);
// every host batch submitted so far has been read by its DMA
int wait_copies(Stream* s);
int drain_stream(Stream* s);
int ingest_pairs_device(Stream* s, const uint32_t* d_pairs, uint64_t nedges, uint32_t min_depth,
                        uint32_t depth_cap);
uint64_t stream_launch_edges(const Stream* s);
int ensure_pair_scratch(Stream* s, PairScratch& ws, uint64_t nedges);
// the grouping scratch for `nedges` pairs carved from one buffer at `base` (256-byte aligned
// pieces; base = nullptr only sizes it): sets ws's pointers and class layout, returns the bytes
size_t carve_pair_scratch(const Stream* s, uint64_t nedges, char* base, PairScratch& ws);
// append a device segment of m pairs to a launch
void add_pair_segment(PairSegs& sg, uint32_t& nblk, const uint32_t* p, uint64_t m);
// the grouping kernels over `sg` with scratch `ws` on s's gstream, then the aggregation kernel
// on s's stream (set: ws is s->ws (0) or s->ws_alt (1), the events of the pipeline); set -1:
// every kernel on s's stream, no events (a caller's scratch)
int run_pair_pipeline(Stream* s, const PairSegs& sg, uint32_t nblk, uint64_t nedges,
                      PairScratch& ws, hipEvent_t input_read, int set);
void launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t stream);
PartParams make_part_params(uint32_t n, int local_bits = -1);
int autotune_stream(Stream* s, const uint32_t* d_pairs, uint64_t nedges, gs_tuning* best);
void free_pair_scratch(PairScratch& ws);
int launch_zero_store(Store* st, hipStream_t stream);
int device_copy(void* dst, const void* src, size_t bytes);
int hash_known_answer(const uint64_t* keys, const uint64_t* seeds, uint64_t count, uint64_t* out);

// launchers implemented in boruvka.hip
int boruvka_query(Store* st, const uint32_t* start_idx_host, uint32_t* label_out,
                  uint32_t* num_cc, uint32_t* rounds, uint32_t* forest_out, uint64_t forest_cap,
                  uint64_t* forest_len);
void free_query_scratch(Store* st);
// batched Boruvka over the members of a group (status per member: GS_OK or out of samples)
int group_query_cc(Group* g, const uint8_t* member_mask, uint32_t* labels, uint32_t* num_cc,
                   uint32_t* rounds, int32_t* status);
void free_group_query(Group* g);
int prepare_group_query(Group* g);
// the stream's fill buffers and grouping scratch for its launch size, allocated now
int prepare_stream(Stream* s);
int k_forests(Store* st, uint32_t k, uint32_t* edges_out, uint64_t edge_cap, uint64_t* counts);

// geometry helper shared with the host façade (include/sketch.h:63-77,198-216)
void geometry(uint32_t n, double factor, uint32_t cols_per_sample, int l0, uint64_t out[5]);

}  // namespace gs

// opaque handles of the C ABI wrap the internal objects
struct gs_store {
  gs::Store impl;
};
struct gs_stream {
  gs::Stream impl;
};
struct gs_group {
  gs::Group impl;
};
