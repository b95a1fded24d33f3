// Sketch update kernels for gfx950.
//
// Two entry paths share the aggregation kernels:
//   * pre-grouped vertex batches (the contract of the reference's sketchUpdate kernels,
//     src/cuda_kernel.cu:152-257): the host turns each batch into work items;
//   * raw undirected pairs (edge batches): the device groups directed entries by vertex in three
//     kernels (bucket_count, bucket_scatter, bucket_sort) and emits the work items itself.
// Sketches of up to 64 columns under the default rule use aggregate_kernel; wider ones and the
// CubeSketch rule (src/sketch.cpp:66-83) use wide_aggregate_kernel (lane = column).
//
// aggregate_kernel: one wave64 per work item (<= kChunk destinations of one vertex). For each
// destination the wave hashes the edge index once for the checksum (XXH3 with the sketch seed)
// and once per column (seed + 5c), XORs {idx, checksum} into a per-wave LDS delta laid out like
// the device rows (depth-major), then flushes only the non-zero buckets to HBM: a plain
// read-xor-write when the item owns its vertex in this launch, 64-bit atomic XOR otherwise. The
// deterministic bucket is a wave XOR-reduction. Depths >= the item's LDS window go straight to
// HBM with atomics (probability 2^-9 per column and update for items of 32 or more entries).
#include <chrono>
#include <cstring>

#include "gs_internal.h"

namespace gs {

constexpr int kWave = 64;
constexpr int kAggWaves = 4;                  // waves per aggregation workgroup
constexpr uint32_t kChunk = 1024;             // max destinations per work item (the 11 item classes hold up to 2047)
constexpr uint32_t kWideChunk = 1024;         // the same for wide stores
constexpr int kPartThreads = 512;             // K1/K2 workgroup size
constexpr int kPartEdges = 4096;              // undirected edges per K1/K2 workgroup
constexpr int kSortThreads = 1024;            // K4 workgroup size (one workgroup per bucket)
constexpr int kSortBatch = 16;                // entries per thread loaded together in K4
constexpr uint32_t kEmpty = 0xFFFFFFFFu;

__host__ __device__ __forceinline__ uint32_t item_class(uint32_t len) {
  return 31u - (uint32_t)__builtin_clz(len);
}

// Orders this wave's LDS accesses across lanes: a wave's LDS instructions execute in program
// order, so only the compiler has to be kept from moving accesses across this point (a memory
// fence would also wait for the wave's outstanding global loads).
__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_wave_barrier();
  asm volatile("" ::: "memory");
}

// set lanes of `mask` below this lane
__device__ __forceinline__ uint32_t lane_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ void atomic_xor_u64(uint64_t* p, uint64_t v) {
  __hip_atomic_fetch_xor(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// subgraph filter of the min-cut path (src/mc_gpu_sketch_alg.cpp:139-178): an edge belongs to
// subgraphs 0..depth, depth = get_index_depth(idx, seed, depth_cap)
__device__ __forceinline__ bool edge_selected(uint64_t idx, uint64_t seed, uint32_t min_depth,
                                              uint32_t depth_cap) {
  if (min_depth == 0) return true;
  return depth_of(xxh3_len8(idx, seed), depth_cap) >= min_depth;
}

__global__ void fill_u32_kernel(uint32_t* p, uint64_t n, uint32_t v) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

void launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t stream) {
  const unsigned grid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 1024));
  hipLaunchKernelGGL(fill_u32_kernel, dim3(grid), dim3(256), 0, stream, p, n, v);
}

__global__ void fill_u64_kernel(uint64_t* p, uint64_t n, uint64_t v) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// ------------------------------------------------------------------------------------------
// Vertex grouping for raw pairs. Submissions are appended to the stream's fill buffer (a copy,
// plus the subgraph filter when one is given); then, once per fill buffer:
//   K1 bucket_count  : per partition block, the directed entries per bucket of hashed vertex ids
//                      (LDS histogram, written out per block: no global atomics); out-of-range
//                      pairs are counted as dropped (reported at the flush) and become the no-op
//                      pair (0, 0), as self loops are;
//   K2 bucket_offsets: per bucket, the exclusive prefix of the block counts and the total;
//   K3 bucket_scatter: every block writes its entries at its own offsets (LDS ranks);
//   K4 bucket_sort   : one workgroup per bucket sorts its entries by vertex (LDS counting sort
//                      over the bucket's local ids) and emits the work items.
// Vertex ids are spread over buckets by a bijective multiplicative hash on nbits bits, so the
// skewed bit patterns of R-MAT ids do not pile up in one bucket.
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ uint32_t hashed_id(const PartParams& pp, uint32_t v) {
  return (v * pp.mul) & pp.mask;
}
__device__ __forceinline__ uint32_t bucket_of(const PartParams& pp, uint32_t v) {
  return hashed_id(pp, v) >> pp.local_bits;
}
// Partition block of workgroup `wg` of a K1/K3 launch of nblk workgroups. Workgroups are dealt to
// the 8 XCDs round-robin (wg mod 8), and consecutive partition blocks write adjacent runs of
// every bucket (K3) and adjacent counters of the histogram (K1): each XCD therefore takes one
// contiguous range of blocks, so those partial cache lines fill up in one L2 instead of eight.
// A bijection of [0, nblk): XCD x owns nblk / 8 blocks, one more if x < nblk mod 8.
__device__ __forceinline__ uint32_t xcd_block(uint32_t wg, uint32_t nblk) {
  constexpr uint32_t kXcds = 8;
  const uint32_t x = wg % kXcds, i = wg / kXcds, q = nblk / kXcds, r = nblk % kXcds;
  return x * q + min(x, r) + i;
}

constexpr uint64_t kZeroCopyEdges = 1ULL << 18;  // device batches at least this long: in place
// this workgroup's pairs: [e0, e1) of segment `p` (uniform per workgroup)
__device__ __forceinline__ void block_pairs(const PairSegs& sg, uint32_t b, const uint2*& p,
                                            uint64_t& e0, uint64_t& e1) {
  uint32_t i = 0;
  while (i + 1 < sg.nseg && sg.blk0[i + 1] <= b) i++;
  p = sg.ptr[i];
  e0 = (uint64_t)(b - sg.blk0[i]) * kPartEdges;
  e1 = min(sg.n[i], e0 + kPartEdges);
}

// copy of a submission into the fill buffer; with a subgraph filter (min_depth > 0) in-range
// pairs outside it become (0, 0). With member_n > 0 the pairs are one group member's, in its own
// ids: in-range pairs (both ids < member_n) are offset by the member's base into the group's
// vertex space, the others become (kBadVertex, kBadVertex), dropped and counted by K1 like any
// out-of-range pair. src and dst may be the same buffer.
constexpr uint32_t kBadVertex = 0xFFFFFFFFu;
__global__ void append_kernel(const uint2* src, uint2* dst, uint64_t nedges, uint32_t n,
                              uint64_t seed, uint32_t min_depth, uint32_t depth_cap,
                              uint32_t base, uint32_t member_n) {
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < nedges;
       e += (uint64_t)gridDim.x * blockDim.x) {
    uint2 uv = src[e];
    if (member_n)
      uv = uv.x < member_n && uv.y < member_n ? make_uint2(uv.x + base, uv.y + base)
                                              : make_uint2(kBadVertex, kBadVertex);
    if (min_depth && uv.x < n && uv.y < n &&
        !edge_selected(concat_pair(uv.x, uv.y), seed, min_depth, depth_cap))
      uv = make_uint2(0, 0);
    dst[e] = uv;
  }
}

// Member segments copied into a fill buffer by one batched host submission: each segment's
// pairs (in the member's ids) get the member's base added in place, a pair with an id at or past
// the member's vertex count becomes (kBadVertex, kBadVertex). One workgroup row per segment.
struct AppendSegs {
  uint64_t off[kAppendSegs];  // first pair of the segment in the fill buffer
  uint32_t m[kAppendSegs];
  uint32_t base[kAppendSegs];
  uint32_t n[kAppendSegs];
};

__global__ void append_segs_kernel(uint2* __restrict__ buf, AppendSegs sg) {
  const uint32_t k = blockIdx.y;
  uint2* p = buf + sg.off[k];
  const uint32_t m = sg.m[k], base = sg.base[k], n = sg.n[k];
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x) {
    const uint2 uv = p[e];
    p[e] = uv.x < n && uv.y < n ? make_uint2(uv.x + base, uv.y + base)
                                : make_uint2(kBadVertex, kBadVertex);
  }
}

// The K pairs of this thread in a K1/K3 workgroup (edges e0 + t + k * kPartThreads below e1), all
// loads issued before any is used, then (groups) the members of their endpoints, also together.
// ok[k]: the slot holds a pair of the launch; a pair the grouping drops (an endpoint outside the
// vertex space or, in a group of stores, endpoints in two members) comes back as
// (kBadVertex, kBadVertex).
template <int K>
__device__ __forceinline__ void load_block_pairs(const uint2* __restrict__ pairs, uint64_t e0,
                                                 uint64_t e1, uint32_t n,
                                                 const uint32_t* __restrict__ vgraph, uint2 (&uv)[K],
                                                 bool (&ok)[K]) {
  const uint32_t t = threadIdx.x;
#pragma unroll
  for (int k = 0; k < K; k++) {
    const uint64_t e = e0 + t + (uint64_t)k * kPartThreads;
    ok[k] = e < e1;
    uv[k] = pairs[ok[k] ? e : e0];  // e0 < e1: every workgroup holds at least one pair
  }
  if (vgraph) {
    uint32_t ga[K], gb[K];
#pragma unroll
    for (int k = 0; k < K; k++) {  // clamped ids: an out-of-range pair is dropped by its ids
      ga[k] = vgraph[min(uv[k].x, n - 1u)];
      gb[k] = vgraph[min(uv[k].y, n - 1u)];
    }
#pragma unroll
    for (int k = 0; k < K; k++)
      if (ga[k] != gb[k]) uv[k] = make_uint2(kBadVertex, kBadVertex);
  }
#pragma unroll
  for (int k = 0; k < K; k++)
    if (uv[k].x >= n || uv[k].y >= n) uv[k] = make_uint2(kBadVertex, kBadVertex);
}

// dropped pairs are counted (reported at the next flush) and skipped here and in K3, never
// written: the pairs may be the caller's own device memory
__global__ __launch_bounds__(kPartThreads) void bucket_count_kernel(
    PairSegs sg, uint32_t n, const uint32_t* __restrict__ vgraph, PartParams pp, uint32_t nblk,
    uint32_t* __restrict__ hist, uint32_t* __restrict__ counters) {
  extern __shared__ uint32_t h[];
  const uint32_t t = threadIdx.x;
  if (blockIdx.x == 0 && t < kClassBase + kClasses && t != 3)
    counters[t] = 0;  // class counts of this launch; [3] persists until a flush
  for (uint32_t i = t; i < pp.nbuckets; i += kPartThreads) h[i] = 0;
  __syncthreads();
  const uint32_t blk = xcd_block(blockIdx.x, nblk);
  const uint2* pairs;
  uint64_t e0, e1;
  block_pairs(sg, blk, pairs, e0, e1);
  constexpr int K = kPartEdges / kPartThreads;
  uint2 uv[K];
  bool ok[K];
  load_block_pairs<K>(pairs, e0, e1, n, vgraph, uv, ok);
#pragma unroll
  for (int k = 0; k < K; k++) {
    if (!ok[k]) continue;
    if (uv[k].x == kBadVertex) {  // never index past the store: drop and report at flush
      atomicAdd(&counters[3], 1u);
      continue;
    }
    if (uv[k].x == uv[k].y) continue;
    atomicAdd(&h[bucket_of(pp, uv[k].x)], 1u);
    atomicAdd(&h[bucket_of(pp, uv[k].y)], 1u);
  }
  __syncthreads();
  for (uint32_t i = t; i < pp.nbuckets; i += kPartThreads) hist[(size_t)i * nblk + blk] = h[i];
}

// exclusive prefix of counts[0..nb) into s_off (LDS; may be counts itself), returns the total;
// blockDim threads (a multiple of 64). Each thread sums a run of counts, the runs are scanned
// within each wave by shuffles, and the wave totals (s_part, one per wave) once more.
__device__ uint32_t block_exclusive_scan(const uint32_t* counts, uint32_t nb, uint32_t* s_off,
                                         uint32_t* s_part) {
  const uint32_t t = threadIdx.x, nt = blockDim.x, lane = t & 63, wid = t >> 6;
  const uint32_t per = (nb + nt - 1) / nt;
  const uint32_t b0 = min(nb, t * per), b1 = min(nb, b0 + per);
  uint32_t sum = 0;
  for (uint32_t b = b0; b < b1; b++) {
    const uint32_t c = counts[b];
    s_off[b] = sum;
    sum += c;
  }
  uint32_t x = sum;  // inclusive scan over the wave
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = (uint32_t)__shfl_up((int)x, o);
    if (lane >= (uint32_t)o) x += y;
  }
  if (lane == 63) s_part[wid] = x;
  __syncthreads();
  uint32_t base = x - sum, total = 0;
  for (uint32_t w = 0; w < nt / 64; w++) {
    const uint32_t p = s_part[w];
    base += w < wid ? p : 0u;
    total += p;
  }
  for (uint32_t b = b0; b < b1; b++) s_off[b] += base;
  __syncthreads();
  return total;
}

// exclusive prefix of in[0..n) into out (global memory), blockDim threads (a multiple of 64):
// each thread scans a run of consecutive values
__device__ void global_exclusive_scan(const uint32_t* in, uint32_t n, uint32_t* out,
                                      uint32_t* s_part) {
  const uint32_t t = threadIdx.x, nt = blockDim.x, lane = t & 63, wid = t >> 6;
  const uint32_t per = (n + nt - 1) / nt;
  const uint32_t b0 = min(n, t * per), b1 = min(n, b0 + per);
  uint32_t sum = 0;
  for (uint32_t b = b0; b < b1; b++) sum += in[b];
  uint32_t x = sum;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = (uint32_t)__shfl_up((int)x, o);
    if (lane >= (uint32_t)o) x += y;
  }
  if (lane == 63) s_part[wid] = x;
  __syncthreads();
  uint32_t run = x - sum;
  for (uint32_t w = 0; w < wid; w++) run += s_part[w];
  for (uint32_t b = b0; b < b1; b++) {
    const uint32_t c = in[b];
    out[b] = run;
    run += c;
  }
}

// one workgroup per bucket: the block counts of the bucket become exclusive offsets (in place)
// and its total is written. The counts go in rounds of kOffThreads x 16: each thread holds a run
// of 16 consecutive counts in registers (four 16-byte loads when the row is 16-byte aligned),
// scans it, the runs are scanned across the wave by shuffles and across the waves through LDS,
// and the offsets go back the same way; a round's carry feeds the next.
constexpr int kOffThreads = 256;
constexpr int kOffRun = 16;
__global__ __launch_bounds__(kOffThreads) void bucket_offsets_kernel(uint32_t* hist, uint32_t nblk,
                                                                     uint32_t* __restrict__ totals) {
  __shared__ uint32_t s_wave[kOffThreads / 64];
  uint32_t* col = hist + (size_t)blockIdx.x * nblk;
  const uint32_t t = threadIdx.x, lane = t & 63, wid = t >> 6;
  const bool vec = (((uintptr_t)col) & 15u) == 0;
  uint32_t carry = 0;
  for (uint32_t r0 = 0; r0 < nblk; r0 += kOffThreads * kOffRun) {
    const uint32_t b0 = r0 + t * kOffRun;
    uint32_t v[kOffRun];
    if (vec && b0 + kOffRun <= nblk) {
#pragma unroll
      for (int q = 0; q < kOffRun / 4; q++) {
        const uint4 x = reinterpret_cast<const uint4*>(col + b0)[q];
        v[4 * q] = x.x;
        v[4 * q + 1] = x.y;
        v[4 * q + 2] = x.z;
        v[4 * q + 3] = x.w;
      }
    } else {
#pragma unroll
      for (int q = 0; q < kOffRun; q++) v[q] = b0 + q < nblk ? col[b0 + q] : 0u;
    }
    uint32_t sum = 0;
#pragma unroll
    for (int q = 0; q < kOffRun; q++) {
      const uint32_t c = v[q];
      v[q] = sum;
      sum += c;
    }
    uint32_t x = sum;  // inclusive scan of the run totals over the wave
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = (uint32_t)__shfl_up((int)x, o);
      if (lane >= (uint32_t)o) x += y;
    }
    if (lane == 63) s_wave[wid] = x;
    __syncthreads();
    uint32_t base = carry + x - sum, round_total = 0;
#pragma unroll
    for (int w = 0; w < kOffThreads / 64; w++) {
      const uint32_t p = s_wave[w];
      base += (uint32_t)w < wid ? p : 0u;
      round_total += p;
    }
    __syncthreads();  // s_wave is rewritten by the next round
    if (vec && b0 + kOffRun <= nblk) {
#pragma unroll
      for (int q = 0; q < kOffRun / 4; q++)
        reinterpret_cast<uint4*>(col + b0)[q] =
            make_uint4(base + v[4 * q], base + v[4 * q + 1], base + v[4 * q + 2], base + v[4 * q + 3]);
    } else {
#pragma unroll
      for (int q = 0; q < kOffRun; q++)
        if (b0 + q < nblk) col[b0 + q] = base + v[q];
    }
    carry += round_total;
  }
  if (t == 0) totals[blockIdx.x] = carry;
}

// one workgroup: the start of every bucket's region (exclusive prefix of the totals)
__global__ __launch_bounds__(kOffThreads) void bucket_starts_kernel(const uint32_t* __restrict__ totals,
                                                                    uint32_t nbuckets,
                                                                    uint32_t* __restrict__ bstart) {
  __shared__ uint32_t s_part[kOffThreads / 64];
  global_exclusive_scan(totals, nbuckets, bstart, s_part);
}

// A directed entry of a bucket: the vertex (from which the sort takes its local id, the low
// local_bits of the hashed id) and the other endpoint. kPacked (local_bits + nbits <= 32): one
// u32, local id << nbits | other, half the bytes through K3 and K4; else a uint2 (vertex, other).
__host__ __device__ __forceinline__ bool packed_entries_fit(const PartParams& pp) {
  return pp.local_bits + pp.nbits <= 32;
}

// The block's entries are first placed in LDS in bucket order, then written out so that
// consecutive lanes write consecutive positions of one bucket's run (coalesced stores). kPacked:
// the stage holds each entry as its packed word and its bucket (6 bytes: 3 workgroups per CU);
// otherwise as the pair (vertex, other) from which the bucket is recomputed.
template <bool kPacked>
__global__ __launch_bounds__(kPartThreads) void bucket_scatter_kernel(
    PairSegs sg, uint32_t n, const uint32_t* __restrict__ vgraph, PartParams pp, uint32_t nblk,
    const uint32_t* __restrict__ bstart,
    const uint32_t* __restrict__ hist, uint2* __restrict__ entries) {
  extern __shared__ __align__(16) uint32_t lds_u32[];
  constexpr uint32_t kStage = 2 * kPartEdges;  // entries of a block
  uint2* s_stage = reinterpret_cast<uint2*>(lds_u32);        // !kPacked: kStage pairs
  uint32_t* s_word = lds_u32;                                 // kPacked: kStage words
  uint16_t* s_bkt = reinterpret_cast<uint16_t*>(lds_u32 + kStage);  // kPacked: and buckets
  uint32_t* s_base = lds_u32 + (kPacked ? kStage + kStage / 2 : 2 * kStage);  // nbuckets
  uint32_t* s_cnt = s_base + pp.nbuckets;  // nbuckets: block counts, then offsets
  uint32_t* s_loc = s_cnt;
  __shared__ uint32_t s_part[kPartThreads];
  const uint32_t t = threadIdx.x;
  const uint32_t blk = xcd_block(blockIdx.x, nblk);
  // this block's start inside every bucket (loads issued before the pairs are processed)
  for (uint32_t i = t; i < pp.nbuckets; i += kPartThreads) {
    s_cnt[i] = 0;
    s_base[i] = bstart[i] + hist[(size_t)i * nblk + blk];
  }
  __syncthreads();

  const uint2* pairs;
  uint64_t e0, e1;
  block_pairs(sg, blk, pairs, e0, e1);
  constexpr int K = kPartEdges / kPartThreads;
  uint32_t bk[2 * K], rk[2 * K];
  uint2 uvs[K];
  bool okk[K];
  load_block_pairs<K>(pairs, e0, e1, n, vgraph, uvs, okk);
#pragma unroll
  for (int k = 0; k < K; k++) {
    bk[2 * k] = bk[2 * k + 1] = kEmpty;
    const uint2 uv = uvs[k];
    if (!okk[k] || uv.x == kBadVertex || uv.x == uv.y) continue;
    bk[2 * k] = bucket_of(pp, uv.x);
    bk[2 * k + 1] = bucket_of(pp, uv.y);
    rk[2 * k] = atomicAdd(&s_cnt[bk[2 * k]], 1u);
    rk[2 * k + 1] = atomicAdd(&s_cnt[bk[2 * k + 1]], 1u);
  }
  __syncthreads();
  const uint32_t total = block_exclusive_scan(s_cnt, pp.nbuckets, s_loc, s_part);
  const uint32_t lmask = (1u << pp.local_bits) - 1u;
#pragma unroll
  for (int k = 0; k < K; k++) {
    if (bk[2 * k] == kEmpty) continue;
    const uint2 uv = uvs[k];
    const uint32_t i0 = s_loc[bk[2 * k]] + rk[2 * k], i1 = s_loc[bk[2 * k + 1]] + rk[2 * k + 1];
    if (kPacked) {
      s_word[i0] = ((hashed_id(pp, uv.x) & lmask) << pp.nbits) | uv.y;
      s_bkt[i0] = (uint16_t)bk[2 * k];
      s_word[i1] = ((hashed_id(pp, uv.y) & lmask) << pp.nbits) | uv.x;
      s_bkt[i1] = (uint16_t)bk[2 * k + 1];
    } else {
      s_stage[i0] = uv;
      s_stage[i1] = make_uint2(uv.y, uv.x);
    }
  }
  __syncthreads();
  for (uint32_t i = t; i < total; i += kPartThreads) {
    if (kPacked) {
      const uint32_t b = s_bkt[i];
      reinterpret_cast<uint32_t*>(entries)[s_base[b] + (i - s_loc[b])] = s_word[i];
    } else {
      const uint2 en = s_stage[i];
      const uint32_t b = bucket_of(pp, en.x);
      entries[s_base[b] + (i - s_loc[b])] = en;
    }
  }
}

// LDS bytes of bucket_scatter_kernel's dynamic part
static size_t scatter_lds_bytes(uint32_t nbuckets, bool packed) {
  const size_t stage = 2 * (size_t)kPartEdges;
  return (packed ? stage * 6 : stage * sizeof(uint2)) + sizeof(uint32_t) * 2 * nbuckets;
}

// member of a group that owns global vertex v: the last base <= v (bases ascending, bases[0] = 0)
__device__ __forceinline__ uint32_t graph_of(const uint32_t* __restrict__ bases, uint32_t ngraphs,
                                             uint32_t v) {
  uint32_t lo = 0, hi = ngraphs;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (bases[mid] <= v) lo = mid;
    else hi = mid;
  }
  return lo;
}

// LDS atomicAdd of `inc` at cnt[key] for every active lane, returning each lane's old value as
// if the lanes had added one after another; the lanes holding the first active lane's key share
// one atomic (a hub vertex fills whole waves with one key, which would otherwise serialize on
// one LDS address)
__device__ __forceinline__ uint32_t wave_aggregated_add(uint32_t* cnt, uint32_t key, uint32_t inc) {
  const uint32_t k0 = __builtin_amdgcn_readfirstlane(key);
  const uint64_t same = __ballot(key == k0);
  const uint32_t below = lane_below(same);
  if (key != k0) return atomicAdd(&cnt[key], inc);
  uint32_t b0 = 0;
  if (below == 0) b0 = atomicAdd(&cnt[k0], inc * (uint32_t)__popcll(same));
  // the leader is the lowest set lane of `same`
  return (uint32_t)__builtin_amdgcn_readlane((int)b0, __builtin_ctzll(same)) + inc * below;
}

constexpr uint32_t kSortTile = kSortBatch * kSortThreads;  // entries per output tile
constexpr uint32_t kTileBins = kSortThreads;               // local ids the tiled output handles

// dynamic LDS of bucket_sort_kernel for 2^local_bits local ids
static size_t sort_lds_bytes(uint32_t local_bits) {
  const size_t bins = (size_t)1 << local_bits;
  if (bins > kTileBins) return sizeof(uint32_t) * 3 * bins;
  return sizeof(uint32_t) * (5 * bins + 2 * kSortTile);
}

template <bool kPacked>
__global__ __launch_bounds__(kSortThreads) void bucket_sort_kernel(
    const uint2* __restrict__ entries, PartParams pp, const uint32_t* __restrict__ counts,
    const uint32_t* __restrict__ bstart,
    uint32_t* __restrict__ sorted, WorkItem* __restrict__ items, ItemClasses cls, uint32_t chunk,
    uint32_t* __restrict__ class_counts, const uint32_t* __restrict__ bases, uint32_t ngraphs) {
  extern __shared__ __align__(16) uint32_t lds_u32[];
  const uint32_t bins = 1u << pp.local_bits;
  uint32_t* s_cnt = lds_u32;              // per local id: count
  uint32_t* s_off = lds_u32 + bins;       // per local id: offset inside the bucket, then cursor
  __shared__ uint32_t s_part[kSortThreads];
  __shared__ uint32_t s_ccnt[kClasses], s_cbase[kClasses];
  const uint32_t t = threadIdx.x, b = blockIdx.x;
  // this bucket's region of the entries: [bstart[b], bstart[b] + counts[b])
  for (uint32_t i = t; i < bins; i += kSortThreads) s_cnt[i] = 0;
  const uint32_t base = bstart[b];
  const uint32_t cnt = counts[b];
  // an empty bucket has no items; every load below is clamped to [0, cnt - 1] (cnt is uniform
  // over the workgroup, so no barrier is skipped by part of it)
  if (cnt == 0) return;
  __syncthreads();
  const uint32_t lmask = bins - 1;
  const uint32_t omask = kPacked ? (1u << pp.nbits) - 1u : 0u;
  const uint32_t* pe = reinterpret_cast<const uint32_t*>(entries);
  // entry i of the bucket: its local id and the other endpoint
  auto entry = [&](uint32_t i, uint32_t& key, uint32_t& val) {
    if (kPacked) {
      const uint32_t e = pe[base + i];
      key = e >> pp.nbits;
      val = e & omask;
    } else {
      const uint2 e = entries[base + i];
      key = hashed_id(pp, e.x) & lmask;
      val = e.y;
    }
  };
  // Count pass: entries per local id. LDS atomics are aggregated per wave for the key of the
  // wave's first lane (a hub vertex fills whole waves with one key, which would otherwise
  // serialize on one LDS address). The entries are loaded kCountBatch per thread at a time, and
  // the next batch is in flight while one is counted (clamped loads: no branch before the wait).
  constexpr int kCountBatch = 16;
  auto key_of = [&](uint32_t i) -> uint32_t {
    uint32_t k, v;
    entry(min(i, cnt - 1), k, v);
    return k;
  };
  uint32_t kb[kCountBatch];
#pragma unroll
  for (int k = 0; k < kCountBatch; k++) kb[k] = key_of(t + k * kSortThreads);
  for (uint32_t c0 = 0; c0 < cnt; c0 += kCountBatch * kSortThreads) {
    uint32_t nk[kCountBatch];
#pragma unroll
    for (int k = 0; k < kCountBatch; k++)
      nk[k] = key_of(c0 + (kCountBatch + k) * kSortThreads + t);
#pragma unroll
    for (int k = 0; k < kCountBatch; k++) {
      if (c0 + k * kSortThreads + t >= cnt) break;
      const uint32_t key = kb[k];
      const uint32_t k0 = __builtin_amdgcn_readfirstlane(key);
      const uint64_t same = __ballot(key == k0);
      if (key != k0)
        atomicAdd(&s_cnt[key], 1u);
      else if (lane_below(same) == 0)
        atomicAdd(&s_cnt[k0], (uint32_t)__popcll(same));
    }
#pragma unroll
    for (int k = 0; k < kCountBatch; k++) kb[k] = nk[k];
  }
  __syncthreads();
  // entry offsets per local id (exclusive prefix over the bucket)
  const uint32_t per = (bins + kSortThreads - 1) / kSortThreads;
  const uint32_t l0 = min(bins, t * per), l1 = min(bins, l0 + per);
  uint32_t se = 0;
  for (uint32_t l = l0; l < l1; l++) {
    s_off[l] = se;
    se += s_cnt[l];
  }
  if (t < kClasses) s_ccnt[t] = 0;
  __syncthreads();
  s_part[t] = se;
  __syncthreads();
  for (uint32_t o = 1; o < kSortThreads; o <<= 1) {
    uint32_t add = t >= o ? s_part[t - o] : 0;
    __syncthreads();
    s_part[t] += add;
    __syncthreads();
  }
  const uint32_t ebefore = t ? s_part[t - 1] : 0;
  // work items go to per-length-class regions (class = floor(log2(len))) so the aggregation
  // kernel can hand them out largest first; count per class, reserve, then write
  for (uint32_t l = l0; l < l1; l++) {
    const uint32_t d = s_cnt[l];
    for (uint32_t q = 0; q < d; q += chunk) atomicAdd(&s_ccnt[item_class(min(chunk, d - q))], 1u);
  }
  __syncthreads();
  if (t < kClasses) {
    s_cbase[t] = s_ccnt[t] ? atomicAdd(&class_counts[t], s_ccnt[t]) : 0;
    s_ccnt[t] = 0;
  }
  __syncthreads();
  for (uint32_t l = l0; l < l1; l++) {
    const uint32_t d = s_cnt[l];
    const uint32_t eo = base + ebefore + s_off[l];
    s_off[l] = ebefore + s_off[l];
    if (!d) continue;
    const uint32_t v = (((b << pp.local_bits) | l) * pp.inv) & pp.mask;
    const uint32_t nit = (d + chunk - 1) / chunk;
    const uint32_t g = ngraphs > 1 ? graph_of(bases, ngraphs, v) : 0u;
    for (uint32_t q = 0; q < nit; q++) {
      WorkItem w;
      w.vertex = v;
      w.graph = g;
      const uint32_t ln = min(chunk, d - q * chunk);
      w.len = ln | (nit == 1 ? kItemExclusive : 0u);
      w.start = eo + q * chunk;
      const uint32_t k = item_class(ln);
      items[cls.off[k] + s_cbase[k] + atomicAdd(&s_ccnt[k], 1u)] = w;
    }
  }
  __syncthreads();
  if (bins > kTileBins) {
    // wide local ids: every destination goes straight to its place (scattered 4-byte stores)
    for (uint32_t i0 = t; i0 < cnt; i0 += kSortTile) {
      uint32_t ka[kSortBatch], va[kSortBatch];
#pragma unroll
      for (int k = 0; k < kSortBatch; k++) entry(min(i0 + k * kSortThreads, cnt - 1), ka[k], va[k]);
#pragma unroll
      for (int k = 0; k < kSortBatch; k++) {
        if (i0 + k * kSortThreads >= cnt) break;
        sorted[base + wave_aggregated_add(s_off, ka[k], 1u)] = va[k];
      }
    }
    return;
  }
  // Tiles of kSortTile entries: each tile is counting-sorted by local id in LDS, every id
  // reserves its run of the bucket's output (the id's cursor s_off advances by the tile's count),
  // and the tile is written out in LDS order, so consecutive lanes store consecutive positions of
  // one id's run instead of scattering 4-byte stores over the whole bucket. An entry's rank among
  // the tile's entries of its id is the old value its counting atomic returns, so one atomic pass
  // serves both; the next tile's entries load while this one is sorted and written.
  // per id: {start of its run in the tile, start of its run in the bucket}, one 64-bit word so
  // that an entry reads both at once; the tile in id order as {destination, position in the
  // bucket}, one 64-bit store and one load per entry (2 * bins words lie before: 8-byte aligned)
  uint2* t_og = reinterpret_cast<uint2*>(s_off + bins);
  uint2* st = t_og + bins;
  uint32_t* t_cnt = reinterpret_cast<uint32_t*>(st + kSortTile);  // per id: entries in this tile
  __shared__ uint32_t s_wsum[kSortThreads / 64];
  const uint32_t lane = t & 63, wid = t >> 6;
  for (uint32_t i = t; i < bins; i += kSortThreads) t_cnt[i] = 0;
  uint32_t key[kSortBatch], val[kSortBatch];
#pragma unroll
  for (int k = 0; k < kSortBatch; k++) entry(min(t + k * kSortThreads, cnt - 1), key[k], val[k]);
  __syncthreads();
  for (uint32_t tile0 = 0; tile0 < cnt; tile0 += kSortTile) {
    const uint32_t i0 = tile0 + t;
    uint32_t rk[kSortBatch];
#pragma unroll
    for (int k = 0; k < kSortBatch; k++)
      rk[k] = i0 + k * kSortThreads < cnt ? wave_aggregated_add(t_cnt, key[k], 1u) : 0u;
    uint32_t nkey[kSortBatch], nval[kSortBatch];
    const uint32_t n0 = min(tile0 + kSortTile, cnt - 1) + t;  // next tile (clamped loads)
#pragma unroll
    for (int k = 0; k < kSortBatch; k++) entry(min(n0 + k * kSortThreads, cnt - 1), nkey[k], nval[k]);
    __syncthreads();
    // exclusive scan of the tile counts (bins <= kSortThreads: one id per thread)
    const uint32_t c = t < bins ? t_cnt[t] : 0u;
    uint32_t x = c;
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = (uint32_t)__shfl_up((int)x, o);
      if (lane >= (uint32_t)o) x += y;
    }
    if (lane == 63) s_wsum[wid] = x;
    __syncthreads();
    if (t < bins) {
      uint32_t excl = x - c;
      for (uint32_t w = 0; w < wid; w++) excl += s_wsum[w];
      t_og[t] = make_uint2(excl, s_off[t]);
      s_off[t] += c;
      t_cnt[t] = 0;  // for the next tile's atomics (after two more barriers)
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kSortBatch; k++) {
      if (i0 + k * kSortThreads >= cnt) break;
      const uint2 og = t_og[key[k]];
      st[og.x + rk[k]] = make_uint2(val[k], og.y + rk[k]);
    }
    __syncthreads();
    const uint32_t ntile = min(kSortTile, cnt - tile0);
    for (uint32_t j = t; j < ntile; j += kSortThreads) {
      const uint2 e = st[j];
      sorted[base + e.y] = e.x;
    }
#pragma unroll
    for (int k = 0; k < kSortBatch; k++) {
      key[k] = nkey[k];
      val[k] = nval[k];
    }
  }
}

// ------------------------------------------------------------------------------------------
// K5: aggregation, one wave64 per work item (see the file comment).
//
// Lane layout: lane = 8*e + s with e = entry slot of a group of 8 consecutive destinations and
// s = column slot; the lane handles columns s, s+8, ..., s+8(T-1) (those < cols) of its entry.
// Its column keys and LDS addresses therefore stay in registers for the whole item, and an
// entry's staged values are read from LDS once for T columns (8 lanes read one address).
//
// Per-wave LDS delta: rows of `cols` 16-byte buckets {alpha, gamma}; row d holds depth d
// (d < dl). A depth is the ctz of the column's low 16 hash bits with a sentinel bit at the item's
// window depth dli, so it comes out clamped to the window: depth dli stands for every depth >=
// dli and addresses a trash row just past the window (cleared after each item). A lane branches
// only when one of its columns reads dli (probability 2^-dli), to find that pair's exact depth
// (the full hash for a 16-bit tie) and XOR it into HBM with 64-bit atomics.
// LDS banks: gfx950 serves a wave's 64-bit LDS atomics 16 lanes (entry slots 2q and 2q+1) at a
// time over 32 four-byte banks, and a row of 8k buckets spans the banks a whole number of times,
// so the two entry slots of a pass hit the same banks for one column whatever their depths. Odd
// entry slots therefore XOR gamma first and alpha second: in each instruction one slot of the
// pair writes alpha words and the other gamma words, 32 distinct banks, and no two lanes of a
// pass ever share an address (lanes of different passes may, at depth 0 mostly: that costs
// little and needs no private rows). The stage holds each entry's {idx, checksum} in that
// order already, so an LDS address is one 24-bit multiply-add and the second one an XOR.
// After the delta, the chunk's stage: for entry j, first = idx (j even) or checksum (j odd) and
// second the other one, then mix(swap(idx)); 64 of each.
// ------------------------------------------------------------------------------------------
struct AggLayout {
  uint32_t cols, dl;
  __host__ __device__ AggLayout(uint32_t cols_, uint32_t dl_) : cols(cols_), dl(dl_) {}
  // the window's dl depth rows and one more (the trash row of a window of dl depths)
  __host__ __device__ uint32_t delta_bytes() const { return (dl + 1) * cols * 16u; }
  __host__ __device__ uint32_t wave_bytes() const { return delta_bytes() + 3 * 64 * 8; }
};
constexpr int kMaxColSteps = 8;  // T <= 8: up to 64 columns
// 64-slot rounds per flush step, and windows of at most defer_rounds x 64 slots flush deferred
// (that many rounds of stores stay live through the next item, 8 VGPRs each; a deferred window
// is one flush step). A single store runs at 6 waves per SIMD with 2 and 2 (4 flush rounds
// would cost a wave per SIMD); groups of stores see mostly short items and keep 4 and 4.
template <bool kGroup>
constexpr int flush_rounds() { return kGroup ? 4 : 2; }
template <bool kGroup>
constexpr int defer_rounds() { return kGroup ? 4 : 2; }
static_assert(defer_rounds<false>() <= flush_rounds<false>() && defer_rounds<true>() <= flush_rounds<true>(),
              "a deferred window must fit one flush step");
constexpr uint32_t kMaxLdsDepths = 16;

// 64-bit XOR at an LDS byte address (kept as a plain 32-bit address so that no base is re-added
// per access)
typedef __attribute__((address_space(3))) unsigned long long lds_u64_t;
__device__ __forceinline__ void lds_xor_u64(uint32_t addr, uint64_t v) {
  __hip_atomic_fetch_xor(reinterpret_cast<lds_u64_t*>((uintptr_t)addr), (unsigned long long)v,
                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Waves per SIMD the register allocation of aggregate_kernel must allow (the occupancy its
// register use reaches on gfx950 without spills).
template <bool kGroup, int T>
constexpr int agg_min_waves() {
  return kGroup ? (T <= 3 ? 4 : T <= 7 ? 3 : 2) : (T <= 1 ? 7 : T <= 3 ? 6 : T == 4 ? 5 : T <= 7 ? 4 : 3);
}

// kGroup: items belong to the members of a group (descs[item.graph]); the LDS delta is sized for
// the widest member (max_cols) and laid out with the current member's column count, which is
// reloaded (with its geometry and keys) when a wave moves to an item of another member (the
// delta is all zero between items). Otherwise every item belongs to the store `st`.
// T = ceil(columns / 8) of the widest store the launch covers. kFull: a single store whose
// column count is exactly 8 T (every column slot of every step is a column), so the T depth
// hashes of a lane need no guard and run as independent chains before the LDS updates.
template <bool kGroup, int T, bool kFull>
__global__ __launch_bounds__(kAggWaves* kWave, (agg_min_waves<kGroup, T>())) void aggregate_kernel(
    StoreView st, const GraphDesc* __restrict__ descs, uint32_t max_cols,
    const WorkItem* __restrict__ items, ItemClasses cls,
    const uint32_t* __restrict__ class_counts, const uint32_t* __restrict__ dsts, uint32_t dl,
    int exclusive_ok, uint32_t min_depth, uint32_t depth_cap,
    unsigned long long* __restrict__ traffic) {
  extern __shared__ __align__(16) unsigned char lds[];
  const AggLayout L(kGroup ? max_cols : st.cols, dl);
  // measurement (traffic != nullptr): device buckets read-modify-written by flushes (delta slots
  // and deterministic buckets) and (vertex, bucket) updates XORed straight into HBM
  uint32_t n_flush = 0, n_deep = 0;
  uint32_t cols = st.cols;
  uint32_t bpc = st.bpc;
  uint64_t seed = st.seed;
  uint32_t vbase = 0;
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint32_t wid = threadIdx.x / kWave;
  const uint32_t le = lane >> 3, ls = lane & 7;
  const uint32_t wbase = wid * L.wave_bytes();  // byte offset of this wave's delta
  uint64_t* sfirst = reinterpret_cast<uint64_t*>(lds + wbase + L.delta_bytes());
  uint64_t* ssecond = sfirst + kWave;
  uint64_t* smix = ssecond + kWave;
  const uint32_t odd = lane & 1;  // of the staging lane's entry, and of the entry slot (le & 1)
  // where this lane stages its entry's idx and checksum (swapped for odd entries)
  uint64_t* const widx = odd ? ssecond + lane : sfirst + lane;
  uint64_t* const wchk = odd ? sfirst + lane : ssecond + lane;

  uint64_t colk[T];
  uint64_t chk_key = 0;  // the checksum's key (sketch seed) in the split form
  // LDS address of the half of this lane's depth-0 bucket it XORs first (gamma for odd entry
  // slots); depth d adds d rows
  const uint32_t lbase = (uint32_t)(uintptr_t)lds + wbase;
  uint32_t addr1[T];
  auto set_columns = [&]() {
    chk_key = xxh3_seed_key(seed);
#pragma unroll
    for (int t = 0; t < T; t++) {
      const uint32_t c = ls + 8 * t;
      colk[t] = c < cols ? xxh3_seed_key(seed + 5ULL * c) : 0;
      addr1[t] = lbase + 16u * c + 8u * (le & 1);
    }
  };
  if (!kGroup) set_columns();
  {
    uint4* z = reinterpret_cast<uint4*>(lds + wbase);
    for (uint32_t i = lane; i < L.delta_bytes() / 16; i += kWave) z[i] = make_uint4(0, 0, 0, 0);
  }
  wave_sync_lds();
  uint32_t cur_graph = 0xFFFFFFFFu;

  // items are handed out largest class first: virtual index -> (class, index in class)
  uint32_t ccount[kClasses];
  uint32_t nitems = 0;
#pragma unroll
  for (int k = 0; k < kClasses; k++) {
    ccount[k] = __builtin_amdgcn_readfirstlane(class_counts[k]);
    nitems += ccount[k];
  }
  // A descriptor is fetched as one word per lane (lanes 0..3 hold vertex, len, start, graph), a
  // vector load whose wait falls where a field is first read (readlane), one item later. The
  // class holding virtual index vi is found with wave-uniform selects, no branches.
  auto item_word = [&](uint32_t vi) -> uint32_t {
    uint64_t off = 0;
    bool found = false;
#pragma unroll
    for (int k = kClasses - 1; k >= 0; k--) {
      const bool here = !found && (k == 0 || vi < ccount[k]);
      off = here ? cls.off[k] + vi : off;
      found = found || here;
      vi = found ? vi : vi - ccount[k];
    }
    return reinterpret_cast<const uint32_t*>(items + off)[lane & 3];
  };
  auto field = [](uint32_t word, int k) -> uint32_t {
    return (uint32_t)__builtin_amdgcn_readlane((int)word, k);
  };
  const uint32_t gw = __builtin_amdgcn_readfirstlane(blockIdx.x * kAggWaves + wid);
  const uint32_t nw = gridDim.x * kAggWaves;

  // software pipeline: while item i is hashed and flushed, the descriptor of item i+2 and the
  // first chunk of destinations of item i+1 are in flight; inside an item the next chunk of
  // destinations is loaded while the current one is processed. Lanes past an item's length load
  // its first entry (unconditional loads: no branch joins force an early wait; never used).
  auto first_chunk = [&](uint32_t ln, uint32_t sp) -> uint32_t {
    return dsts[(uint64_t)sp + (lane < ln ? lane : 0u)];
  };
  // deferred flush of the previous exclusive item (its window fits kDeferRounds x 64 slots)
  constexpr int kDeferRounds = defer_rounds<kGroup>();
  constexpr int kFlushRounds = flush_rounds<kGroup>();
  bool pending = false;
  Bucket16* prows = nullptr;
  uint64_t pa[kDeferRounds], pg[kDeferRounds];
  Bucket16 pcur[kDeferRounds];
  auto complete_pending = [&]() {
    if (!pending) return;
#pragma unroll
    for (int r = 0; r < kDeferRounds; r++)
      if (pa[r] | pg[r]) prows[r * kWave + lane] = Bucket16{pcur[r].alpha ^ pa[r], pcur[r].gamma ^ pg[r]};
    pending = false;
  };
  const uint32_t last = nitems ? nitems - 1 : 0;
  uint32_t wcur = item_word(min(gw, last));
  uint32_t wnext = item_word(min(gw + nw, last));
  uint32_t dcur = gw < nitems ? first_chunk(field(wcur, 1) & ~kItemExclusive, field(wcur, 2)) : 0u;
  for (uint32_t it = gw; it < nitems; it += nw) {
    const uint32_t w = wcur;
    // fields of this item (loaded two items ago) and of the next one (loaded one item ago), read
    // before any new load is issued
    const uint32_t v_g = field(w, 0);
    const uint32_t wlen = field(w, 1);
    const uint64_t start = field(w, 2);
    const uint32_t next_len = field(wnext, 1) & ~kItemExclusive, next_start = field(wnext, 2);
    const uint32_t wafter = item_word(min(it + 2 * nw, last));
    const uint32_t dfirst_next = first_chunk(next_len, next_start);
    if (kGroup) {
      const uint32_t g = field(w, 3);
      if (g != cur_graph) {  // the wave moves to another member: its geometry and column keys
        cur_graph = g;
        const GraphDesc& gd = descs[g];
        st.rows = gd.rows;
        st.det = gd.det;
        st.n = gd.n;
        st.vstride = gd.vstride;
        cols = gd.cols;
        bpc = gd.bpc;
        seed = gd.seed;
        vbase = gd.base;
        set_columns();
      }
    }
    const uint32_t v = v_g - vbase;
    const uint32_t len = wlen & ~kItemExclusive;
    const bool excl = exclusive_ok && (wlen & kItemExclusive);
    // LDS depth window of this item: floor(log2 len) + 4 depths (at most dl); the deeper pairs,
    // about 1.5 per item, go to HBM atomics, and the flush scans only the window
    const uint32_t dli = min(dl, 35u - (uint32_t)__builtin_clz(len));
    const uint32_t islots = dli * cols;
    const uint32_t row_bytes = 16u * cols;
    const uint32_t sent = 1u << dli;                 // dli <= 16: within the exact low bits
    const uint32_t cap16 = 1u << (bpc < 16u ? bpc : 16u);
    const uint32_t dtail = bpc - kRowDepths, dbase = kRowDepths * cols - kRowDepths;
    uint64_t det_a = 0, det_g = 0;
    Bucket16* vrows = st.rows + (size_t)v * st.vstride;

    for (uint32_t j0 = 0; j0 < len; j0 += kWave) {
      const uint32_t cnt = min((uint32_t)kWave, len - j0);
      // next chunk (clamped to the item's last entry: an unconditional load)
      const uint32_t dn = dsts[start + min(j0 + kWave + lane, len - 1)];
      if (lane < cnt) {
        const uint32_t other = dcur - vbase;
        uint64_t idx = concat_pair(v, other);
        uint64_t mx = xxh3_mix(xxh3_swap(idx));  // shared by the checksum and every column
        uint64_t chk = xxh3_finish(mx ^ chk_key);  // = xxh3_len8(idx, seed)
        // subgraph filter: an update outside the store's subgraph becomes the (0, 0) no-op
        if (min_depth && depth_of(chk, depth_cap) < min_depth) idx = chk = mx = 0;
        *widx = idx;
        *wchk = chk;
        smix[lane] = mx;
        det_a ^= idx;
        det_g ^= chk;
      }
      dcur = dn;
      wave_sync_lds();
      const uint32_t ngroups = (cnt + 7) >> 3;
      for (uint32_t g = 0; g < ngroups; g++) {
        const uint32_t j = 8 * g + le;
        if (j < cnt) {
          const uint64_t v1 = sfirst[j], v2 = ssecond[j], m = smix[j];
          // the low 16 hash bits of each column and its depth clamped to the window: the sentinel
          // bit at dli makes a depth >= dli read as dli, whose LDS row is the item's trash row
          uint32_t h[T], dd[T];
#pragma unroll
          for (int t = 0; t < T; t++) {
            h[t] = hash_low16(m ^ colk[t]);
            dd[t] = (uint32_t)__builtin_ctz(h[t] | sent);
          }
          uint32_t dmax = 0;
#pragma unroll
          for (int t = 0; t < T; t++) {
            if (kFull || ls + 8 * t < cols) {
              const uint32_t off = mad24_lo(dd[t], row_bytes, addr1[t]);
              lds_xor_u64(off, v1);
              lds_xor_u64(off ^ 8u, v2);
              dmax = max(dmax, dd[t]);
            }
          }
          if (dmax == dli) {  // rare: depths >= the window straight to HBM (probability 2^-dli)
            const uint64_t a = (le & 1) ? v2 : v1, gm = (le & 1) ? v1 : v2;
#pragma unroll
            for (int t = 0; t < T; t++) {
              const uint32_t c = ls + 8 * t;
              if ((kFull || c < cols) && dd[t] == dli) {
                uint32_t d = (uint32_t)__builtin_ctz(h[t] | cap16);
                if (d == 16 && bpc > 16) d = depth_of(xxh3_finish(m ^ colk[t]), bpc);
                if (d < bpc) {
                  // bucket_slot with the deep runs' start folded in: depth rows below
                  // kRowDepths, then run c of B - kRowDepths buckets
                  Bucket16* b = &vrows[d < kRowDepths ? d * cols + c : c * dtail + d + dbase];
                  atomic_xor_u64(&b->alpha, a);
                  atomic_xor_u64(&b->gamma, gm);
                  n_deep++;
                }
              }
            }
          }
        }
      }
      wave_sync_lds();
    }
    dcur = dfirst_next;
    wcur = wnext;
    wnext = wafter;
    // the trash row took the LDS updates of the deep pairs
    if (lane < cols) reinterpret_cast<Bucket16*>(lds + wbase + dli * row_bytes)[lane] = Bucket16{0, 0};
    det_a = wave_xor_reduce(det_a);
    det_g = wave_xor_reduce(det_g);
    if (lane == 0 && (det_a | det_g)) {
      Bucket16* b = &st.det[v];
      atomic_xor_u64(&b->alpha, det_a);
      atomic_xor_u64(&b->gamma, det_g);
    }
    if (traffic && (det_a | det_g)) n_flush++;
    wave_sync_lds();
    // flush, kFlushRounds x 64 slots at a time: all LDS reads first, then the HBM reads of the
    // non-zero slots together, one wait, then the stores (exclusive item: plain read-xor-write;
    // otherwise 64-bit atomics, which need no read). Slot s = depth * cols + column is in LDS
    // row `depth` and is device bucket s of the vertex (the window lies in the depth rows of
    // bucket_slot). An exclusive item whose window fits one step leaves its stores pending: they
    // complete after the next item, so the reads' latency is hidden behind that item's work (the
    // vertex is this item's alone; its deep buckets, written by atomics, lie outside the window).
    Bucket16* drow = reinterpret_cast<Bucket16*>(lds + wbase);
    const bool defer = excl && islots <= kDeferRounds * kWave;
    if (!defer) complete_pending();
    for (uint32_t s0 = 0; s0 < islots; s0 += kFlushRounds * kWave) {
      uint64_t xa[kFlushRounds], xg[kFlushRounds];
#pragma unroll
      for (int r = 0; r < kFlushRounds; r++) {
        const uint32_t sl = s0 + r * kWave + lane;
        const Bucket16 b = sl < islots ? drow[sl] : Bucket16{0, 0};
        xa[r] = b.alpha;
        xg[r] = b.gamma;
      }
#pragma unroll
      for (int r = 0; r < kFlushRounds; r++) {
        const uint32_t sl = s0 + r * kWave + lane;
        if (sl < islots) drow[sl] = Bucket16{0, 0};
      }
      if (traffic) {
#pragma unroll
        for (int r = 0; r < kFlushRounds; r++) n_flush += (uint32_t)__popcll(__ballot((xa[r] | xg[r]) != 0));
      }
      if (defer) {
        complete_pending();  // the previous item's stores, its reads long since arrived
#pragma unroll
        for (int r = 0; r < kDeferRounds; r++) {
          pa[r] = xa[r];
          pg[r] = xg[r];
          pcur[r] = vrows[(xa[r] | xg[r]) ? r * kWave + lane : 0u];
        }
        prows = vrows;
        pending = true;
      } else if (excl) {
        // unconditional loads (lanes with nothing to flush read the vertex's first bucket) so the
        // reads issue together and one wait covers them
        Bucket16 cur[kFlushRounds];
#pragma unroll
        for (int r = 0; r < kFlushRounds; r++)
          cur[r] = vrows[(xa[r] | xg[r]) ? s0 + r * kWave + lane : 0u];
#pragma unroll
        for (int r = 0; r < kFlushRounds; r++)
          if (xa[r] | xg[r])
            vrows[s0 + r * kWave + lane] = Bucket16{cur[r].alpha ^ xa[r], cur[r].gamma ^ xg[r]};
      } else {
#pragma unroll
        for (int r = 0; r < kFlushRounds; r++)
          if (xa[r] | xg[r]) {
            Bucket16* b = &vrows[s0 + r * kWave + lane];
            atomic_xor_u64(&b->alpha, xa[r]);
            atomic_xor_u64(&b->gamma, xg[r]);
          }
      }
    }
    wave_sync_lds();
  }
  complete_pending();
  if (traffic) {
    for (int o = 32; o > 0; o >>= 1) n_deep += __shfl_xor(n_deep, o);
    if (lane == 0) {
      atomicAdd(&traffic[0], (unsigned long long)n_flush);
      atomicAdd(&traffic[1], (unsigned long long)n_deep);
    }
  }
}

// ------------------------------------------------------------------------------------------
// K5 for wide sketches: more than 64 columns (several columns per sample) or the CubeSketch
// rule (src/sketch.cpp:66-83: an update of depth d toggles depths 0..d of the column).
//
// One wave per (work item, group of up to 64 columns), lane = column. The item's destinations
// are staged 64 at a time (edge index, checksum, mix(swap(idx)) in LDS, read back as broadcasts)
// and each lane hashes every entry for its own column. The LDS delta is private per lane:
// depth-major planes alpha[d][64] and gamma[d][64] of u64 (8-byte lane stride: no bank
// conflicts, no contention), so a pair is one 24-bit multiply-add and two ds_xor_b64.
// CubeSketch: a depth-d entry XORs only depth d (the window's last depth when d lies beyond it)
// and the flush takes the suffix XOR over depths, which is the rule's column exactly; depths
// past the window (rare) go to HBM with atomics, for the CubeSketch rule every one of them
// up to d. The flush writes a lane's non-zero depths (consecutive columns: one coalesced row of
// atomics per depth). The column group 0 wave also owns the deterministic bucket.
// ------------------------------------------------------------------------------------------
constexpr int kWideWaves = 4;
__host__ __device__ __forceinline__ uint32_t wide_wave_bytes(uint32_t dl) {
  return dl * kWave * 16u + 3u * kWave * 8u;
}

template <bool kL0>
__global__ __launch_bounds__(kWideWaves* kWave) void wide_aggregate_kernel(
    StoreView st, const WorkItem* __restrict__ items, ItemClasses cls,
    const uint32_t* __restrict__ class_counts, const uint32_t* __restrict__ dsts, uint32_t dl,
    uint32_t gcols, uint32_t ngroups, uint32_t min_depth, uint32_t depth_cap,
    unsigned long long* __restrict__ traffic) {
  extern __shared__ __align__(16) unsigned char lds[];
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint32_t wid = threadIdx.x / kWave;
  const uint32_t wbase = wid * wide_wave_bytes(dl);
  uint64_t* pa = reinterpret_cast<uint64_t*>(lds + wbase);  // alpha plane [dl][64]
  uint64_t* pg = pa + dl * kWave;                            // gamma plane [dl][64]
  uint64_t* sidx = pg + dl * kWave;
  uint64_t* schk = sidx + kWave;
  uint64_t* smix = schk + kWave;
  const uint32_t abase = (uint32_t)(uintptr_t)lds + wbase + 8u * lane;
  const uint32_t gplane = dl * kWave * 8u;
  for (uint32_t i = lane; i < dl * kWave; i += kWave) pa[i] = pg[i] = 0;
  wave_sync_lds();
  uint32_t n_flush = 0, n_deep = 0;

  uint32_t ccount[kClasses];
  uint32_t nitems = 0;
#pragma unroll
  for (int k = 0; k < kClasses; k++) {
    ccount[k] = __builtin_amdgcn_readfirstlane(class_counts[k]);
    nitems += ccount[k];
  }
  const uint32_t units = nitems * ngroups;
  const uint32_t gw = __builtin_amdgcn_readfirstlane(blockIdx.x * kWideWaves + wid);
  const uint32_t nw = gridDim.x * kWideWaves;
  for (uint32_t u = gw; u < units; u += nw) {
    // item u / ngroups in largest-class-first order, column group u % ngroups
    uint32_t vi = u / ngroups;
    const uint32_t grp = u - vi * ngroups;
    uint64_t off = 0;
    bool found = false;
#pragma unroll
    for (int k = kClasses - 1; k >= 0; k--) {
      const bool here = !found && (k == 0 || vi < ccount[k]);
      off = here ? cls.off[k] + vi : off;
      found = found || here;
      vi = found ? vi : vi - ccount[k];
    }
    const WorkItem wi = items[off];
    const uint32_t v = __builtin_amdgcn_readfirstlane(wi.vertex);
    const uint32_t len = __builtin_amdgcn_readfirstlane(wi.len) & ~kItemExclusive;
    const uint32_t start = __builtin_amdgcn_readfirstlane(wi.start);
    const uint32_t c = grp * gcols + lane;
    const bool active = lane < gcols && c < st.cols;
    const uint64_t colk = active ? xxh3_seed_key(st.seed + 5ULL * c) : 0;
    const uint32_t dli = min(dl, 35u - (uint32_t)__builtin_clz(len));
    Bucket16* vrows = st.rows + (size_t)v * st.vstride;
    uint64_t det_a = 0, det_g = 0;
    for (uint32_t j0 = 0; j0 < len; j0 += kWave) {
      const uint32_t cnt = min((uint32_t)kWave, len - j0);
      if (lane < cnt) {
        const uint64_t idx0 = concat_pair(v, dsts[(uint64_t)start + j0 + lane]);
        uint64_t idx = idx0, chk = xxh3_len8(idx0, st.seed);
        if (min_depth && depth_of(chk, depth_cap) < min_depth) idx = chk = 0;
        sidx[lane] = idx;
        schk[lane] = chk;
        smix[lane] = xxh3_mix(xxh3_swap(idx));
        det_a ^= idx;
        det_g ^= chk;
      }
      wave_sync_lds();
      if (active) {
        // the next entry's staged values are read while this one is hashed (broadcast reads)
        uint64_t na = sidx[0], ng = schk[0], nm = smix[0];
        for (uint32_t j = 0; j < cnt; j++) {
          const uint64_t a = na, gm = ng, x = nm ^ colk;
          const uint32_t jn = min(j + 1, cnt - 1);
          na = sidx[jn];
          ng = schk[jn];
          nm = smix[jn];
          uint32_t d = depth_low16(x, st.bpc);
          if (kL0) {
            // every entry of depth >= the window's last depth lands there (the suffix XOR of
            // the flush spreads it over depths 0..dli-1); the rare branch below settles the rest
            const uint32_t ad = abase + min(d, dli - 1) * (kWave * 8u);
            lds_xor_u64(ad, a);
            lds_xor_u64(ad + gplane, gm);
          } else if (d < dli) {
            const uint32_t ad = abase + d * (kWave * 8u);
            lds_xor_u64(ad, a);
            lds_xor_u64(ad + gplane, gm);
          }
          if (d >= dli) {  // rare: past the window (and the 16-bit tie)
            if (d == 16 && st.bpc > 16) d = depth_of(xxh3_finish(x), st.bpc);
            if (d < st.bpc) {
              for (uint32_t k = kL0 ? dli : d; k <= d; k++) {
                Bucket16* b = &vrows[bucket_slot(k, c, st.cols, st.bpc)];
                atomic_xor_u64(&b->alpha, a);
                atomic_xor_u64(&b->gamma, gm);
                n_deep++;
              }
            } else if (kL0) {  // depth B updates nothing: take the window entry back out
              const uint32_t ad = abase + (dli - 1) * (kWave * 8u);
              lds_xor_u64(ad, a);
              lds_xor_u64(ad + gplane, gm);
            }
          }
        }
      }
      wave_sync_lds();
    }
    if (grp == 0) {
      det_a = wave_xor_reduce(det_a);
      det_g = wave_xor_reduce(det_g);
      if (lane == 0 && (det_a | det_g)) {
        Bucket16* b = &st.det[v];
        atomic_xor_u64(&b->alpha, det_a);
        atomic_xor_u64(&b->gamma, det_g);
      }
      if (traffic && lane == 0 && (det_a | det_g)) n_flush++;
    }
    // flush, deepest depth first (the CubeSketch suffix runs that way)
    uint64_t ra = 0, rg = 0;
    for (uint32_t k = dli; k-- > 0;) {
      const uint32_t s = k * kWave + lane;
      uint64_t xa = pa[s], xg = pg[s];
      pa[s] = 0;
      pg[s] = 0;
      if (kL0) {
        ra ^= xa;
        rg ^= xg;
        xa = ra;
        xg = rg;
      }
      if (active && (xa | xg)) {
        Bucket16* b = &vrows[bucket_slot(k, c, st.cols, st.bpc)];
        atomic_xor_u64(&b->alpha, xa);
        atomic_xor_u64(&b->gamma, xg);
        n_flush++;
      }
    }
    wave_sync_lds();
  }
  if (traffic) {
    for (int o = 32; o > 0; o >>= 1) {
      n_deep += __shfl_xor(n_deep, o);
      n_flush += __shfl_xor(n_flush, o);
    }
    if (lane == 0) {
      atomicAdd(&traffic[0], (unsigned long long)n_flush);
      atomicAdd(&traffic[1], (unsigned long long)n_deep);
    }
  }
}

// ------------------------------------------------------------------------------------------
// launch helpers
// ------------------------------------------------------------------------------------------
// depths kept in the LDS delta of wide_aggregate_kernel by default; deeper updates
// (probability 2^-9 per column) go to HBM directly
constexpr uint32_t kDefaultLdsDepths = 9;

static size_t agg_lds_bytes(uint32_t cols, uint32_t dl) {
  return (size_t)kAggWaves * AggLayout(cols, dl).wave_bytes();
}

// Default LDS window of aggregate_kernel: the most depths (up to 12) that still leave 6
// workgroups per CU (its register use allows 6 waves per SIMD), and at least 9. Every depth
// more halves the pairs that go to HBM atomics (2^-dl per column) while the flush scans one row
// more per item: 12 depths measured fastest for 24 columns (R-MAT n=2^17 launches of 2^25
// edges; 13 costs a workgroup per CU).
static uint32_t default_agg_depths(uint32_t cols) {
  constexpr size_t kLdsBytes = 160 * 1024, kWorkgroups = 6;
  uint32_t dl = 9;
  while (dl < 12 && agg_lds_bytes(cols, dl + 1) * kWorkgroups <= kLdsBytes) dl++;
  return dl;
}

// persistent grid: as many workgroups as can be resident at once (the occupancy calculator
// accounts for LDS and registers), so the static round-robin over work items never waits on a
// late-starting workgroup
static int agg_grid(int device, const void* kernel, size_t lds_bytes, int waves = kAggWaves) {
  static int cu = 0;
  if (cu == 0) {
    if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cu <= 0)
      cu = 256;
  }
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, waves * kWave, lds_bytes) !=
          hipSuccess ||
      per_cu <= 0) {
    (void)hipGetLastError();
    per_cu = 1;
  }
  return cu * per_cu;
}

// Dynamic LDS above the default 64 KB cap needs an explicit per-kernel opt-in; request exactly
// what a launch needs (static + dynamic must fit the CU's LDS) and report a refusal as an error.
static int ensure_dyn_lds(const void* kernel, size_t bytes) {
  static std::mutex mu;
  static std::vector<std::pair<const void*, size_t>> granted;
  if (bytes <= 64 * 1024) return GS_OK;
  std::lock_guard<std::mutex> lk(mu);
  for (auto& g : granted)
    if (g.first == kernel && g.second >= bytes) return GS_OK;
  hipError_t e = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(GS_ERR_UNSUPPORTED, "kernel needs %zu bytes of LDS: %s", bytes, hipGetErrorString(e));
  }
  granted.emplace_back(kernel, bytes);
  return GS_OK;
}

// one instantiation of the aggregation kernel per (group, T = ceil(columns / 8))
struct AggArgs {
  StoreView v;
  const GraphDesc* descs;
  uint32_t cols;
  const WorkItem* items;
  ItemClasses cls;
  const uint32_t* class_counts;
  const uint32_t* dsts;
  uint32_t dl;
  int exclusive_ok;
  uint32_t min_depth, depth_cap;
  unsigned long long* traffic;
};
struct AggKernel {
  const void* fn;
  void (*launch)(dim3, dim3, size_t, hipStream_t, const AggArgs&);
};
template <bool G, int T, bool F>
static void agg_launch(dim3 grid, dim3 block, size_t lds, hipStream_t s, const AggArgs& a) {
  hipLaunchKernelGGL((aggregate_kernel<G, T, F>), grid, block, lds, s, a.v, a.descs, a.cols,
                     a.items, a.cls, a.class_counts, a.dsts, a.dl, a.exclusive_ok, a.min_depth,
                     a.depth_cap, a.traffic);
}
template <bool G, int T, bool F>
static AggKernel agg_kernel_entry() {
  return AggKernel{reinterpret_cast<const void*>(&aggregate_kernel<G, T, F>), &agg_launch<G, T, F>};
}
// [group / single store with partial column steps / single store with full steps][T - 1]
static const AggKernel& agg_kernel_for(bool group, uint32_t cols) {
  static const AggKernel table[3][kMaxColSteps] = {
      {agg_kernel_entry<false, 1, false>(), agg_kernel_entry<false, 2, false>(),
       agg_kernel_entry<false, 3, false>(), agg_kernel_entry<false, 4, false>(),
       agg_kernel_entry<false, 5, false>(), agg_kernel_entry<false, 6, false>(),
       agg_kernel_entry<false, 7, false>(), agg_kernel_entry<false, 8, false>()},
      {agg_kernel_entry<false, 1, true>(), agg_kernel_entry<false, 2, true>(),
       agg_kernel_entry<false, 3, true>(), agg_kernel_entry<false, 4, true>(),
       agg_kernel_entry<false, 5, true>(), agg_kernel_entry<false, 6, true>(),
       agg_kernel_entry<false, 7, true>(), agg_kernel_entry<false, 8, true>()},
      {agg_kernel_entry<true, 1, false>(), agg_kernel_entry<true, 2, false>(),
       agg_kernel_entry<true, 3, false>(), agg_kernel_entry<true, 4, false>(),
       agg_kernel_entry<true, 5, false>(), agg_kernel_entry<true, 6, false>(),
       agg_kernel_entry<true, 7, false>(), agg_kernel_entry<true, 8, false>()}};
  const uint32_t t = (cols + 7) / 8;
  const int kind = group ? 2 : (cols == 8 * t ? 1 : 0);
  return table[kind][t - 1];
}

// wide sketches (more than 64 columns, or the CubeSketch rule): wide_aggregate_kernel
static int launch_wide_aggregate(Stream* s, const WorkItem* d_items, const ItemClasses& cls,
                                 const uint32_t* class_counts, const uint32_t* d_dsts,
                                 uint64_t updates, uint32_t min_depth, uint32_t depth_cap) {
  Store* st = s->store;
  const StoreView& v = st->view;
  const bool l0 = (st->flags & GS_STORE_CUBESKETCH) != 0;
  uint32_t dl = std::min(v.bpc, s->tune_dl ? s->tune_dl : kDefaultLdsDepths);
  dl = std::min(dl, kMaxLdsDepths);
  const uint32_t ngroups = (v.cols + kWave - 1) / kWave;
  const uint32_t gcols = (v.cols + ngroups - 1) / ngroups;
  const size_t lds = (size_t)kWideWaves * wide_wave_bytes(dl);
  const void* fn = l0 ? reinterpret_cast<const void*>(&wide_aggregate_kernel<true>)
                      : reinterpret_cast<const void*>(&wide_aggregate_kernel<false>);
  int rc_lds = ensure_dyn_lds(fn, lds);
  if (rc_lds) return rc_lds;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (s->profile) {
    GS_HIP(hipEventCreate(&e0));
    GS_HIP(hipEventCreate(&e1));
    GS_HIP(hipEventRecord(e0, s->stream));
    if (!s->d_traffic) {
      GS_HIP(hipMalloc(&s->d_traffic, 2 * sizeof(unsigned long long)));
      GS_HIP(hipMemsetAsync(s->d_traffic, 0, 2 * sizeof(unsigned long long), s->stream));
    }
  }
  const dim3 grid(agg_grid(st->device, fn, lds, kWideWaves)), block(kWideWaves * kWave);
  unsigned long long* tr = s->profile ? s->d_traffic : nullptr;
  if (l0)
    hipLaunchKernelGGL(wide_aggregate_kernel<true>, grid, block, lds, s->stream, v, d_items, cls,
                       class_counts, d_dsts, dl, gcols, ngroups, min_depth, depth_cap, tr);
  else
    hipLaunchKernelGGL(wide_aggregate_kernel<false>, grid, block, lds, s->stream, v, d_items, cls,
                       class_counts, d_dsts, dl, gcols, ngroups, min_depth, depth_cap, tr);
  GS_HIP(hipGetLastError());
  if (s->profile) {
    GS_HIP(hipEventRecord(e1, s->stream));
    s->ev_pairs.emplace_back(e0, e1);
    s->ev_group.emplace_back(nullptr, nullptr);
    s->agg_updates += updates;
  }
  return GS_OK;
}

static int launch_aggregate(Stream* s, const WorkItem* d_items, const ItemClasses& cls,
                            const uint32_t* class_counts, const uint32_t* d_dsts,
                            uint64_t updates, uint32_t min_depth = 0, uint32_t depth_cap = 0) {
  Group* grp = s->group;
  if (stream_wide(s))
    return launch_wide_aggregate(s, d_items, cls, class_counts, d_dsts, updates, min_depth,
                                 depth_cap);
  StoreView v{};
  uint32_t cols, dl;
  int device = 0;
  int exclusive_ok = 1;
  if (grp) {
    cols = grp->max_cols;
    dl = std::min(grp->min_bpc, s->tune_dl ? s->tune_dl : default_agg_depths(cols));
    device = grp->members[0]->device;
    for (Store* m : grp->members) {  // plain stores only if the group is every member's only writer
      std::lock_guard<std::mutex> lk(m->mu);
      if (m->streams.size() != 1) exclusive_ok = 0;
    }
  } else {
    v = s->store->view;
    cols = v.cols;
    dl = std::min(v.bpc, s->tune_dl ? s->tune_dl : default_agg_depths(cols));
    device = s->store->device;
    std::lock_guard<std::mutex> lk(s->store->mu);
    exclusive_ok = s->store->streams.size() == 1 ? 1 : 0;
  }
  // the window stays in the depth rows of bucket_slot, where a flush slot is the device bucket
  dl = std::min(dl, kRowDepths);
  GS_CHECK(cols >= 1 && cols <= 8 * kMaxColSteps, GS_ERR_UNSUPPORTED,
           "sketches with %u columns (more than %d) are not supported by the update kernel", cols,
           8 * kMaxColSteps);
  const size_t lds = agg_lds_bytes(cols, dl);
  GS_CHECK(lds <= 160 * 1024, GS_ERR_UNSUPPORTED, "sketch too wide for the LDS window (%zu B)", lds);
  const AggKernel& kern = agg_kernel_for(grp != nullptr, cols);
  int rc_lds = ensure_dyn_lds(kern.fn, lds);
  if (rc_lds) return rc_lds;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (s->profile) {
    GS_HIP(hipEventCreate(&e0));
    GS_HIP(hipEventCreate(&e1));
    GS_HIP(hipEventRecord(e0, s->stream));
  }
  const dim3 grid(agg_grid(device, kern.fn, lds)), block(kAggWaves * kWave);
  if (s->profile && !s->d_traffic) {
    GS_HIP(hipMalloc(&s->d_traffic, 2 * sizeof(unsigned long long)));
    GS_HIP(hipMemsetAsync(s->d_traffic, 0, 2 * sizeof(unsigned long long), s->stream));
  }
  AggArgs a{v, grp ? grp->d_descs : nullptr, cols, d_items, cls, class_counts, d_dsts, dl,
            exclusive_ok, grp ? 0u : min_depth, grp ? 0u : depth_cap,
            s->profile ? s->d_traffic : nullptr};
  kern.launch(grid, block, lds, s->stream, a);
  GS_HIP(hipGetLastError());
  if (s->profile) {
    GS_HIP(hipEventRecord(e1, s->stream));
    s->ev_pairs.emplace_back(e0, e1);
    s->ev_group.emplace_back(nullptr, nullptr);
    s->agg_updates += updates;
  }
  return GS_OK;
}

int launch_vertex_batches(Stream* s, const uint32_t* d_class_counts, const WorkItem* d_items,
                          uint32_t nitems, const uint32_t* d_dsts, uint64_t updates,
                          uint32_t min_depth, uint32_t depth_cap) {
  if (nitems == 0) return GS_OK;
  ItemClasses cls{};  // every item in class 0
  return launch_aggregate(s, d_items, cls, d_class_counts, d_dsts, updates, min_depth, depth_cap);
}

void free_pair_scratch(PairScratch& ws) {
  (void)hipFree(ws.d_entries);
  (void)hipFree(ws.d_sorted);
  (void)hipFree(ws.d_items);
  (void)hipFree(ws.d_hist);
  (void)hipFree(ws.d_totals);
  (void)hipFree(ws.d_counters);
  ws = PairScratch{};
}

PartParams make_part_params(uint32_t n, int local_bits) {
  PartParams p{};
  uint32_t nbits = 1;
  while (nbits < 32 && (1ULL << nbits) < n) nbits++;
  int lb = local_bits >= 0 ? local_bits : (int)nbits - 8;
  p.nbits = nbits;
  p.local_bits = (uint32_t)std::min(13, std::min((int)nbits, std::max(0, lb)));
  p.nbuckets = 1u << (nbits - p.local_bits);
  p.mask = nbits >= 32 ? 0xFFFFFFFFu : (uint32_t)((1ULL << nbits) - 1);
  p.mul = 0x9E3779B1u;  // odd: a bijection modulo 2^nbits
  uint32_t inv = p.mul;  // Newton iteration for the inverse modulo 2^32
  for (int i = 0; i < 5; i++) inv *= 2u - p.mul * inv;
  p.inv = inv;
  return p;
}


static uint32_t part_blocks(uint64_t nedges) {
  return (uint32_t)((nedges + kPartEdges - 1) / kPartEdges);
}

// Item capacity of every length class for a launch of `cap` edges over `verts` vertices: classes
// that may hold full items (len >= 64, the smallest item size a stream can be tuned to; items of
// one vertex may repeat) are bounded by the entries alone; shorter items are a vertex's last, at
// most one per vertex. Returns the total.
static uint64_t item_class_layout(uint64_t verts, uint64_t cap, ItemClasses& cls) {
  uint64_t icap = 0;
  for (int k = 0; k < kClasses; k++) {
    cls.off[k] = icap;
    uint64_t ck = (2 * cap) >> k;
    if (k < 6) ck = std::min<uint64_t>(verts, ck);
    icap += ck + 16;
  }
  return icap;
}

size_t carve_pair_scratch(const Stream* s, uint64_t nedges, char* base, PairScratch& ws) {
  const PartParams& pp = stream_part(s);
  const uint64_t verts = std::min<uint64_t>((uint64_t)stream_vertices(s), 2 * nedges);
  ws = PairScratch{};
  const uint64_t icap = item_class_layout(verts, nedges, ws.cls);
  const uint64_t hist_words = (uint64_t)pp.nbuckets * (part_blocks(nedges) + kMaxSegs);
  size_t off = 0;
  auto take = [&](size_t bytes) -> char* {
    char* p = base ? base + off : nullptr;
    off += (bytes + 255) & ~(size_t)255;
    return p;
  };
  ws.d_counters = reinterpret_cast<uint32_t*>(take(sizeof(uint32_t) * kCounterWords));
  ws.d_entries = reinterpret_cast<uint2*>(take(sizeof(uint2) * 2 * nedges));
  ws.d_sorted = reinterpret_cast<uint32_t*>(take(sizeof(uint32_t) * 2 * nedges));
  ws.d_items = reinterpret_cast<WorkItem*>(take(sizeof(WorkItem) * icap));
  ws.d_hist = reinterpret_cast<uint32_t*>(take(sizeof(uint32_t) * hist_words));
  ws.d_totals = reinterpret_cast<uint32_t*>(take(sizeof(uint32_t) * 2 * pp.nbuckets));
  ws.cap_edges = nedges;
  ws.cap_items = icap;
  ws.cap_hist = hist_words;
  ws.cap_buckets = pp.nbuckets;
  return off;
}

// (scratch being replaced may still be read by a grouping on gstream or an aggregation on stream)
static int sync_pipeline(Stream* s) {
  GS_HIP(hipStreamSynchronize(s->stream));
  if (s->gstream) GS_HIP(hipStreamSynchronize(s->gstream));
  return GS_OK;
}

int ensure_pair_scratch(Stream* s, PairScratch& ws, uint64_t nedges) {
  const PartParams& pp = stream_part(s);
  if (!ws.d_counters) {
    GS_HIP(hipMalloc(&ws.d_counters, sizeof(uint32_t) * kCounterWords));
    GS_HIP(hipMemsetAsync(ws.d_counters, 0, sizeof(uint32_t) * kCounterWords,
                          s->gstream ? s->gstream : s->stream));
  }
  const uint64_t hist_words = (uint64_t)pp.nbuckets * (part_blocks(nedges) + kMaxSegs);
  if (ws.cap_buckets < pp.nbuckets || ws.cap_hist < hist_words) {
    if (int rc = sync_pipeline(s)) return rc;
    (void)hipFree(ws.d_hist);
    (void)hipFree(ws.d_totals);
    ws.d_hist = ws.d_totals = nullptr;
    const uint32_t nb = std::max(ws.cap_buckets, pp.nbuckets);
    const uint64_t hw = std::max(ws.cap_hist, hist_words);
    GS_HIP(hipMalloc(&ws.d_hist, sizeof(uint32_t) * hw));
    GS_HIP(hipMalloc(&ws.d_totals, sizeof(uint32_t) * 2 * nb));
    ws.cap_buckets = nb;
    ws.cap_hist = hw;
  }
  if (ws.cap_edges < nedges) {
    if (int rc = sync_pipeline(s)) return rc;
    (void)hipFree(ws.d_entries);
    (void)hipFree(ws.d_sorted);
    (void)hipFree(ws.d_items);
    ws.d_entries = nullptr;
    ws.d_sorted = nullptr;
    ws.d_items = nullptr;
    uint64_t cap = nedges;
    const uint64_t verts = std::min<uint64_t>((uint64_t)stream_vertices(s), 2 * cap);
    const uint64_t icap = item_class_layout(verts, cap, ws.cls);
    GS_HIP(hipMalloc(&ws.d_entries, sizeof(uint2) * 2 * cap));
    GS_HIP(hipMalloc(&ws.d_sorted, sizeof(uint32_t) * 2 * cap));
    GS_HIP(hipMalloc(&ws.d_items, sizeof(WorkItem) * icap));
    ws.cap_edges = cap;
    ws.cap_items = icap;
  }
  return GS_OK;
}

// Edges per launch: the tuned value, else 256 edges per vertex of the stream's vertex space, at
// least 2^16 and at most kLaunchEdges (2^27: 2^19 vertices). Longer items change fewer buckets
// per update and pay their per-item costs less often: at n = 2^17, 2^25-edge launches ingest
// 5.6 % faster than 2^23 (halving the CubeSketch rule's buckets per update), and a group of 256
// graphs with 4.9 M vertices in all ingests 58 % faster at 2^27 than at 2^25. The fill buffers
// and scratch are sized for it (about 40 B per edge of launch: 5.4 GB at 2^27).
uint64_t stream_launch_edges(const Stream* s) {
  if (s->tune_launch) return s->tune_launch;
  const uint64_t per_vertex = 256ULL * stream_vertices(s);
  return std::min(kLaunchEdges, std::max<uint64_t>(1ULL << 16, per_vertex));
}

// the fill buffers, sized for the stream's launch size (grown on demand, never while pending)
static int ensure_pending(Stream* s) {
  const uint64_t want = stream_launch_edges(s);
  if (s->pend_cap >= want && s->d_pend[0]) {
    // both sets follow the stream's current grouping parameters (a retune may need more
    // bucket histograms than a set was made with)
    if (s->ws_alt.d_counters) {
      int rc = ensure_pair_scratch(s, s->ws_alt, s->pend_cap);
      if (rc) return rc;
    }
    return ensure_pair_scratch(s, s->ws, s->pend_cap);
  }
  GS_CHECK(s->pend_n == 0, GS_ERR_INVALID, "launch size changed with updates pending");
  if (int rc = sync_pipeline(s)) return rc;
  if (s->copy_stream) GS_HIP(hipStreamSynchronize(s->copy_stream));
  for (int b = 0; b < 2; b++) {
    (void)hipFree(s->d_pend[b]);
    s->d_pend[b] = nullptr;
    s->pend_used[b] = false;
  }
  for (int b = 0; b < 2; b++)
    if (hipMalloc(&s->d_pend[b], want * 2 * sizeof(uint32_t)) != hipSuccess)
      return fail(GS_ERR_OOM, "edge buffers of %llu edges unavailable", (unsigned long long)want);
  s->pend_cap = want;
  s->pend_buf = 0;
  if (s->ws_alt.d_counters) {  // the second set grows on its next use
    int rc = ensure_pair_scratch(s, s->ws_alt, want);
    if (rc) return rc;
  }
  return ensure_pair_scratch(s, s->ws, want);
}

int prepare_stream(Stream* s) {
  std::lock_guard<std::mutex> lk(s->mu);
  return ensure_pending(s);
}

// `m` pairs at src into the fill buffer behind what is pending (src may be that very place: a
// host batch copied there; then only a subgraph filter needs a pass)
static int append_pairs(Stream* s, const uint32_t* src, uint64_t m, uint32_t min_depth,
                        uint32_t depth_cap, bool host_copy, const MemberIds& mid = MemberIds{}) {
  GS_CHECK(!s->group || min_depth == 0, GS_ERR_INVALID, "no subgraph filter on group streams");
  uint2* dst = reinterpret_cast<uint2*>(s->d_pend[s->pend_buf]) + s->fill_n;
  if (reinterpret_cast<const uint2*>(src) != dst || min_depth || mid.n) {
    const uint64_t seed = s->group ? 0 : s->store->view.seed;
    const unsigned grid = (unsigned)std::min<uint64_t>((m + 255) / 256, 4096);
    hipLaunchKernelGGL(append_kernel, dim3(grid), dim3(256), 0, host_copy ? s->copy_stream : s->gstream,
                       reinterpret_cast<const uint2*>(src), dst, m, stream_vertices(s), seed,
                       min_depth, depth_cap, mid.base, mid.n);
    GS_HIP(hipGetLastError());
  }
  if (host_copy) GS_HIP(hipEventRecord(s->copied, s->copy_stream));
  s->fill_n += m;
  s->pend_n += m;
  return GS_OK;
}

// host copies (copy stream) may write the fill buffer once the launch that last read it is done
static int wait_fill_buffer(Stream* s) {
  if (s->pend_used[s->pend_buf])
    GS_HIP(hipStreamWaitEvent(s->copy_stream, s->pend_free[s->pend_buf], 0));
  return GS_OK;
}

// K1..K5 over the segments of `sg` (nblk K1/K3 workgroups, nedges pairs) on the stream's HIP
// stream, with the grouping scratch `ws` (sized for nedges by ensure_pair_scratch, or carved
// from a caller's fixed buffer). `input_read` (may be null) is recorded once the pairs have been
// read (after K3).
int run_pair_pipeline(Stream* s, const PairSegs& sg, uint32_t nblk, uint64_t nedges,
                      PairScratch& ws, hipEvent_t input_read, int set) {
  const bool piped = set >= 0 && s->gstream;
  hipStream_t gst = piped ? s->gstream : s->stream;  // the grouping kernels' stream
  const PartParams& pp = stream_part(s);
  const uint32_t nverts = stream_vertices(s);
  const size_t sort_lds = sort_lds_bytes(pp.local_bits);
  int rc;
  const bool packed = packed_entries_fit(pp);
  const void* scatter_fn = packed ? (const void*)bucket_scatter_kernel<true>
                                  : (const void*)bucket_scatter_kernel<false>;
  const void* sort_fn = packed ? (const void*)bucket_sort_kernel<true>
                               : (const void*)bucket_sort_kernel<false>;
  if ((rc = ensure_dyn_lds((const void*)bucket_count_kernel, sizeof(uint32_t) * pp.nbuckets)) ||
      (rc = ensure_dyn_lds(scatter_fn, scatter_lds_bytes(pp.nbuckets, packed))) ||
      (rc = ensure_dyn_lds(sort_fn, sort_lds)))
    return rc;
  // the scratch set's previous aggregation must be done before its grouping overwrites it
  if (piped && s->agg_used[set]) GS_HIP(hipStreamWaitEvent(gst, s->aggregated[set], 0));
  hipEvent_t g0 = nullptr, g1 = nullptr;  // around the grouping kernels (measurement only)
  if (s->profile) {
    GS_HIP(hipEventCreate(&g0));
    GS_HIP(hipEventCreate(&g1));
    GS_HIP(hipEventRecord(g0, gst));
  }
  const uint32_t* vgraph = s->group ? s->group->d_vgraph : nullptr;
  hipLaunchKernelGGL(bucket_count_kernel, dim3(nblk), dim3(kPartThreads),
                     sizeof(uint32_t) * pp.nbuckets, gst, sg, nverts, vgraph, pp, nblk,
                     ws.d_hist, ws.d_counters);
  GS_HIP(hipGetLastError());
  uint32_t* const bstart = ws.d_totals + pp.nbuckets;  // bucket region starts
  hipLaunchKernelGGL(bucket_offsets_kernel, dim3(pp.nbuckets), dim3(kOffThreads), 0, gst,
                     ws.d_hist, nblk, ws.d_totals);
  GS_HIP(hipGetLastError());
  hipLaunchKernelGGL(bucket_starts_kernel, dim3(1), dim3(kOffThreads), 0, gst, ws.d_totals,
                     pp.nbuckets, bstart);
  GS_HIP(hipGetLastError());
  if (packed)
    hipLaunchKernelGGL(bucket_scatter_kernel<true>, dim3(nblk), dim3(kPartThreads),
                       scatter_lds_bytes(pp.nbuckets, true), gst, sg, nverts, vgraph, pp,
                       nblk, bstart, ws.d_hist, ws.d_entries);
  else
    hipLaunchKernelGGL(bucket_scatter_kernel<false>, dim3(nblk), dim3(kPartThreads),
                       scatter_lds_bytes(pp.nbuckets, false), gst, sg, nverts, vgraph, pp,
                       nblk, bstart, ws.d_hist, ws.d_entries);
  GS_HIP(hipGetLastError());
  if (input_read) GS_HIP(hipEventRecord(input_read, gst));
  const uint32_t chunk = s->tune_chunk ? s->tune_chunk : (stream_wide(s) ? kWideChunk : kChunk);
  const uint32_t* bases = s->group ? (const uint32_t*)s->group->d_bases : nullptr;
  const uint32_t ngraphs = s->group ? (uint32_t)s->group->members.size() : 1u;
  if (packed)
    hipLaunchKernelGGL(bucket_sort_kernel<true>, dim3(pp.nbuckets), dim3(kSortThreads), sort_lds,
                       gst, ws.d_entries, pp, ws.d_totals, bstart, ws.d_sorted, ws.d_items,
                       ws.cls, chunk, ws.d_counters + kClassBase, bases, ngraphs);
  else
    hipLaunchKernelGGL(bucket_sort_kernel<false>, dim3(pp.nbuckets), dim3(kSortThreads), sort_lds,
                       gst, ws.d_entries, pp, ws.d_totals, bstart, ws.d_sorted, ws.d_items,
                       ws.cls, chunk, ws.d_counters + kClassBase, bases, ngraphs);
  GS_HIP(hipGetLastError());
  if (g1) GS_HIP(hipEventRecord(g1, gst));
  if (piped) {  // the aggregation waits for this grouping only
    GS_HIP(hipEventRecord(s->grouped[set], gst));
    GS_HIP(hipStreamWaitEvent(s->stream, s->grouped[set], 0));
  }
  const size_t nev = s->ev_pairs.size();
  rc = launch_aggregate(s, ws.d_items, ws.cls, ws.d_counters + kClassBase, ws.d_sorted,
                        2 * nedges);
  if (!rc && piped) {
    GS_HIP(hipEventRecord(s->aggregated[set], s->stream));
    s->agg_used[set] = true;
  }
  if (g0) {
    if (s->ev_pairs.size() == nev + 1) {
      s->ev_group.back() = {g0, g1};
    } else {
      (void)hipEventDestroy(g0);
      (void)hipEventDestroy(g1);
    }
  }
  return rc;
}

void add_pair_segment(PairSegs& sg, uint32_t& nblk, const uint32_t* p, uint64_t m) {
  sg.ptr[sg.nseg] = reinterpret_cast<const uint2*>(p);
  sg.n[sg.nseg] = m;
  sg.blk0[sg.nseg] = nblk;
  nblk += part_blocks(m);
  sg.nseg++;
  sg.blk0[sg.nseg] = nblk;
}

// K1..K5 over the fill buffer and the device segments; the other buffer becomes the fill buffer
static int process_pending(Stream* s) {
  if (s->pend_n == 0) return GS_OK;
  const uint64_t nedges = s->pend_n;
  const int b = s->pend_buf;
  PairSegs sg{};
  uint32_t nblk = 0;
  if (s->fill_n) add_pair_segment(sg, nblk, s->d_pend[b], s->fill_n);
  for (const auto& z : s->zsegs) add_pair_segment(sg, nblk, z.first, z.second);
  if (s->copies_pending) {  // every host copy into the buffer has landed
    GS_HIP(hipStreamWaitEvent(s->gstream, s->copied, 0));
    s->copies_pending = false;
  }
  // the scratch set: the one the previous launch did not use, unless that is the second set,
  // not yet allocated, and the first is already free (a stream that launches now and then
  // never allocates the second)
  int set = s->ws_next;
  if (set == 1 && !s->ws_alt.d_counters) {
    const hipError_t q = s->agg_used[0] ? hipEventQuery(s->aggregated[0]) : hipSuccess;
    if (q == hipSuccess) {
      set = 0;
    } else {
      if (q != hipErrorNotReady) GS_HIP(q);
      int rc = ensure_pair_scratch(s, s->ws_alt, s->pend_cap);
      if (rc) return rc;
    }
  }
  s->ws_next = set ^ 1;
  // the fill buffer is free again once K3 has read it
  int rc = run_pair_pipeline(s, sg, nblk, nedges, set ? s->ws_alt : s->ws, s->pend_free[b], set);
  s->pend_used[b] = true;
  s->pend_n = 0;
  s->fill_n = 0;
  s->zsegs.clear();
  s->pend_buf ^= 1;
  return rc;
}

int submit_pairs_device(Stream* s, const uint32_t* d_pairs, uint64_t nedges, uint32_t min_depth,
                        uint32_t depth_cap) {
  if (nedges == 0) return GS_OK;
  std::lock_guard<std::mutex> lk(s->mu);
  int rc = ensure_pending(s);
  if (rc) return rc;
  // (the launch that last read the fill buffer ran on this same stream). Long unfiltered
  // batches are read in place by the grouping kernels (the caller keeps them valid until a
  // flush or sync); the rest is copied into the fill buffer.
  for (uint64_t e0 = 0; e0 < nedges;) {
    const uint64_t m = std::min(nedges - e0, s->pend_cap - s->pend_n);
    if (min_depth == 0 && m >= kZeroCopyEdges && s->zsegs.size() + 1 < (size_t)kMaxSegs) {
      s->zsegs.emplace_back(d_pairs + 2 * e0, m);
      s->pend_n += m;
    } else if ((rc = append_pairs(s, d_pairs + 2 * e0, m, min_depth, depth_cap, false))) {
      return rc;
    }
    e0 += m;
    if (s->pend_n == s->pend_cap && (rc = process_pending(s))) return rc;
  }
  return GS_OK;
}

// Host batches: copied on the copy stream straight into the fill buffer (from the caller's
// memory when it is pinned, else through the stream's pinned staging), then appended in place on
// the compute stream. The call returns once the caller's memory has been read. Pairs naming a
// vertex outside the store are dropped by the grouping and reported by the next flush, as for
// device batches, whatever memory they came from.
int submit_pairs_host(Stream* s, const uint32_t* pairs, uint64_t nedges, uint32_t min_depth,
                      uint32_t depth_cap, const MemberIds& mid, bool wait_read) {
  if (nedges == 0) return GS_OK;
  std::lock_guard<std::mutex> lk(s->mu);
  int rc = ensure_pending(s);
  if (rc) return rc;
  hipPointerAttribute_t attr{};
  bool pinned = false;
  if (hipPointerGetAttributes(&attr, pairs) == hipSuccess)
    pinned = attr.type == hipMemoryTypeHost;
  else
    (void)hipGetLastError();
  for (uint64_t e0 = 0; e0 < nedges;) {
    const uint64_t m = std::min(nedges - e0, s->pend_cap - s->pend_n);
    const size_t bytes = m * 2 * sizeof(uint32_t);
    const uint32_t* src = pairs + 2 * e0;
    if (!pinned) {
      if ((rc = ensure_stage(s, bytes, false))) return rc;
      std::memcpy(s->h_stage, src, bytes);
      src = s->h_stage;
    }
    // the fill buffer may still be read by the launch that processed it two buffers back
    if ((rc = wait_fill_buffer(s))) return rc;
    uint32_t* dst = s->d_pend[s->pend_buf] + 2 * s->fill_n;
    GS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s->copy_stream));
    if (!pinned) {
      GS_HIP(hipEventRecord(s->stage_free, s->copy_stream));
      s->stage_pending = true;
    }
    if ((rc = append_pairs(s, dst, m, min_depth, depth_cap, true, mid))) return rc;  // records `copied`
    s->copies_pending = true;
    e0 += m;
    if (s->pend_n == s->pend_cap && (rc = process_pending(s))) return rc;
  }
  if (pinned && wait_read) {  // the caller may reuse its buffer once the DMA has read it
    // poll first: the copy of a batch takes ~0.15 ms, and a blocking wait's wake-up delay would
    // otherwise idle the link between one submission and the next; block if it runs long
    const auto t0 = std::chrono::steady_clock::now();
    hipError_t q;
    while ((q = hipEventQuery(s->copied)) == hipErrorNotReady &&
           std::chrono::steady_clock::now() - t0 < std::chrono::milliseconds(2)) {
    }
    if (q == hipErrorNotReady) q = hipEventSynchronize(s->copied);
    GS_HIP(q);
  }
  return GS_OK;
}

// Several members' host batches in one call, without waiting for their DMAs: each chunk is
// copied into the fill buffer as in submit_pairs_host, but the rebase runs as one append launch
// per up to kAppendSegs chunks (and before each launch of the pipeline) instead of one per batch.
int submit_member_batches_host(Stream* s, uint32_t count, const MemberIds* mids,
                               const uint32_t* const* pairs, const uint64_t* nedges) {
  std::lock_guard<std::mutex> lk(s->mu);
  int rc = ensure_pending(s);
  if (rc) return rc;
  AppendSegs sg;
  uint32_t nseg = 0;
  uint64_t seg_max = 0;
  auto append = [&]() -> int {
    if (nseg == 0) return GS_OK;
    const unsigned gx = (unsigned)std::min<uint64_t>((seg_max + 255) / 256, 1024);
    hipLaunchKernelGGL(append_segs_kernel, dim3(gx, nseg), dim3(256), 0, s->copy_stream,
                       reinterpret_cast<uint2*>(s->d_pend[s->pend_buf]), sg);
    GS_HIP(hipGetLastError());
    GS_HIP(hipEventRecord(s->copied, s->copy_stream));
    s->copies_pending = true;
    nseg = 0;
    seg_max = 0;
    return GS_OK;
  };
  if ((rc = wait_fill_buffer(s))) return rc;
  for (uint32_t i = 0; i < count; i++) {
    if (nedges[i] == 0) continue;
    hipPointerAttribute_t attr{};
    bool pinned = false;
    if (hipPointerGetAttributes(&attr, pairs[i]) == hipSuccess)
      pinned = attr.type == hipMemoryTypeHost;
    else
      (void)hipGetLastError();
    for (uint64_t e0 = 0; e0 < nedges[i];) {
      const uint64_t m = std::min(nedges[i] - e0, s->pend_cap - s->pend_n);
      const size_t bytes = m * 2 * sizeof(uint32_t);
      const uint32_t* src = pairs[i] + 2 * e0;
      if (!pinned) {
        if ((rc = ensure_stage(s, bytes, false))) return rc;
        std::memcpy(s->h_stage, src, bytes);
        src = s->h_stage;
      }
      GS_HIP(hipMemcpyAsync(s->d_pend[s->pend_buf] + 2 * s->fill_n, src, bytes,
                            hipMemcpyHostToDevice, s->copy_stream));
      if (!pinned) {
        GS_HIP(hipEventRecord(s->stage_free, s->copy_stream));
        s->stage_pending = true;
      }
      sg.off[nseg] = s->fill_n;
      sg.m[nseg] = (uint32_t)m;
      sg.base[nseg] = mids[i].base;
      sg.n[nseg] = mids[i].n;
      seg_max = std::max(seg_max, m);
      s->fill_n += m;
      s->pend_n += m;
      e0 += m;
      if (++nseg == kAppendSegs && (rc = append())) return rc;
      if (s->pend_n == s->pend_cap) {
        if ((rc = append()) || (rc = process_pending(s)) || (rc = wait_fill_buffer(s))) return rc;
      }
    }
  }
  return append();
}

int wait_copies(Stream* s) {
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->copy_stream) GS_HIP(hipStreamSynchronize(s->copy_stream));
  return GS_OK;
}

int drain_stream(Stream* s) {
  std::lock_guard<std::mutex> lk(s->mu);
  return process_pending(s);
}

int ingest_pairs_device(Stream* s, const uint32_t* d_pairs, uint64_t nedges, uint32_t min_depth,
                        uint32_t depth_cap) {
  GS_CHECK(nedges < (1ULL << 31), GS_ERR_INVALID, "edge batch too large (%llu)",
           (unsigned long long)nedges);
  int rc = submit_pairs_device(s, d_pairs, nedges, min_depth, depth_cap);
  if (rc) return rc;
  return drain_stream(s);
}

// ------------------------------------------------------------------------------------------
// Autotuner: times the whole ingest (grouping + aggregation) of a device-resident sample batch
// for every candidate (LDS depth window, destinations per work item, bucket bits). Each
// candidate applies the batch twice, which toggles every update back out, so the sketches end
// as they started. The fastest candidate (best of two timings) stays set on the stream.
// ------------------------------------------------------------------------------------------
int autotune_stream(Stream* s, const uint32_t* d_pairs, uint64_t nedges, gs_tuning* best) {
  GS_CHECK(nedges > 0, GS_ERR_INVALID, "autotune needs a sample batch");
  const uint32_t min_bpc = s->group ? s->group->min_bpc : s->store->view.bpc;
  const PartParams base = s->group ? s->group->part : s->store->part;
  std::vector<uint32_t> dls, chunks = {256, 512, 1024}, bits;
  for (uint32_t d : {9u, 10u, 11u, 12u})
    if (d <= min_bpc) dls.push_back(d);
  if (dls.empty()) dls.push_back(min_bpc);
  for (int b = (int)base.local_bits - 1; b <= (int)base.local_bits + 1; b++)
    if (b >= 0 && b <= 13 && b <= (int)base.nbits && (base.nbits - b) <= 13) bits.push_back((uint32_t)b);
  hipEvent_t e0, e1;
  GS_HIP(hipEventCreate(&e0));
  GS_HIP(hipEventCreate(&e1));
  const bool prof = s->profile;
  s->profile = false;
  gs_tuning win{0, 0, 0, s->tune_launch, 1e30f};
  int rc = GS_OK;
  for (uint32_t b : bits)
    for (uint32_t c : chunks)
      for (uint32_t d : dls) {
        s->tune_dl = d;
        s->tune_chunk = c;
        s->tune_bucket_bits = b;
        s->tuned_part = make_part_params(stream_vertices(s), (int)b);
        float t_best = 1e30f;
        for (int rep = 0; rep < 2 && rc == GS_OK; rep++) {
          if (hipEventRecord(e0, s->stream) != hipSuccess) rc = fail(GS_ERR_HIP, "event record");
          if (!rc) rc = ingest_pairs_device(s, d_pairs, nedges, 0, 0);
          if (!rc && hipEventRecord(e1, s->stream) != hipSuccess) rc = fail(GS_ERR_HIP, "event record");
          if (!rc) rc = ingest_pairs_device(s, d_pairs, nedges, 0, 0);  // toggles it back out
          if (!rc && hipEventSynchronize(e1) != hipSuccess) rc = fail(GS_ERR_HIP, "event sync");
          float ms = 0;
          if (!rc && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = fail(GS_ERR_HIP, "elapsed");
          if (!rc) t_best = std::min(t_best, ms);
        }
        if (rc) break;
        if (t_best < win.kernel_ms) win = gs_tuning{d, c, b, s->tune_launch, t_best};
      }
  (void)hipStreamSynchronize(s->stream);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  s->profile = prof;
  if (rc) {
    s->tune_dl = s->tune_chunk = s->tune_bucket_bits = 0;
    return rc;
  }
  s->tune_dl = win.lds_depths;
  s->tune_chunk = win.item_updates;
  s->tune_bucket_bits = win.bucket_bits;
  s->tuned_part = make_part_params(stream_vertices(s), (int)win.bucket_bits);
  if (best) *best = win;
  return GS_OK;
}

__global__ void copy16_kernel(uint4* __restrict__ dst, const uint4* __restrict__ src, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

int device_copy(void* dst, const void* src, size_t bytes) {
  GS_CHECK(bytes % 16 == 0 && ((uintptr_t)dst % 16) == 0 && ((uintptr_t)src % 16) == 0,
           GS_ERR_INVALID, "gs_device_copy: 16-byte aligned sizes and pointers only");
  uint64_t n = bytes / 16;
  unsigned grid = (unsigned)std::min<uint64_t>((n + 255) / 256, 65536);
  if (grid == 0) return GS_OK;
  hipLaunchKernelGGL(copy16_kernel, dim3(grid), dim3(256), 0, 0, reinterpret_cast<uint4*>(dst),
                     reinterpret_cast<const uint4*>(src), n);
  GS_HIP(hipGetLastError());
  GS_HIP(hipDeviceSynchronize());
  return GS_OK;
}

int launch_zero_store(Store* st, hipStream_t stream) {
  const StoreView& v = st->view;
  uint64_t words = (uint64_t)v.n * v.vstride * 2;
  hipLaunchKernelGGL(fill_u64_kernel, dim3(2048), dim3(256), 0, stream,
                     reinterpret_cast<uint64_t*>(v.rows), words, 0ULL);
  GS_HIP(hipGetLastError());
  hipLaunchKernelGGL(fill_u64_kernel, dim3(256), dim3(256), 0, stream,
                     reinterpret_cast<uint64_t*>(v.det), (uint64_t)v.n * 2, 0ULL);
  GS_HIP(hipGetLastError());
  return GS_OK;
}

// known-answer kernel: both the direct and the split form, so a mismatch between them shows
__global__ void hash_kat_kernel(const uint64_t* keys, const uint64_t* seeds, uint64_t n,
                                uint64_t* out, uint32_t* split_mismatch) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t h = xxh3_len8(keys[i], seeds[i]);
    uint64_t h2 = xxh3_finish(xxh3_mix(xxh3_swap(keys[i])) ^ xxh3_seed_key(seeds[i]));
    if (h != h2) atomicAdd(split_mismatch, 1u);
    out[i] = h;
  }
}

int hash_known_answer(const uint64_t* keys, const uint64_t* seeds, uint64_t count, uint64_t* out) {
  if (count == 0) return GS_OK;
  uint64_t *dk = nullptr, *ds = nullptr, *dout = nullptr;
  uint32_t* dmis = nullptr;
  uint32_t mis = 0;
  int rc = GS_OK;
  auto cleanup = [&] { (void)hipFree(dk); (void)hipFree(ds); (void)hipFree(dout); (void)hipFree(dmis); };
  if (hipMalloc(&dk, 8 * count) != hipSuccess || hipMalloc(&ds, 8 * count) != hipSuccess ||
      hipMalloc(&dout, 8 * count) != hipSuccess || hipMalloc(&dmis, 4) != hipSuccess) {
    cleanup();
    return fail(GS_ERR_OOM, "hash_known_answer: allocation failed");
  }
  if (hipMemcpy(dk, keys, 8 * count, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(ds, seeds, 8 * count, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(dmis, 0, 4) != hipSuccess) {
    cleanup();
    return fail(GS_ERR_HIP, "hash_known_answer: copy failed");
  }
  hipLaunchKernelGGL(hash_kat_kernel, dim3((unsigned)std::min<uint64_t>((count + 255) / 256, 4096)),
                     dim3(256), 0, 0, dk, ds, count, dout, dmis);
  if (hipGetLastError() != hipSuccess ||
      hipMemcpy(out, dout, 8 * count, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(&mis, dmis, 4, hipMemcpyDeviceToHost) != hipSuccess) {
    cleanup();
    return fail(GS_ERR_HIP, "hash_known_answer: kernel failed");
  }
  cleanup();
  if (mis) rc = fail(GS_ERR_INVALID, "split hash form disagrees on %u keys", mis);
  return rc;
}

}  // namespace gs
