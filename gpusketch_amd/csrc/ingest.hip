// Sketch update kernels for gfx950.
//
// Two entry paths share one aggregation kernel:
//   * pre-grouped vertex batches (the contract of the reference's sketchUpdate kernels,
//     src/cuda_kernel.cu:152-257): the host turns each batch into work items;
//   * raw undirected pairs (edge batches): the device groups directed entries by vertex in three
//     kernels (bucket_count, bucket_scatter, bucket_sort) and emits the work items itself.
//
// aggregate_kernel: one wave64 per work item (<= kChunk destinations of one vertex). For each
// destination the wave hashes the edge index once for the checksum (XXH3 with the sketch seed)
// and once per column (seed + 5c), XORs {idx, checksum} into a per-wave LDS delta laid out like
// the device rows (depth-major), then flushes only the non-zero buckets to HBM: a plain
// read-xor-write when the item owns its vertex in this launch, 64-bit atomic XOR otherwise. The
// deterministic bucket is a wave XOR-reduction. Depths >= the LDS window go straight to HBM with
// atomics (probability 2^-16 per column and update).
#include "gs_internal.h"

namespace gs {

constexpr int kWave = 64;
constexpr int kAggWaves = 4;                  // waves per aggregation workgroup
constexpr uint32_t kChunk = 512;              // max destinations per work item
constexpr int kPartThreads = 512;             // K1/K2 workgroup size
constexpr int kPartEdges = 4096;              // undirected edges per K1/K2 workgroup
constexpr int kSortThreads = 1024;            // K3 workgroup size (one workgroup per bucket)
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

__global__ void fill_u64_kernel(uint64_t* p, uint64_t n, uint64_t v) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// ------------------------------------------------------------------------------------------
// Vertex grouping for raw pairs, in two levels with few global atomics:
//   K1 append_count  : one per submission: copies the pairs into the stream's fill buffer (in
//                      place for host batches), drops out-of-range pairs (counted, reported at
//                      the flush), self loops and pairs outside the subgraph filter (written as
//                      the no-op pair (0, 0)), and adds the batch's directed entries per bucket of
//                      hashed vertex ids to the fill buffer's bucket counts (LDS histogram, one
//                      non-returning global add per workgroup and bucket);
//   K2 bucket_scatter: once per fill buffer: entries written into their bucket's region (LDS
//                      rank + one returning global add per workgroup and bucket);
//   K3 bucket_sort   : one workgroup per bucket sorts its entries by vertex (LDS counting sort
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
__device__ __forceinline__ bool entry_ok(uint2 uv, uint32_t n) {
  return uv.x < n && uv.y < n && uv.x != uv.y;
}

// src and dst may be the same buffer (each pair is read and written by one thread)
__global__ __launch_bounds__(kPartThreads) void append_count_kernel(
    const uint2* src, uint2* dst, uint64_t nedges, uint32_t n, PartParams pp, uint64_t seed,
    uint32_t min_depth, uint32_t depth_cap, uint32_t* __restrict__ counts_cur,
    uint32_t* __restrict__ counts_next, uint32_t* __restrict__ cursors,
    uint32_t* __restrict__ counters) {
  extern __shared__ uint32_t hist[];
  const uint32_t t = threadIdx.x;
  if (blockIdx.x == 0) {
    // the next fill buffer's counts and this one's scatter cursors start at zero; the class
    // counters belong to the launch that will process this fill buffer
    for (uint32_t i = t; i < pp.nbuckets; i += kPartThreads) { counts_next[i] = 0; cursors[i] = 0; }
    if (t < kClassBase + kClasses && t != 3) counters[t] = 0;  // [3] persists until a flush
  }
  for (uint32_t i = t; i < pp.nbuckets; i += kPartThreads) hist[i] = 0;
  __syncthreads();
  const uint64_t e0 = (uint64_t)blockIdx.x * kPartEdges;
#pragma unroll
  for (int k = 0; k < kPartEdges / kPartThreads; k++) {
    uint64_t e = e0 + t + (uint64_t)k * kPartThreads;
    if (e >= nedges) continue;
    uint2 uv = src[e];
    bool ok = true;
    if (uv.x >= n || uv.y >= n) {  // never index past the store: drop and report at flush
      atomicAdd(&counters[3], 1u);
      ok = false;
    } else if (uv.x == uv.y || !edge_selected(concat_pair(uv.x, uv.y), seed, min_depth, depth_cap)) {
      ok = false;
    }
    dst[e] = ok ? uv : make_uint2(0, 0);
    if (!ok) continue;
    atomicAdd(&hist[bucket_of(pp, uv.x)], 1u);
    atomicAdd(&hist[bucket_of(pp, uv.y)], 1u);
  }
  __syncthreads();
  for (uint32_t i = t; i < pp.nbuckets; i += kPartThreads)
    if (hist[i]) __hip_atomic_fetch_add(&counts_cur[i], hist[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// exclusive prefix of counts[0..nb) into s_off (LDS), returns the total; blockDim threads
__device__ uint32_t block_exclusive_scan(const uint32_t* counts, uint32_t nb, uint32_t* s_off,
                                         uint32_t* s_part) {
  const uint32_t t = threadIdx.x, nt = blockDim.x;
  const uint32_t per = (nb + nt - 1) / nt;
  const uint32_t b0 = min(nb, t * per), b1 = min(nb, b0 + per);
  uint32_t sum = 0;
  for (uint32_t b = b0; b < b1; b++) { s_off[b] = sum; sum += counts[b]; }
  s_part[t] = sum;
  __syncthreads();
  for (uint32_t o = 1; o < nt; o <<= 1) {
    uint32_t add = t >= o ? s_part[t - o] : 0;
    __syncthreads();
    s_part[t] += add;
    __syncthreads();
  }
  uint32_t base = t ? s_part[t - 1] : 0;
  for (uint32_t b = b0; b < b1; b++) s_off[b] += base;
  uint32_t total = s_part[nt - 1];
  __syncthreads();
  return total;
}

__global__ __launch_bounds__(kPartThreads) void bucket_scatter_kernel(
    const uint32_t* __restrict__ pairs, uint64_t nedges, uint32_t n, PartParams pp,
    const uint32_t* __restrict__ counts, uint32_t* __restrict__ cursors,
    uint2* __restrict__ entries) {
  extern __shared__ uint32_t lds_u32[];
  uint32_t* s_off = lds_u32;                   // nbuckets: region start of every bucket
  uint32_t* s_hist = lds_u32 + pp.nbuckets;    // nbuckets: this block's count, then its base
  __shared__ uint32_t s_part[kPartThreads];
  const uint32_t t = threadIdx.x;
  for (uint32_t i = t; i < pp.nbuckets; i += kPartThreads) s_hist[i] = 0;
  block_exclusive_scan(counts, pp.nbuckets, s_off, s_part);

  const uint64_t e0 = (uint64_t)blockIdx.x * kPartEdges;
  constexpr int K = kPartEdges / kPartThreads;
  uint32_t bk[2 * K], rk[2 * K];
  uint2 uvs[K];
#pragma unroll
  for (int k = 0; k < K; k++) {
    uint64_t e = e0 + t + (uint64_t)k * kPartThreads;
    bk[2 * k] = bk[2 * k + 1] = kEmpty;
    uvs[k] = make_uint2(0, 0);
    if (e >= nedges) continue;
    uint2 uv = reinterpret_cast<const uint2*>(pairs)[e];
    if (!entry_ok(uv, n)) continue;
    uvs[k] = uv;
    bk[2 * k] = bucket_of(pp, uv.x);
    bk[2 * k + 1] = bucket_of(pp, uv.y);
    rk[2 * k] = atomicAdd(&s_hist[bk[2 * k]], 1u);
    rk[2 * k + 1] = atomicAdd(&s_hist[bk[2 * k + 1]], 1u);
  }
  __syncthreads();
  for (uint32_t i = t; i < pp.nbuckets; i += kPartThreads) {
    uint32_t c = s_hist[i];
    if (c) s_hist[i] = s_off[i] + __hip_atomic_fetch_add(&cursors[i], c, __ATOMIC_RELAXED,
                                                          __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; k++) {
    if (bk[2 * k] == kEmpty) continue;
    const uint2 uv = uvs[k];
    entries[s_hist[bk[2 * k]] + rk[2 * k]] = uv;
    entries[s_hist[bk[2 * k + 1]] + rk[2 * k + 1]] = make_uint2(uv.y, uv.x);
  }
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

__global__ __launch_bounds__(kSortThreads) void bucket_sort_kernel(
    const uint2* __restrict__ entries, PartParams pp, const uint32_t* __restrict__ counts,
    uint32_t* __restrict__ sorted, WorkItem* __restrict__ items, ItemClasses cls, uint32_t chunk,
    uint32_t* __restrict__ class_counts, const uint32_t* __restrict__ bases, uint32_t ngraphs) {
  extern __shared__ uint32_t lds_u32[];
  const uint32_t bins = 1u << pp.local_bits;
  uint32_t* s_cnt = lds_u32;              // per local id: count
  uint32_t* s_off = lds_u32 + bins;       // per local id: offset inside the bucket, then cursor
  __shared__ uint32_t s_part[kSortThreads];
  __shared__ uint32_t s_ccnt[kClasses], s_cbase[kClasses];
  const uint32_t t = threadIdx.x, b = blockIdx.x;
  // region start of this bucket = sum of the counts of the buckets before it
  uint32_t acc = 0;
  for (uint32_t i = t; i < b; i += kSortThreads) acc += counts[i];
  s_part[t] = acc;
  for (uint32_t i = t; i < bins; i += kSortThreads) s_cnt[i] = 0;
  __syncthreads();
  for (uint32_t o = kSortThreads / 2; o > 0; o >>= 1) {
    if (t < o) s_part[t] += s_part[t + o];
    __syncthreads();
  }
  const uint32_t base = s_part[0];
  const uint32_t cnt = counts[b];
  __syncthreads();
  const uint32_t lmask = bins - 1;
  for (uint32_t i = t; i < cnt; i += kSortThreads)
    atomicAdd(&s_cnt[hashed_id(pp, entries[base + i].x) & lmask], 1u);
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
  for (uint32_t i = t; i < cnt; i += kSortThreads) {
    uint2 e = entries[base + i];
    uint32_t pos = atomicAdd(&s_off[hashed_id(pp, e.x) & lmask], 1u);
    sorted[base + pos] = e.y;
  }
}

// ------------------------------------------------------------------------------------------
// K4: aggregation, one wave64 per work item (see the file comment).
//
// Lane layout: lane = 8*e + s with e = entry slot of a group of 8 consecutive destinations and
// s = column slot; the lane handles columns s, s+8, ..., s+8(T-1) (those < cols) of its entry.
// Its column keys therefore stay in registers for the whole item, and an entry's staged values
// are read from LDS once for T columns (8 lanes read one address: a broadcast).
//
// Depth 0 receives half of all (update, column) pairs; those are XORed into per-lane registers
// and merged into the LDS delta once per item. Depths 1..dl-1 are XORed into the LDS delta, so
// the 8 lanes sharing a column rarely hit one LDS address in the same instruction. Depths >= dl
// (probability 2^-dl per pair) go to HBM with 64-bit atomics.
//
// Per-wave LDS (64-bit words): the delta as structure of arrays (alpha[dslots], gamma[dslots],
// slot = depth * cols + column) and the chunk's stage (edge index, checksum, mix(swap(idx)), 64
// each).
// ------------------------------------------------------------------------------------------
struct AggLayout {
  uint32_t cols, dl, dslots;
  __host__ __device__ AggLayout(uint32_t cols_, uint32_t dl_) : cols(cols_), dl(dl_), dslots(dl_ * cols_) {}
  __host__ __device__ uint32_t wave_words() const { return 2 * dslots + 3 * 64; }
};
constexpr int kMaxColSteps = 8;  // T <= 8: up to 64 columns
constexpr int kFlushRounds = 4;  // 64-slot rounds per flush step

// kGroup: items belong to the members of a group (descs[item.graph]); the LDS layout is sized
// for the widest member (max_cols) and the per-member geometry is reloaded when a wave moves to
// an item of another member. Otherwise every item belongs to the store `st`.
// T = ceil(columns / 8) of the widest store the launch covers.
template <bool kGroup, int T>
__global__ __launch_bounds__(kAggWaves* kWave) void aggregate_kernel(
    StoreView st, const GraphDesc* __restrict__ descs, uint32_t max_cols,
    const WorkItem* __restrict__ items, ItemClasses cls,
    const uint32_t* __restrict__ class_counts, const uint32_t* __restrict__ dsts, uint32_t dl,
    int exclusive_ok, uint32_t min_depth, uint32_t depth_cap) {
  extern __shared__ __align__(16) unsigned char lds[];
  const AggLayout L(kGroup ? max_cols : st.cols, dl);
  uint32_t cols = st.cols;
  uint32_t bpc = st.bpc;
  uint64_t seed = st.seed;
  uint32_t vbase = 0;
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint32_t wid = threadIdx.x / kWave;
  const uint32_t le = lane >> 3, ls = lane & 7;
  uint64_t* da = reinterpret_cast<uint64_t*>(lds) + (size_t)wid * L.wave_words();
  uint64_t* dg = da + L.dslots;
  uint64_t* sidx = dg + L.dslots;
  uint64_t* schk = sidx + kWave;
  uint64_t* smix = schk + kWave;

  uint64_t colk[T];
  uint64_t acc_a[T], acc_g[T];
#pragma unroll
  for (int t = 0; t < T; t++) {
    const uint32_t c = ls + 8 * t;
    colk[t] = (!kGroup && c < cols) ? xxh3_seed_key(seed + 5ULL * c) : 0;
    acc_a[t] = acc_g[t] = 0;
  }
  for (uint32_t e = lane; e < L.dslots; e += kWave) { da[e] = 0; dg[e] = 0; }
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
  // destinations is loaded while the current one is processed
  // the first chunk of an item's destinations; lanes past its length load its first entry (an
  // unconditional load, so no branch joins force an early wait; those values are never used)
  auto first_chunk = [&](uint32_t ln, uint32_t sp) -> uint32_t {
    return dsts[(uint64_t)sp + (lane < ln ? lane : 0u)];
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
#pragma unroll
        for (int t = 0; t < T; t++) {
          const uint32_t c = ls + 8 * t;
          colk[t] = c < cols ? xxh3_seed_key(seed + 5ULL * c) : 0;
        }
      }
    }
    const uint32_t v = v_g - vbase;
    const uint32_t len = wlen & ~kItemExclusive;
    const bool excl = exclusive_ok && (wlen & kItemExclusive);
    // LDS depth window of this item: floor(log2 len) + 4 depths (at most dl); the deeper pairs,
    // about 1.5 per item, go to HBM atomics, and the flush scans only the window
    const uint32_t dli = min(dl, 35u - (uint32_t)__builtin_clz(len));
    const uint32_t islots = dli * cols;
    uint64_t det_a = 0, det_g = 0;
    Bucket16* vrows = st.rows + (size_t)v * st.vstride;

    for (uint32_t j0 = 0; j0 < len; j0 += kWave) {
      const uint32_t cnt = min((uint32_t)kWave, len - j0);
      // next chunk (clamped to the item's last entry: an unconditional load)
      const uint32_t dn = dsts[start + min(j0 + kWave + lane, len - 1)];
      if (lane < cnt) {
        const uint32_t other = dcur - vbase;
        uint64_t idx = concat_pair(v, other);
        uint64_t chk = xxh3_len8(idx, seed);
        // subgraph filter: an update outside the store's subgraph becomes the (0, 0) no-op
        if (min_depth && depth_of(chk, depth_cap) < min_depth) idx = chk = 0;
        sidx[lane] = idx;
        schk[lane] = chk;
        smix[lane] = xxh3_mix(xxh3_swap(idx));
        det_a ^= idx;
        det_g ^= chk;
      }
      dcur = dn;
      wave_sync_lds();
      const uint32_t ngroups = (cnt + 7) >> 3;
      for (uint32_t g = 0; g < ngroups; g++) {
        const uint32_t j = 8 * g + le;
        if (j < cnt) {
          const uint64_t a = sidx[j], gm = schk[j], m = smix[j];
#pragma unroll
          for (int t = 0; t < T; t++) {
            const uint32_t c = ls + 8 * t;
            if (c < cols) {
              const uint32_t d = depth_fast(m ^ colk[t], bpc);
              if (d == 0) {
                acc_a[t] ^= a;
                acc_g[t] ^= gm;
              } else if (d < dli) {
                const uint32_t slot = mul24_lo(d, cols) + c;
                atomicXor(reinterpret_cast<unsigned long long*>(&da[slot]), (unsigned long long)a);
                atomicXor(reinterpret_cast<unsigned long long*>(&dg[slot]), (unsigned long long)gm);
              } else if (d < bpc) {  // rare: deep buckets straight to HBM
                Bucket16* b = &vrows[(size_t)d * cols + c];
                atomic_xor_u64(&b->alpha, a);
                atomic_xor_u64(&b->gamma, gm);
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
    // depth-0 registers into the LDS delta (slot = column): the 8 lanes of a column slot XOR into
    // one address, once per item
#pragma unroll
    for (int t = 0; t < T; t++) {
      const uint32_t c = ls + 8 * t;
      if (c < cols && (acc_a[t] | acc_g[t])) {
        atomicXor(reinterpret_cast<unsigned long long*>(&da[c]), (unsigned long long)acc_a[t]);
        atomicXor(reinterpret_cast<unsigned long long*>(&dg[c]), (unsigned long long)acc_g[t]);
      }
      acc_a[t] = acc_g[t] = 0;
    }
    det_a = wave_xor_reduce(det_a);
    det_g = wave_xor_reduce(det_g);
    if (lane == 0 && (det_a | det_g)) {
      Bucket16* b = &st.det[v];
      atomic_xor_u64(&b->alpha, det_a);
      atomic_xor_u64(&b->gamma, det_g);
    }
    wave_sync_lds();
    // flush, kFlushRounds x 64 slots at a time: all LDS reads first, then the HBM reads of the
    // non-zero slots together, one wait, then the stores (exclusive item: plain read-xor-write;
    // otherwise 64-bit atomics, which need no read)
    for (uint32_t s0 = 0; s0 < islots; s0 += kFlushRounds * kWave) {
      uint64_t xa[kFlushRounds], xg[kFlushRounds];
#pragma unroll
      for (int r = 0; r < kFlushRounds; r++) {
        const uint32_t sl = s0 + r * kWave + lane;
        xa[r] = sl < islots ? da[sl] : 0;
        xg[r] = sl < islots ? dg[sl] : 0;
      }
#pragma unroll
      for (int r = 0; r < kFlushRounds; r++) {
        const uint32_t sl = s0 + r * kWave + lane;
        if (sl < islots) {
          da[sl] = 0;
          dg[sl] = 0;
        }
      }
      if (excl) {
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
}

// ------------------------------------------------------------------------------------------
// launch helpers
// ------------------------------------------------------------------------------------------
// depths kept in the LDS delta; deeper updates (probability 2^-12 per column) go to HBM directly
static uint32_t lds_window(const StoreView& v) { return v.bpc < 12 ? v.bpc : 12; }

static size_t agg_lds_bytes(uint32_t cols, uint32_t dl) {
  const AggLayout L(cols, dl);
  return (size_t)kAggWaves * L.wave_words() * sizeof(uint64_t);
}

// persistent grid: as many workgroups as can be resident at once (the occupancy calculator
// accounts for LDS and registers), so the static round-robin over work items never waits on a
// late-starting workgroup
static int agg_grid(int device, const void* kernel, size_t lds_bytes) {
  static int cu = 0;
  if (cu == 0) {
    if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cu <= 0)
      cu = 256;
  }
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kAggWaves * kWave, lds_bytes) !=
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
};
struct AggKernel {
  const void* fn;
  void (*launch)(dim3, dim3, size_t, hipStream_t, const AggArgs&);
};
template <bool G, int T>
static void agg_launch(dim3 grid, dim3 block, size_t lds, hipStream_t s, const AggArgs& a) {
  hipLaunchKernelGGL((aggregate_kernel<G, T>), grid, block, lds, s, a.v, a.descs, a.cols, a.items,
                     a.cls, a.class_counts, a.dsts, a.dl, a.exclusive_ok, a.min_depth, a.depth_cap);
}
template <bool G, int T>
static AggKernel agg_kernel_entry() {
  return AggKernel{reinterpret_cast<const void*>(&aggregate_kernel<G, T>), &agg_launch<G, T>};
}
static const AggKernel& agg_kernel_for(bool group, uint32_t cols) {
  static const AggKernel table[2][kMaxColSteps] = {
      {agg_kernel_entry<false, 1>(), agg_kernel_entry<false, 2>(), agg_kernel_entry<false, 3>(),
       agg_kernel_entry<false, 4>(), agg_kernel_entry<false, 5>(), agg_kernel_entry<false, 6>(),
       agg_kernel_entry<false, 7>(), agg_kernel_entry<false, 8>()},
      {agg_kernel_entry<true, 1>(), agg_kernel_entry<true, 2>(), agg_kernel_entry<true, 3>(),
       agg_kernel_entry<true, 4>(), agg_kernel_entry<true, 5>(), agg_kernel_entry<true, 6>(),
       agg_kernel_entry<true, 7>(), agg_kernel_entry<true, 8>()}};
  const uint32_t t = (cols + 7) / 8;
  return table[group ? 1 : 0][t - 1];
}

static int launch_aggregate(Stream* s, const WorkItem* d_items, const ItemClasses& cls,
                            const uint32_t* class_counts, const uint32_t* d_dsts,
                            uint64_t updates, uint32_t min_depth = 0, uint32_t depth_cap = 0) {
  Group* grp = s->group;
  StoreView v{};
  uint32_t cols, dl;
  int device = 0;
  int exclusive_ok = 1;
  if (grp) {
    cols = grp->max_cols;
    dl = std::min(grp->min_bpc, s->tune_dl ? s->tune_dl : 12u);
    device = grp->members[0]->device;
    for (Store* m : grp->members) {  // plain stores only if the group is every member's only writer
      std::lock_guard<std::mutex> lk(m->mu);
      if (m->streams.size() != 1) exclusive_ok = 0;
    }
  } else {
    v = s->store->view;
    cols = v.cols;
    dl = s->tune_dl ? std::min(s->tune_dl, v.bpc) : lds_window(v);
    device = s->store->device;
    std::lock_guard<std::mutex> lk(s->store->mu);
    exclusive_ok = s->store->streams.size() == 1 ? 1 : 0;
  }
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
  AggArgs a{v, grp ? grp->d_descs : nullptr, cols, d_items, cls, class_counts, d_dsts, dl,
            exclusive_ok, grp ? 0u : min_depth, grp ? 0u : depth_cap};
  kern.launch(grid, block, lds, s->stream, a);
  GS_HIP(hipGetLastError());
  if (s->profile) {
    GS_HIP(hipEventRecord(e1, s->stream));
    s->ev_pairs.emplace_back(e0, e1);
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
  (void)hipFree(ws.d_counts);
  (void)hipFree(ws.d_cursors);
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

int ensure_pair_scratch(Stream* s, uint64_t nedges) {
  PairScratch& ws = s->ws;
  const PartParams& pp = stream_part(s);
  if (ws.cap_buckets < pp.nbuckets) {
    GS_HIP(hipStreamSynchronize(s->stream));
    (void)hipFree(ws.d_counts);
    (void)hipFree(ws.d_cursors);
    (void)hipFree(ws.d_counters);
    ws.d_counts = ws.d_cursors = ws.d_counters = nullptr;
    GS_HIP(hipMalloc(&ws.d_counts, sizeof(uint32_t) * 2 * pp.nbuckets));
    GS_HIP(hipMalloc(&ws.d_cursors, sizeof(uint32_t) * pp.nbuckets));
    GS_HIP(hipMalloc(&ws.d_counters, sizeof(uint32_t) * kCounterWords));
    GS_HIP(hipMemsetAsync(ws.d_counts, 0, sizeof(uint32_t) * 2 * pp.nbuckets, s->stream));
    GS_HIP(hipMemsetAsync(ws.d_counters, 0, sizeof(uint32_t) * kCounterWords, s->stream));
    ws.cap_buckets = pp.nbuckets;
    ws.cur_buckets = pp.nbuckets;
    ws.parity = 0;
  }
  if (ws.cur_buckets != pp.nbuckets) {  // re-tuned: both count buffers must start at zero
    GS_HIP(hipMemsetAsync(ws.d_counts, 0, sizeof(uint32_t) * 2 * ws.cap_buckets, s->stream));
    ws.cur_buckets = pp.nbuckets;
    ws.parity = 0;
  }
  if (ws.cap_edges < nedges) {
    GS_HIP(hipStreamSynchronize(s->stream));
    (void)hipFree(ws.d_entries);
    (void)hipFree(ws.d_sorted);
    (void)hipFree(ws.d_items);
    ws.d_entries = nullptr;
    ws.d_sorted = nullptr;
    ws.d_items = nullptr;
    uint64_t cap = nedges;
    // class k holds items of length >= 2^k, at most one shorter-than-chunk item per vertex
    const uint64_t verts = std::min<uint64_t>((uint64_t)stream_vertices(s), 2 * cap);
    uint64_t icap = 0;
    // classes that may hold full items (len >= 64, the smallest item size a stream can be tuned
    // to; items of one vertex may repeat) are bounded by the entries alone; shorter items are a
    // vertex's last, at most one per vertex
    for (int k = 0; k < kClasses; k++) {
      ws.cls.off[k] = icap;
      uint64_t ck = (2 * cap) >> k;
      if (k < 6) ck = std::min<uint64_t>(verts, ck);
      icap += ck + 16;
    }
    GS_HIP(hipMalloc(&ws.d_entries, sizeof(uint2) * 2 * cap));
    GS_HIP(hipMalloc(&ws.d_sorted, sizeof(uint32_t) * 2 * cap));
    GS_HIP(hipMalloc(&ws.d_items, sizeof(WorkItem) * icap));
    ws.cap_edges = cap;
    ws.cap_items = icap;
  }
  return GS_OK;
}

uint64_t stream_launch_edges(const Stream* s) {
  return s->tune_launch ? (uint64_t)s->tune_launch : kLaunchEdges;
}

// the fill buffers, sized for the stream's launch size (grown on demand, never while pending)
static int ensure_pending(Stream* s) {
  const uint64_t want = stream_launch_edges(s);
  if (s->pend_cap >= want && s->d_pend[0]) return ensure_pair_scratch(s, s->pend_cap);
  GS_CHECK(s->pend_n == 0, GS_ERR_INVALID, "launch size changed with updates pending");
  GS_HIP(hipStreamSynchronize(s->stream));
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
  return ensure_pair_scratch(s, want);
}

// K1 over `m` pairs at src, into the fill buffer behind what is pending
static int append_pairs(Stream* s, const uint32_t* src, uint64_t m, uint32_t min_depth,
                        uint32_t depth_cap) {
  PairScratch& ws = s->ws;
  const PartParams& pp = stream_part(s);
  const uint32_t nverts = stream_vertices(s);
  const uint64_t seed = s->group ? 0 : s->store->view.seed;
  GS_CHECK(!s->group || min_depth == 0, GS_ERR_INVALID, "no subgraph filter on group streams");
  uint32_t* counts_cur = ws.d_counts + (size_t)ws.parity * pp.nbuckets;
  uint32_t* counts_next = ws.d_counts + (size_t)(ws.parity ^ 1) * pp.nbuckets;
  int rc = ensure_dyn_lds((const void*)append_count_kernel, sizeof(uint32_t) * pp.nbuckets);
  if (rc) return rc;
  uint2* dst = reinterpret_cast<uint2*>(s->d_pend[s->pend_buf]) + s->pend_n;
  const uint32_t nblk = (uint32_t)((m + kPartEdges - 1) / kPartEdges);
  hipLaunchKernelGGL(append_count_kernel, dim3(nblk), dim3(kPartThreads),
                     sizeof(uint32_t) * pp.nbuckets, s->stream, reinterpret_cast<const uint2*>(src),
                     dst, m, nverts, pp, seed, min_depth, depth_cap, counts_cur, counts_next,
                     ws.d_cursors, ws.d_counters);
  GS_HIP(hipGetLastError());
  s->pend_n += m;
  return GS_OK;
}

// K2..K4 over the fill buffer; the other buffer becomes the fill buffer
static int process_pending(Stream* s) {
  if (s->pend_n == 0) return GS_OK;
  PairScratch& ws = s->ws;
  const PartParams& pp = stream_part(s);
  const uint32_t nverts = stream_vertices(s);
  const uint64_t nedges = s->pend_n;
  const int b = s->pend_buf;
  uint32_t* counts_cur = ws.d_counts + (size_t)ws.parity * pp.nbuckets;
  const uint32_t nblk = (uint32_t)((nedges + kPartEdges - 1) / kPartEdges);
  const size_t sort_lds = sizeof(uint32_t) * 3 * ((size_t)1 << pp.local_bits);
  int rc;
  if ((rc = ensure_dyn_lds((const void*)bucket_scatter_kernel, sizeof(uint32_t) * 2 * pp.nbuckets)) ||
      (rc = ensure_dyn_lds((const void*)bucket_sort_kernel, sort_lds)))
    return rc;
  hipLaunchKernelGGL(bucket_scatter_kernel, dim3(nblk), dim3(kPartThreads),
                     sizeof(uint32_t) * 2 * pp.nbuckets, s->stream, s->d_pend[b], nedges, nverts,
                     pp, counts_cur, ws.d_cursors, ws.d_entries);
  GS_HIP(hipGetLastError());
  GS_HIP(hipEventRecord(s->pend_free[b], s->stream));  // the fill buffer has been read
  s->pend_used[b] = true;
  hipLaunchKernelGGL(bucket_sort_kernel, dim3(pp.nbuckets), dim3(kSortThreads), sort_lds,
                     s->stream, ws.d_entries, pp, counts_cur, ws.d_sorted, ws.d_items, ws.cls,
                     s->tune_chunk ? s->tune_chunk : kChunk, ws.d_counters + kClassBase,
                     s->group ? (const uint32_t*)s->group->d_bases : (const uint32_t*)nullptr,
                     s->group ? (uint32_t)s->group->members.size() : 1u);
  GS_HIP(hipGetLastError());
  ws.parity ^= 1;
  s->pend_n = 0;
  s->pend_buf ^= 1;
  return launch_aggregate(s, ws.d_items, ws.cls, ws.d_counters + kClassBase, ws.d_sorted,
                          2 * nedges);
}

int submit_pairs_device(Stream* s, const uint32_t* d_pairs, uint64_t nedges, uint32_t min_depth,
                        uint32_t depth_cap) {
  if (nedges == 0) return GS_OK;
  std::lock_guard<std::mutex> lk(s->mu);
  int rc = ensure_pending(s);
  if (rc) return rc;
  // (the launches that read the fill buffer earlier ran on this same stream)
  for (uint64_t e0 = 0; e0 < nedges;) {
    const uint64_t m = std::min(nedges - e0, s->pend_cap - s->pend_n);
    if ((rc = append_pairs(s, d_pairs + 2 * e0, m, min_depth, depth_cap))) return rc;
    e0 += m;
    if (s->pend_n == s->pend_cap && (rc = process_pending(s))) return rc;
  }
  return GS_OK;
}

// Host batches: copied on the copy stream straight into the fill buffer (from the caller's
// memory when it is pinned, else through the stream's pinned staging, range-checked on the way),
// then appended in place on the compute stream. The call returns once the caller's memory has
// been read.
int submit_pairs_host(Stream* s, const uint32_t* pairs, uint64_t nedges, uint32_t min_depth,
                      uint32_t depth_cap) {
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
  const uint32_t n = stream_vertices(s);
  for (uint64_t e0 = 0; e0 < nedges;) {
    const uint64_t m = std::min(nedges - e0, s->pend_cap - s->pend_n);
    const size_t bytes = m * 2 * sizeof(uint32_t);
    const uint32_t* src = pairs + 2 * e0;
    if (!pinned) {
      if ((rc = ensure_stage(s, bytes, false))) return rc;
      uint32_t bad = 0;
      for (uint64_t i = 0; i < 2 * m; i++) {
        const uint32_t v = src[i];
        bad |= v >= n;
        s->h_stage[i] = v;
      }
      if (bad) {
        for (uint64_t i = 0; i < 2 * m; i++)
          GS_CHECK(src[i] < n, GS_ERR_INVALID, "edge %llu: vertex %u out of range",
                   (unsigned long long)(e0 + i / 2), src[i]);
      }
      src = s->h_stage;
    }
    // the fill buffer may still be read by the launch that processed it two buffers back
    if (s->pend_used[s->pend_buf])
      GS_HIP(hipStreamWaitEvent(s->copy_stream, s->pend_free[s->pend_buf], 0));
    uint32_t* dst = s->d_pend[s->pend_buf] + 2 * s->pend_n;
    GS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s->copy_stream));
    GS_HIP(hipEventRecord(s->copied, s->copy_stream));
    GS_HIP(hipStreamWaitEvent(s->stream, s->copied, 0));
    if (!pinned) {
      GS_HIP(hipEventRecord(s->stage_free, s->copy_stream));
      s->stage_pending = true;
    }
    if ((rc = append_pairs(s, dst, m, min_depth, depth_cap))) return rc;
    e0 += m;
    if (s->pend_n == s->pend_cap && (rc = process_pending(s))) return rc;
  }
  if (pinned) GS_HIP(hipEventSynchronize(s->copied));  // the caller may reuse its buffer
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
  for (uint32_t d : {8u, 10u, 12u, 14u, 16u})
    if (d <= min_bpc) dls.push_back(d);
  if (dls.empty()) dls.push_back(min_bpc);
  for (int b = (int)base.local_bits - 1; b <= (int)base.local_bits + 1; b++)
    if (b >= 0 && b <= 13 && b <= (int)base.nbits && (base.nbits - b) <= 14) bits.push_back((uint32_t)b);
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
