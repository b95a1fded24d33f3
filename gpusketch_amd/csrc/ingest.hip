// Sketch update kernels for gfx950.
//
// Two entry paths share one aggregation kernel:
//   * pre-grouped vertex batches (the contract of the reference's sketchUpdate kernels,
//     src/cuda_kernel.cu:152-257): the host turns each batch into work items;
//   * raw undirected pairs (edge batches): the device groups directed entries by vertex in three
//     kernels (bin_rank, scan_items, scatter) and emits the work items itself.
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
constexpr uint32_t kChunk = 256;              // max destinations per work item
constexpr uint32_t kSmallItem = 0;            // items this short skip the LDS delta (0 = never)
constexpr int kBinThreads = 256;
constexpr int kBinEdgesPerThread = 4;         // 1024 edges = 2048 entries per bin workgroup
constexpr int kBinHash = 4096;                // LDS hash slots (> entries per workgroup)
constexpr int kScanThreads = 256;
constexpr int kScanPerThread = 8;             // 2048 vertices per scan tile
constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr uint32_t kNoRank = 0xFFFFFFFFu;
constexpr uint64_t kFlagAgg = 1ULL << 62;
constexpr uint64_t kFlagPrefix = 2ULL << 62;
constexpr uint64_t kValueMask = (1ULL << 62) - 1;
constexpr int kEntryBits = 34;                // packed scan value: items << 34 | entries

__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ void atomic_xor_u64(uint64_t* p, uint64_t v) {
  __hip_atomic_fetch_xor(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t wave_xor_reduce(uint64_t x) {
  // butterfly over the 64 lanes with ds_swizzle-free shuffles
  for (int off = 32; off >= 1; off >>= 1) {
    uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
    lo ^= (uint32_t)__shfl_xor((int)lo, off, kWave);
    hi ^= (uint32_t)__shfl_xor((int)hi, off, kWave);
    x = ((uint64_t)hi << 32) | lo;
  }
  return x;
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
// K1: rank every directed entry within its vertex. A workgroup aggregates its 2048 entries per
// vertex in an LDS hash table first, so a hub vertex costs one global atomic per workgroup
// instead of one per entry. Workgroup 0 also resets the scan state used by K2.
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kBinThreads) void bin_rank_kernel(
    const uint32_t* __restrict__ pairs, uint64_t nedges, uint32_t n, uint32_t* __restrict__ deg,
    uint32_t* __restrict__ rank, uint64_t seed, uint32_t min_depth, uint32_t depth_cap,
    uint64_t* __restrict__ tile_status, uint32_t ntiles, uint32_t* __restrict__ counters) {
  __shared__ uint32_t hkey[kBinHash];
  __shared__ uint32_t hcnt[kBinHash];
  const uint32_t t = threadIdx.x;
  if (blockIdx.x == 0) {
    for (uint32_t i = t; i < ntiles; i += kBinThreads) tile_status[i] = 0;
    if (t == 0) { counters[0] = 0; counters[1] = 0; counters[2] = 0; }
  }
  for (uint32_t i = t; i < kBinHash; i += kBinThreads) { hkey[i] = kEmpty; hcnt[i] = 0; }
  __syncthreads();

  const uint64_t e0 = (uint64_t)blockIdx.x * (kBinThreads * kBinEdgesPerThread);
  uint32_t slot[2 * kBinEdgesPerThread];
  uint32_t lrank[2 * kBinEdgesPerThread];
#pragma unroll
  for (int k = 0; k < kBinEdgesPerThread; k++) {
    uint64_t e = e0 + t + (uint64_t)k * kBinThreads;
    slot[2 * k] = slot[2 * k + 1] = kEmpty;
    if (e >= nedges) continue;
    uint2 uv = reinterpret_cast<const uint2*>(pairs)[e];
    if (uv.x >= n || uv.y >= n) {  // never index past the store: drop and report at flush
      atomicAdd(&counters[3], 1u);
      continue;
    }
    if (uv.x == uv.y) continue;  // a self loop toggles the same index twice: no effect
    if (!edge_selected(concat_pair(uv.x, uv.y), seed, min_depth, depth_cap)) continue;
    uint32_t ends[2] = {uv.x, uv.y};
#pragma unroll
    for (int s = 0; s < 2; s++) {
      uint32_t v = ends[s];
      uint32_t h = (v * 0x9E3779B1u) >> 20;  // 12-bit slot
      for (;;) {
        uint32_t k0 = hkey[h];
        if (k0 == v) break;
        if (k0 == kEmpty) {
          uint32_t old = atomicCAS(&hkey[h], kEmpty, v);
          if (old == kEmpty || old == v) break;
        }
        h = (h + 1) & (kBinHash - 1);
      }
      slot[2 * k + s] = h;
      lrank[2 * k + s] = atomicAdd(&hcnt[h], 1u);
    }
  }
  __syncthreads();
  for (uint32_t i = t; i < kBinHash; i += kBinThreads) {
    uint32_t key = hkey[i];
    if (key != kEmpty) hcnt[i] = atomicAdd(&deg[key], hcnt[i]);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kBinEdgesPerThread; k++) {
    uint64_t e = e0 + t + (uint64_t)k * kBinThreads;
    if (e >= nedges) continue;
    uint2 r;
    r.x = slot[2 * k] == kEmpty ? kNoRank : hcnt[slot[2 * k]] + lrank[2 * k];
    r.y = slot[2 * k + 1] == kEmpty ? kNoRank : hcnt[slot[2 * k + 1]] + lrank[2 * k + 1];
    reinterpret_cast<uint2*>(rank)[e] = r;
  }
}

// ------------------------------------------------------------------------------------------
// K2: single-pass exclusive scan of the per-vertex counts (decoupled look-back over tiles of
// 2048 vertices) producing each vertex's offset and its work items; resets the counts.
// The scanned value packs (items << 34) | entries.
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kScanThreads) void scan_items_kernel(
    uint32_t n, uint32_t* __restrict__ deg, uint32_t* __restrict__ off,
    WorkItem* __restrict__ items, uint32_t chunk, uint64_t* tile_status, uint32_t ntiles,
    uint32_t* counters) {
  __shared__ uint32_t s_tile;
  __shared__ uint64_t s_scan[kScanThreads];
  __shared__ uint64_t s_excl;
  const uint32_t t = threadIdx.x;
  if (t == 0) s_tile = atomicAdd(&counters[0], 1u);
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint64_t v0 = (uint64_t)tile * (kScanThreads * kScanPerThread) + (uint64_t)t * kScanPerThread;

  uint32_t d[kScanPerThread];
  uint64_t mine = 0;
#pragma unroll
  for (int k = 0; k < kScanPerThread; k++) {
    uint64_t v = v0 + k;
    d[k] = v < n ? deg[v] : 0;
    uint64_t it = d[k] ? (d[k] + chunk - 1) / chunk : 0;
    mine += (it << kEntryBits) | d[k];
  }
  // block-wide inclusive scan (Hillis-Steele in LDS)
  s_scan[t] = mine;
  __syncthreads();
  for (uint32_t o = 1; o < kScanThreads; o <<= 1) {
    uint64_t add = t >= o ? s_scan[t - o] : 0;
    __syncthreads();
    s_scan[t] += add;
    __syncthreads();
  }
  const uint64_t agg = s_scan[kScanThreads - 1];
  if (t == 0) {
    uint64_t excl = 0;
    if (tile == 0) {
      __hip_atomic_store(&tile_status[0], kFlagPrefix | agg, __ATOMIC_RELEASE,
                         __HIP_MEMORY_SCOPE_AGENT);
    } else {
      __hip_atomic_store(&tile_status[tile], kFlagAgg | agg, __ATOMIC_RELEASE,
                         __HIP_MEMORY_SCOPE_AGENT);
      int64_t j = (int64_t)tile - 1;
      for (;;) {
        uint64_t st;
        do {
          st = __hip_atomic_load(&tile_status[j], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        } while ((st >> 62) == 0);
        excl += st & kValueMask;
        if ((st & kFlagPrefix) != 0) break;
        j--;
      }
      __hip_atomic_store(&tile_status[tile], kFlagPrefix | (excl + agg), __ATOMIC_RELEASE,
                         __HIP_MEMORY_SCOPE_AGENT);
    }
    s_excl = excl;
    if (tile == ntiles - 1) {
      uint64_t total = excl + agg;
      counters[1] = (uint32_t)(total >> kEntryBits);
      counters[2] = (uint32_t)(total & ((1ULL << kEntryBits) - 1));
    }
  }
  __syncthreads();
  uint64_t run = s_excl + s_scan[t] - mine;
  uint64_t e_run = run & ((1ULL << kEntryBits) - 1);
  uint64_t i_run = run >> kEntryBits;
#pragma unroll
  for (int k = 0; k < kScanPerThread; k++) {
    uint64_t v = v0 + k;
    if (v >= n) break;
    off[v] = (uint32_t)e_run;
    uint32_t dv = d[k];
    if (dv) {
      deg[v] = 0;
      uint32_t nit = (dv + chunk - 1) / chunk;
      for (uint32_t q = 0; q < nit; q++) {
        WorkItem w;
        w.vertex = (uint32_t)v;
        uint32_t len = min(chunk, dv - q * chunk);
        w.len = len | (nit == 1 ? kItemExclusive : 0u);
        w.start = e_run + (uint64_t)q * chunk;
        items[i_run + q] = w;
      }
      i_run += nit;
      e_run += dv;
    }
  }
}

// ------------------------------------------------------------------------------------------
// K3: place every directed entry at off[vertex] + rank.
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void scatter_kernel(const uint32_t* __restrict__ pairs,
                                                      uint64_t nedges,
                                                      const uint32_t* __restrict__ rank,
                                                      const uint32_t* __restrict__ off,
                                                      uint32_t* __restrict__ sorted) {
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < nedges;
       e += (uint64_t)gridDim.x * blockDim.x) {
    uint2 r = reinterpret_cast<const uint2*>(rank)[e];
    if (r.x == kNoRank) continue;
    uint2 uv = reinterpret_cast<const uint2*>(pairs)[e];
    sorted[(uint64_t)off[uv.x] + r.x] = uv.y;
    sorted[(uint64_t)off[uv.y] + r.y] = uv.x;
  }
}

// ------------------------------------------------------------------------------------------
// K4: aggregation (see the file comment). Dynamic LDS per wave: delta[dl * cols] buckets +
// stage[64] x {idx, checksum, mix(swap(idx))}; shared table of per-column seed keys.
// ------------------------------------------------------------------------------------------
struct StageEntry {
  uint64_t idx, chk, mixed;
};

__global__ __launch_bounds__(kAggWaves* kWave) void aggregate_kernel(
    StoreView st, const WorkItem* __restrict__ items, uint32_t nitems_host,
    const uint32_t* __restrict__ nitems_dev, const uint32_t* __restrict__ dsts, uint32_t dl,
    int exclusive_ok) {
  extern __shared__ __align__(16) unsigned char lds[];
  const uint32_t cols = st.cols;
  const uint32_t bpc = st.bpc;
  const uint32_t dslots = dl * cols;
  uint64_t* colkey = reinterpret_cast<uint64_t*>(lds);
  unsigned char* wave_base = lds + ((cols * 8 + 15) & ~15u);
  const uint32_t wave_bytes = dslots * 16 + kWave * sizeof(StageEntry);
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint32_t wid = threadIdx.x / kWave;
  Bucket16* delta = reinterpret_cast<Bucket16*>(wave_base + wid * wave_bytes);
  StageEntry* stage = reinterpret_cast<StageEntry*>(wave_base + wid * wave_bytes + dslots * 16);

  for (uint32_t c = threadIdx.x; c < cols; c += blockDim.x) colkey[c] = xxh3_seed_key(st.seed + 5ULL * c);
  for (uint32_t s = lane; s < dslots; s += kWave) delta[s] = Bucket16{0, 0};
  __syncthreads();

  const uint32_t nitems = nitems_dev ? *nitems_dev : nitems_host;
  const uint32_t gw = blockIdx.x * kAggWaves + wid;
  const uint32_t nw = gridDim.x * kAggWaves;
  const uint64_t seed = st.seed;
  const uint32_t q64 = kWave / cols, r64 = kWave % cols;  // 64 = q*cols + r

  for (uint32_t it = gw; it < nitems; it += nw) {
    const WorkItem w = items[it];
    const uint32_t v = __builtin_amdgcn_readfirstlane(w.vertex);
    const uint32_t len = __builtin_amdgcn_readfirstlane(w.len & ~kItemExclusive);
    const bool excl = exclusive_ok && (w.len & kItemExclusive);
    const uint64_t start = w.start;
    uint64_t det_a = 0, det_g = 0;
    Bucket16* vrows = st.rows + (size_t)v * st.vstride;

    for (uint32_t j0 = 0; j0 < len; j0 += kWave) {
      const uint32_t cnt = min((uint32_t)kWave, len - j0);
      if (lane < cnt) {
        uint32_t other = dsts[start + j0 + lane];
        uint64_t idx = concat_pair(v, other);
        uint64_t chk = xxh3_len8(idx, seed);
        stage[lane] = StageEntry{idx, chk, xxh3_mix(xxh3_swap(idx))};
        det_a ^= idx;
        det_g ^= chk;
      }
      wave_sync_lds();
      const uint32_t nwork = cnt * cols;
      uint32_t j = lane / cols, c = lane - j * cols;
      for (uint32_t i = lane; i < nwork; i += kWave) {
        const StageEntry e = stage[j];
        const uint64_t h = xxh3_finish(e.mixed ^ colkey[c]);
        const uint32_t d = depth_of(h, bpc);
        if (d < dl) {
          Bucket16* b = &delta[d * cols + c];
          atomicXor(reinterpret_cast<unsigned long long*>(&b->alpha), (unsigned long long)e.idx);
          atomicXor(reinterpret_cast<unsigned long long*>(&b->gamma), (unsigned long long)e.chk);
        } else if (d < bpc) {
          Bucket16* b = &vrows[d * cols + c];
          atomic_xor_u64(&b->alpha, e.idx);
          atomic_xor_u64(&b->gamma, e.chk);
        }
        j += q64;
        c += r64;
        if (c >= cols) { c -= cols; j++; }
      }
      wave_sync_lds();
    }
    det_a = wave_xor_reduce(det_a);
    det_g = wave_xor_reduce(det_g);
    if (lane == 0 && (det_a | det_g)) {
      Bucket16* b = &st.det[v];
      if (excl) {
        Bucket16 cur = *b;
        cur.alpha ^= det_a;
        cur.gamma ^= det_g;
        *b = cur;
      } else {
        atomic_xor_u64(&b->alpha, det_a);
        atomic_xor_u64(&b->gamma, det_g);
      }
    }
    for (uint32_t s = lane; s < dslots; s += kWave) {
      Bucket16 x = delta[s];
      if (x.alpha | x.gamma) {
        Bucket16* g = &vrows[s];
        if (excl) {
          Bucket16 cur = *g;
          cur.alpha ^= x.alpha;
          cur.gamma ^= x.gamma;
          *g = cur;
        } else {
          atomic_xor_u64(&g->alpha, x.alpha);
          atomic_xor_u64(&g->gamma, x.gamma);
        }
        delta[s] = Bucket16{0, 0};
      }
    }
    wave_sync_lds();
  }
}

// ------------------------------------------------------------------------------------------
// launch helpers
// ------------------------------------------------------------------------------------------
static uint32_t lds_window(const StoreView& v) { return v.bpc < 16 ? v.bpc : 16; }

static size_t agg_lds_bytes(const StoreView& v) {
  uint32_t dl = lds_window(v);
  size_t wave_bytes = (size_t)dl * v.cols * 16 + kWave * sizeof(StageEntry);
  return ((v.cols * 8 + 15) & ~15u) + kAggWaves * wave_bytes;
}

static int agg_grid(int device) {
  static int cu = 0;
  if (cu == 0) {
    if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cu <= 0)
      cu = 256;
  }
  return cu * 8;  // persistent: 8 workgroups per CU (LDS-limited occupancy decides residency)
}

static int launch_aggregate(Stream* s, const WorkItem* d_items, uint32_t nitems_host,
                            const uint32_t* nitems_dev, const uint32_t* d_dsts, uint64_t updates) {
  Store* st = s->store;
  const StoreView& v = st->view;
  size_t lds = agg_lds_bytes(v);
  GS_CHECK(lds <= 160 * 1024, GS_ERR_UNSUPPORTED, "sketch too wide for the LDS window (%zu B)", lds);
  static std::once_flag once;
  std::call_once(once, [] {
    (void)hipFuncSetAttribute((const void*)aggregate_kernel,
                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  });
  int exclusive_ok = 1;
  {
    std::lock_guard<std::mutex> lk(st->mu);
    exclusive_ok = st->streams.size() == 1 ? 1 : 0;
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (s->profile) {
    GS_HIP(hipEventCreate(&e0));
    GS_HIP(hipEventCreate(&e1));
    GS_HIP(hipEventRecord(e0, s->stream));
  }
  hipLaunchKernelGGL(aggregate_kernel, dim3(agg_grid(st->device)), dim3(kAggWaves * kWave), lds,
                     s->stream, v, d_items, nitems_host, nitems_dev, d_dsts, lds_window(v),
                     exclusive_ok);
  GS_HIP(hipGetLastError());
  if (s->profile) {
    GS_HIP(hipEventRecord(e1, s->stream));
    s->ev_pairs.emplace_back(e0, e1);
    s->agg_updates += updates;
  }
  return GS_OK;
}

int launch_vertex_batches(Stream* s, const uint32_t*, const WorkItem* d_items, uint32_t nitems,
                          const uint32_t* d_dsts) {
  if (nitems == 0) return GS_OK;
  return launch_aggregate(s, d_items, nitems, nullptr, d_dsts, 0);
}

void free_pair_scratch(PairScratch& ws) {
  hipFree(ws.d_pairs);
  hipFree(ws.d_rank);
  hipFree(ws.d_sorted);
  hipFree(ws.d_items);
  hipFree(ws.d_deg);
  hipFree(ws.d_off);
  hipFree(ws.d_tile_status);
  hipFree(ws.d_counters);
  ws = PairScratch{};
}

static uint32_t scan_tiles(uint32_t n) {
  const uint32_t per = kScanThreads * kScanPerThread;
  return (n + per - 1) / per;
}

int ensure_pair_scratch(Stream* s, uint64_t nedges) {
  PairScratch& ws = s->ws;
  const uint32_t n = s->store->view.n;
  if (ws.cap_n < n) {
    hipFree(ws.d_deg);
    hipFree(ws.d_off);
    hipFree(ws.d_tile_status);
    hipFree(ws.d_counters);
    ws.d_deg = ws.d_off = nullptr;
    ws.d_tile_status = nullptr;
    ws.d_counters = nullptr;
    GS_HIP(hipMalloc(&ws.d_deg, sizeof(uint32_t) * n));
    GS_HIP(hipMalloc(&ws.d_off, sizeof(uint32_t) * n));
    GS_HIP(hipMalloc(&ws.d_tile_status, sizeof(uint64_t) * scan_tiles(n)));
    GS_HIP(hipMalloc(&ws.d_counters, sizeof(uint32_t) * 4));
    GS_HIP(hipMemsetAsync(ws.d_deg, 0, sizeof(uint32_t) * n, s->stream));
    GS_HIP(hipMemsetAsync(ws.d_counters, 0, sizeof(uint32_t) * 4, s->stream));
    ws.cap_n = n;
  }
  if (ws.cap_edges < nedges) {
    GS_HIP(hipStreamSynchronize(s->stream));
    hipFree(ws.d_pairs);
    hipFree(ws.d_rank);
    hipFree(ws.d_sorted);
    hipFree(ws.d_items);
    ws.d_pairs = ws.d_rank = ws.d_sorted = nullptr;
    ws.d_items = nullptr;
    uint64_t cap = nedges;
    uint64_t icap = std::min<uint64_t>((uint64_t)n, 2 * cap) + (2 * cap) / kChunk + 16;
    GS_HIP(hipMalloc(&ws.d_pairs, sizeof(uint32_t) * 2 * cap));
    GS_HIP(hipMalloc(&ws.d_rank, sizeof(uint32_t) * 2 * cap));
    GS_HIP(hipMalloc(&ws.d_sorted, sizeof(uint32_t) * 2 * cap));
    GS_HIP(hipMalloc(&ws.d_items, sizeof(WorkItem) * icap));
    ws.cap_edges = cap;
    ws.cap_items = icap;
  }
  return GS_OK;
}

int ingest_pairs_device(Stream* s, const uint32_t* d_pairs, uint64_t nedges, uint32_t min_depth,
                        uint32_t depth_cap) {
  if (nedges == 0) return GS_OK;
  GS_CHECK(nedges < (1ULL << 31), GS_ERR_INVALID, "edge batch too large (%llu)",
           (unsigned long long)nedges);
  int rc = ensure_pair_scratch(s, nedges);
  if (rc) return rc;
  PairScratch& ws = s->ws;
  const StoreView& v = s->store->view;
  const uint32_t ntiles = scan_tiles(v.n);
  const uint32_t per_block = kBinThreads * kBinEdgesPerThread;
  const uint32_t nbin = (uint32_t)((nedges + per_block - 1) / per_block);
  hipLaunchKernelGGL(bin_rank_kernel, dim3(nbin), dim3(kBinThreads), 0, s->stream, d_pairs,
                     nedges, v.n, ws.d_deg, ws.d_rank, v.seed, min_depth, depth_cap,
                     ws.d_tile_status, ntiles, ws.d_counters);
  GS_HIP(hipGetLastError());
  hipLaunchKernelGGL(scan_items_kernel, dim3(ntiles), dim3(kScanThreads), 0, s->stream, v.n,
                     ws.d_deg, ws.d_off, ws.d_items, kChunk, ws.d_tile_status, ntiles,
                     ws.d_counters);
  GS_HIP(hipGetLastError());
  const uint32_t nsc = (uint32_t)std::min<uint64_t>((nedges + 255) / 256, 65535ULL * 4);
  hipLaunchKernelGGL(scatter_kernel, dim3(nsc), dim3(256), 0, s->stream, d_pairs, nedges,
                     ws.d_rank, ws.d_off, ws.d_sorted);
  GS_HIP(hipGetLastError());
  return launch_aggregate(s, ws.d_items, 0, ws.d_counters + 1, ws.d_sorted, 2 * nedges);
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
  auto cleanup = [&] { hipFree(dk); hipFree(ds); hipFree(dout); hipFree(dmis); };
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
