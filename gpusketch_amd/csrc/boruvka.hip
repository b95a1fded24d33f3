// Borůvka emulation on the device (CCSketchAlg::boruvka_emulation, src/cc_sketch_alg.cpp:491-540,
// and the k-forest variant of src/mc_sketch_alg.cpp:738-920).
//
// Per round r:
//   r == 0 : every vertex samples its own sketch at column col0 (run_round_zero :250-271);
//   r >= 1 : acc_kernel XORs column r + the deterministic bucket of every vertex into the
//            accumulator of its supernode root (Sketch::range_merge, src/sketch.cpp:165-190),
//            sample_kernel samples each root's accumulator (Sketch::sample :110-132) and clears
//            it for the next round;
//   a GOOD sample is united into a lock-free union-find (larger root linked under the smaller,
//   so the root of every set is its smallest vertex); FAIL or a successful union marks the round
//   modified (sample_supernode :164-193); compress_kernel publishes the new supernode roots.
// The partition after every round is a function of the sketch contents only, so it equals the
// reference's; which of two equivalent sampled edges enters the spanning forest may differ.
#include <algorithm>
#include <cstring>

#include "gs_internal.h"

namespace gs {

constexpr int kQThreads = 256;
// query flags: [0] modified this round, [1] out of samples, [2] forest edges, [3] giant
// supernode of the round, [kFlagBig + k] big supernode k (see pick_big), [kFlagDone] the
// partition is final (or out of samples): every later round's kernels return at once,
// [kFlagRound] index of the round being run, [kFlagRounds] rounds the query took
constexpr uint32_t kNoGiant = 0xFFFFFFFFu;
// big supernodes per round (pick_big) and the share of kPickSamples sampled vertices one
// must hold: 2 of 256 at n = 2^17 is about 1000 members (R-MAT, plain CC query: 1.10 ms against
// 1.22 ms with 8 of at least 4 samples; 1024 samples cost more to pick than they save)
constexpr int kBig = 64;
constexpr int kPickSamples = 256;
constexpr uint32_t kMinBig = 2;
constexpr int kFlagBig = 4;
constexpr int kFlagDone = kFlagBig + kBig;
constexpr int kFlagRound = kFlagDone + 1;
constexpr int kFlagRounds = kFlagDone + 2;
constexpr int kFlagRoots = kFlagDone + 3;  // components, counted after the last round
constexpr int kFlagWords = kFlagDone + 4;
// Vertices per wave of the per-vertex kernels of the single-store query (one lane per vertex for
// the vertex attributes, the wave's vertices then taken a few at a time, one bucket per lane).
// Fewer vertices per wave put more waves in flight (64 per wave is 2 waves per SIMD at n = 2^17)
// but more of them on the same hot rows and words at once. CC query at n = 2^17 (R-MAT), per
// (round 0, accumulate, sample): 64/64/64 0.93 ms, 16/64/16 0.89, 16/32/16 0.92, 4/64/64 0.92;
// while every wave still added to one shared forest counter, 16 and fewer were up to 2.7x slower.
constexpr uint32_t kR0Verts = 16;
constexpr uint32_t kAccVerts = 64;
constexpr uint32_t kSampleVerts = 16;
static_assert(kR0Verts <= 64 && kAccVerts <= 64 && kSampleVerts <= 64, "a lane per vertex");

// blocks of kQThreads threads for waves of `per_wave` vertices over n vertices
static inline dim3 wave_grid(uint32_t n, uint32_t per_wave) {
  const uint64_t waves = ((uint64_t)n + per_wave - 1) / per_wave;
  return dim3((uint32_t)((waves * 64 + kQThreads - 1) / kQThreads));
}

__device__ __forceinline__ uint32_t load_parent(uint32_t* p, uint32_t x) {
  return __hip_atomic_load(&p[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t find_root(uint32_t* parent, uint32_t x) {
  uint32_t p = load_parent(parent, x);
  while (p != x) {
    x = p;
    p = load_parent(parent, x);
  }
  return x;
}

// per-vertex query state, meaningful at supernode roots
constexpr uint32_t kVZero = 1;    // sampled ZERO this round (no edge leaves the supernode)
constexpr uint32_t kVMerged = 2;  // absorbed another supernode this round
constexpr uint32_t kVDone = 4;    // ZERO and not merged since: the supernode can never change

// A root's state for the next round: a root that sampled ZERO and absorbed nothing has no
// leaving edge and stays as it is in every later round. settle(settle(f)) == settle(f).
__device__ __forceinline__ uint32_t settle(uint32_t f) {
  return ((f & (kVZero | kVDone)) && !(f & kVMerged)) ? kVDone : 0u;
}

// find_root where no union runs during the kernel (compress_kernel): plain loads, served by the
// caches, instead of the device-coherent loads the concurrent unions of sample_kernel need
__device__ __forceinline__ uint32_t find_root_settled(const uint32_t* parent, uint32_t x) {
  uint32_t p = parent[x];
  while (p != x) {
    x = p;
    p = parent[x];
  }
  return x;
}

__device__ __forceinline__ bool unite(uint32_t* parent, uint32_t a, uint32_t b) {
  for (;;) {
    a = find_root(parent, a);
    b = find_root(parent, b);
    if (a == b) return false;
    if (a > b) { uint32_t t = a; a = b; b = t; }
    uint32_t expected = b;
    if (__hip_atomic_compare_exchange_strong(&parent[b], &expected, a, __ATOMIC_RELAXED,
                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      return true;
  }
}

// result of sampling: 0 GOOD, 1 ZERO, 2 FAIL
// load(d) reads depth d of the sampled column
template <class Load>
__device__ __forceinline__ int sample_column(const Bucket16 det, Load load, uint32_t bpc,
                                             uint64_t seed, uint64_t* idx) {
  if (det.alpha == 0 && det.gamma == 0) return 1;
  if (det.gamma == xxh3_len8(det.alpha, seed)) { *idx = det.alpha; return 0; }
  for (uint32_t d = 0; d < bpc; d++) {
    Bucket16 b = load(d);
    if (b.gamma == xxh3_len8(b.alpha, seed)) { *idx = b.alpha; return 0; }
  }
  return 2;
}

// One sample per active lane of the wave (self: the lane's root). The words every lane of the
// grid would otherwise write -- the round's modified flag, the forest counter, the merged flag of
// a root that absorbs many others (a hub) -- are written once per wave, or only when not yet set.
__device__ __forceinline__ void handle_sample(int res, uint64_t idx, uint32_t self, uint32_t n,
                                              uint32_t* parent, uint32_t* vflags,
                                              uint32_t* flags, uint32_t* forest,
                                              uint64_t forest_cap) {
  const uint32_t a = (uint32_t)(idx >> 32), b = (uint32_t)idx;
  bool merged = false;
  if (res == 1) {
    atomicOr(&vflags[self], kVZero);
  } else if (res == 0 && a < n && b < n && a != b && unite(parent, a, b)) {
    merged = true;
    // the new root is the smaller of the two roots; find it again after the link
    const uint32_t r = find_root(parent, a);
    if (!(__hip_atomic_load(&vflags[r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & kVMerged))
      atomicOr(&vflags[r], kVMerged);
  }
  const uint32_t lane = __lane_id();
  const uint64_t active = __ballot(1);
  const uint64_t mm = __ballot(merged);
  const uint32_t leader = (uint32_t)__builtin_ctzll(active);
  const bool modified = (__ballot(res == 2) | mm) != 0;
  uint32_t base = 0;
  if (lane == leader) {
    if (modified && !__hip_atomic_load(&flags[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      __hip_atomic_store(&flags[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // (a query that returns no forest records none: no counter, the grid's one shared atomic)
    if (mm && forest_cap) base = atomicAdd(&flags[2], (uint32_t)__builtin_popcountll(mm));
  }
  base = __shfl(base, (int)leader);
  if (merged) {
    const uint32_t pos = base + (uint32_t)__builtin_popcountll(mm & ((1ULL << lane) - 1));
    if (pos < forest_cap) {
      forest[2 * (uint64_t)pos] = min(a, b);
      forest[2 * (uint64_t)pos + 1] = max(a, b);
    }
  }
}

__global__ void init_query_kernel(uint32_t n, uint32_t* parent, uint32_t* comp, uint32_t* vflags,
                                  uint32_t* flags) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (flags && t < (uint32_t)kFlagWords) flags[t] = 0;
  for (uint32_t v = t; v < n; v += gridDim.x * blockDim.x) {
    parent[v] = v;
    comp[v] = v;
    vflags[v] = 0;
  }
}

// Rounds are enqueued several at a time without looking at the host; the round ends on the
// device (the last workgroup of compress_kernel): a round that modified nothing leaves the
// partition final (the query took `round` rounds); a modified round needs another, and past the
// last sample that is out of samples (the reference throws when it samples past num_samples,
// src/sketch.cpp:110-114); otherwise the next round starts with its modified flag clear.
__device__ __forceinline__ void end_round(uint32_t* flags, uint32_t num_samples) {
  const uint32_t round = flags[kFlagRound];
  if (!flags[0]) {
    flags[kFlagDone] = 1;
    flags[kFlagRounds] = round;
  } else if (round + 1 >= num_samples) {
    flags[1] = 1;
    flags[kFlagDone] = 1;
    flags[kFlagRounds] = round + 1;
  } else {
    flags[kFlagRound] = round + 1;
    flags[0] = 0;
  }
}

__global__ void zero_acc_kernel(Bucket16* acc, uint64_t count) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < count;
       i += (uint64_t)gridDim.x * blockDim.x)
    acc[i] = Bucket16{0, 0};
}

// Sketch::sample (src/sketch.cpp:110-132) over buckets held one per lane: lanes 0..bpc-1 hold
// depths 0..bpc-1 of the sampled column (bpc = B, or fewer for a first pass over the shallow
// depths, which then answers FAIL for "none of these"), lane bpc the deterministic bucket. ZERO if that bucket is
// empty, GOOD on it if it is good, else GOOD on the first good depth, else FAIL. Every lane takes
// part; the result and *idx are wave-uniform.
__device__ __forceinline__ int wave_sample(const Bucket16 b, uint32_t lane, uint32_t bpc,
                                           uint64_t seed, uint64_t* idx) {
  const bool in = lane <= bpc;
  const uint64_t good = __ballot(in && b.gamma == xxh3_len8(b.alpha, seed));
  const uint64_t det_zero = __ballot(lane == bpc && b.alpha == 0 && b.gamma == 0);
  if (det_zero) return 1;
  int l = -1;
  if ((good >> bpc) & 1) {
    l = (int)bpc;
  } else {
    const uint64_t dm = good & ((1ULL << bpc) - 1);
    if (dm) l = __builtin_ctzll(dm);
  }
  if (l < 0) return 2;
  *idx = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(b.alpha >> 32), l) << 32) |
         (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b.alpha, l);
  return 0;
}

// Sample slots of a sketch with cps columns per sample: slot j < sw = cps*B of sample r is
// column r*cps + j % cps at depth j / cps (in the depth rows the cps columns of one depth are
// adjacent in memory), slot sw is the deterministic bucket. Offset of slot j inside a vertex's
// rows, and its position in the scan order of Sketch::sample (src/sketch.cpp:121-128: column by
// column, depths ascending).
__device__ __forceinline__ size_t slot_offset(const StoreView& st, uint32_t cps, uint32_t r,
                                              uint32_t j) {
  const uint32_t d = j / cps;
  return bucket_slot(d, r * cps + (j - d * cps), st.cols, st.bpc);
}
__device__ __forceinline__ uint32_t slot_scan_pos(uint32_t cps, uint32_t bpc, uint32_t j) {
  const uint32_t d = j / cps;
  return (j - d * cps) * bpc + d;
}

// Sketch::sample over a row of sw + 1 slots in storage order (slot sw the deterministic bucket)
// that may be wider than a wave: ZERO or GOOD on the deterministic bucket first, else GOOD on the
// good slot earliest in scan order (a wave minimum of scan positions), else FAIL. The
// deterministic bucket and 4 x 64 slots are loaded together. load(j) reads slot j; the result
// and *idx are wave-uniform.
template <class Load>
__device__ __forceinline__ int wave_sample_wide(Load load, uint32_t lane, uint32_t sw,
                                                uint32_t cps, uint32_t bpc, uint64_t seed,
                                                uint64_t* idx) {
  const Bucket16 det = load(sw);
  uint32_t best = 0xFFFFFFFFu;  // this lane's lowest scan position of a good slot
  uint64_t best_idx = 0;
  for (uint32_t j0 = 0; j0 < sw; j0 += 4 * 64) {
    Bucket16 b[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const uint32_t j = j0 + q * 64 + lane;
      b[q] = load(j < sw ? j : 0u);
    }
    if (j0 == 0) {
      if (det.alpha == 0 && det.gamma == 0) return 1;
      if (det.gamma == xxh3_len8(det.alpha, seed)) {
        *idx = det.alpha;
        return 0;
      }
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const uint32_t j = j0 + q * 64 + lane;
      if (j < sw && b[q].gamma == xxh3_len8(b[q].alpha, seed)) {
        const uint32_t p = slot_scan_pos(cps, bpc, j);
        if (p < best) {
          best = p;
          best_idx = b[q].alpha;
        }
      }
    }
  }
  uint32_t m = best;
  for (int o = 32; o > 0; o >>= 1) m = min(m, (uint32_t)__shfl_xor((int)m, o));
  if (m == 0xFFFFFFFFu) return 2;
  const int l = __builtin_ctzll(__ballot(best == m));
  *idx = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(best_idx >> 32), l) << 32) |
         (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)best_idx, l);
  return 0;
}

// Round 0, sketches with several columns per sample: every vertex samples its own sample col0,
// one vertex at a time per wave.
// rec_res (test hook, else null): every vertex's sample result (0 GOOD, 1 ZERO, 2 FAIL) and
// index are recorded there instead of being merged
__global__ __launch_bounds__(kQThreads) void round0_wide_kernel(StoreView st, uint32_t cps,
                                                                uint32_t col0, uint32_t* parent,
                                                                uint32_t* vflags, uint32_t* flags,
                                                                uint32_t* forest,
                                                                uint64_t forest_cap,
                                                                int32_t* rec_res,
                                                                uint64_t* rec_idx) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t v0 = (blockIdx.x * blockDim.x + threadIdx.x) / 64 * kR0Verts;
  if (v0 >= st.n) return;
  const uint32_t cnt = min(kR0Verts, st.n - v0);
  const uint32_t sw = cps * st.bpc;
  int my_res = -1;
  uint64_t my_idx = 0;
  for (uint32_t i = 0; i < cnt; i++) {
    const uint32_t vi = v0 + i;
    const Bucket16* vrows = st.rows + (size_t)vi * st.vstride;
    auto load = [&](uint32_t j) -> Bucket16 {
      return j < sw ? vrows[slot_offset(st, cps, col0, j)] : st.det[vi];
    };
    uint64_t idx = 0;
    const int res = wave_sample_wide(load, lane, sw, cps, st.bpc, st.seed, &idx);
    if (lane == i) {
      my_res = res;
      my_idx = idx;
    }
  }
  if (my_res >= 0) {
    if (rec_res) {
      rec_res[v0 + lane] = my_res;
      rec_idx[v0 + lane] = my_idx;
    } else {
      handle_sample(my_res, my_idx, v0 + lane, st.n, parent, vflags, flags, forest, forest_cap);
    }
  }
}

// Round 0: every vertex samples its own column col0 (run_round_zero, src/cc_sketch_alg.cpp:
// 250-271). A wave takes kR0Verts consecutive vertices, four at a time, one bucket per lane. Each
// bucket of the column's depth rows is a 128-byte line of its own, so the first pass reads the
// deterministic bucket and depths 0..15 only; a vertex with no good bucket among them reads its
// deeper depths in a second pass (the scan order stays the reference's).
constexpr uint32_t kRound0Depths = 16;
// (rec_res, rec_idx: the test hook of round0_wide_kernel)
__global__ __launch_bounds__(kQThreads) void round0_kernel(StoreView st, uint32_t col0,
                                                           uint32_t* parent, uint32_t* vflags,
                                                           uint32_t* flags, uint32_t* forest,
                                                           uint64_t forest_cap, int32_t* rec_res,
                                                           uint64_t* rec_idx) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t v0 = (blockIdx.x * blockDim.x + threadIdx.x) / 64 * kR0Verts;
  if (v0 >= st.n) return;
  const uint32_t cnt = min(kR0Verts, st.n - v0);
  const uint32_t nd = min(st.bpc, kRound0Depths);  // depths of the first pass, det in lane nd
  int my_res = -1;  // the sample of vertex v0 + lane, handled by its own lane at the end
  uint64_t my_idx = 0;
  for (uint32_t i0 = 0; i0 < cnt; i0 += 4) {
    Bucket16 bv[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const uint32_t vi = v0 + min(i0 + q, cnt - 1);
      bv[q] = lane > nd    ? Bucket16{0, 0}
              : lane < nd ? st.rows[(size_t)vi * st.vstride + bucket_slot(lane, col0, st.cols, st.bpc)]
                          : st.det[vi];
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
      if (i0 + q >= cnt) break;
      uint64_t idx = 0;
      int res = wave_sample(bv[q], lane, nd, st.seed, &idx);
      if (res == 2 && nd < st.bpc) {  // wave-uniform: depths nd..B-1, one per lane
        const uint32_t dd = nd + lane;
        const Bucket16 b2 = dd < st.bpc
            ? st.rows[(size_t)(v0 + i0 + q) * st.vstride + bucket_slot(dd, col0, st.cols, st.bpc)]
            : Bucket16{0, 0};
        const uint64_t good = __ballot(dd < st.bpc && b2.gamma == xxh3_len8(b2.alpha, st.seed));
        if (good) {
          const int l = __builtin_ctzll(good);
          idx = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(b2.alpha >> 32), l) << 32) |
                (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b2.alpha, l);
          res = 0;
        }
      }
      if (lane == i0 + q) {
        my_res = res;
        my_idx = idx;
      }
    }
  }
  if (my_res >= 0) {
    if (rec_res) {
      rec_res[v0 + lane] = my_res;
      rec_idx[v0 + lane] = my_idx;
    } else {
      handle_sample(my_res, my_idx, v0 + lane, st.n, parent, vflags, flags, forest, forest_cap);
    }
  }
}

// Big supernodes of a round, picked from kPickSamples sampled vertices (pick_big): up to kBig live
// supernodes holding at least kMinBig of the samples each (flags[4 + k], kNoGiant when unused),
// largest first. Their column sums are formed from per-block partial sums (acc_kernel sums them in
// LDS) instead of atomics into one accumulator row (which thousands of waves would contend for:
// with 16 such supernodes, the medium supernodes of rounds 2-3 of an R-MAT query took most of
// their accumulation time in contended atomics). When the store's
// updates were all symmetric (edge pairs: every update XORs the same bucket of both endpoints)
// the XOR of a column over all vertices is zero, so any supernode's sum is the XOR over all
// vertices outside it: the supernode holding the most samples is the giant (flags[3]) and its
// members are not even read, when it holds at least kMinGiant samples and more than the finished
// supernodes do (those are read for the complement and are skipped otherwise).
constexpr uint32_t kMinGiant = 8;

// pick_big: one workgroup of kPickSamples threads, after the round's compress
__device__ void pick_big(uint32_t n, const uint32_t* parent, const uint32_t* vflags, int symmetric,
                         uint32_t* flags) {
  __shared__ uint32_t s_r[kPickSamples];
  __shared__ uint32_t s_cand[kPickSamples];
  __shared__ uint32_t s_ncand, s_max, s_done;
  const uint32_t t = threadIdx.x;
  const uint32_t v = (uint32_t)(((uint64_t)t * n) / kPickSamples + n / (2 * kPickSamples));
  uint32_t r = find_root_settled(parent, v);
  if (settle(vflags[r]) & kVDone) r = kNoGiant;
  s_r[t] = r;
  if (t == 0) {
    s_ncand = 0;
    s_max = 0;
    s_done = 0;
  }
  __syncthreads();
  // samples in this supernode, counted at its first sample only; a candidate key orders by
  // count, then by sample position (unique)
  uint32_t cnt = 0, first = t;
  for (uint32_t i = 0; i < kPickSamples; i++) {
    const bool same = s_r[i] == r;
    cnt += same;
    if (same && i < first) first = i;
  }
  const uint32_t key = (r != kNoGiant && first == t && cnt >= kMinBig) ? (cnt << 16) | t : 0u;
  if (key) {
    s_cand[atomicAdd(&s_ncand, 1u)] = key;  // at most kPickSamples / kMinBig
    atomicMax(&s_max, key);
  }
  if (r == kNoGiant) atomicAdd(&s_done, 1u);
  __syncthreads();
  // the giant: the candidate of most samples (over all candidates)
  const uint32_t gkey = s_max;
  const bool giant = symmetric && gkey && (gkey >> 16) >= kMinGiant && (gkey >> 16) > s_done;
  const uint32_t groot = giant ? s_r[gkey & 0xFFFFu] : kNoGiant;
  if (t == 0) flags[3] = groot;
  if (t < (uint32_t)kBig) flags[kFlagBig + t] = kNoGiant;
  __syncthreads();
  // the big supernodes: the other candidates of most samples, each placed at its rank (keys are
  // unique: samples, then first sample position)
  const uint32_t nc = s_ncand;
  if (t < nc) {
    const uint32_t x = s_cand[t];
    if (!(giant && x == gkey)) {
      uint32_t rank = 0;
      for (uint32_t u = 0; u < nc; u++) {
        const uint32_t y = s_cand[u];
        rank += y > x && !(giant && y == gkey);
      }
      if (rank < (uint32_t)kBig) flags[kFlagBig + rank] = s_r[x & 0xFFFFu];
    }
  }
}

// Rounds >= 1: the sample sums (sample r: its cps columns, depths 0..B-1, plus the deterministic
// bucket; w = cps*B + 1 slots) of the live supernodes. A wave takes 64 consecutive vertices and
// walks the ones it must read, four at a time; lane l holds slot 64*blockIdx.y + l (so one
// vertex is one load instruction, and its row of atomics is one instruction over consecutive
// addresses). Big supernodes accumulate into LDS rows of the block (LDS atomics) and leave
// per-block partials bpart[block][k]; with a giant, every vertex outside it (finished supernodes
// included) also adds into cpart[block] (registers, the block's waves summed through LDS); the
// other supernodes take atomics.
__global__ __launch_bounds__(kQThreads) void acc_kernel(StoreView st, uint32_t cps,
                                                        uint32_t* __restrict__ parent,
                                                        const uint32_t* __restrict__ comp,
                                                        const uint32_t* __restrict__ vflags,
                                                        const uint32_t* __restrict__ flags,
                                                        uint64_t* __restrict__ acc, uint64_t plane,
                                                        Bucket16* __restrict__ cpart,
                                                        Bucket16* __restrict__ bpart) {
  if (flags[kFlagDone]) return;
  const uint32_t col = flags[kFlagRound];  // round r >= 1 samples sample r
  const uint32_t sw = cps * st.bpc, w = sw + 1;
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t slot = blockIdx.y * 64 + lane;
  // the previous round's unions, compressed: every vertex links straight to its root again
  // (nothing reads `parent` before this round's sample_kernel)
  if (blockIdx.y == 0) {
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u < st.n) parent[u] = comp[u];
  }
  const uint32_t gwave = (blockIdx.x * blockDim.x + threadIdx.x) / 64;
  const uint32_t v0 = gwave * kAccVerts;
  const uint32_t giant = flags[3];
  // the big supernodes' sums of this block: LDS rows of rw slots, one per big supernode
  // (alpha plane then gamma plane, kBig x rw words each: a row of LDS atomics is consecutive
  // 8-byte words, no bank conflicts)
  extern __shared__ uint64_t s_big[];
  __shared__ uint32_t s_bigr[kBig];
  const uint32_t rw = min(64u, w);
  for (uint32_t i = threadIdx.x; i < 2 * kBig * rw; i += blockDim.x) s_big[i] = 0;
  if (threadIdx.x < (uint32_t)kBig) s_bigr[threadIdx.x] = flags[kFlagBig + threadIdx.x];
  __syncthreads();
  const bool split = giant != kNoGiant;
  const uint32_t v = v0 + lane;
  const bool inr = lane < kAccVerts && v < st.n;
  const uint32_t r = inr ? comp[v] : 0;
  const bool done = inr && (vflags[r] & kVDone);
  const bool rd = inr && (split ? r != giant : !done);  // this vertex's column is read
  const bool act = rd && !done;                         // and belongs to a live supernode
  int kb = -1;
  for (int k = 0; k < kBig; k++)
    kb = (kb < 0 && r == s_bigr[k]) ? k : kb;
  const uint64_t rdm = __ballot(rd), actm = __ballot(act);
  const bool lj = slot < w;
  const uint32_t jj = lj ? slot : 0;
  const bool isdet = jj == sw;
  const size_t soff = isdet ? 0 : slot_offset(st, cps, col, jj);
  uint64_t ca = 0, cg = 0;
  uint64_t m = rdm;
  while (m) {
    uint32_t ii[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      ii[q] = m ? (uint32_t)__builtin_ctzll(m) : ii[0];
      m &= m ? m - 1 : 0;
    }
    Bucket16 bv[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const uint32_t vi = v0 + ii[q];
      bv[q] = isdet ? st.det[vi] : st.rows[(size_t)vi * st.vstride + soff];
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
      if (q > 0 && ii[q] == ii[0]) break;  // padding of the last group
      const uint32_t i = ii[q];
      const uint64_t xa = lj ? bv[q].alpha : 0, xg = lj ? bv[q].gamma : 0;
      if (split) {
        ca ^= xa;
        cg ^= xg;
      }
      if (!((actm >> i) & 1)) continue;
      const int kbi = __builtin_amdgcn_readlane(kb, i);
      if (kbi >= 0) {
        if (lj && (xa | xg)) {
          uint64_t* dst = &s_big[(uint32_t)kbi * rw + lane];
          __hip_atomic_fetch_xor(dst, xa, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          __hip_atomic_fetch_xor(dst + kBig * rw, xg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
      } else if (lj && (xa | xg)) {
        uint64_t* dst = acc + (size_t)__builtin_amdgcn_readlane((int)r, i) * w + jj;
        __hip_atomic_fetch_xor(dst, xa, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_fetch_xor(dst + plane, xg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
  }
  // the block's partial sums (one per slot): waves 1.. hand theirs to wave 0 through LDS
  __shared__ Bucket16 s_red[kQThreads / 64 - 1][64];
  const uint32_t wv = threadIdx.x / 64;
  auto block_sum = [&](uint64_t& a, uint64_t& g) {
    if (wv > 0) s_red[wv - 1][lane] = Bucket16{a, g};
    __syncthreads();
    if (wv == 0)
      for (int q = 0; q < kQThreads / 64 - 1; q++) {
        a ^= s_red[q][lane].alpha;
        g ^= s_red[q][lane].gamma;
      }
    __syncthreads();
  };
  if (split) {
    block_sum(ca, cg);
    if (wv == 0 && lj) cpart[(size_t)blockIdx.x * w + slot] = Bucket16{ca, cg};
  }
  __syncthreads();  // every wave's LDS sums are in
  for (uint32_t i = threadIdx.x; i < kBig * rw; i += blockDim.x) {
    const uint32_t k = i / rw, sl = blockIdx.y * 64 + (i - k * rw);
    if (s_bigr[k] != kNoGiant && sl < w)
      bpart[((size_t)blockIdx.x * kBig + k) * w + sl] = Bucket16{s_big[i], s_big[kBig * rw + i]};
  }
}

// accumulators of the big supernodes (blockIdx.y = k < kBig) and of the giant (blockIdx.y =
// kBig), bucket blockIdx.x: XOR of the block partials of acc_kernel (for the giant: the XOR over
// all vertices being zero, the partials of everything outside it)
__global__ __launch_bounds__(kQThreads) void big_acc_kernel(const uint32_t* __restrict__ flags,
                                                            const Bucket16* __restrict__ cpart,
                                                            const Bucket16* __restrict__ bpart,
                                                            uint32_t nparts, uint32_t w,
                                                            uint64_t* __restrict__ acc,
                                                            uint64_t plane) {
  const uint32_t k = blockIdx.y;
  const uint32_t root = k < (uint32_t)kBig ? flags[kFlagBig + k] : flags[3];
  if (flags[kFlagDone] || root == kNoGiant) return;
  const uint32_t j = blockIdx.x;
  uint64_t a = 0, g = 0;
  for (uint32_t i = threadIdx.x; i < nparts; i += kQThreads) {
    const Bucket16 b = k < (uint32_t)kBig ? bpart[((size_t)i * kBig + k) * w + j] : cpart[(size_t)i * w + j];
    a ^= b.alpha;
    g ^= b.gamma;
  }
  a = wave_xor_reduce(a);
  g = wave_xor_reduce(g);
  __shared__ uint64_t sa[kQThreads / 64], sg[kQThreads / 64];
  if ((threadIdx.x & 63) == 0) {
    sa[threadIdx.x / 64] = a;
    sg[threadIdx.x / 64] = g;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < kQThreads / 64; q++) {
      a ^= sa[q];
      g ^= sg[q];
    }
    acc[(size_t)root * w + j] = a;
    acc[plane + (size_t)root * w + j] = g;
  }
}

// Every live root samples its accumulator (bucket bpc = the deterministic bucket) and clears it
// for the next round. A wave takes the roots among kSampleVerts consecutive vertices, four at a
// time, one bucket per lane.
__global__ __launch_bounds__(kQThreads) void sample_kernel(StoreView st,
                                                           const uint32_t* __restrict__ comp,
                                                           uint64_t* __restrict__ acc,
                                                           uint64_t plane,
                                                           uint32_t* parent, uint32_t* vflags,
                                                           uint32_t* flags, uint32_t* forest,
                                                           uint64_t forest_cap) {
  if (flags[kFlagDone]) return;
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t v0 = (blockIdx.x * blockDim.x + threadIdx.x) / 64 * kSampleVerts;
  const uint32_t v = v0 + lane;
  const uint32_t w = st.bpc + 1;
  uint64_t m = __ballot(lane < kSampleVerts && v < st.n && comp[v] == v && !(vflags[v] & kVDone));
  int my_res = -1;  // the sample of root v0 + lane, handled by its own lane at the end
  uint64_t my_idx = 0;
  while (m) {
    uint32_t ii[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      ii[q] = m ? (uint32_t)__builtin_ctzll(m) : ii[0];
      m &= m ? m - 1 : 0;
    }
    Bucket16 bv[4];
#pragma unroll
    for (int q = 0; q < 4; q++)
      bv[q] = lane < w ? Bucket16{acc[(size_t)(v0 + ii[q]) * w + lane],
                                  acc[plane + (size_t)(v0 + ii[q]) * w + lane]}
                       : Bucket16{0, 0};
#pragma unroll
    for (int q = 0; q < 4; q++) {
      if (q > 0 && ii[q] == ii[0]) break;  // padding of the last group
      const uint32_t root = v0 + ii[q];
      uint64_t idx = 0;
      const int res = wave_sample(bv[q], lane, st.bpc, st.seed, &idx);
      if (lane < w) {
        acc[(size_t)root * w + lane] = 0;
        acc[plane + (size_t)root * w + lane] = 0;
      }
      if (lane == ii[q]) {
        my_res = res;
        my_idx = idx;
      }
    }
  }
  if (my_res >= 0) handle_sample(my_res, my_idx, v, st.n, parent, vflags, flags, forest, forest_cap);
}

// sample_kernel for rows of w = cps*B + 1 slots: one root at a time per wave
__global__ __launch_bounds__(kQThreads) void sample_wide_kernel(StoreView st, uint32_t cps,
                                                                const uint32_t* __restrict__ comp,
                                                                uint64_t* __restrict__ acc,
                                                                uint64_t plane,
                                                                uint32_t* parent, uint32_t* vflags,
                                                                uint32_t* flags, uint32_t* forest,
                                                                uint64_t forest_cap) {
  if (flags[kFlagDone]) return;
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t v0 = (blockIdx.x * blockDim.x + threadIdx.x) / 64 * kSampleVerts;
  const uint32_t v = v0 + lane;
  const uint32_t sw = cps * st.bpc, w = sw + 1;
  uint64_t m = __ballot(lane < kSampleVerts && v < st.n && comp[v] == v && !(vflags[v] & kVDone));
  int my_res = -1;
  uint64_t my_idx = 0;
  while (m) {
    const uint32_t i = (uint32_t)__builtin_ctzll(m);
    m &= m - 1;
    uint64_t* row = acc + (size_t)(v0 + i) * w;
    auto load = [&](uint32_t j) -> Bucket16 { return Bucket16{row[j], row[plane + j]}; };
    uint64_t idx = 0;
    const int res = wave_sample_wide(load, lane, sw, cps, st.bpc, st.seed, &idx);
    for (uint32_t j = lane; j < w; j += 64) row[j] = row[plane + j] = 0;
    if (lane == i) {
      my_res = res;
      my_idx = idx;
    }
  }
  if (my_res >= 0) handle_sample(my_res, my_idx, v, st.n, parent, vflags, flags, forest, forest_cap);
}

// Publishes the roots after a round's unions and settles the root states (settle).
// Workgroup 0 also ends the round (end_round) and, when another round follows, picks its big
// supernodes (pick_big): one launch instead of three per round. It reads nothing the other
// workgroups write but root states, and settle() maps a root state before and after its
// workgroup settled it to the same answer. Other workgroups may see the done flag it sets and
// return: the done flag follows a round without unions (compress would change no root) or runs
// out of samples (the partition is discarded).
__global__ __launch_bounds__(kQThreads) void compress_kernel(uint32_t n, uint32_t* parent,
                                                             uint32_t* comp, uint32_t* vflags,
                                                             uint32_t* flags, uint32_t num_samples,
                                                             int symmetric) {
  static_assert(kQThreads == kPickSamples, "pick_big runs on one compress workgroup");
  if (flags[kFlagDone]) return;
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) {
    const uint32_t r = find_root_settled(parent, v);
    comp[v] = r;
    if (r == v) vflags[v] = settle(vflags[v]);
  }
  if (blockIdx.x != 0) return;
  __shared__ uint32_t s_done;
  __syncthreads();  // every lane of the workgroup has read the done flag
  if (threadIdx.x == 0) {
    end_round(flags, num_samples);
    s_done = flags[kFlagDone];
  }
  __syncthreads();
  if (!s_done) pick_big(n, parent, vflags, symmetric, flags);
}

// components of the final partition: roots (comp[v] == v) counted per wave, one atomic each
// (enqueued after every batch of rounds: counts only once the partition is final, which happens
// in exactly one of them)
__global__ __launch_bounds__(kQThreads) void count_roots_kernel(uint32_t n,
                                                                const uint32_t* __restrict__ comp,
                                                                const uint32_t* flags,
                                                                uint32_t* count) {
  if (!flags[kFlagDone]) return;
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t m = __ballot(v < n && comp[v] == v);
  if ((threadIdx.x & 63) == 0 && m) atomicAdd(count, (uint32_t)__builtin_popcountll(m));
}

__global__ __launch_bounds__(kQThreads) void finalize_parent_kernel(uint32_t n, uint32_t* parent,
                                                                    const uint32_t* comp,
                                                                    const uint32_t* __restrict__ flags) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n && !(flags && flags[kFlagDone])) parent[v] = comp[v];
}

void free_query_scratch(Store* st) {
  (void)hipFree(st->d_parent);
  (void)hipFree(st->d_comp);
  (void)hipFree(st->d_vflags);
  (void)hipFree(st->d_acc);
  (void)hipFree(st->d_cpart);
  (void)hipFree(st->d_bpart);
  st->d_cpart = st->d_bpart = nullptr;
  (void)hipFree(st->d_flags);
  (void)hipFree(st->d_forest);
  if (st->h_flags) (void)hipHostFree(st->h_flags);
  if (st->h_labels) (void)hipHostFree(st->h_labels);
  st->h_labels = nullptr;
  st->d_parent = st->d_comp = st->d_vflags = st->d_flags = st->d_forest = st->h_flags = nullptr;
  st->d_acc = nullptr;
  st->forest_cap = 0;
  if (st->qs) {
    Stream* s = st->qs;
    destroy_stream_resources(*s);  // its HIP stream is the store's query stream, kept
    delete s;
    st->qs = nullptr;
  }
}

static int ensure_query_scratch(Store* st, uint64_t forest_pairs) {
  const uint32_t n = st->view.n;
  if (!st->d_parent) {
    GS_HIP(hipMalloc(&st->d_parent, sizeof(uint32_t) * n));
    GS_HIP(hipMalloc(&st->d_comp, sizeof(uint32_t) * n));
    GS_HIP(hipMalloc(&st->d_vflags, sizeof(uint32_t) * n));
    const uint64_t w = (uint64_t)st->cols_per_sample * st->view.bpc + 1;  // slots per sample
    GS_HIP(hipMalloc(&st->d_acc, sizeof(Bucket16) * (uint64_t)n * w));
    const uint64_t nparts = wave_grid(n, kAccVerts).x;  // acc_kernel blocks
    GS_HIP(hipMalloc(&st->d_cpart, sizeof(Bucket16) * w * nparts));
    GS_HIP(hipMalloc(&st->d_bpart, sizeof(Bucket16) * w * nparts * kBig));
    GS_HIP(hipMalloc(&st->d_flags, sizeof(uint32_t) * kFlagWords));
    GS_HIP(hipHostMalloc((void**)&st->h_flags, sizeof(uint32_t) * kFlagWords, hipHostMallocDefault));
    GS_HIP(hipHostMalloc((void**)&st->h_labels, sizeof(uint32_t) * n, hipHostMallocDefault));
    uint64_t cnt = (uint64_t)n * w;
    hipLaunchKernelGGL(zero_acc_kernel, dim3(2048), dim3(256), 0, st->qstream, st->d_acc, cnt);
    GS_HIP(hipGetLastError());
  }
  if (st->forest_cap < forest_pairs) {
    GS_HIP(hipStreamSynchronize(st->qstream));
    (void)hipFree(st->d_forest);
    st->d_forest = nullptr;
    GS_HIP(hipMalloc(&st->d_forest, sizeof(uint32_t) * 2 * forest_pairs));
    st->forest_cap = forest_pairs;
  }
  if (!st->qs) {
    Stream* s = new Stream();
    s->store = st;
    s->stream = st->qstream;  // queries and their forest toggles share one in-order stream
    s->own_stream = false;
    int rc = init_stream_objects(s);
    if (rc) {
      destroy_stream_resources(*s);
      delete s;
      return rc;
    }
    st->qs = s;
  }
  return GS_OK;
}

// One Borůvka pass. Forest edges are appended at d_forest[2*forest_base...]; returns the pass's
// forest size in *forest_len.
// Rounds are enqueued speculatively and the host reads the flags once per batch of rounds: a
// round after the final one only launches kernels that return at once (about 8 us of device
// time), a host round trip between batches costs several times that. The first batch covers
// the rounds the store's previous pass took plus one (at least kRoundsAhead), later batches
// kRoundsAhead rounds each.
constexpr uint32_t kRoundsAhead = 4;
// tail (optional): enqueued after every batch of rounds, before the host waits for the flags --
// the root count and a copy of the labels into the store's pinned staging, valid after the
// batch that ends the pass
struct PassTail {
  bool count_roots = false;
  bool labels = false;
};
static int boruvka_pass(Store* st, uint32_t col0, uint64_t forest_base, uint32_t* rounds_out,
                        uint64_t* forest_len, bool record_forest = true, PassTail tail = {}) {
  const StoreView& v = st->view;
  const uint32_t n = v.n;
  const dim3 grid((n + kQThreads - 1) / kQThreads), block(kQThreads);
  const uint32_t cps = st->cols_per_sample;
  const uint32_t w = cps * v.bpc + 1;
  const bool wide = cps > 1;  // several columns per sample: the one-vertex-per-wave samplers
  const dim3 g_r0 = wave_grid(n, kR0Verts), g_acc = wave_grid(n, kAccVerts),
             g_smp = wave_grid(n, kSampleVerts);
  const uint32_t nparts = g_acc.x;
  hipStream_t q = st->qstream;
  uint32_t* forest = record_forest ? st->d_forest + 2 * forest_base : nullptr;
  uint64_t cap = record_forest ? st->forest_cap - forest_base : 0;
  if (col0 >= st->num_samples) return fail(GS_ERR_OUT_OF_SAMPLES, "sketch out of samples");
  hipLaunchKernelGGL(init_query_kernel, grid, block, 0, q, n, st->d_parent, st->d_comp,
                     st->d_vflags, st->d_flags);
  // round 0
  if (wide)
    hipLaunchKernelGGL(round0_wide_kernel, g_r0, block, 0, q, v, cps, col0, st->d_parent,
                       st->d_vflags, st->d_flags, forest, cap, nullptr, nullptr);
  else
    hipLaunchKernelGGL(round0_kernel, g_r0, block, 0, q, v, col0 * cps, st->d_parent,
                       st->d_vflags, st->d_flags, forest, cap, nullptr, nullptr);
  const int symmetric = st->symmetric.load() ? 1 : 0;
  hipLaunchKernelGGL(compress_kernel, grid, block, 0, q, n, st->d_parent, st->d_comp,
                     st->d_vflags, st->d_flags, st->num_samples, symmetric);
  GS_HIP(hipGetLastError());
  // the per-root accumulators as two planes of n x w words (alpha, then gamma): a row of atomics
  // (or loads) over a root's w slots is w consecutive 8-byte words
  uint64_t* const acc = reinterpret_cast<uint64_t*>(st->d_acc);
  const uint64_t plane = (uint64_t)n * w;
  uint32_t batch = std::max(kRoundsAhead, std::min(st->last_rounds + 1, st->num_samples));
  for (uint32_t enqueued = 1;; enqueued += batch, batch = kRoundsAhead) {
    for (uint32_t k = 0; k < batch; k++) {
      hipLaunchKernelGGL(acc_kernel, dim3(g_acc.x, (w + 63) / 64), block,
                         sizeof(Bucket16) * kBig * std::min(64u, w), q, v, cps,
                         st->d_parent, st->d_comp, st->d_vflags, st->d_flags, acc, plane,
                         st->d_cpart, st->d_bpart);
      hipLaunchKernelGGL(big_acc_kernel, dim3(w, kBig + 1), dim3(kQThreads), 0, q, st->d_flags,
                         st->d_cpart, st->d_bpart, nparts, w, acc, plane);
      if (wide)
        hipLaunchKernelGGL(sample_wide_kernel, g_smp, block, 0, q, v, cps, st->d_comp, acc, plane,
                           st->d_parent, st->d_vflags, st->d_flags, forest, cap);
      else
        hipLaunchKernelGGL(sample_kernel, g_smp, block, 0, q, v, st->d_comp, acc, plane,
                           st->d_parent, st->d_vflags, st->d_flags, forest, cap);
      hipLaunchKernelGGL(compress_kernel, grid, block, 0, q, n, st->d_parent, st->d_comp,
                         st->d_vflags, st->d_flags, st->num_samples, symmetric);
    }
    if (tail.count_roots)  // flags[kFlagRoots] is zero until the final batch's count
      hipLaunchKernelGGL(count_roots_kernel, grid, block, 0, q, n, st->d_comp, st->d_flags,
                         st->d_flags + kFlagRoots);
    GS_HIP(hipGetLastError());
    GS_HIP(hipMemcpyAsync(st->h_flags, st->d_flags, sizeof(uint32_t) * kFlagWords,
                          hipMemcpyDeviceToHost, q));
    if (tail.labels)
      GS_HIP(hipMemcpyAsync(st->h_labels, st->d_comp, sizeof(uint32_t) * n,
                            hipMemcpyDeviceToHost, q));
    GS_HIP(hipStreamSynchronize(q));
    if (st->h_flags[kFlagDone]) break;
    GS_CHECK(enqueued <= st->num_samples + kRoundsAhead, GS_ERR_INVALID,
             "Boruvka rounds did not settle");
  }
  const uint32_t round = st->h_flags[kFlagRounds];
  st->last_rounds = round;
  if (st->h_flags[1])
    return fail(GS_ERR_OUT_OF_SAMPLES,
                "This sketch (seed=%llu, max samples=%u) cannot be sampled more times (cur idx=%u)!",
                (unsigned long long)v.seed, st->num_samples, round);
  if (rounds_out) *rounds_out = round;
  uint64_t fl = st->h_flags[2];
  GS_CHECK(fl <= cap, GS_ERR_INVALID, "forest capacity exceeded");
  *forest_len = fl;
  return GS_OK;
}

static int sync_all_streams(Store* st) { return sync_store(st); }

int boruvka_query(Store* st, const uint32_t*, uint32_t* label_out, uint32_t* num_cc,
                  uint32_t* rounds, uint32_t* forest_out, uint64_t forest_cap,
                  uint64_t* forest_len) {
  int rc = sync_all_streams(st);
  if (rc) return rc;
  rc = ensure_query_scratch(st, st->view.n);
  if (rc) return rc;
  uint64_t fl = 0;
  const bool record = forest_out && forest_cap;
  const bool count = num_cc || (forest_len && !record);
  PassTail tail;
  tail.count_roots = count;
  tail.labels = label_out != nullptr;  // through the store's pinned staging, one DMA
  rc = boruvka_pass(st, 0, 0, rounds, &fl, record, tail);
  if (rc) return rc;
  const uint32_t n = st->view.n;
  if (label_out) std::memcpy(label_out, st->h_labels, sizeof(uint32_t) * n);
  if (count) {
    const uint32_t c = st->h_flags[kFlagRoots];
    if (num_cc) *num_cc = c;
    if (!record) fl = n - c;  // a spanning forest has one edge per union: n - components
  }
  if (forest_len) *forest_len = fl;
  if (forest_out && forest_cap) {
    uint64_t m = std::min(fl, forest_cap);
    GS_HIP(hipMemcpy(forest_out, st->d_forest, sizeof(uint32_t) * 2 * m, hipMemcpyDeviceToHost));
  }
  return GS_OK;
}

int k_forests(Store* st, uint32_t k, uint32_t* edges_out, uint64_t edge_cap, uint64_t* counts) {
  int rc = sync_all_streams(st);
  if (rc) return rc;
  const uint32_t n = st->view.n;
  rc = ensure_query_scratch(st, (uint64_t)k * n);
  if (rc) return rc;
  uint64_t total = 0;
  int err = GS_OK;
  for (uint32_t i = 0; i < k; i++) {
    uint64_t fl = 0;
    err = boruvka_pass(st, i, total, nullptr, &fl);
    if (err) break;
    counts[i] = fl;
    // toggle this forest out of both endpoints' sketches (filter_sf_edges, src/mc_sketch_alg.cpp:837-888)
    err = ingest_pairs_device(st->qs, st->d_forest + 2 * total, fl, 0, 0);
    if (err) break;
    total += fl;
  }
  // toggle every forest found back in, also after a failure, so the sketches are unchanged
  if (total) {
    int rc2 = ingest_pairs_device(st->qs, st->d_forest, total, 0, 0);
    if (!err) err = rc2;
  }
  GS_HIP(hipStreamSynchronize(st->qstream));
  if (err) return err;
  GS_CHECK(total <= edge_cap, GS_ERR_INVALID, "edge_cap %llu too small for %llu forest edges",
           (unsigned long long)edge_cap, (unsigned long long)total);
  if (total) GS_HIP(hipMemcpy(edges_out, st->d_forest, sizeof(uint32_t) * 2 * total, hipMemcpyDeviceToHost));
  return GS_OK;
}

// ------------------------------------------------------------------------------------------
// Batched Boruvka over the members of a group (virtual vertex space, member of global vertex V
// in vgraph[V]). Every member runs its own rounds: a member whose round modified nothing is done
// (its partition is final), one that still changes after its last column is out of samples.
// Rounds >= 1 are one captured HIP graph: the round number lives in device memory, so the same
// graph replays round after round and the host only reads back whether any member is active.
// gstate: [0,G) modified this round, [G,2G) status (0 active, 1 done, 2 out of samples),
// [2G,3G) rounds used, [3G] round counter, [3G+1] any member active after this round, [3G+2] the
// number M of vertices of the queried members. The per-vertex kernels walk only those: vmap
// lists their global ids member after member, and a fixed grid strides over [0, M) (M read from
// device memory), so one captured round serves any selection of members and a query of a few
// small members costs about what their own single-store queries would.
// ------------------------------------------------------------------------------------------
constexpr uint32_t kSelGrid = 2048;  // workgroups of the per-vertex kernels of a group query
__device__ __forceinline__ uint32_t sel_count(const uint32_t* gstate, uint32_t G) {
  return gstate[3 * G + 2];
}
constexpr uint32_t kGActive = 0, kGDone = 1, kGOutOfSamples = 2;

__device__ __forceinline__ void handle_group_sample(int res, uint64_t idx, uint32_t self,
                                                    const GraphDesc& d, uint32_t* parent,
                                                    uint32_t* vflags, uint32_t* gmod) {
  if (res == 2) {
    *gmod = 1;
  } else if (res == 1) {
    atomicOr(&vflags[self], kVZero);
  } else {
    const uint32_t a = (uint32_t)(idx >> 32), b = (uint32_t)idx;
    if (a < d.n && b < d.n && a != b && unite(parent, a + d.base, b + d.base)) {
      const uint32_t r = find_root(parent, a + d.base);
      if (!(__hip_atomic_load(&vflags[r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & kVMerged))
        atomicOr(&vflags[r], kVMerged);
      if (!__hip_atomic_load(gmod, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        __hip_atomic_store(gmod, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

__global__ void g_begin_kernel(uint32_t* gstate, uint32_t G) {
  if (threadIdx.x == 0 && blockIdx.x == 0) gstate[3 * G + 1] = 0;
}

// the queried members' vertices of this thread (grid-stride over [0, M))
#define FOR_SELECTED(V, vmap, gstate, G)                                                        \
  for (uint32_t _t = blockIdx.x * blockDim.x + threadIdx.x, _m = sel_count(gstate, G); _t < _m; \
       _t += gridDim.x * blockDim.x)                                                            \
    if (const uint32_t V = vmap[_t]; true)

__global__ void g_init_kernel(const uint32_t* __restrict__ vmap, uint32_t G,
                              const uint32_t* __restrict__ gstate, uint32_t* parent,
                              uint32_t* comp, uint32_t* vflags) {
  FOR_SELECTED(V, vmap, gstate, G) {
    parent[V] = V;
    comp[V] = V;
    vflags[V] = 0;
  }
}

// vmap: block row j lists queried member j's vertices (sel: its base and first position)
__global__ void g_vmap_kernel(const uint32_t* __restrict__ sel, uint32_t* __restrict__ vmap) {
  const uint32_t j = blockIdx.x, base = sel[2 * j], p0 = sel[2 * j + 1], p1 = sel[2 * j + 3];
  for (uint32_t t = blockIdx.y * blockDim.x + threadIdx.x; p0 + t < p1; t += gridDim.y * blockDim.x)
    vmap[p0 + t] = base + t;
}

__global__ __launch_bounds__(kQThreads) void g_round0_kernel(
    const GraphDesc* __restrict__ descs, const uint32_t* __restrict__ vgraph,
    const uint32_t* __restrict__ vmap, uint32_t G, uint32_t col0, uint32_t* parent,
    uint32_t* vflags, uint32_t* gstate) {
  FOR_SELECTED(V, vmap, gstate, G) {
  const uint32_t g = vgraph[V];
  if (gstate[G + g] != kGActive) continue;
  const GraphDesc d = descs[g];
  const uint32_t v = V - d.base;
  uint64_t idx = 0;
  const Bucket16* vrow = d.rows + (size_t)v * d.vstride;
  const int res = sample_column(
      d.det[v], [&](uint32_t k) { return vrow[bucket_slot(k, col0, d.cols, d.bpc)]; }, d.bpc,
      d.seed, &idx);
  handle_group_sample(res, idx, V, d, parent, vflags, &gstate[g]);
  }
}

// Per member and round: the supernode whose sum is formed from the rest of the member (the
// group form of the giant, see pick_big_kernel): the root holding more than half of 8 evenly
// spaced sampled vertices, on a member whose vertex sketches XOR to zero (pair-fed), else none.
// Any supernode's sum equals the XOR of every vertex outside it then, so the choice only decides
// which vertices are read.
constexpr int kGiantSamples = 8;
__global__ void g_pick_giant_kernel(const GraphDesc* __restrict__ descs, uint32_t G,
                                    const uint32_t* __restrict__ comp,
                                    const uint32_t* __restrict__ vflags,
                                    const uint32_t* __restrict__ gsym,
                                    const uint32_t* __restrict__ gstate,
                                    uint32_t* __restrict__ giant) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G) return;
  uint32_t pick = kNoGiant;
  if (gstate[G + g] == kGActive && gsym[g]) {
    const GraphDesc d = descs[g];
    uint32_t r[kGiantSamples];
    for (int i = 0; i < kGiantSamples; i++)
      r[i] = comp[d.base + (uint32_t)(((uint64_t)i * d.n) / kGiantSamples + d.n / (2 * kGiantSamples))];
    for (int i = 0; i < kGiantSamples; i++) {
      int c = 0;
      for (int j = 0; j < kGiantSamples; j++) c += r[j] == r[i];
      if (2 * c > kGiantSamples) pick = r[i];
    }
    if (pick != kNoGiant && (vflags[pick] & kVDone)) pick = kNoGiant;
  }
  giant[g] = pick;
}

// The group form of acc_kernel. A wave takes 64 consecutive queried vertices (lane = vertex for
// their attributes) and walks the ones it must read, four at a time, one bucket per lane (bucket
// j < B of the round's column, lane B the deterministic bucket: one vertex is one load instruction
// and its row of atomics one instruction over consecutive addresses). A vertex of a live
// supernode adds its column into that supernode's sum; with a giant picked for its member
// (g_pick_giant_kernel), a vertex inside it is not read at all and every other vertex of the
// member (finished supernodes included) also adds into the giant's sum. Both sums are held in
// registers while consecutive vertices share the target (the giant: every read vertex of the
// member) and go out as one row of atomics when it changes.
__global__ __launch_bounds__(kQThreads) void g_acc_kernel(
    const GraphDesc* __restrict__ descs, const uint32_t* __restrict__ vgraph,
    const uint32_t* __restrict__ vmap, uint32_t G, const uint32_t* __restrict__ comp,
    const uint32_t* __restrict__ vflags, const uint32_t* __restrict__ gstate,
    const uint32_t* __restrict__ giant, uint64_t* __restrict__ acc, uint64_t plane,
    uint32_t acc_stride) {
  const uint32_t col = gstate[3 * G];
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t M = sel_count(gstate, G);
  // running sums: of supernode sr (row width sw) and of giant gr (row width gw)
  uint32_t sr = kNoGiant, sw = 0, gr = kNoGiant, gw = 0;
  uint64_t sa = 0, sg = 0, ga = 0, gg = 0;
  auto flush_row = [&](uint32_t row, uint32_t w, uint64_t& a, uint64_t& g) {
    if (row != kNoGiant && lane < w && (a | g)) {
      uint64_t* dst = acc + (size_t)row * acc_stride + lane;
      __hip_atomic_fetch_xor(dst, a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_xor(dst + plane, g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    a = g = 0;
  };
  // whole waves stride over [0, M) (the ballots below need every lane of the wave)
  for (uint32_t t0 = (blockIdx.x * blockDim.x + threadIdx.x) & ~63u; t0 < M;
       t0 += gridDim.x * blockDim.x) {
    bool inm = t0 + lane < M;
    const uint32_t V = inm ? vmap[t0 + lane] : 0u;
    uint32_t g = 0, r = 0, gt = kNoGiant;
    bool done = false;
    if (inm) {
      g = vgraph[V];
      inm = gstate[G + g] == kGActive;
    }
    if (inm) {
      r = comp[V];
      done = (vflags[r] & kVDone) != 0;
      gt = giant[g];
    }
    const bool ingiant = inm && gt != kNoGiant && r == gt;
    const bool act = inm && !done && !ingiant;         // into its own supernode's sum
    const bool cmp = inm && gt != kNoGiant && !ingiant;  // into its member's giant
    const uint64_t actm = __ballot(act), cmpm = __ballot(cmp);
    uint64_t m = actm | cmpm;
    while (m) {
      uint32_t ii[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        ii[q] = m ? (uint32_t)__builtin_ctzll(m) : ii[0];
        m &= m ? m - 1 : 0;
      }
      Bucket16 bv[4];
      uint32_t wq[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const uint32_t Vi = __builtin_amdgcn_readlane((int)V, ii[q]);
        const uint32_t gi = __builtin_amdgcn_readlane((int)g, ii[q]);
        const GraphDesc& d = descs[gi];
        const uint32_t v = Vi - d.base, bpc = d.bpc;
        wq[q] = bpc + 1;
        bv[q] = lane < bpc    ? d.rows[(size_t)v * d.vstride + bucket_slot(lane, col, d.cols, bpc)]
                : lane == bpc ? d.det[v]
                              : Bucket16{0, 0};
      }
#pragma unroll
      for (int q = 0; q < 4; q++) {
        if (q > 0 && ii[q] == ii[0]) break;  // padding of the last group
        const uint32_t i = ii[q];
        if ((actm >> i) & 1) {
          const uint32_t ri = __builtin_amdgcn_readlane((int)r, i);
          if (ri != sr) {
            flush_row(sr, sw, sa, sg);
            sr = ri;
            sw = wq[q];
          }
          sa ^= bv[q].alpha;
          sg ^= bv[q].gamma;
        }
        if ((cmpm >> i) & 1) {
          const uint32_t gi = __builtin_amdgcn_readlane((int)gt, i);
          if (gi != gr) {
            flush_row(gr, gw, ga, gg);
            gr = gi;
            gw = wq[q];
          }
          ga ^= bv[q].alpha;
          gg ^= bv[q].gamma;
        }
      }
    }
  }
  flush_row(sr, sw, sa, sg);
  flush_row(gr, gw, ga, gg);
}

__global__ __launch_bounds__(kQThreads) void g_sample_kernel(
    const GraphDesc* __restrict__ descs, const uint32_t* __restrict__ vgraph,
    const uint32_t* __restrict__ vmap, uint32_t G, const uint32_t* __restrict__ comp,
    uint64_t* __restrict__ acc, uint64_t plane, uint32_t acc_stride, uint32_t* parent,
    uint32_t* vflags, uint32_t* gstate) {
  FOR_SELECTED(V, vmap, gstate, G) {
  if (comp[V] != V || (vflags[V] & kVDone)) continue;
  const uint32_t g = vgraph[V];
  if (gstate[G + g] != kGActive) continue;
  const GraphDesc d = descs[g];
  uint64_t* a = acc + (size_t)V * acc_stride;
  uint64_t idx = 0;
  auto load = [&](uint32_t k) { return Bucket16{a[k], a[plane + k]}; };
  const int res = sample_column(load(d.bpc), load, d.bpc, d.seed, &idx);
  for (uint32_t j = 0; j <= d.bpc; j++) a[j] = a[plane + j] = 0;
  handle_group_sample(res, idx, V, d, parent, vflags, &gstate[g]);
  }
}

// After the last round: the queried members' labels in member-local ids (the smallest vertex of
// each component, as the single-store query gives them) and each member's component count (its
// roots). The selected vertices are listed member after member, so a wave's lanes mostly share
// one member: one atomic per member present in the wave, with the count of its roots there.
__global__ __launch_bounds__(kQThreads) void g_labels_kernel(const GraphDesc* __restrict__ descs,
                                                             const uint32_t* __restrict__ vgraph,
                                                             const uint32_t* __restrict__ vmap,
                                                             uint32_t G, uint32_t* gstate,
                                                             uint32_t* comp, uint32_t* ncc) {
  const uint32_t lane = threadIdx.x & 63;
  FOR_SELECTED(V, vmap, gstate, G) {
    const uint32_t g = vgraph[V];
    const uint32_t c = comp[V];
    comp[V] = c - descs[g].base;
    const bool root = c == V;
    uint64_t pending = __ballot(root);
    while (pending) {  // wave-uniform: one member per pass
      const uint32_t l = (uint32_t)__builtin_ctzll(pending);
      const uint32_t gl = (uint32_t)__shfl((int)g, (int)l);
      const uint64_t same = __ballot(root && g == gl);
      if (lane == l) atomicAdd(&ncc[gl], (uint32_t)__builtin_popcountll(same));
      pending &= ~same;
    }
  }
}

// compress_kernel and finalize_parent_kernel over the queried members' vertices
__global__ __launch_bounds__(kQThreads) void g_compress_kernel(const uint32_t* __restrict__ vmap,
                                                               uint32_t G,
                                                               const uint32_t* __restrict__ gstate,
                                                               uint32_t* parent, uint32_t* comp,
                                                               uint32_t* vflags) {
  FOR_SELECTED(v, vmap, gstate, G) {
    const uint32_t r = find_root(parent, v);
    comp[v] = r;
    if (r == v) {
      const uint32_t f = vflags[v];
      vflags[v] = settle(f);
    }
  }
}
__global__ __launch_bounds__(kQThreads) void g_finalize_kernel(const uint32_t* __restrict__ vmap,
                                                               uint32_t G,
                                                               const uint32_t* __restrict__ gstate,
                                                               uint32_t* parent,
                                                               const uint32_t* comp) {
  FOR_SELECTED(v, vmap, gstate, G) parent[v] = comp[v];
}

// after a round: settle every active member and advance the round counter (one workgroup)
__global__ void g_settle_kernel(const GraphDesc* __restrict__ descs, uint32_t G,
                                uint32_t* gstate) {
  const uint32_t round = gstate[3 * G];
  uint32_t active = 0;
  for (uint32_t g = threadIdx.x; g < G; g += blockDim.x) {
    if (gstate[G + g] != kGActive) continue;
    if (!gstate[g]) {
      gstate[G + g] = kGDone;  // nothing changed: the partition is final
      gstate[2 * G + g] = round;
    } else if (round + 1 >= descs[g].samples) {
      gstate[G + g] = kGOutOfSamples;
      gstate[2 * G + g] = round + 1;
    } else {
      active = 1;
    }
    gstate[g] = 0;
  }
  if (active) atomicOr(&gstate[3 * G + 1], 1u);
  __syncthreads();
  if (threadIdx.x == 0) gstate[3 * G] = round + 1;
}

void free_group_query(Group* g) {
  (void)hipFree(g->d_parent);
  (void)hipFree(g->d_comp);
  (void)hipFree(g->d_vflags);
  (void)hipFree(g->d_gstate);
  (void)hipFree(g->d_acc);
  (void)hipFree(g->d_giant);
  (void)hipFree(g->d_vmap);
  g->d_giant = g->d_gsym = g->d_vmap = nullptr;  // d_gsym lies inside d_gstate
  if (g->h_gstate) (void)hipHostFree(g->h_gstate);
  if (g->round_exec) (void)hipGraphExecDestroy(g->round_exec);
  if (g->qstream) (void)hipStreamDestroy(g->qstream);
  if (g->ev_fence) (void)hipEventDestroy(g->ev_fence);
  g->qstream = nullptr;
  g->ev_fence = nullptr;
  g->fenced = false;
  g->d_parent = g->d_comp = g->d_vflags = g->d_gstate = g->h_gstate = nullptr;
  g->d_acc = nullptr;
  g->round_exec = nullptr;
}

// the kernels of one round >= 1, launched on `q` (eagerly or under capture)
static void launch_group_round(Group* g, hipStream_t q) {
  const uint32_t G = (uint32_t)g->members.size();
  const dim3 grid(kSelGrid), block(kQThreads);
  hipLaunchKernelGGL(g_begin_kernel, dim3(1), dim3(64), 0, q, g->d_gstate, G);
  hipLaunchKernelGGL(g_pick_giant_kernel, dim3((G + 255) / 256), dim3(256), 0, q, g->d_descs, G,
                     g->d_comp, g->d_vflags, g->d_gsym, g->d_gstate, g->d_giant);
  hipLaunchKernelGGL(g_acc_kernel, grid, block, 0, q, g->d_descs, g->d_vgraph, g->d_vmap, G,
                     g->d_comp, g->d_vflags, g->d_gstate, g->d_giant,
                     reinterpret_cast<uint64_t*>(g->d_acc), (uint64_t)g->n_total * g->acc_stride,
                     (uint32_t)g->acc_stride);
  hipLaunchKernelGGL(g_sample_kernel, grid, block, 0, q, g->d_descs, g->d_vgraph, g->d_vmap, G,
                     g->d_comp, reinterpret_cast<uint64_t*>(g->d_acc),
                     (uint64_t)g->n_total * g->acc_stride, (uint32_t)g->acc_stride, g->d_parent,
                     g->d_vflags, g->d_gstate);
  hipLaunchKernelGGL(g_compress_kernel, grid, block, 0, q, g->d_vmap, G, g->d_gstate, g->d_parent,
                     g->d_comp, g->d_vflags);
  hipLaunchKernelGGL(g_settle_kernel, dim3(1), dim3(1024), 0, q, g->d_descs, G, g->d_gstate);
  hipLaunchKernelGGL(g_finalize_kernel, grid, block, 0, q, g->d_vmap, G, g->d_gstate, g->d_parent,
                     g->d_comp);
}

// words of d_gstate / h_gstate: the state (3G + 3), the selection (2 per member + 2), per
// member whether its vertex sketches XOR to zero (d_gsym points there), then per member its
// component count (g_labels_kernel)
static size_t gstate_words(uint32_t G) { return 3 * (size_t)G + 3 + 2 * ((size_t)G + 1) + 2 * (size_t)G; }
static size_t gstate_ncc(uint32_t G) { return 3 * (size_t)G + 3 + 2 * ((size_t)G + 1) + G; }

// the batched query's scratch and its captured round, made once per group
int prepare_group_query(Group* g) {
  const uint32_t N = g->n_total, G = (uint32_t)g->members.size();
  if (!g->qstream) GS_HIP(hipStreamCreateWithFlags(&g->qstream, hipStreamNonBlocking));
  hipStream_t q = g->qstream;
  if (!g->d_parent) {
    g->acc_stride = g->max_bpc + 1;
    GS_HIP(hipMalloc(&g->d_parent, sizeof(uint32_t) * N));
    GS_HIP(hipMalloc(&g->d_comp, sizeof(uint32_t) * N));
    GS_HIP(hipMalloc(&g->d_vflags, sizeof(uint32_t) * N));
    GS_HIP(hipMalloc(&g->d_gstate, sizeof(uint32_t) * gstate_words(G)));
    GS_HIP(hipHostMalloc((void**)&g->h_gstate, sizeof(uint32_t) * gstate_words(G),
                         hipHostMallocDefault));
    GS_HIP(hipMalloc(&g->d_vmap, sizeof(uint32_t) * N));
    GS_HIP(hipMalloc(&g->d_acc, sizeof(Bucket16) * (size_t)N * g->acc_stride));
    GS_HIP(hipMalloc(&g->d_giant, sizeof(uint32_t) * G));
    g->d_gsym = g->d_gstate + 3 * (size_t)G + 3 + 2 * ((size_t)G + 1);  // part of d_gstate
    hipLaunchKernelGGL(zero_acc_kernel, dim3(2048), dim3(256), 0, q, g->d_acc,
                       (uint64_t)N * g->acc_stride);
    GS_HIP(hipGetLastError());
    GS_HIP(hipStreamSynchronize(q));
    // capture one round once; replays reuse the same scratch pointers
    hipGraph_t graph = nullptr;
    GS_HIP(hipStreamBeginCapture(q, hipStreamCaptureModeThreadLocal));
    launch_group_round(g, q);
    GS_HIP(hipStreamEndCapture(q, &graph));
    hipError_t e = hipGraphInstantiate(&g->round_exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    GS_HIP(e);
  }
  return GS_OK;
}

int group_query_cc(Group* g, const uint8_t* mask, uint32_t* labels, uint32_t* num_cc,
                   uint32_t* rounds, int32_t* status) {
  const uint32_t G = (uint32_t)g->members.size();
  int prc = prepare_group_query(g);
  if (prc) return prc;
  // the query runs on its own stream: after the group stream's fence (gs_group_fence), or after
  // everything (the caller synchronized the members' writers)
  hipStream_t q = g->qstream;
  if (g->fenced) GS_HIP(hipStreamWaitEvent(q, g->ev_fence, 0));
  // host-initialised state: nothing modified, members outside the mask already done, round 0;
  // the queried members' bases and positions in vmap; which members' sketches XOR to zero (the
  // giant of such a member is summed from the rest)
  std::fill(g->h_gstate, g->h_gstate + gstate_words(G), 0u);
  uint32_t* sel = g->h_gstate + 3 * (size_t)G + 3;
  uint32_t* sym = sel + 2 * ((size_t)G + 1);
  uint32_t nsel = 0, M = 0;
  for (uint32_t i = 0; i < G; i++) {
    sym[i] = g->members[i]->symmetric.load() ? 1u : 0u;
    const bool on = !mask || mask[i];
    g->h_gstate[G + i] = on ? kGActive : kGDone;
    if (!on) continue;
    sel[2 * nsel] = g->bases[i];
    sel[2 * nsel + 1] = M;
    M += g->members[i]->view.n;
    nsel++;
  }
  sel[2 * nsel + 1] = M;
  g->h_gstate[3 * G + 2] = M;
  GS_HIP(hipMemcpyAsync(g->d_gstate, g->h_gstate, gstate_words(G) * sizeof(uint32_t),
                        hipMemcpyHostToDevice, q));
  if (nsel) {
    hipLaunchKernelGGL(g_vmap_kernel, dim3(nsel, 16), dim3(256), 0, q, g->d_gstate + 3 * G + 3,
                       g->d_vmap);
  }
  const dim3 grid(kSelGrid), block(kQThreads);
  hipLaunchKernelGGL(g_init_kernel, grid, block, 0, q, g->d_vmap, G, g->d_gstate, g->d_parent,
                     g->d_comp, g->d_vflags);
  hipLaunchKernelGGL(g_begin_kernel, dim3(1), dim3(64), 0, q, g->d_gstate, G);
  hipLaunchKernelGGL(g_round0_kernel, grid, block, 0, q, g->d_descs, g->d_vgraph, g->d_vmap, G, 0u,
                     g->d_parent, g->d_vflags, g->d_gstate);
  hipLaunchKernelGGL(g_compress_kernel, grid, block, 0, q, g->d_vmap, G, g->d_gstate, g->d_parent,
                     g->d_comp, g->d_vflags);
  hipLaunchKernelGGL(g_settle_kernel, dim3(1), dim3(1024), 0, q, g->d_descs, G, g->d_gstate);
  hipLaunchKernelGGL(g_finalize_kernel, grid, block, 0, q, g->d_vmap, G, g->d_gstate, g->d_parent,
                     g->d_comp);
  GS_HIP(hipGetLastError());
  uint32_t max_samples = 0;
  for (Store* m : g->members) max_samples = std::max(max_samples, m->num_samples);
  // replay rounds, two per host check: a round with no active member does nothing
  for (uint32_t done = 0; done <= max_samples + 1; done += 2) {
    GS_HIP(hipMemcpyAsync(&g->h_gstate[3 * G + 1], &g->d_gstate[3 * G + 1], sizeof(uint32_t),
                          hipMemcpyDeviceToHost, q));
    GS_HIP(hipStreamSynchronize(q));
    if (!g->h_gstate[3 * G + 1]) break;
    GS_HIP(hipGraphLaunch(g->round_exec, q));
    GS_HIP(hipGraphLaunch(g->round_exec, q));
  }
  // local labels and component counts on the device, then the state, the counts and the labels
  // of the queried members (one copy per run of consecutive queried members)
  uint32_t* d_ncc = g->d_gstate + gstate_ncc(G);
  hipLaunchKernelGGL(g_labels_kernel, grid, block, 0, q, g->d_descs, g->d_vgraph, g->d_vmap, G,
                     g->d_gstate, g->d_comp, d_ncc);
  GS_HIP(hipGetLastError());
  GS_HIP(hipMemcpyAsync(g->h_gstate, g->d_gstate, sizeof(uint32_t) * (3 * (size_t)G + 2),
                        hipMemcpyDeviceToHost, q));
  GS_HIP(hipMemcpyAsync(g->h_gstate + gstate_ncc(G), d_ncc, sizeof(uint32_t) * G,
                        hipMemcpyDeviceToHost, q));
  for (uint32_t i = 0; i < G;) {
    if (mask && !mask[i]) {
      i++;
      continue;
    }
    uint32_t j = i + 1;
    while (j < G && (!mask || mask[j])) j++;
    const uint32_t b0 = g->bases[i], b1 = g->bases[j - 1] + g->members[j - 1]->view.n;
    GS_HIP(hipMemcpyAsync(labels + b0, g->d_comp + b0, sizeof(uint32_t) * (b1 - b0),
                          hipMemcpyDeviceToHost, q));
    i = j;
  }
  GS_HIP(hipStreamSynchronize(q));
  int rc = GS_OK;
  for (uint32_t i = 0; i < G; i++) {
    const uint32_t n = g->members[i]->view.n;
    if (mask && !mask[i]) {  // not queried: its labels are left as they are
      if (num_cc) num_cc[i] = n;
      if (rounds) rounds[i] = 0;
      if (status) status[i] = GS_OK;
      continue;
    }
    const uint32_t st = g->h_gstate[G + i];
    if (num_cc) num_cc[i] = g->h_gstate[gstate_ncc(G) + i];
    if (rounds) rounds[i] = g->h_gstate[2 * G + i];
    const int32_t code = st == kGOutOfSamples ? GS_ERR_OUT_OF_SAMPLES
                         : st == kGDone       ? GS_OK
                                              : GS_ERR_INVALID;
    if (status) status[i] = code;
    if (code != GS_OK && !(mask && !mask[i]) && rc == GS_OK) rc = code;
  }
  if (rc == GS_ERR_OUT_OF_SAMPLES)
    return fail(rc, "a member of the group ran out of samples (see the per-member status)");
  if (rc) return fail(rc, "batched query did not converge");
  return GS_OK;
}

}  // namespace gs

extern "C" int gs_query_cc(gs_store* h, uint32_t* label_out, uint32_t* num_cc, uint32_t* rounds,
                           uint32_t* forest_out, uint64_t forest_cap, uint64_t* forest_len) {
  GS_CHECK(h, GS_ERR_INVALID, "null store");
  return gs::boruvka_query(&h->impl, nullptr, label_out, num_cc, rounds,
                           forest_out, forest_cap, forest_len);
}

extern "C" int gs_k_spanning_forests(gs_store* h, uint32_t k, uint32_t* edges_out,
                                     uint64_t edge_cap, uint64_t* counts) {
  GS_CHECK(h && counts && (edges_out || edge_cap == 0), GS_ERR_INVALID, "null argument");
  GS_CHECK(k >= 1, GS_ERR_INVALID, "k must be >= 1");
  return gs::k_forests(&h->impl, k, edges_out, edge_cap, counts);
}

// Test hook: Sketch::sample of every vertex sketch at sample index `sample_idx`, computed by the
// round-0 samplers of the device Borůvka (the scan the query runs), recorded instead of merged.
using namespace gs;
extern "C" int gs_sample_vertices(gs_store* h, uint32_t sample_idx, int32_t* result,
                                  uint64_t* idx_out) {
  GS_CHECK(h && result && idx_out, GS_ERR_INVALID, "null argument");
  Store* st = &h->impl;
  GS_CHECK(sample_idx < st->num_samples, GS_ERR_OUT_OF_SAMPLES,
           "sample index %u past the sketch's %u samples", sample_idx, st->num_samples);
  int rc = sync_store(st);
  if (rc) return rc;
  const StoreView& v = st->view;
  const uint32_t n = v.n;
  int32_t* d_res = nullptr;
  uint64_t* d_idx = nullptr;
  if (hipMalloc((void**)&d_res, sizeof(int32_t) * n) != hipSuccess ||
      hipMalloc((void**)&d_idx, sizeof(uint64_t) * n) != hipSuccess) {
    (void)hipFree(d_res);
    return fail(GS_ERR_OOM, "gs_sample_vertices: %u results", n);
  }
  hipStream_t q = st->qstream;
  const dim3 g_r0 = wave_grid(n, kR0Verts), block(kQThreads);
  if (st->cols_per_sample > 1)
    hipLaunchKernelGGL(round0_wide_kernel, g_r0, block, 0, q, v, st->cols_per_sample, sample_idx,
                       nullptr, nullptr, nullptr, nullptr, 0ull, d_res, d_idx);
  else
    hipLaunchKernelGGL(round0_kernel, g_r0, block, 0, q, v, sample_idx, nullptr, nullptr,
                       nullptr, nullptr, 0ull, d_res, d_idx);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(q);
  if (e == hipSuccess) e = hipMemcpy(result, d_res, sizeof(int32_t) * n, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(idx_out, d_idx, sizeof(uint64_t) * n, hipMemcpyDeviceToHost);
  (void)hipFree(d_res);
  (void)hipFree(d_idx);
  GS_HIP(e);
  return GS_OK;
}
