/*
 * CPU oracle for the l0-sampling sketch engine -- TEST INFRASTRUCTURE ONLY.
 *
 * This file restates, in plain C, the reference algorithm that the HIP path must reproduce
 * bit-for-bit. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it, and only as the checker (or the timed CPU baseline). The product library never links it.
 *
 * Every function cites the reference file:line it follows. Pinning (what checks each function
 * against something other than this restatement) is listed in the header comment of
 * oracle/oracle.py; no function here is left unpinned.
 *
 * Bucket arrays in this file use the REFERENCE layout: per vertex, num_columns*bkt_per_col
 * buckets ordered [column*bkt_per_col + depth], then the deterministic bucket last
 * (src/sketch.cpp:8-37). A bucket is two uint64 words {alpha, gamma} (include/bucket.h:7-12,
 * vec_hash_t taken as 64-bit).
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

enum { ORC_GOOD = 0, ORC_ZERO = 1, ORC_FAIL = 2, ORC_OUT_OF_SAMPLES = -1 };

/* ---------------------------------------------------------------------------------------------
 * a3: XXH3_64bits_withSeed for an 8-byte input (include/types.h:8-9 binds vec_hash/col_hash to
 * it). Restated from the published XXH3 short-input path (len 4..8 + rrmxmx). The two secret
 * words are bytes 8..15 and 16..23 of XXH3's default 192-byte secret read little-endian.
 * ------------------------------------------------------------------------------------------- */
static const uint64_t XXH3_SECRET_8 = 0x1cad21f72c81017cULL;
static const uint64_t XXH3_SECRET_16 = 0xdb979083e96dd4deULL;
static const uint64_t XXH3_RRMXMX_MUL = 0x9FB21C651E98DF25ULL;

static inline uint64_t rotl64(uint64_t x, unsigned r) { return (x << r) | (x >> (64u - r)); }
static inline uint32_t bswap32(uint32_t x) {
  return (x >> 24) | ((x >> 8) & 0xff00u) | ((x << 8) & 0xff0000u) | (x << 24);
}

uint64_t orc_xxh3_len8(uint64_t key, uint64_t seed) {
  uint64_t s = seed ^ ((uint64_t)bswap32((uint32_t)seed) << 32);
  uint64_t bitflip = (XXH3_SECRET_8 ^ XXH3_SECRET_16) - s;
  /* input1 = low word, input2 = high word of the little-endian key; input64 = input2 + (input1<<32) */
  uint64_t x = ((key >> 32) | (key << 32)) ^ bitflip;
  x ^= rotl64(x, 49) ^ rotl64(x, 24);
  x *= XXH3_RRMXMX_MUL;
  x ^= (x >> 35) + 8u; /* + len */
  x *= XXH3_RRMXMX_MUL;
  x ^= x >> 28;
  return x;
}

/* a4: Bucket_Boruvka::get_index_depth (include/bucket.h:54-59) */
uint32_t orc_index_depth(uint64_t idx, uint64_t seed_and_col, uint32_t max_depth) {
  uint64_t h = orc_xxh3_len8(idx, seed_and_col);
  h |= (1ULL << max_depth);
  return (uint32_t)__builtin_ctzll(h);
}

/* a2: concat_pairing_fn (src/util.cpp:45-51) */
uint64_t orc_concat_pair(uint32_t i, uint32_t j) {
  if (i > j) { uint32_t t = i; i = j; j = t; }
  return ((uint64_t)i << 32) | j;
}

/* ---------------------------------------------------------------------------------------------
 * a1: geometry (include/sketch.h:63-77, 189-216; src/sketch.cpp:8-37). out[] =
 * {vector_len, bkt_per_col, num_samples, num_columns, num_buckets}.
 * ------------------------------------------------------------------------------------------- */
void orc_geometry(uint32_t n, double factor, uint32_t cols_per_sample, int l0, uint64_t *out) {
  uint64_t len = (uint64_t)ceil((double)n * (uint32_t)(n - 1u) / 2);
  uint64_t bpc = (uint64_t)(ceil(log2((double)len)) + 1);
  if (bpc % 16 == 0) bpc += 1;
  double div = l0 ? (log2(3.0) - 1) : (1 - 0.263034406);
  uint64_t s = (uint64_t)ceil(factor * log2((double)n) / div);
  if (s < 18) s = 18;
  uint64_t cols = s * cols_per_sample;
  out[0] = len;
  out[1] = bpc;
  out[2] = s;
  out[3] = cols;
  out[4] = cols * bpc + 1;
}

/* a7: Sketch::update -- non-L0 rule src/sketch.cpp:85-99, L0 rule src/sketch.cpp:66-83.
 * column_seed(c) = seed + 5c, checksum_seed = seed (include/sketch.h:189-190). */
void orc_sketch_update(uint64_t *bk, uint32_t bpc, uint32_t cols, uint64_t seed, uint64_t idx,
                       int l0) {
  uint64_t nb = (uint64_t)cols * bpc + 1;
  uint64_t chk = orc_xxh3_len8(idx, seed);
  bk[2 * (nb - 1)] ^= idx;
  bk[2 * (nb - 1) + 1] ^= chk;
  for (uint32_t c = 0; c < cols; c++) {
    uint32_t d = orc_index_depth(idx, seed + 5ULL * c, bpc);
    if (d < bpc) {
      if (l0) {
        for (uint32_t j = 0; j <= d; j++) {
          uint64_t b = (uint64_t)c * bpc + j;
          bk[2 * b] ^= idx;
          bk[2 * b + 1] ^= chk;
        }
      } else {
        uint64_t b = (uint64_t)c * bpc + d;
        bk[2 * b] ^= idx;
        bk[2 * b + 1] ^= chk;
      }
    }
  }
}

/* a14: CCSketchAlg::apply_update_batch's sketch effect (src/cc_sketch_alg.cpp:133-145): every
 * dst of a vertex batch toggles concat(src, dst) in src's sketch. XOR makes the per-thread delta
 * + merge of the reference equal to direct updates. */
void orc_apply_vertex_batches(uint64_t *store, uint32_t bpc, uint32_t cols, uint64_t seed,
                              const uint32_t *src, const uint64_t *sizes, const uint64_t *offsets,
                              uint32_t nbatches, const uint32_t *dsts, int l0) {
  uint64_t nb = (uint64_t)cols * bpc + 1;
  for (uint32_t b = 0; b < nbatches; b++) {
    uint64_t *sk = store + 2 * nb * src[b];
    for (uint64_t i = 0; i < sizes[b]; i++)
      orc_sketch_update(sk, bpc, cols, seed, orc_concat_pair(src[b], dsts[offsets[b] + i]), l0);
  }
}

/* One vertex's sketch (any storage) after toggling concat(src, dsts[i]) for i < count -- the
 * per-vertex batch of src/cc_sketch_alg.cpp:139-141 without the store around it. */
void orc_update_vertex(uint64_t *bk, uint32_t bpc, uint32_t cols, uint64_t seed, uint32_t src,
                       const uint32_t *dsts, uint64_t count, int l0) {
  for (uint64_t i = 0; i < count; i++)
    orc_sketch_update(bk, bpc, cols, seed, orc_concat_pair(src, dsts[i]), l0);
}

/* Both endpoints of each undirected update, as the driver inserts {u,v} and {v,u}
 * (include/graph_sketch_driver.h:171-172) and CCSketchAlg::update does (src/cc_sketch_alg.cpp:154-160). */
void orc_apply_edges(uint64_t *store, uint32_t bpc, uint32_t cols, uint64_t seed,
                     const uint32_t *pairs, uint64_t nedges, int l0) {
  uint64_t nb = (uint64_t)cols * bpc + 1;
  for (uint64_t e = 0; e < nedges; e++) {
    uint32_t u = pairs[2 * e], v = pairs[2 * e + 1];
    uint64_t idx = orc_concat_pair(u, v);
    orc_sketch_update(store + 2 * nb * u, bpc, cols, seed, idx, l0);
    orc_sketch_update(store + 2 * nb * v, bpc, cols, seed, idx, l0);
  }
}

/* a9: Sketch::sample (src/sketch.cpp:110-132); is_good (include/bucket.h:65-67). Returns the
 * result code; ORC_OUT_OF_SAMPLES mirrors OutOfSamplesException. */
int orc_sample(const uint64_t *bk, uint32_t bpc, uint32_t num_samples, uint32_t cps,
               uint64_t seed, uint32_t sample_idx, uint64_t *idx_out) {
  uint64_t nb = (uint64_t)num_samples * cps * bpc + 1;
  *idx_out = 0;
  if (sample_idx >= num_samples) return ORC_OUT_OF_SAMPLES;
  uint64_t first_col = (uint64_t)sample_idx * cps;
  const uint64_t *det = bk + 2 * (nb - 1);
  if (det[0] == 0 && det[1] == 0) return ORC_ZERO;
  if (det[1] == orc_xxh3_len8(det[0], seed)) { *idx_out = det[0]; return ORC_GOOD; }
  for (uint32_t i = 0; i < cps; i++)
    for (uint32_t j = 0; j < bpc; j++) {
      const uint64_t *b = bk + 2 * ((i + first_col) * bpc + j);
      if (b[1] == orc_xxh3_len8(b[0], seed)) { *idx_out = b[0]; return ORC_GOOD; }
    }
  return ORC_FAIL;
}

/* ---------------------------------------------------------------------------------------------
 * a20: serial union-find. Roots are picked by a fixed rule here (smaller id wins); the reference
 * picks them by random priority (include/dsu.h:101-174), which changes root ids but never the
 * partition, and every caller below only exposes the partition.
 * ------------------------------------------------------------------------------------------- */
static uint32_t uf_find(uint32_t *p, uint32_t x) {
  while (p[x] != x) { p[x] = p[p[x]]; x = p[x]; }
  return x;
}
static int uf_union(uint32_t *p, uint32_t a, uint32_t b) {
  a = uf_find(p, a); b = uf_find(p, b);
  if (a == b) return 0;
  if (a < b) p[b] = a; else p[a] = b;
  return 1;
}
/* canonical label = smallest vertex id in the component */
static void uf_labels(uint32_t *p, uint32_t n, uint32_t *label) {
  for (uint32_t v = 0; v < n; v++) label[v] = uf_find(p, v);
}

/* Components of an explicit edge list: the "DSU over the net graph" check. */
void orc_dsu_components(uint32_t n, const uint32_t *pairs, uint64_t m, uint32_t *label) {
  uint32_t *p = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
  for (uint32_t v = 0; v < n; v++) p[v] = v;
  for (uint64_t e = 0; e < m; e++) uf_union(p, pairs[2 * e], pairs[2 * e + 1]);
  uf_labels(p, n, label);
  free(p);
}

/* ---------------------------------------------------------------------------------------------
 * a19: Borůvka emulation (src/cc_sketch_alg.cpp:162-540), serial restatement. Round 0 samples each
 * vertex sketch at its current sample index (run_round_zero :250-271); round r>=1 XOR-merges
 * column r (+ deterministic bucket) of every member of each supernode into a zeroed sketch and
 * samples it (perform_boruvka_round :273-407 via Sketch::range_merge, src/sketch.cpp:165-190).
 * FAIL and successful merges mark the round modified; the loop stops at the first unmodified
 * round (:517-534). start_idx[v] is each vertex sketch's sample_idx at query start (0 for CC; the
 * MC forest loop keeps it running, src/mc_sketch_alg.cpp:738-774). Returns 0, or
 * ORC_OUT_OF_SAMPLES. forest_out (may be NULL) receives merged edges as (min,max) pairs.
 * ------------------------------------------------------------------------------------------- */
int orc_boruvka(const uint64_t *store, uint32_t n, uint32_t bpc, uint32_t num_samples, uint32_t cps,
                uint64_t seed, uint32_t *start_idx, uint32_t *label_out, uint32_t *rounds_out,
                uint32_t *forest_out, uint64_t *forest_len) {
  uint64_t cols = (uint64_t)num_samples * cps;
  uint64_t nb = cols * bpc + 1;
  uint32_t *p = (uint32_t *)malloc(sizeof(uint32_t) * n);
  uint32_t *root = (uint32_t *)malloc(sizeof(uint32_t) * n);
  uint64_t *acc = (uint64_t *)calloc((size_t)n * 2 * nb, sizeof(uint64_t));
  uint8_t *has = (uint8_t *)calloc(n, 1);
  uint64_t nf = 0;
  int rc = 0;
  for (uint32_t v = 0; v < n; v++) p[v] = v;
  uint32_t round = 0;
  for (;;) {
    int modified = 0;
    uint64_t *samp_idx = (uint64_t *)malloc(sizeof(uint64_t) * n);
    int *samp_res = (int *)malloc(sizeof(int) * n);
    uint32_t nsup = 0;
    if (round == 0) {
      for (uint32_t v = 0; v < n; v++) {
        samp_res[v] = orc_sample(store + 2 * nb * v, bpc, num_samples, cps, seed, start_idx[v],
                                 &samp_idx[v]);
        start_idx[v]++;
      }
      nsup = n;
      for (uint32_t v = 0; v < n; v++) root[v] = v;
    } else {
      for (uint32_t v = 0; v < n; v++) root[v] = uf_find(p, v);
      memset(has, 0, n);
      if (round + 1 > num_samples) { rc = ORC_OUT_OF_SAMPLES; free(samp_idx); free(samp_res); break; }
      uint64_t b0 = (uint64_t)round * cps * bpc, nbk = (uint64_t)cps * bpc;
      for (uint32_t v = 0; v < n; v++) {
        uint32_t r = root[v];
        uint64_t *a = acc + 2 * nb * r;
        const uint64_t *s = store + 2 * nb * v;
        if (!has[r]) { memset(a, 0, sizeof(uint64_t) * 2 * nb); has[r] = 1; }
        a[2 * (nb - 1)] ^= s[2 * (nb - 1)];
        a[2 * (nb - 1) + 1] ^= s[2 * (nb - 1) + 1];
        for (uint64_t i = 0; i < nbk; i++) {
          a[2 * (b0 + i)] ^= s[2 * (b0 + i)];
          a[2 * (b0 + i) + 1] ^= s[2 * (b0 + i) + 1];
        }
      }
      for (uint32_t r = 0; r < n; r++) {
        if (!has[r]) { samp_res[r] = -2; continue; }
        samp_res[r] = orc_sample(acc + 2 * nb * r, bpc, num_samples, cps, seed, round, &samp_idx[r]);
        nsup++;
      }
    }
    (void)nsup;
    for (uint32_t r = 0; r < n; r++) {
      int res = samp_res[r];
      if (res == -2) continue;
      if (res == ORC_OUT_OF_SAMPLES) { rc = ORC_OUT_OF_SAMPLES; break; }
      if (res == ORC_FAIL) { modified = 1; continue; }
      if (res == ORC_GOOD) {
        uint32_t a = (uint32_t)(samp_idx[r] >> 32), b = (uint32_t)samp_idx[r];
        if (a >= n || b >= n) continue; /* cannot occur for a good bucket of a valid stream */
        if (uf_union(p, a, b)) {
          modified = 1;
          if (forest_out) { forest_out[2 * nf] = a < b ? a : b; forest_out[2 * nf + 1] = a < b ? b : a; }
          nf++;
        }
      }
    }
    free(samp_idx);
    free(samp_res);
    if (rc) break;
    if (!modified) break;
    round++;
  }
  if (rounds_out) *rounds_out = round;
  if (forest_len) *forest_len = nf;
  uf_labels(p, n, label_out);
  free(p); free(root); free(acc); free(has);
  return rc;
}

/* a18: k disjoint spanning forests (MCSketchAlg::calc_disjoint_spanning_forests,
 * src/mc_sketch_alg.cpp:889-920): forest i is a Borůvka pass whose round 0 continues each vertex
 * sketch's sample index (no reset between forests); its edges are then toggled out of both
 * endpoints' sketches (filter_sf_edges :837-888) and all forests are toggled back in at the end.
 * store is modified during the call and restored before return. edges_out holds up to
 * k*(n-1) (min,max) pairs; counts[i] is forest i's size. */
int orc_k_forests(uint64_t *store, uint32_t n, uint32_t bpc, uint32_t num_samples, uint32_t cps,
                  uint64_t seed, uint32_t k, uint32_t *edges_out, uint64_t *counts) {
  uint32_t *start = (uint32_t *)calloc(n, sizeof(uint32_t));
  uint32_t *label = (uint32_t *)malloc(sizeof(uint32_t) * n);
  uint64_t total = 0;
  int rc = 0;
  uint32_t cols = num_samples * cps;
  for (uint32_t i = 0; i < k; i++) {
    uint32_t rounds;
    uint64_t len = 0;
    rc = orc_boruvka(store, n, bpc, num_samples, cps, seed, start, label, &rounds,
                     edges_out + 2 * total, &len);
    counts[i] = len;
    if (rc) break;
    for (uint64_t e = 0; e < len; e++) {
      uint32_t a = edges_out[2 * (total + e)], b = edges_out[2 * (total + e) + 1];
      uint64_t idx = orc_concat_pair(a, b);
      orc_sketch_update(store + 2 * ((uint64_t)cols * bpc + 1) * a, bpc, cols, seed, idx, 0);
      orc_sketch_update(store + 2 * ((uint64_t)cols * bpc + 1) * b, bpc, cols, seed, idx, 0);
    }
    total += len;
  }
  for (uint64_t e = 0; e < total; e++) {
    uint32_t a = edges_out[2 * e], b = edges_out[2 * e + 1];
    uint64_t idx = orc_concat_pair(a, b);
    orc_sketch_update(store + 2 * ((uint64_t)cols * bpc + 1) * a, bpc, cols, seed, idx, 0);
    orc_sketch_update(store + 2 * ((uint64_t)cols * bpc + 1) * b, bpc, cols, seed, idx, 0);
  }
  free(start); free(label);
  return rc;
}

/* ---------------------------------------------------------------------------------------------
 * Exact global minimum cut by Stoer–Wagner (textbook, O(n^3) on an adjacency matrix): the
 * independent check for the min-cut answers (the reference calls VieCut at
 * src/mc_sketch_alg.cpp:966-1003; its planted configs are checked against the exact value).
 * Multi-edges add weight. side_out[v] = 1 for the vertices on one side of a minimum cut.
 * Returns UINT64_MAX for n < 2.
 * ------------------------------------------------------------------------------------------- */
uint64_t orc_stoer_wagner(uint32_t n, const uint32_t *pairs, uint64_t m, uint8_t *side_out) {
  if (n < 2) return UINT64_MAX;
  uint64_t *w = (uint64_t *)calloc((size_t)n * n, sizeof(uint64_t));
  for (uint64_t e = 0; e < m; e++) {
    uint32_t a = pairs[2 * e], b = pairs[2 * e + 1];
    if (a == b) continue;
    w[(size_t)a * n + b]++;
    w[(size_t)b * n + a]++;
  }
  uint32_t *members = (uint32_t *)malloc(sizeof(uint32_t) * n); /* merge tree: next pointer */
  uint32_t *head_of = (uint32_t *)malloc(sizeof(uint32_t) * n);
  uint32_t *alive = (uint32_t *)malloc(sizeof(uint32_t) * n);
  uint64_t *key = (uint64_t *)malloc(sizeof(uint64_t) * n);
  uint8_t *added = (uint8_t *)malloc(n);
  for (uint32_t v = 0; v < n; v++) { members[v] = UINT32_MAX; head_of[v] = v; alive[v] = v; }
  uint32_t na = n;
  uint64_t best = UINT64_MAX;
  uint32_t best_head = 0;
  uint8_t *best_side = (uint8_t *)calloc(n, 1);
  while (na > 1) {
    memset(added, 0, n);
    for (uint32_t i = 0; i < na; i++) key[alive[i]] = 0;
    uint32_t prev = UINT32_MAX, last = UINT32_MAX;
    for (uint32_t it = 0; it < na; it++) {
      uint32_t sel = UINT32_MAX;
      for (uint32_t i = 0; i < na; i++) {
        uint32_t v = alive[i];
        if (!added[v] && (sel == UINT32_MAX || key[v] > key[sel])) sel = v;
      }
      added[sel] = 1;
      prev = last;
      last = sel;
      for (uint32_t i = 0; i < na; i++) {
        uint32_t v = alive[i];
        if (!added[v]) key[v] += w[(size_t)sel * n + v];
      }
    }
    if (key[last] < best) {
      best = key[last];
      best_head = last;
      memset(best_side, 0, n);
      for (uint32_t x = last; x != UINT32_MAX; x = members[x]) best_side[x] = 1;
    }
    /* merge last into prev */
    for (uint32_t i = 0; i < na; i++) {
      uint32_t v = alive[i];
      w[(size_t)prev * n + v] += w[(size_t)last * n + v];
      w[(size_t)v * n + prev] = w[(size_t)prev * n + v];
    }
    w[(size_t)prev * n + prev] = 0;
    uint32_t t = prev;
    while (members[t] != UINT32_MAX) t = members[t];
    members[t] = last;
    for (uint32_t i = 0; i < na; i++)
      if (alive[i] == last) { alive[i] = alive[na - 1]; break; }
    na--;
  }
  (void)best_head; (void)head_of;
  if (side_out) memcpy(side_out, best_side, n);
  free(w); free(members); free(head_of); free(alive); free(key); free(added); free(best_side);
  return best;
}

/* ---------------------------------------------------------------------------------------------
 * Bytes model: number of distinct (vertex, bucket) slots a batch of undirected updates touches
 * (each slot costs one 16-byte read and one 16-byte write). Buckets are counted in reference
 * numbering; the deterministic bucket counts once per touched vertex.
 * ------------------------------------------------------------------------------------------- */
static int cmp_u64(const void *a, const void *b) {
  uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
  return x < y ? -1 : x > y;
}
uint64_t orc_distinct_slots(uint32_t bpc, uint32_t cols, uint64_t seed, const uint32_t *pairs,
                            uint64_t nedges) {
  uint64_t nb = (uint64_t)cols * bpc + 1;
  uint64_t cap = nedges * 2 * (cols + 1);
  uint64_t *keys = (uint64_t *)malloc(sizeof(uint64_t) * (cap ? cap : 1));
  uint64_t m = 0;
  for (uint64_t e = 0; e < nedges; e++) {
    uint32_t u = pairs[2 * e], v = pairs[2 * e + 1];
    uint64_t idx = orc_concat_pair(u, v);
    uint32_t ends[2] = {u, v};
    for (int s = 0; s < 2; s++) {
      keys[m++] = (uint64_t)ends[s] * nb + (nb - 1);
      for (uint32_t c = 0; c < cols; c++) {
        uint32_t d = orc_index_depth(idx, seed + 5ULL * c, bpc);
        if (d < bpc) keys[m++] = (uint64_t)ends[s] * nb + (uint64_t)c * bpc + d;
      }
    }
  }
  qsort(keys, m, sizeof(uint64_t), cmp_u64);
  uint64_t d = 0;
  for (uint64_t i = 0; i < m; i++)
    if (i == 0 || keys[i] != keys[i - 1]) d++;
  free(keys);
  return d;
}

/* ---------------------------------------------------------------------------------------------
 * CPU baseline: the reference CPU update path (CCSketchAlg::apply_update_batch,
 * src/cc_sketch_alg.cpp:133-145): each worker thread zeroes its delta sketch, applies
 * Sketch::update for each dst of a pre-grouped vertex batch, then merges the delta into the
 * vertex sketch under that sketch's mutex. Batches are taken round-robin by threads. The store
 * is in reference layout; l0 selects the CubeSketch rule. Returns elapsed seconds (monotonic
 * clock).
 * ------------------------------------------------------------------------------------------- */
typedef struct {
  uint64_t *store;
  uint32_t bpc, cols;
  uint64_t seed;
  const uint32_t *src;
  const uint64_t *sizes, *offsets;
  const uint32_t *dsts;
  uint32_t nbatches;
  uint32_t tid, nthreads;
  pthread_mutex_t *locks;
  uint32_t nlocks;
  int l0;
} cpu_job_t;

static void *cpu_worker(void *arg) {
  cpu_job_t *j = (cpu_job_t *)arg;
  uint64_t nb = (uint64_t)j->cols * j->bpc + 1;
  uint64_t *delta = (uint64_t *)malloc(sizeof(uint64_t) * 2 * nb);
  for (uint32_t b = j->tid; b < j->nbatches; b += j->nthreads) {
    memset(delta, 0, sizeof(uint64_t) * 2 * nb);
    uint32_t s = j->src[b];
    for (uint64_t i = 0; i < j->sizes[b]; i++)
      orc_sketch_update(delta, j->bpc, j->cols, j->seed,
                        orc_concat_pair(s, j->dsts[j->offsets[b] + i]), j->l0);
    pthread_mutex_t *lk = &j->locks[s % j->nlocks];
    pthread_mutex_lock(lk);
    uint64_t *sk = j->store + 2 * nb * s;
    for (uint64_t i = 0; i < 2 * nb; i++) sk[i] ^= delta[i];
    pthread_mutex_unlock(lk);
  }
  free(delta);
  return NULL;
}

/* The directed entries of m undirected pairs (u->v for every pair in order, then v->u) grouped
 * by source vertex in that order (a counting sort), counts[v] = the entries of v: the update
 * lists GraphSketchDriver's gutters hand to apply_update_batch (include/graph_sketch_driver.h:
 * 171-172,204-208) before they are cut into batches. Returns -1 if a vertex id is >= n. */
int orc_group_by_vertex(const uint32_t *pairs, uint64_t m, uint32_t n, uint64_t *counts,
                        uint32_t *dst) {
  memset(counts, 0, sizeof(uint64_t) * n);
  for (uint64_t i = 0; i < 2 * m; i++) {
    if (pairs[i] >= n) return -1;
    counts[pairs[i]]++;
  }
  uint64_t *pos = (uint64_t *)malloc(sizeof(uint64_t) * n);
  if (!pos) return -1;
  uint64_t s = 0;
  for (uint32_t v = 0; v < n; v++) {
    pos[v] = s;
    s += counts[v];
  }
  for (uint64_t i = 0; i < m; i++) dst[pos[pairs[2 * i]]++] = pairs[2 * i + 1];
  for (uint64_t i = 0; i < m; i++) dst[pos[pairs[2 * i + 1]]++] = pairs[2 * i];
  free(pos);
  return 0;
}

double orc_cpu_baseline(uint64_t *store, uint32_t bpc, uint32_t cols, uint64_t seed,
                        const uint32_t *src, const uint64_t *sizes, const uint64_t *offsets,
                        uint32_t nbatches, const uint32_t *dsts, uint32_t nthreads, int l0) {
  if (nthreads < 1) nthreads = 1;
  uint32_t nlocks = 4096;
  pthread_mutex_t *locks = (pthread_mutex_t *)malloc(sizeof(pthread_mutex_t) * nlocks);
  for (uint32_t i = 0; i < nlocks; i++) pthread_mutex_init(&locks[i], NULL);
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nthreads);
  cpu_job_t *jobs = (cpu_job_t *)malloc(sizeof(cpu_job_t) * nthreads);
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  for (uint32_t t = 0; t < nthreads; t++) {
    cpu_job_t j = {store, bpc, cols, seed, src, sizes, offsets, dsts, nbatches, t, nthreads, locks, nlocks, l0};
    jobs[t] = j;
    pthread_create(&th[t], NULL, cpu_worker, &jobs[t]);
  }
  for (uint32_t t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  for (uint32_t i = 0; i < nlocks; i++) pthread_mutex_destroy(&locks[i]);
  free(locks); free(th); free(jobs);
  return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}
