// Seeded synthetic dynamic-graph streams (host code). The reference generates its streams with
// un-vendored tools (StreamingUtilities' generators, tools/stream_generators/*.cpp); these are
// this library's own. Every stream has toggle semantics: an update of a live edge deletes it, so
// each generator guarantees that deletions only hit live edges and inserts never duplicate a
// live edge, and it never emits self loops.
#include <sched.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <thread>
#include <unordered_set>
#include <vector>

#include "../../include/gpusketch.h"

namespace {

inline uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

struct Rng {  // xoshiro256**
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    for (auto& x : s) x = splitmix64(seed);
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (next() >> 11) * 0x1.0p-53; }  // [0,1)
  uint64_t below(uint64_t m) { return m ? next() % m : 0; }
};

inline uint64_t key_of(uint32_t u, uint32_t v) {
  return u < v ? ((uint64_t)u << 32) | v : ((uint64_t)v << 32) | u;
}

// all pairs (i<j) of [lo, lo+n) chosen independently with probability p, lexicographic order
void bernoulli_pairs(uint32_t lo, uint32_t n, double p, Rng& rng, std::vector<uint64_t>& out) {
  if (n < 2 || p <= 0) return;
  if (p >= 1) {
    for (uint32_t i = 0; i < n; i++)
      for (uint32_t j = i + 1; j < n; j++) out.push_back(key_of(lo + i, lo + j));
    return;
  }
  const double lq = std::log1p(-p);
  uint64_t i = 0, j = 0;  // current pair (i, j); start before (0, 1)
  bool first = true;
  for (;;) {
    double u = rng.uniform();
    uint64_t skip = (uint64_t)std::floor(std::log1p(-u) / lq);
    uint64_t adv = skip + (first ? 1 : 1);
    if (first) { i = 0; j = 0; first = false; }
    j += adv;
    while (j >= n) {
      j = j - n + i + 2;
      i++;
      if (i >= (uint64_t)n - 1) return;
    }
    out.push_back(key_of(lo + (uint32_t)i, lo + (uint32_t)j));
  }
}

void shuffle(std::vector<uint64_t>& v, Rng& rng) {
  for (size_t k = v.size(); k > 1; k--) std::swap(v[k - 1], v[rng.below(k)]);
}

// Append `extra` events to a base sequence: event e goes right after base position pos[e]
// (pos in [0, base.size()]), events at equal positions keep their given order.
std::vector<uint64_t> interleave(const std::vector<uint64_t>& base,
                                 const std::vector<std::pair<uint64_t, uint64_t>>& extra) {
  // extra: (position, edge key); stable by position
  std::vector<std::pair<uint64_t, uint64_t>> ev(extra);
  std::stable_sort(ev.begin(), ev.end(),
                   [](const auto& a, const auto& b) { return a.first < b.first; });
  std::vector<uint64_t> out;
  out.reserve(base.size() + ev.size());
  size_t k = 0;
  for (size_t i = 0; i <= base.size(); i++) {
    while (k < ev.size() && ev[k].first == i) out.push_back(ev[k++].second);
    if (i < base.size()) out.push_back(base[i]);
  }
  return out;
}

void emit(const std::vector<uint64_t>& keys, uint32_t* pairs) {
  for (size_t i = 0; i < keys.size(); i++) {
    pairs[2 * i] = (uint32_t)(keys[i] >> 32);
    pairs[2 * i + 1] = (uint32_t)keys[i];
  }
}

}  // namespace

extern "C" int gs_gen_erdos_renyi(uint32_t n, double p, double del_frac, uint64_t seed,
                                  uint32_t* pairs, uint64_t* count) {
  if (!count || n < 2 || p < 0 || p > 1 || del_frac < 0 || del_frac > 1) return GS_ERR_INVALID;
  Rng rng(seed);
  std::vector<uint64_t> ins;
  bernoulli_pairs(0, n, p, rng, ins);
  shuffle(ins, rng);
  const uint64_t m = ins.size();
  const uint64_t d = (uint64_t)std::llround(del_frac * (double)m);
  // choose d distinct insert positions; each is deleted after a uniform later position
  std::vector<uint64_t> pos(m);
  for (uint64_t i = 0; i < m; i++) pos[i] = i;
  std::vector<std::pair<uint64_t, uint64_t>> dels;
  dels.reserve(d);
  for (uint64_t k = 0; k < d; k++) {
    uint64_t j = k + rng.below(m - k);
    std::swap(pos[k], pos[j]);
    uint64_t p0 = pos[k];
    uint64_t after = p0 + 1 + rng.below(m - p0);  // in [p0+1, m]
    dels.push_back({after, ins[p0]});
  }
  if (!pairs) {
    *count = m + d;
    return GS_OK;
  }
  if (*count < m + d) return GS_ERR_INVALID;
  std::vector<uint64_t> all = interleave(ins, dels);
  emit(all, pairs);
  *count = all.size();
  return GS_OK;
}

// ---- R-MAT, generated in parallel ------------------------------------------------------------
// The stream is a pure function of the seed, whatever the thread count:
//   * candidate i (i = 0, 1, ...) is an R-MAT edge drawn from its own counter-seeded generator
//     (RmatGen; a self loop redraws from the same generator, at most 64 times, after which the
//     candidate is skipped);
//   * a candidate is an insertion when its edge never appeared among the candidates before it
//     (first occurrence wins: duplicates are dropped, no edge is inserted twice). The check is
//     sharded over threads by the 64-bit word of the edge's bit in a dense bitset (or by a hash
//     of the edge when the bitset would be too large), each shard walking its candidates in
//     index order, so the outcome is the sequential one;
//   * each insertion j (j-th kept edge) is deleted with probability del_frac / (1 - del_frac),
//     at a position drawn uniformly from the insertions after it (as the Erdős–Rényi stream does);
//   * the stream is the insertions in order with the deletions placed behind them, cut to
//     `count` updates.
namespace {

constexpr int kGenShardBits = 10;
constexpr uint64_t kGenChunk = 1ULL << 24;  // candidates per dedupe round
constexpr uint64_t kNoKey = ~0ULL;

inline uint64_t mix_index(uint64_t seed, uint64_t i, uint64_t salt) {
  uint64_t s = seed ^ (i * 0xD1B54A32D192ED03ULL) ^ salt;
  return splitmix64(s);
}

unsigned gen_threads() {
  if (const char* e = std::getenv("GPUSKETCH_GEN_THREADS")) {
    const int t = std::atoi(e);
    if (t >= 1) return (unsigned)std::min(t, 256);
  }
  cpu_set_t set;
  unsigned n = 0;
  if (sched_getaffinity(0, sizeof(set), &set) == 0) n = (unsigned)CPU_COUNT(&set);
  if (n == 0) n = std::max(1u, std::thread::hardware_concurrency());
  return std::min(n, 64u);
}

// run f(t) for t in [0, nt) on nt threads
template <class F>
void parallel_for_threads(unsigned nt, F&& f) {
  std::vector<std::thread> th;
  th.reserve(nt);
  for (unsigned t = 1; t < nt; t++) th.emplace_back([&f, t] { f(t); });
  f(0);
  for (auto& x : th) x.join();
}

// R-MAT edge sampler: two levels of the recursion per table lookup. The 16 outcomes of two levels
// (quadrants a = (0,0), b = (0,1), c = (1,0), d = (1,1) of u's and v's bit at each level) get
// round(p * 2^16) of the 65536 entries of a table indexed by 16 random bits (largest remainders
// settle the sum), so a pair of levels costs one lookup; an odd last level compares 32 random
// bits with the thresholds directly.
struct RmatGen {
  uint32_t scale;
  uint64_t seed, ta, tab, tabc;
  std::vector<uint8_t> pair_table;  // 16 random bits -> (u bit pair << 2) | v bit pair
  RmatGen(uint32_t sc, uint64_t sd, double a, double b, double c)
      : scale(sc), seed(sd), ta((uint64_t)(a * 4294967296.0)), tab((uint64_t)((a + b) * 4294967296.0)),
        tabc((uint64_t)((a + b + c) * 4294967296.0)), pair_table(1u << 16) {
    const double pq[4] = {a, b, c, std::max(0.0, 1 - a - b - c)};
    const uint32_t qu[4] = {0, 0, 1, 1}, qv[4] = {0, 1, 0, 1};
    uint32_t cnt[16];
    double rem[16];
    uint32_t used = 0;
    for (int o = 0; o < 16; o++) {
      const double x = pq[o >> 2] * pq[o & 3] * 65536.0;
      cnt[o] = (uint32_t)std::floor(x);
      rem[o] = x - cnt[o];
      used += cnt[o];
    }
    while (used < 65536) {  // largest remainder first (ties: lowest outcome)
      int best = 0;
      for (int o = 1; o < 16; o++)
        if (rem[o] > rem[best]) best = o;
      cnt[best]++;
      rem[best] = -1;
      used++;
    }
    uint32_t pos = 0;
    for (int o = 0; o < 16; o++) {
      const uint32_t q1 = o >> 2, q2 = o & 3;
      const uint8_t code = (uint8_t)((((qu[q1] << 1) | qu[q2]) << 2) | ((qv[q1] << 1) | qv[q2]));
      for (uint32_t k = 0; k < cnt[o]; k++) pair_table[pos++] = code;
    }
  }
  // candidate i's edge key, or kNoKey when it only drew self loops
  uint64_t key(uint64_t i) const {
    uint64_t s = mix_index(seed, i, 0x52A7C0DEULL);
    const uint32_t npairs = scale >> 1;
    for (int tries = 0; tries < 64; tries++) {
      uint32_t u = 0, v = 0;
      uint64_t r = 0;
      for (uint32_t k = 0; k < npairs; k++) {
        if ((k & 3) == 0) r = splitmix64(s);
        const uint8_t code = pair_table[(r >> (16 * (k & 3))) & 0xFFFFu];
        u = (u << 2) | (code >> 2);
        v = (v << 2) | (code & 3u);
      }
      if (scale & 1) {
        const uint64_t x = splitmix64(s) & 0xFFFFFFFFu;
        u = (u << 1) | (uint32_t)(x >= tab);
        v = (v << 1) | (uint32_t)((x >= ta && x < tab) || x >= tabc);
      }
      if (u != v) return key_of(u, v);
    }
    return kNoKey;
  }
};

}  // namespace

extern "C" int gs_gen_rmat(uint32_t scale, double a, double b, double c, uint64_t count,
                           double del_frac, uint64_t seed, uint32_t* pairs) {
  if (!pairs || scale < 1 || scale > 30 || a < 0 || b < 0 || c < 0 || a + b + c > 1 ||
      del_frac < 0 || del_frac >= 1)
    return GS_ERR_INVALID;
  if (count == 0) return GS_OK;
  const RmatGen gen(scale, seed, a, b, c);
  const unsigned nt = gen_threads();
  const uint32_t nshards = 1u << kGenShardBits;
  const bool dense = 2 * scale <= 34;
  // every 64-bit word of the bitset belongs to one shard, so shard threads never share one;
  // calloc: the zero pages are mapped lazily, by the shard threads that touch them
  const size_t nwords = dense ? (((uint64_t)1 << (2 * scale)) / 64 + 1) : 1;
  std::unique_ptr<uint64_t, decltype(&std::free)> bits_mem(
      static_cast<uint64_t*>(std::calloc(nwords, sizeof(uint64_t))), &std::free);
  if (!bits_mem) return GS_ERR_OOM;
  uint64_t* const bits = bits_mem.get();
  std::vector<std::unordered_set<uint64_t>> seen(dense ? 0 : nshards);
  auto slot_of = [scale](uint64_t k) -> uint64_t { return ((k >> 32) << scale) | (k & 0xFFFFFFFFu); };
  auto shard_of = [&](uint64_t k) -> uint32_t {
    const uint64_t w = dense ? (slot_of(k) >> 6) : k;
    return (uint32_t)((w * 0x9E3779B97F4A7C15ULL) >> (64 - kGenShardBits));
  };
  // kept insertions: enough that, with deletions added, the stream reaches `count` (with a
  // margin of several standard deviations; retried with more if not)
  const double q = del_frac / (1 - del_frac);
  uint64_t target = (uint64_t)std::ceil((double)count * (1 - del_frac) +
                                        6 * std::sqrt((double)count) + 64);
  target = std::min(target, count);
  std::vector<uint64_t> ins;
  ins.reserve(target);
  const uint64_t chunk = std::min<uint64_t>(kGenChunk, std::max<uint64_t>(1ULL << 16, 2 * target));
  std::vector<uint64_t> keys(chunk), order(chunk);
  std::vector<uint8_t> kept(chunk);
  std::vector<uint64_t> cnt((size_t)nt * nshards), tkept(nt);
  uint64_t next = 0, stalled = 0;
  for (;;) {
    // insertions: dedupe rounds over candidate chunks until `target` are kept
    while (ins.size() < target) {
      const uint64_t base = next, m = chunk;
      next += m;
      auto range = [&](unsigned t, uint64_t& lo, uint64_t& hi) {
        lo = m * t / nt;
        hi = m * (t + 1) / nt;
      };
      parallel_for_threads(nt, [&](unsigned t) {
        uint64_t lo, hi;
        range(t, lo, hi);
        uint64_t* ct = &cnt[(size_t)t * nshards];
        std::fill(ct, ct + nshards, 0);
        for (uint64_t i = lo; i < hi; i++) {
          keys[i] = gen.key(base + i);
          if (keys[i] != kNoKey) ct[shard_of(keys[i])]++;
        }
      });
      // stable partition of the chunk's candidates by shard
      uint64_t run = 0;
      std::vector<uint64_t> shard_start(nshards + 1);
      for (uint32_t sh = 0; sh < nshards; sh++) {
        shard_start[sh] = run;
        for (unsigned t = 0; t < nt; t++) {
          const uint64_t x = cnt[(size_t)t * nshards + sh];
          cnt[(size_t)t * nshards + sh] = run;
          run += x;
        }
      }
      shard_start[nshards] = run;
      parallel_for_threads(nt, [&](unsigned t) {
        uint64_t lo, hi;
        range(t, lo, hi);
        uint64_t* ct = &cnt[(size_t)t * nshards];
        for (uint64_t i = lo; i < hi; i++) {
          kept[i] = 0;
          if (keys[i] != kNoKey) order[ct[shard_of(keys[i])]++] = i;
        }
      });
      // shards in candidate order: the first occurrence of an edge is kept
      std::atomic<uint32_t> next_shard{0};
      parallel_for_threads(nt, [&](unsigned) {
        for (uint32_t sh; (sh = next_shard.fetch_add(1)) < nshards;) {
          const uint64_t p0 = shard_start[sh], p1 = shard_start[sh + 1];
          constexpr uint64_t kAhead = 16;  // bitset words requested ahead of their test
          if (dense)
            for (uint64_t p = p0; p < std::min(p1, p0 + kAhead); p++)
              __builtin_prefetch(&bits[slot_of(keys[order[p]]) >> 6], 1);
          for (uint64_t p = p0; p < p1; p++) {
            const uint64_t i = order[p], k = keys[i];
            if (dense) {
              if (p + kAhead < p1) __builtin_prefetch(&bits[slot_of(keys[order[p + kAhead]]) >> 6], 1);
              const uint64_t slot = slot_of(k);
              uint64_t& w = bits[slot >> 6];
              const uint64_t bit = 1ULL << (slot & 63);
              if (!(w & bit)) {
                w |= bit;
                kept[i] = 1;
              }
            } else if (seen[sh].insert(k).second) {
              kept[i] = 1;
            }
          }
        }
      });
      // append the kept candidates in index order (up to target)
      parallel_for_threads(nt, [&](unsigned t) {
        uint64_t lo, hi, c0 = 0;
        range(t, lo, hi);
        for (uint64_t i = lo; i < hi; i++) c0 += kept[i];
        tkept[t] = c0;
      });
      const uint64_t have = ins.size();
      uint64_t add = 0;
      for (unsigned t = 0; t < nt; t++) add += tkept[t];
      stalled = add ? 0 : stalled + 1;
      if (stalled > 64) return GS_ERR_INVALID;  // the R-MAT distribution is exhausted
      const uint64_t take = std::min(add, target - have);
      ins.resize(have + take);
      uint64_t off = have;
      std::vector<uint64_t> toff(nt);
      for (unsigned t = 0; t < nt; t++) {
        toff[t] = off;
        off += tkept[t];
      }
      parallel_for_threads(nt, [&](unsigned t) {
        uint64_t lo, hi, o = toff[t];
        range(t, lo, hi);
        for (uint64_t i = lo; i < hi && o < have + take; i++)
          if (kept[i]) ins[o++] = keys[i];
      });
    }
    // deletions: insertion j is deleted with probability q, behind insertion pos in (j, I]
    // (pos = I: at the very end); grouped by pos, in insertion order within one pos
    const uint64_t I = ins.size();
    const uint64_t tdel = q >= 1 ? ~0ULL : (uint64_t)(q * 18446744073709551616.0);
    std::vector<std::vector<std::pair<uint64_t, uint64_t>>> tdels(nt);
    parallel_for_threads(nt, [&](unsigned t) {
      const uint64_t lo = I * t / nt, hi = I * (t + 1) / nt;
      auto& out = tdels[t];
      for (uint64_t j = lo; j < hi; j++) {
        uint64_t s = mix_index(seed, j, 0xDE1E7EULL);
        if (splitmix64(s) >= tdel) continue;
        const uint64_t pos = j + 1 + splitmix64(s) % (I - j);  // in [j+1, I]
        out.push_back({pos, j});
      }
    });
    uint64_t D = 0;
    for (auto& v : tdels) D += v.size();
    if (I + D < count) {  // short of the requested length: more insertions and redo
      if (target >= count) return GS_ERR_INVALID;
      target = std::min(count, target + (count - I - D) * 2 + 64);
      continue;
    }
    // all deletions sorted by (pos, j): bucketed by pos range, then sorted per bucket
    std::vector<std::vector<std::pair<uint64_t, uint64_t>>> buck(nt);
    {
      std::vector<std::vector<std::vector<std::pair<uint64_t, uint64_t>>>> part(
          nt, std::vector<std::vector<std::pair<uint64_t, uint64_t>>>(nt));
      std::vector<uint64_t> lo_of(nt);  // first insertion of each merge segment
      for (unsigned bkt = 0; bkt < nt; bkt++) lo_of[bkt] = I * bkt / nt;
      parallel_for_threads(nt, [&](unsigned t) {
        for (auto& e : tdels[t]) {  // the segment holding insertion pos - 1
          const unsigned bkt =
              (unsigned)(std::upper_bound(lo_of.begin(), lo_of.end(), e.first - 1) - lo_of.begin()) - 1;
          part[t][bkt].push_back(e);
        }
        tdels[t].clear();
        tdels[t].shrink_to_fit();
      });
      parallel_for_threads(nt, [&](unsigned bkt) {
        auto& out = buck[bkt];
        for (unsigned t = 0; t < nt; t++) out.insert(out.end(), part[t][bkt].begin(), part[t][bkt].end());
        std::sort(out.begin(), out.end());
      });
    }
    // merge: segment b covers the insertions [I*b/nt, I*(b+1)/nt) and the deletions whose pos
    // lies in (I*b/nt, I*(b+1)/nt] -- exactly bucket b. Deletions at pos follow insertion pos-1.
    std::vector<uint64_t> out0(nt);
    {
      uint64_t run = 0;
      for (unsigned bkt = 0; bkt < nt; bkt++) {
        out0[bkt] = run;
        run += (I * (bkt + 1) / nt - I * bkt / nt) + buck[bkt].size();
      }
    }
    parallel_for_threads(nt, [&](unsigned bkt) {
      const uint64_t lo = I * bkt / nt, hi = I * (bkt + 1) / nt;
      uint64_t o = out0[bkt];
      size_t d = 0;
      const auto& dl = buck[bkt];
      auto put = [&](uint64_t k) {
        if (o < count) {
          pairs[2 * o] = (uint32_t)(k >> 32);
          pairs[2 * o + 1] = (uint32_t)k;
        }
        o++;
      };
      for (uint64_t j = lo; j < hi; j++) {
        put(ins[j]);
        while (d < dl.size() && dl[d].first == j + 1) put(ins[dl[d++].second]);
      }
      while (d < dl.size()) put(ins[dl[d++].second]);  // pos == I at the end of the last bucket
    });
    return GS_OK;
  }
}

extern "C" int gs_gen_planted_cut(uint32_t n, double p, uint32_t cut, double churn_frac,
                                  uint64_t seed, uint32_t* pairs, uint64_t* count) {
  if (!count || n < 4 || p < 0 || p > 1 || churn_frac < 0) return GS_ERR_INVALID;
  const uint32_t h = n / 2;
  if ((uint64_t)cut > (uint64_t)h * (n - h)) return GS_ERR_INVALID;
  Rng rng(seed);
  std::vector<uint64_t> base;
  bernoulli_pairs(0, h, p, rng, base);  // keys come out in increasing order
  bernoulli_pairs(h, n - h, p, rng, base);
  auto in_base = [&](uint64_t key) { return std::binary_search(base.begin(), base.end(), key); };
  // crossing edges are never inside a half, so only their own duplicates need a check
  std::unordered_set<uint64_t> crossing;
  while (crossing.size() < cut) {
    uint32_t x = (uint32_t)rng.below(h), y = h + (uint32_t)rng.below(n - h);
    crossing.insert(key_of(x, y));
  }
  const uint64_t m = base.size() + crossing.size();
  const uint64_t nchurn = (uint64_t)std::llround(churn_frac * (double)m);
  const uint64_t all_pairs = (uint64_t)n * (n - 1) / 2;
  const uint64_t churn = std::min(nchurn, all_pairs > m ? all_pairs - m : 0);
  std::vector<std::pair<uint64_t, uint64_t>> extra;
  extra.reserve(2 * churn);
  std::unordered_set<uint64_t> used;
  used.reserve(2 * churn);
  for (uint64_t k = 0; k < churn;) {
    uint32_t x = (uint32_t)rng.below(n), y = (uint32_t)rng.below(n);
    if (x == y) continue;
    uint64_t key = key_of(x, y);
    if (in_base(key) || crossing.count(key) || !used.insert(key).second) continue;
    uint64_t p1 = rng.below(m + 1), p2 = rng.below(m + 1);
    if (p1 > p2) std::swap(p1, p2);
    extra.push_back({p1, key});  // insert ...
    extra.push_back({p2, key});  // ... and its later deletion (stable order keeps insert first)
    k++;
  }
  if (!pairs) {
    *count = m + extra.size();
    return GS_OK;
  }
  for (uint64_t key : crossing) base.push_back(key);
  shuffle(base, rng);
  if (*count < m + extra.size()) return GS_ERR_INVALID;
  std::vector<uint64_t> all = interleave(base, extra);
  emit(all, pairs);
  *count = all.size();
  return GS_OK;
}
