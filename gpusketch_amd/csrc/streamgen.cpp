// Seeded synthetic dynamic-graph streams (host code). The reference generates its streams with
// un-vendored tools (StreamingUtilities' generators, tools/stream_generators/*.cpp); these are
// this library's own. Every stream has toggle semantics: an update of a live edge deletes it, so
// each generator guarantees that deletions only hit live edges and inserts never duplicate a
// live edge, and it never emits self loops.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
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

// live-edge membership for R-MAT: dense bitset over the key space when small enough
struct LiveSet {
  std::vector<uint64_t> bits;
  std::unordered_set<uint64_t> hs;
  uint32_t scale;
  bool dense;
  explicit LiveSet(uint32_t sc) : scale(sc), dense(2 * sc <= 34) {
    if (dense) bits.assign(((uint64_t)1 << (2 * sc)) / 64 + 1, 0);
  }
  uint64_t slot(uint64_t k) const { return ((k >> 32) << scale) | (k & 0xFFFFFFFFu); }
  bool test(uint64_t k) const {
    if (!dense) return hs.count(k) != 0;
    uint64_t s = slot(k);
    return (bits[s >> 6] >> (s & 63)) & 1;
  }
  void set(uint64_t k) {
    if (!dense) { hs.insert(k); return; }
    uint64_t s = slot(k);
    bits[s >> 6] |= 1ULL << (s & 63);
  }
  void clear(uint64_t k) {
    if (!dense) { hs.erase(k); return; }
    uint64_t s = slot(k);
    bits[s >> 6] &= ~(1ULL << (s & 63));
  }
};

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

extern "C" int gs_gen_rmat(uint32_t scale, double a, double b, double c, uint64_t count,
                           double del_frac, uint64_t seed, uint32_t* pairs) {
  if (!pairs || scale < 1 || scale > 30 || a < 0 || b < 0 || c < 0 || a + b + c > 1 ||
      del_frac < 0 || del_frac >= 1)
    return GS_ERR_INVALID;
  Rng rng(seed);
  LiveSet live(scale);
  std::vector<uint64_t> list;
  list.reserve(count);
  const uint64_t ta = (uint64_t)(a * 4294967296.0), tab = (uint64_t)((a + b) * 4294967296.0),
                 tabc = (uint64_t)((a + b + c) * 4294967296.0);
  const uint64_t tdel = (uint64_t)(del_frac * 18446744073709551616.0);
  for (uint64_t i = 0; i < count; i++) {
    bool del = !list.empty() && rng.next() < tdel;
    if (!del) {
      uint64_t key = 0;
      int tries = 0;
      for (; tries < 4096; tries++) {
        uint32_t u = 0, v = 0;
        uint64_t r = 0;
        for (uint32_t l = 0; l < scale; l++) {
          if ((l & 1) == 0) r = rng.next();
          uint64_t x = (l & 1) ? (r >> 32) : (r & 0xFFFFFFFFu);
          uint32_t bu = x >= tab, bv = (x >= ta && x < tab) || x >= tabc;
          u = (u << 1) | bu;
          v = (v << 1) | bv;
        }
        if (u == v) continue;
        key = key_of(u, v);
        if (!live.test(key)) break;
      }
      if (tries == 4096) {
        if (list.empty()) return GS_ERR_INVALID;
        del = true;  // saturated: turn this update into a deletion
      } else {
        live.set(key);
        list.push_back(key);
        pairs[2 * i] = (uint32_t)(key >> 32);
        pairs[2 * i + 1] = (uint32_t)key;
        continue;
      }
    }
    uint64_t j = rng.below(list.size());
    uint64_t key = list[j];
    list[j] = list.back();
    list.pop_back();
    live.clear(key);
    pairs[2 * i] = (uint32_t)(key >> 32);
    pairs[2 * i + 1] = (uint32_t)key;
  }
  return GS_OK;
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
