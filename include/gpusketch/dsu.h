// Union-find (the reference's include/dsu.h:1-174): DisjointSetUnion for one thread and
// DisjointSetUnion_MT, lock-free with compare-and-swap links. Roots are ordered by a random
// priority per element (ties by smaller id), so merge order does not skew tree shapes; the
// priorities come from a seeded generator (the reference seeds from the clock), pass a seed to
// make a run repeatable.
#pragma once
#include <atomic>
#include <chrono>
#include <cstdint>
#include <random>
#include <utility>

template <class T>
struct DSUMergeRet {
  bool merged;  // a link was made
  T root;       // the surviving root
  T child;      // the root linked under it
};

namespace gs_detail {
inline uint64_t clock_seed() {
  return (uint64_t)std::chrono::duration_cast<std::chrono::microseconds>(
             std::chrono::high_resolution_clock::now().time_since_epoch())
      .count();
}
}  // namespace gs_detail

template <class T>
class DisjointSetUnion {
  T n;
  T* parent;
  T* priority;

  // a becomes the surviving root: higher priority, ties to the smaller id
  inline void order_edge(T& a, T& b) const {
    if (priority[a] < priority[b] || (priority[a] == priority[b] && a > b)) std::swap(a, b);
  }

 public:
  explicit DisjointSetUnion(T n, uint64_t seed = gs_detail::clock_seed())
      : n(n), parent(new T[n]), priority(new T[n]) {
    std::mt19937_64 gen(seed);
    for (T i = 0; i < n; i++) {
      parent[i] = i;
      priority[i] = (T)gen();
    }
  }
  ~DisjointSetUnion() {
    delete[] parent;
    delete[] priority;
  }
  DisjointSetUnion(const DisjointSetUnion& o) : n(o.n), parent(new T[o.n]), priority(new T[o.n]) {
    for (T i = 0; i < n; i++) {
      parent[i] = o.parent[i];
      priority[i] = o.priority[i];
    }
  }
  DisjointSetUnion(DisjointSetUnion&& o) noexcept : n(o.n), parent(o.parent), priority(o.priority) {
    o.n = 0;
    o.parent = o.priority = nullptr;
  }
  DisjointSetUnion& operator=(const DisjointSetUnion&) = delete;

  // root of u, halving the path on the way up
  inline T find_root(T u) {
    while (parent[u] != u) {
      parent[u] = parent[parent[u]];
      u = parent[u];
    }
    return u;
  }

  inline DSUMergeRet<T> merge(T u, T v) {
    T a = find_root(u), b = find_root(v);
    if (a == b) return {false, 0, 0};
    order_edge(a, b);
    parent[b] = a;
    return {true, a, b};
  }

  inline void reset() {
    for (T i = 0; i < n; i++) parent[i] = i;
  }
  T size() const { return n; }
};

template <class T>
class DisjointSetUnion_MT {
  T n;
  std::atomic<T>* parent;
  T* priority;

  inline void order_edge(T& a, T& b) const {
    if (priority[a] < priority[b] || (priority[a] == priority[b] && a > b)) std::swap(a, b);
  }

 public:
  explicit DisjointSetUnion_MT(T n, uint64_t seed = gs_detail::clock_seed())
      : n(n), parent(new std::atomic<T>[n]), priority(new T[n]) {
    std::mt19937_64 gen(seed);
    for (T i = 0; i < n; i++) {
      parent[i].store(i, std::memory_order_relaxed);
      priority[i] = (T)gen();
    }
  }
  ~DisjointSetUnion_MT() {
    delete[] parent;
    delete[] priority;
  }
  DisjointSetUnion_MT(const DisjointSetUnion_MT& o)
      : n(o.n), parent(new std::atomic<T>[o.n]), priority(new T[o.n]) {
    for (T i = 0; i < n; i++) {
      parent[i].store(o.parent[i].load());
      priority[i] = o.priority[i];
    }
  }
  DisjointSetUnion_MT& operator=(const DisjointSetUnion_MT&) = delete;

  // root of u; compresses with benign races (any stored value is an ancestor)
  inline T find_root(T u) {
    T p = parent[u].load(std::memory_order_relaxed);
    while (p != u) {
      T gp = parent[p].load(std::memory_order_relaxed);
      if (gp != p) parent[u].compare_exchange_weak(p, gp, std::memory_order_relaxed);
      u = p;
      p = parent[u].load(std::memory_order_relaxed);
    }
    return u;
  }

  // safe to call concurrently: a link only succeeds while b is still a root
  inline DSUMergeRet<T> merge(T a, T b) {
    for (;;) {
      a = find_root(a);
      b = find_root(b);
      if (a == b) return {false, 0, 0};
      order_edge(a, b);
      T expected = b;
      if (parent[b].compare_exchange_strong(expected, a)) return {true, a, b};
    }
  }

  inline void reset() {
    for (T i = 0; i < n; i++) parent[i].store(i, std::memory_order_relaxed);
  }
  T size() const { return n; }
};
