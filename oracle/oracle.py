"""CPU oracle for the l0-sampling sketch engine -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
this module, and only as the checker (or, for the CPU baseline, the thing timed on the host).
The product library never loads it.

It wraps ``oracle/liboracle.so`` (a C restatement of the reference algorithm, built from
``oracle/oracle.c`` by ``oracle/Makefile``). Bucket arrays use the reference layout: per vertex,
``num_columns * bkt_per_col`` buckets ordered ``[column * bkt_per_col + depth]`` followed by the
deterministic bucket, each bucket two uint64 ``{alpha, gamma}``.

How each function is pinned (checked against something other than this restatement):

* ``xxh3_len8`` -- golden vectors from the xxhash library (python ``xxhash`` 3.8.1 wrapping
  libxxhash 0.8.2) in ``tests/golden/xxh3_len8.json``; the 8-byte closed form also matches the
  library on 20,000 random (key, seed) pairs (``tests/golden/make_golden.py``).
* ``geometry`` -- the sizing table in ``tests/golden/geometry.json`` (computed from the formulas
  in ``include/sketch.h:63-77,198-216`` with Python's IEEE doubles) plus the closed-form values
  quoted for the reference's configs (n=8192: B=26, C=18; n=2^17: B=34, C=24, 817 buckets).
* ``index_depth`` / ``concat_pair`` -- closed forms (ctz of the hash with bit B forced; the
  concat pairing inverse of ``test/util_test.cpp:4-15``), brute-forced in tests.
* ``sketch_update`` / ``apply_edges`` / ``apply_vertex_batches`` -- an independent pure-Python
  restatement on small cases (tests/test_oracle.py), plus the invariants the reference tests:
  updating an index twice is the identity, order independence, merge linearity
  (``test/sketch_test.cpp:423-458``), index 0 updated 3 times samples GOOD 0 (``:66-80``). The
  CubeSketch rule (``l0=True``, ``src/sketch.cpp:66-83``) is pinned by the same pure-Python
  restatement and by a closed form: each of its columns is the suffix XOR over depths of the
  support-finding column; its geometry by the quoted values (n=2^17: B=34, S=30, C=210).
* ``sample`` -- brute force on tiny vectors: GOOD implies the index is a live coordinate, ZERO
  iff the vector is empty, the crafted FAIL vector of ``test/sketch_test.cpp:14-64``.
* ``boruvka`` / ``dsu_components`` -- the partition equals a union-find over the net graph on
  random streams, and the reference fixture ``multiples_graph_1024`` gives 78 components
  (``test/cc_alg_test.cpp:41-57``).
* ``k_forests`` -- every forest is acyclic and edge-disjoint from the others and uses only
  live edges; the union is a k-certificate (its min cut equals the graph's min cut when < k).
* ``stoer_wagner`` -- closed forms (cycles: 2, paths: 1, complete graphs K_n: n-1, planted cuts)
  and brute force over all 2^n bipartitions for n <= 12.
* ``distinct_slots`` -- brute-force set counting in Python on small batches.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

GOOD, ZERO, FAIL, OUT_OF_SAMPLES = 0, 1, 2, -1

_lib = None


def build() -> None:
    """Compile liboracle.so in place (gcc only)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    u32, u64, dbl, vp, i32 = (ctypes.c_uint32, ctypes.c_uint64, ctypes.c_double,
                              ctypes.c_void_p, ctypes.c_int)
    lib.orc_xxh3_len8.argtypes = [u64, u64]
    lib.orc_xxh3_len8.restype = u64
    lib.orc_index_depth.argtypes = [u64, u64, u32]
    lib.orc_index_depth.restype = u32
    lib.orc_concat_pair.argtypes = [u32, u32]
    lib.orc_concat_pair.restype = u64
    lib.orc_geometry.argtypes = [u32, dbl, u32, i32, vp]
    lib.orc_geometry.restype = None
    lib.orc_sketch_update.argtypes = [vp, u32, u32, u64, u64, i32]
    lib.orc_apply_vertex_batches.argtypes = [vp, u32, u32, u64, vp, vp, vp, u32, vp, i32]
    lib.orc_apply_edges.argtypes = [vp, u32, u32, u64, vp, u64, i32]
    lib.orc_update_vertex.argtypes = [vp, u32, u32, u64, u32, vp, u64, i32]
    lib.orc_sample.argtypes = [vp, u32, u32, u32, u64, u32, vp]
    lib.orc_sample.restype = i32
    lib.orc_dsu_components.argtypes = [u32, vp, u64, vp]
    lib.orc_group_by_vertex.argtypes = [vp, u64, u32, vp, vp]
    lib.orc_group_by_vertex.restype = i32
    lib.orc_boruvka.argtypes = [vp, u32, u32, u32, u32, u64, vp, vp, vp, vp, vp]
    lib.orc_boruvka.restype = i32
    lib.orc_k_forests.argtypes = [vp, u32, u32, u32, u32, u64, u32, vp, vp]
    lib.orc_k_forests.restype = i32
    lib.orc_stoer_wagner.argtypes = [u32, vp, u64, vp]
    lib.orc_stoer_wagner.restype = u64
    lib.orc_distinct_slots.argtypes = [u32, u32, u64, vp, u64]
    lib.orc_distinct_slots.restype = u64
    lib.orc_cpu_baseline.argtypes = [vp, u32, u32, u64, vp, vp, vp, u32, vp, u32, i32]
    lib.orc_cpu_baseline.restype = dbl
    _lib = lib
    return lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


class Geometry:
    """Sketch sizing of ``include/sketch.h:63-77,198-216`` for ``n`` vertices."""

    def __init__(self, n: int, factor: float = 1.0, cols_per_sample: int = 1, l0: bool = False):
        out = np.zeros(5, dtype=np.uint64)
        _load().orc_geometry(n, factor, cols_per_sample, int(l0), _p(out))
        self.n = n
        self.vector_len, self.bkt_per_col, self.num_samples, self.num_columns, self.num_buckets = (
            int(x) for x in out)
        self.cols_per_sample = cols_per_sample

    @classmethod
    def explicit(cls, n: int, bkt_per_col: int, num_samples: int, cols_per_sample: int = 1):
        """Geometry of a Sketch built from a vector length directly (the Sketch constructors of
        include/sketch.h:88-110 take vector_len, not a vertex count)."""
        g = cls.__new__(cls)
        g.n = n
        g.vector_len = None
        g.bkt_per_col = bkt_per_col
        g.num_samples = num_samples
        g.cols_per_sample = cols_per_sample
        g.num_columns = num_samples * cols_per_sample
        g.num_buckets = g.num_columns * bkt_per_col + 1
        return g

    def empty_store(self, n: int | None = None) -> np.ndarray:
        n = self.n if n is None else n
        return np.zeros((n, self.num_buckets, 2), dtype=np.uint64)


def xxh3_len8(key: int, seed: int) -> int:
    return int(_load().orc_xxh3_len8(key & (2**64 - 1), seed & (2**64 - 1)))


def index_depth(idx: int, seed_and_col: int, max_depth: int) -> int:
    return int(_load().orc_index_depth(idx, seed_and_col & (2**64 - 1), max_depth))


def concat_pair(i: int, j: int) -> int:
    return int(_load().orc_concat_pair(i, j))


def sketch_update(bk: np.ndarray, g: Geometry, seed: int, idx: int, l0: bool = False) -> None:
    assert bk.dtype == np.uint64 and bk.flags.c_contiguous
    _load().orc_sketch_update(_p(bk), g.bkt_per_col, g.num_columns, seed, idx, int(l0))


def apply_edges(store: np.ndarray, g: Geometry, seed: int, pairs: np.ndarray,
                l0: bool = False) -> None:
    pairs = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1, 2)
    assert store.dtype == np.uint64 and store.flags.c_contiguous
    assert int(pairs.max(initial=0)) < store.shape[0]
    _load().orc_apply_edges(_p(store), g.bkt_per_col, g.num_columns, seed, _p(pairs),
                            pairs.shape[0], int(l0))


def vertex_sketch(g: Geometry, seed: int, src: int, dsts: np.ndarray, l0: bool = False) -> np.ndarray:
    """The sketch of vertex ``src`` alone after toggling concat(src, d) for every d in dsts."""
    dsts = np.ascontiguousarray(dsts, dtype=np.uint32)
    bk = np.zeros((g.num_buckets, 2), dtype=np.uint64)
    _load().orc_update_vertex(_p(bk), g.bkt_per_col, g.num_columns, seed, src, _p(dsts), dsts.size,
                              int(l0))
    return bk


def apply_vertex_batches(store, g: Geometry, seed, src, sizes, offsets, dsts, l0=False) -> None:
    src = np.ascontiguousarray(src, dtype=np.uint32)
    sizes = np.ascontiguousarray(sizes, dtype=np.uint64)
    offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
    dsts = np.ascontiguousarray(dsts, dtype=np.uint32)
    _load().orc_apply_vertex_batches(_p(store), g.bkt_per_col, g.num_columns, seed, _p(src),
                                     _p(sizes), _p(offsets), len(src), _p(dsts), int(l0))


def sample(bk: np.ndarray, g: Geometry, seed: int, sample_idx: int = 0):
    out = np.zeros(1, dtype=np.uint64)
    bk = np.ascontiguousarray(bk, dtype=np.uint64)
    res = _load().orc_sample(_p(bk), g.bkt_per_col, g.num_samples, g.cols_per_sample, seed,
                             sample_idx, _p(out))
    return int(out[0]), int(res)


def dsu_components(n: int, pairs: np.ndarray) -> np.ndarray:
    pairs = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1, 2)
    label = np.zeros(n, dtype=np.uint32)
    _load().orc_dsu_components(n, _p(pairs), pairs.shape[0], _p(label))
    return label


def boruvka(store: np.ndarray, g: Geometry, seed: int, start_idx: np.ndarray | None = None):
    """Returns (rc, labels (min vertex of component), rounds, forest edges (m,2))."""
    n = store.shape[0]
    lib = _load()
    if start_idx is None:
        start_idx = np.zeros(n, dtype=np.uint32)
    label = np.zeros(n, dtype=np.uint32)
    rounds = np.zeros(1, dtype=np.uint32)
    forest = np.zeros((max(n, 1), 2), dtype=np.uint32)
    flen = np.zeros(1, dtype=np.uint64)
    rc = lib.orc_boruvka(_p(store), n, g.bkt_per_col, g.num_samples, g.cols_per_sample, seed,
                         _p(start_idx), _p(label), _p(rounds), _p(forest), _p(flen))
    return int(rc), label, int(rounds[0]), forest[: int(flen[0])].copy()


def k_forests(store: np.ndarray, g: Geometry, seed: int, k: int):
    """Returns (rc, list of (m_i, 2) forests). store is restored before return."""
    n = store.shape[0]
    edges = np.zeros((max(k * max(n - 1, 1), 1), 2), dtype=np.uint32)
    counts = np.zeros(k, dtype=np.uint64)
    rc = _load().orc_k_forests(_p(store), n, g.bkt_per_col, g.num_samples, g.cols_per_sample,
                               seed, k, _p(edges), _p(counts))
    out, off = [], 0
    for c in counts:
        out.append(edges[off:off + int(c)].copy())
        off += int(c)
    return int(rc), out


def stoer_wagner(n: int, pairs: np.ndarray):
    """Exact global min cut (value, side mask)."""
    pairs = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1, 2)
    side = np.zeros(n, dtype=np.uint8)
    val = _load().orc_stoer_wagner(n, _p(pairs), pairs.shape[0], _p(side))
    return int(val), side


def distinct_slots(g: Geometry, seed: int, pairs: np.ndarray) -> int:
    pairs = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1, 2)
    return int(_load().orc_distinct_slots(g.bkt_per_col, g.num_columns, seed, _p(pairs),
                                          pairs.shape[0]))


def cpu_baseline(store, g: Geometry, seed, src, sizes, offsets, dsts, nthreads: int,
                 l0: bool = False) -> float:
    """Times the reference CPU update path on pre-grouped vertex batches; returns seconds."""
    src = np.ascontiguousarray(src, dtype=np.uint32)
    sizes = np.ascontiguousarray(sizes, dtype=np.uint64)
    offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
    dsts = np.ascontiguousarray(dsts, dtype=np.uint32)
    return float(_load().orc_cpu_baseline(_p(store), g.bkt_per_col, g.num_columns, seed, _p(src),
                                          _p(sizes), _p(offsets), len(src), _p(dsts), nthreads,
                                          int(l0)))


def group_by_vertex(pairs: np.ndarray, n: int, batch_size: int):
    """Splits an undirected update list into pre-grouped vertex batches (both directions), each
    at most ``batch_size`` destinations, the shape GraphSketchDriver hands to
    ``apply_update_batch`` (include/graph_sketch_driver.h:171-172,204-208). The entries of a
    vertex keep the order u->v of every pair, then v->u (orc_group_by_vertex)."""
    pairs = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1, 2)
    m = len(pairs)
    counts = np.zeros(n, dtype=np.uint64)
    dst = np.zeros(2 * m, dtype=np.uint32)
    if _load().orc_group_by_vertex(_p(pairs), m, n, _p(counts), _p(dst)) != 0:
        raise ValueError("vertex id out of range")
    nz = np.nonzero(counts)[0]
    c = counts[nz].astype(np.int64)
    k = (c + batch_size - 1) // batch_size            # batches per vertex
    starts = np.cumsum(c) - c                          # each vertex's first entry
    first = np.cumsum(k) - k                           # each vertex's first batch
    j = np.arange(int(k.sum()), dtype=np.int64) - np.repeat(first, k)  # batch index in vertex
    b_src = np.repeat(nz, k).astype(np.uint32)
    b_off = (np.repeat(starts, k) + j * batch_size).astype(np.uint64)
    b_size = np.minimum(batch_size, np.repeat(c, k) - j * batch_size).astype(np.uint64)
    return b_src, b_size, b_off, dst
