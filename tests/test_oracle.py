"""Pins the CPU oracle (oracle/oracle.c) against golden vectors, closed forms, brute force and
the reference's own fixtures and known answers. CPU only."""
import itertools
import json
import os
import random
import struct

import numpy as np
import pytest

from oracle import oracle as orc
from tests import refpy

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
M64 = (1 << 64) - 1


def load_fixture_stream():
    """test/res/multiples_graph_1024_stream.data: 12-byte header {u32 n, u64 m}, then 9-byte
    records {u8 type, u32 src, u32 dst}."""
    raw = open(os.path.join(GOLD, "multiples_graph_1024_stream.data"), "rb").read()
    n, m = struct.unpack_from("<IQ", raw, 0)
    rec = np.frombuffer(raw, dtype=np.dtype([("t", "u1"), ("s", "<u4"), ("d", "<u4")]),
                        count=m, offset=12)
    return n, rec


def load_fixture_graph():
    lines = open(os.path.join(GOLD, "multiples_graph_1024.txt")).read().split("\n")
    n, m = map(int, lines[0].split())
    edges = np.array([list(map(int, l.split())) for l in lines[1:1 + m]], dtype=np.uint32)
    return n, edges


# ---------------------------------------------------------------------------------------------
# a3: XXH3 golden vectors from the xxhash library
# ---------------------------------------------------------------------------------------------
def test_xxh3_golden_vectors():
    vec = json.load(open(os.path.join(GOLD, "xxh3_len8.json")))["vectors"]
    assert len(vec) > 500
    for k, s, h in vec:
        assert orc.xxh3_len8(int(k), int(s)) == int(h), (k, s)
        assert refpy.xxh3_len8(int(k), int(s)) == int(h), (k, s)


def test_xxh3_seed_sensitivity_and_key_sensitivity():
    # a dropped term in the seed path or a swapped word order would make these collide
    assert orc.xxh3_len8(1, 0) != orc.xxh3_len8(1 << 32, 0)
    assert orc.xxh3_len8(7, 1) != orc.xxh3_len8(7, 1 << 32)


# ---------------------------------------------------------------------------------------------
# a1: geometry
# ---------------------------------------------------------------------------------------------
def test_geometry_table():
    for g in json.load(open(os.path.join(GOLD, "geometry.json"))):
        o = orc.Geometry(g["n"], g["factor"], g["cols_per_sample"], g["l0"])
        assert (o.vector_len, o.bkt_per_col, o.num_samples, o.num_columns, o.num_buckets) == (
            g["vector_len"], g["bkt_per_col"], g["num_samples"], g["num_columns"],
            g["num_buckets"]), g


def test_geometry_reference_configs():
    g = orc.Geometry(8192)
    assert (g.bkt_per_col, g.num_columns, g.num_buckets) == (26, 18, 469)
    g = orc.Geometry(1 << 17)
    assert (g.bkt_per_col, g.num_columns, g.num_buckets) == (34, 24, 817)
    assert g.num_buckets * 16 == 13072
    g = orc.Geometry(1 << 15, factor=2.7)
    assert (g.bkt_per_col, g.num_columns) == (30, 55)
    # bkt_per_col avoids multiples of 16 (include/sketch.h:198-206): n=2^16 gives 32 -> 33
    assert orc.Geometry(1 << 16).bkt_per_col == 33


# ---------------------------------------------------------------------------------------------
# a2, a4: pairing and depth
# ---------------------------------------------------------------------------------------------
def test_concat_pairing_inverse():
    rng = random.Random(3)
    for _ in range(2000):
        i, j = rng.getrandbits(32), rng.getrandbits(32)
        idx = orc.concat_pair(i, j)
        assert idx == orc.concat_pair(j, i)
        assert (idx >> 32, idx & 0xFFFFFFFF) == (min(i, j), max(i, j))


def test_index_depth_closed_form():
    rng = random.Random(4)
    for _ in range(3000):
        idx, seed = rng.getrandbits(64), rng.getrandbits(64)
        bpc = rng.randrange(1, 40)
        h = refpy.xxh3_len8(idx, seed)
        expect = min(bpc, (h & -h).bit_length() - 1) if h else bpc
        assert orc.index_depth(idx, seed, bpc) == expect


# ---------------------------------------------------------------------------------------------
# a6-a8: update rule, invariants
# ---------------------------------------------------------------------------------------------
def test_update_matches_pure_python():
    rng = random.Random(5)
    g = orc.Geometry(300)
    seed = rng.getrandbits(64)
    bk = np.zeros((g.num_buckets, 2), dtype=np.uint64)
    ref = [[0, 0] for _ in range(g.num_buckets)]
    for _ in range(200):
        u, v = rng.sample(range(300), 2)
        idx = refpy.concat(u, v)
        orc.sketch_update(bk, g, seed, idx)
        refpy.update(ref, g.bkt_per_col, g.num_columns, seed, idx)
    assert bk.tolist() == ref


def test_update_twice_is_identity_and_order_free():
    rng = np.random.default_rng(6)
    g = orc.Geometry(2000)
    seed = 0x1234
    pairs = rng.integers(0, 2000, size=(3000, 2), dtype=np.uint32)
    pairs = pairs[pairs[:, 0] != pairs[:, 1]]
    a = g.empty_store()
    orc.apply_edges(a, g, seed, pairs)
    b = g.empty_store()
    orc.apply_edges(b, g, seed, pairs[rng.permutation(len(pairs))])
    assert np.array_equal(a, b)
    orc.apply_edges(a, g, seed, pairs[::-1])
    assert not a.any()


def test_merge_linearity():
    # sketch(A) xor sketch(B) == sketch(A xor B) (src/sketch.cpp:165-202)
    rng = np.random.default_rng(7)
    g = orc.Geometry(500)
    A = rng.integers(0, 500, size=(800, 2), dtype=np.uint32)
    B = rng.integers(0, 500, size=(800, 2), dtype=np.uint32)
    A, B = A[A[:, 0] != A[:, 1]], B[B[:, 0] != B[:, 1]]
    sa, sb, sab = g.empty_store(), g.empty_store(), g.empty_store()
    orc.apply_edges(sa, g, 9, A)
    orc.apply_edges(sb, g, 9, B)
    orc.apply_edges(sab, g, 9, np.concatenate([A, B]))
    assert np.array_equal(sa ^ sb, sab)


def test_vertex_batches_equal_edge_updates():
    rng = np.random.default_rng(8)
    n = 700
    g = orc.Geometry(n)
    pairs = rng.integers(0, n, size=(5000, 2), dtype=np.uint32)
    pairs = pairs[pairs[:, 0] != pairs[:, 1]]
    a = g.empty_store()
    orc.apply_edges(a, g, 11, pairs)
    src, sizes, offs, dsts = orc.group_by_vertex(pairs, n, 37)
    b = g.empty_store()
    orc.apply_vertex_batches(b, g, 11, src, sizes, offs, dsts)
    assert np.array_equal(a, b)


def test_raw_bucket_merge_twice_gives_zero():
    # test/sketch_test.cpp:423-458
    g = orc.Geometry.explicit(1, orc.Geometry(4096).bkt_per_col, 1)
    seed = 77
    sk = np.zeros((g.num_buckets, 2), dtype=np.uint64)
    for i in range(1024):
        orc.sketch_update(sk, g, seed, i)
    other = np.zeros_like(sk)
    other ^= sk
    idx, res = orc.sample(other, g, seed)
    assert res != orc.ZERO
    if res == orc.GOOD:
        assert 0 <= idx < 1024
    other ^= sk
    assert orc.sample(other, g, seed)[1] == orc.ZERO


# ---------------------------------------------------------------------------------------------
# a9: sampling
# ---------------------------------------------------------------------------------------------
def test_index_zero_three_times_is_good_zero():
    # test/sketch_test.cpp:66-80: Sketch(40, seed, 1 sample, 7 columns), update(0) x3
    g = orc.Geometry.explicit(1, 7, 1, 7)  # bkt_per_col = ceil(log2 40) + 1 = 7
    for seed in (0, 1, 12345, M64):
        sk = np.zeros((g.num_buckets, 2), dtype=np.uint64)
        for _ in range(3):
            orc.sketch_update(sk, g, seed, 0)
        assert orc.sample(sk, g, seed) == (0, orc.GOOD)


def test_sample_zero_and_out_of_samples():
    g = orc.Geometry.explicit(1, 5, 1, 7)  # Sketch(10, seed, 1, 7): bkt_per_col = ceil(log2 10)+1
    sk = np.zeros((g.num_buckets, 2), dtype=np.uint64)
    assert orc.sample(sk, g, 3)[1] == orc.ZERO
    orc.sketch_update(sk, g, 3, 1)
    assert orc.sample(sk, g, 3, sample_idx=1)[1] == orc.OUT_OF_SAMPLES


def crafted_fail_sketch():
    """test/sketch_test.cpp:14-64: a vector whose updates only reach depths <= 1 with >= 2 at
    depth 1 in every column, so that no bucket is good; returns (sketch buckets, geometry, seed)."""
    g = orc.Geometry.explicit(1, 8, 1, 7)  # Sketch(100, seed=0, 1, 7)
    seed, cols, guesses = 0, 7, 8
    vec_idx = [True] * (100 * 100)
    total = 2
    i = 0
    while i < cols:
        d1 = k = u = 0
        restart = False
        while u < total or d1 < 2:
            if not vec_idx[k]:
                k += 1
                continue
            dep = orc.index_depth(k, seed + 5 * i, guesses)
            if dep >= 2:
                vec_idx[k] = False
                i = 0
                restart = True
                break
            if dep == 1:
                d1 += 1
            u += 1
            k += 1
        if restart:
            continue
        if u > total:
            total = u
            i = 0
        elif u == total:
            i += 1
    sk = np.zeros((g.num_buckets, 2), dtype=np.uint64)
    applied = 0
    for x in range(len(vec_idx)):
        if vec_idx[x]:
            orc.sketch_update(sk, g, seed, x)
            applied += 1
            if applied >= total:
                break
    return sk, g, seed


def test_sample_crafted_fail_vector():
    sk, g, seed = crafted_fail_sketch()
    assert sk.any()
    assert orc.sample(sk, g, seed)[1] == orc.FAIL


def test_sample_brute_force_small_vectors():
    rng = random.Random(12)
    g = orc.Geometry.explicit(1, 6, 3, 1)
    for _ in range(300):
        seed = rng.getrandbits(64)
        live = set(rng.sample(range(1, 40), rng.randrange(0, 6)))
        sk = np.zeros((g.num_buckets, 2), dtype=np.uint64)
        ref = [[0, 0] for _ in range(g.num_buckets)]
        for x in live:
            orc.sketch_update(sk, g, seed, x)
            refpy.update(ref, g.bkt_per_col, g.num_columns, seed, x)
        for s in range(g.num_samples):
            idx, res = orc.sample(sk, g, seed, s)
            ridx, rres = refpy.sample(ref, g.bkt_per_col, 1, seed, s)
            assert (idx, {0: "GOOD", 1: "ZERO", 2: "FAIL"}[res]) == (ridx, rres)
            if res == orc.GOOD:
                assert idx in live
            assert (res == orc.ZERO) == (len(live) == 0)


# ---------------------------------------------------------------------------------------------
# CubeSketch (L0_SAMPLING) rule, src/sketch.cpp:66-83, and several columns per sample
# ---------------------------------------------------------------------------------------------
def test_l0_geometry_closed_form():
    # include/sketch.h:208-211: 7 columns per sample, num_samples_div = log2(3) - 1
    g = orc.Geometry(1 << 17, 1.0, 7, True)
    assert (g.bkt_per_col, g.num_samples, g.num_columns) == (34, 30, 210)
    g = orc.Geometry(8192, 1.0, 7, True)
    assert (g.bkt_per_col, g.num_samples, g.num_columns) == (26, 23, 161)
    g = orc.Geometry(300, 1.0, 7, True)  # ceil(log2(300) / 0.585) = 15 < 18
    assert (g.num_samples, g.num_columns) == (18, 126)


def test_l0_update_matches_pure_python():
    rng = random.Random(15)
    g = orc.Geometry(300, 1.0, 7, True)
    seed = rng.getrandbits(64)
    bk = np.zeros((g.num_buckets, 2), dtype=np.uint64)
    ref = [[0, 0] for _ in range(g.num_buckets)]
    for _ in range(150):
        u, v = rng.sample(range(300), 2)
        idx = refpy.concat(u, v)
        orc.sketch_update(bk, g, seed, idx, l0=True)
        refpy.update(ref, g.bkt_per_col, g.num_columns, seed, idx, l0=True)
    assert bk.tolist() == ref


def test_l0_is_the_suffix_xor_of_the_point_rule():
    # depth j of a CubeSketch column holds every update of depth >= j (< B): the XOR of depths
    # j..B-1 of the same column under the support-finding rule; the deterministic bucket agrees
    rng = np.random.default_rng(16)
    n = 400
    g = orc.Geometry(n, 1.0, 7, True)
    pairs = rng.integers(0, n, size=(4000, 2), dtype=np.uint32)
    pairs = pairs[pairs[:, 0] != pairs[:, 1]]
    a, b = g.empty_store(), g.empty_store()
    orc.apply_edges(a, g, 21, pairs, l0=True)
    orc.apply_edges(b, g, 21, pairs, l0=False)
    B, C = g.bkt_per_col, g.num_columns
    pa = a[:, :C * B].reshape(n, C, B, 2)
    pb = b[:, :C * B].reshape(n, C, B, 2)
    suffix = np.bitwise_xor.accumulate(pb[:, :, ::-1], axis=2)[:, :, ::-1]
    assert np.array_equal(pa, suffix)
    assert np.array_equal(a[:, -1], b[:, -1])
    assert pa.any()


def test_l0_sample_brute_force_small_vectors():
    rng = random.Random(17)
    g = orc.Geometry.explicit(1, 6, 3, 7)
    for _ in range(200):
        seed = rng.getrandbits(64)
        live = set(rng.sample(range(1, 60), rng.randrange(0, 8)))
        sk = np.zeros((g.num_buckets, 2), dtype=np.uint64)
        ref = [[0, 0] for _ in range(g.num_buckets)]
        for x in live:
            orc.sketch_update(sk, g, seed, x, l0=True)
            refpy.update(ref, g.bkt_per_col, g.num_columns, seed, x, l0=True)
        assert sk.tolist() == ref
        for s in range(g.num_samples):
            idx, res = orc.sample(sk, g, seed, s)
            ridx, rres = refpy.sample(ref, g.bkt_per_col, 7, seed, s)
            assert (idx, {0: "GOOD", 1: "ZERO", 2: "FAIL"}[res]) == (ridx, rres)
            if res == orc.GOOD:
                assert idx in live
            assert (res == orc.ZERO) == (len(live) == 0)


@pytest.mark.parametrize("cps,l0", [(7, True), (3, False)])
def test_boruvka_several_columns_per_sample_equals_dsu(cps, l0):
    rng = np.random.default_rng(18 + cps)
    n = 1500
    pairs = rng.integers(0, n, size=(1300, 2), dtype=np.uint32)
    pairs = pairs[pairs[:, 0] != pairs[:, 1]]
    dels = pairs[rng.choice(len(pairs), size=len(pairs) // 10, replace=False)]
    stream = np.concatenate([pairs, dels])
    g = orc.Geometry(n, 1.0, cps, l0)
    st = g.empty_store()
    orc.apply_edges(st, g, 99, stream, l0=l0)
    rc, lab, rounds, forest = orc.boruvka(st, g, 99)
    assert rc == 0
    assert np.array_equal(lab, orc.dsu_components(n, _net_edges(stream)))
    assert len(forest) == n - len(np.unique(lab))


# ---------------------------------------------------------------------------------------------
# a19, a20: Borůvka and DSU
# ---------------------------------------------------------------------------------------------
def _net_edges(pairs):
    keys = (pairs.min(1).astype(np.uint64) << 32) | pairs.max(1)
    uniq, cnt = np.unique(keys, return_counts=True)
    live = uniq[cnt % 2 == 1]
    return np.stack([live >> 32, live & 0xFFFFFFFF], 1).astype(np.uint32)


def test_fixture_multiples_graph_has_78_components():
    n, edges = load_fixture_graph()
    lab = orc.dsu_components(n, edges)
    assert len(np.unique(lab)) == 78  # test/cc_alg_test.cpp:56


def test_fixture_stream_net_graph_is_the_text_graph():
    n, rec = load_fixture_stream()
    n2, edges = load_fixture_graph()
    assert n == n2 == 1024
    pairs = np.stack([rec["s"], rec["d"]], 1).astype(np.uint32)
    net = _net_edges(pairs)
    ref = _net_edges(edges)
    assert np.array_equal(net, ref)


def test_boruvka_fixture_stream_78():
    n, rec = load_fixture_stream()
    pairs = np.stack([rec["s"], rec["d"]], 1).astype(np.uint32)
    g = orc.Geometry(n)
    st = g.empty_store()
    orc.apply_edges(st, g, 0xDEADBEEF, pairs)
    rc, lab, rounds, forest = orc.boruvka(st, g, 0xDEADBEEF)
    assert rc == 0
    assert len(np.unique(lab)) == 78
    assert np.array_equal(lab, orc.dsu_components(n, _net_edges(pairs)))
    assert len(forest) == n - 78


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_boruvka_partition_equals_dsu(seed):
    rng = np.random.default_rng(seed)
    n = 3000
    # sparse random graph with many components + a stream with deletions
    m = 2500
    pairs = rng.integers(0, n, size=(m, 2), dtype=np.uint32)
    pairs = pairs[pairs[:, 0] != pairs[:, 1]]
    dels = pairs[rng.choice(len(pairs), size=len(pairs) // 10, replace=False)]
    stream = np.concatenate([pairs, dels])
    g = orc.Geometry(n)
    st = g.empty_store()
    orc.apply_edges(st, g, seed * 1000 + 7, stream)
    rc, lab, rounds, forest = orc.boruvka(st, g, seed * 1000 + 7)
    assert rc == 0
    net = _net_edges(stream)
    assert np.array_equal(lab, orc.dsu_components(n, net))
    # forest edges are live edges and form a spanning forest
    netset = set(map(tuple, net.tolist()))
    assert all(tuple(e) in netset for e in forest.tolist())
    assert len(forest) == n - len(np.unique(lab))


def test_boruvka_out_of_samples_with_tiny_sketch():
    n = 64
    g = orc.Geometry.explicit(n, orc.Geometry(n).bkt_per_col, 1)
    pairs = np.array([[i, i + 1] for i in range(n - 1)], dtype=np.uint32)  # a path needs rounds
    st = g.empty_store()
    orc.apply_edges(st, g, 5, pairs)
    rc, *_ = orc.boruvka(st, g, 5)
    assert rc == orc.OUT_OF_SAMPLES


# ---------------------------------------------------------------------------------------------
# a18: k forests, exact min cut
# ---------------------------------------------------------------------------------------------
def _brute_min_cut(n, pairs):
    best = None
    for mask in range(1, 2 ** (n - 1)):
        side = [(mask >> v) & 1 for v in range(n)]
        c = sum(1 for a, b in pairs if side[a] != side[b])
        best = c if best is None else min(best, c)
    return best


def test_stoer_wagner_closed_forms_and_brute_force():
    cyc = np.array([[i, (i + 1) % 10] for i in range(10)], dtype=np.uint32)
    assert orc.stoer_wagner(10, cyc)[0] == 2
    path = np.array([[i, i + 1] for i in range(9)], dtype=np.uint32)
    assert orc.stoer_wagner(10, path)[0] == 1
    k7 = np.array(list(itertools.combinations(range(7), 2)), dtype=np.uint32)
    assert orc.stoer_wagner(7, k7)[0] == 6
    rng = random.Random(13)
    for _ in range(40):
        n = rng.randrange(3, 11)
        allp = list(itertools.combinations(range(n), 2))
        pairs = rng.sample(allp, rng.randrange(n - 1, len(allp) + 1))
        pairs += rng.sample(allp, 2)  # multi-edges add weight
        val, side = orc.stoer_wagner(n, np.array(pairs, dtype=np.uint32))
        assert val == _brute_min_cut(n, pairs)
        assert 0 < side.sum() < n
        assert sum(1 for a, b in pairs if side[a] != side[b]) == val


def test_k_forests_are_disjoint_forests_of_live_edges():
    rng = np.random.default_rng(21)
    n = 200
    pairs = np.array([e for e in itertools.combinations(range(n), 2) if rng.random() < 0.08],
                     dtype=np.uint32)
    g = orc.Geometry(n)
    st = g.empty_store()
    orc.apply_edges(st, g, 99, pairs)
    before = st.copy()
    rc, forests = orc.k_forests(st, g, 99, 4)
    assert rc == 0
    assert np.array_equal(st, before)  # toggled back in
    netset = set(map(tuple, pairs.tolist()))
    seen = set()
    for f in forests:
        fs = set(map(tuple, f.tolist()))
        assert fs <= netset and not (fs & seen)
        seen |= fs
        lab = orc.dsu_components(n, f)
        assert len(f) == n - len(np.unique(lab))  # acyclic
    # forest 0 spans the graph's components
    assert len(forests[0]) == n - len(np.unique(orc.dsu_components(n, pairs)))


def test_planted_cut_certificate_keeps_small_cut():
    rng = np.random.default_rng(22)
    n, half = 120, 60
    edges = [e for e in itertools.combinations(range(half), 2) if rng.random() < 0.3]
    edges += [(a + half, b + half) for a, b in itertools.combinations(range(half), 2)
              if rng.random() < 0.3]
    cross = set()
    while len(cross) < 3:
        cross.add((int(rng.integers(0, half)), int(rng.integers(half, n))))
    edges += sorted(cross)
    pairs = np.array(edges, dtype=np.uint32)
    assert orc.stoer_wagner(n, pairs)[0] == 3
    g = orc.Geometry(n)
    st = g.empty_store()
    orc.apply_edges(st, g, 5, pairs)
    rc, forests = orc.k_forests(st, g, 5, 5)
    assert rc == 0
    cert = np.concatenate(forests)
    assert orc.stoer_wagner(n, cert)[0] == 3


# ---------------------------------------------------------------------------------------------
# bytes model
# ---------------------------------------------------------------------------------------------
def test_distinct_slots_brute_force():
    rng = np.random.default_rng(30)
    n = 100
    g = orc.Geometry(n)
    pairs = rng.integers(0, n, size=(400, 2), dtype=np.uint32)
    pairs = pairs[pairs[:, 0] != pairs[:, 1]]
    seen = set()
    for u, v in pairs.tolist():
        idx = refpy.concat(u, v)
        for x in (u, v):
            seen.add((x, g.num_buckets - 1))
            for c in range(g.num_columns):
                d = refpy.depth(idx, 33 + 5 * c, g.bkt_per_col)
                if d < g.bkt_per_col:
                    seen.add((x, c * g.bkt_per_col + d))
    assert orc.distinct_slots(g, 33, pairs) == len(seen)


def test_cpu_baseline_matches_direct_updates():
    rng = np.random.default_rng(31)
    n = 400
    g = orc.Geometry(n)
    pairs = rng.integers(0, n, size=(3000, 2), dtype=np.uint32)
    pairs = pairs[pairs[:, 0] != pairs[:, 1]]
    a = g.empty_store()
    orc.apply_edges(a, g, 17, pairs)
    b = g.empty_store()
    src, sizes, offs, dsts = orc.group_by_vertex(pairs, n, 4 * g.num_buckets)
    secs = orc.cpu_baseline(b, g, 17, src, sizes, offs, dsts, 4)
    assert secs > 0
    assert np.array_equal(a, b)
