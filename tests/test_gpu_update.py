"""GPU parity of the sketch update path: every bucket the HIP kernels write must equal the CPU
oracle bit for bit (integer XOR arithmetic, no tolerance)."""
import ctypes
import json
import os

import numpy as np
import pytest

import gpusketch_amd as gs
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def oracle_store(n, seed, pairs):
    g = orc.Geometry(n)
    st = g.empty_store()
    orc.apply_edges(st, g, seed, pairs)
    return st


def test_device_hash_golden_vectors():
    vec = json.load(open(os.path.join(GOLD, "xxh3_len8.json")))["vectors"]
    keys = np.array([int(k) for k, _, _ in vec], dtype=np.uint64)
    seeds = np.array([int(s) for _, s, _ in vec], dtype=np.uint64)
    expect = np.array([int(h) for _, _, h in vec], dtype=np.uint64)
    assert np.array_equal(gs.hash_known_answer(keys, seeds), expect)


def test_device_hash_random_against_oracle():
    rng = np.random.default_rng(1)
    keys = rng.integers(0, 2**63, size=50000, dtype=np.uint64) * 2 + 1
    seeds = rng.integers(0, 2**63, size=50000, dtype=np.uint64)
    out = gs.hash_known_answer(keys, seeds)
    for i in range(0, 50000, 997):
        assert int(out[i]) == orc.xxh3_len8(int(keys[i]), int(seeds[i]))


@pytest.mark.parametrize("n,m,seed", [(64, 300, 1), (1000, 20000, 2), (8192, 150000, 3)])
def test_edge_pairs_match_oracle(n, m, seed):
    rng = np.random.default_rng(seed)
    pairs = rng.integers(0, n, size=(m, 2), dtype=np.uint32)
    key = 0xC0FFEE + seed
    with gs.SketchStore(n, key) as st:
        s = st.stream()
        s.submit_pairs(pairs)
        s.flush()
        got = st.export_buckets()
    assert np.array_equal(got, oracle_store(n, key, pairs))


def test_rmat_hubs_and_deletions_match_oracle():
    # skewed degrees: hub vertices span many work items (atomic flush path)
    pairs = gs.gen_rmat(12, 200000, 5)
    n = 1 << 12
    with gs.SketchStore(n, 77) as st:
        s = st.stream()
        for b in range(0, len(pairs), 65536):  # several device batches
            s.submit_pairs(pairs[b:b + 65536])
        s.flush()
        got = st.export_buckets()
    deg = np.bincount(pairs.ravel(), minlength=n)
    assert deg.max() > 2000
    assert np.array_equal(got, oracle_store(n, 77, pairs))


def test_vertex_batches_match_oracle():
    rng = np.random.default_rng(9)
    n = 2000
    pairs = rng.integers(0, n, size=(40000, 2), dtype=np.uint32)
    pairs = pairs[pairs[:, 0] != pairs[:, 1]]
    g = orc.Geometry(n)
    src, sizes, offs, dsts = orc.group_by_vertex(pairs, n, 4 * g.num_buckets)
    # also a few empty batches, an oversize batch for one vertex and repeated sources
    src = np.concatenate([src, [5, 6]]).astype(np.uint32)
    sizes = np.concatenate([sizes, [0, 0]]).astype(np.uint64)
    offs = np.concatenate([offs, [0, 0]]).astype(np.uint64)
    big = rng.integers(0, n, size=1500).astype(np.uint32)
    big = big[big != 7]
    src = np.concatenate([src, [7]]).astype(np.uint32)
    sizes = np.concatenate([sizes, [len(big)]]).astype(np.uint64)
    offs = np.concatenate([offs, [len(dsts)]]).astype(np.uint64)
    dsts = np.concatenate([dsts, big]).astype(np.uint32)
    with gs.SketchStore(n, 4242) as st:
        s = st.stream()
        s.submit_vertex_batches(src, sizes, offs, dsts)
        s.flush()
        got = st.export_buckets()
    ref = g.empty_store()
    orc.apply_vertex_batches(ref, g, 4242, src, sizes, offs, dsts)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("min_depth", [1, 3])
def test_vertex_batches_subgraph_filter_match_oracle(min_depth):
    # MC routing: a destination reaches subgraph g iff get_index_depth(idx, seed, cap) >= g
    rng = np.random.default_rng(11 + min_depth)
    n, seed, cap = 1500, 909, 20
    pairs = rng.integers(0, n, size=(30000, 2), dtype=np.uint32)
    pairs = pairs[pairs[:, 0] != pairs[:, 1]]
    g = orc.Geometry(n)
    src, sizes, offs, dsts = orc.group_by_vertex(pairs, n, 512)
    keep = np.zeros(len(dsts), dtype=bool)
    for b in range(len(src)):
        for i in range(int(offs[b]), int(offs[b] + sizes[b])):
            idx = orc.concat_pair(int(src[b]), int(dsts[i]))
            keep[i] = orc.index_depth(idx, seed, cap) >= min_depth
    assert 0 < keep.sum() < len(dsts)
    fsrc, fsizes, foffs, fdsts = [], [], [], []
    for b in range(len(src)):
        sel = dsts[int(offs[b]):int(offs[b] + sizes[b])][keep[int(offs[b]):int(offs[b] + sizes[b])]]
        fsrc.append(src[b]); fsizes.append(len(sel)); foffs.append(len(fdsts)); fdsts.extend(sel)
    with gs.SketchStore(n, seed) as st:
        s = st.stream()
        s.submit_vertex_batches(src, sizes, offs, dsts, min_depth=min_depth, depth_cap=cap)
        s.flush()
        got = st.export_buckets()
    ref = g.empty_store()
    orc.apply_vertex_batches(ref, g, seed, fsrc, fsizes, foffs, np.array(fdsts, dtype=np.uint32))
    assert np.array_equal(got, ref)


def test_empty_self_loops_and_toggle_cancel():
    n = 300
    with gs.SketchStore(n, 5) as st:
        s = st.stream()
        s.submit_pairs(np.zeros((0, 2), dtype=np.uint32))
        s.submit_pairs(np.array([[3, 3], [7, 7]], dtype=np.uint32))
        s.flush()
        assert not st.export_buckets().any()
        rng = np.random.default_rng(3)
        pairs = rng.integers(0, n, size=(5000, 2), dtype=np.uint32)
        s.submit_pairs(pairs)
        s.submit_pairs(pairs[::-1])  # every update toggled back out
        s.flush()
        assert not st.export_buckets().any()


def test_out_of_range_vertex_is_dropped_and_reported():
    # host batches (pageable and pinned) behave as device batches: the pairs naming a vertex >= n
    # are skipped, the rest is applied, and the next flush reports the dropped count
    good = np.array([[1, 2], [3, 99], [0, 50]], dtype=np.uint32)
    bad = np.array([[1, 100], [4000000000, 3]], dtype=np.uint32)
    mixed = np.concatenate([good[:2], bad[:1], good[2:], bad[1:]])
    with gs.SketchStore(100, 5) as st:
        s = st.stream()
        s.submit_pairs(mixed)
        with pytest.raises(gs.GsError, match="2 submitted updates"):
            s.flush()
        s.flush()  # reported once
        assert np.array_equal(st.export_buckets(), oracle_store(100, 5, good))
        pin = gs.PinnedBuffer(mixed.shape)
        pin.array[:] = mixed
        s.submit_pairs(pin.array)  # toggles the good pairs back out
        with pytest.raises(gs.GsError, match="dropped"):
            s.flush()
        pin.free()
        assert not st.export_buckets().any()


def test_mixed_submissions_in_one_launch_and_caller_memory_untouched():
    # one launch holds a host batch (copied into the fill buffer), a short device batch (copied)
    # and long device batches (read in place, one of them holding out-of-range pairs); the
    # out-of-range pairs are dropped and reported at the flush, the caller's memory is not written
    rng = np.random.default_rng(13)
    n = 4000
    host = rng.integers(0, n, size=(70000, 2), dtype=np.uint32)
    short = rng.integers(0, n, size=(5000, 2), dtype=np.uint32)
    long1 = rng.integers(0, n, size=(300000, 2), dtype=np.uint32)
    long2 = rng.integers(0, n, size=(280000, 2), dtype=np.uint32)
    bad_rows = rng.choice(len(long2), size=50, replace=False)
    long2_bad = long2.copy()
    long2_bad[bad_rows, 1] = n + 7
    with gs.SketchStore(n, 2024) as st:
        s = st.stream()
        s.set_tuning(launch_edges=1 << 21)  # every submission below lands in one launch
        bufs = [gs.DeviceBuffer(a) for a in (short, long1, long2_bad)]
        s.submit_pairs(host)
        s.submit_pairs_device(bufs[0].offset(0), len(short))
        s.submit_pairs_device(bufs[1].offset(0), len(long1))
        s.submit_pairs_device(bufs[2].offset(0), len(long2_bad))
        with pytest.raises(gs.GsError, match="dropped"):
            s.flush()
        back = np.zeros_like(long2_bad)
        gs.lib.gs_memcpy_d2h(back.ctypes.data_as(ctypes.c_void_p), bufs[2].ptr, back.nbytes)
        got = st.export_buckets()
        for b in bufs:
            b.free()
    assert np.array_equal(back, long2_bad)
    good = np.delete(long2, bad_rows, axis=0)
    assert np.array_equal(got, oracle_store(n, 2024, np.concatenate([host, short, long1, good])))


def test_multiple_streams_share_a_store():
    rng = np.random.default_rng(11)
    n = 1500
    pairs = rng.integers(0, n, size=(60000, 2), dtype=np.uint32)
    with gs.SketchStore(n, 99) as st:
        s1, s2 = st.stream(), st.stream()
        s1.submit_pairs(pairs[:30000])
        s2.submit_pairs(pairs[30000:])
        st.sync()
        got = st.export_buckets()
    assert np.array_equal(got, oracle_store(n, 99, pairs))


def test_export_import_roundtrip():
    rng = np.random.default_rng(12)
    n = 500
    pairs = rng.integers(0, n, size=(8000, 2), dtype=np.uint32)
    ref = oracle_store(n, 31, pairs)
    with gs.SketchStore(n, 31) as st:
        st.import_buckets(ref)
        assert np.array_equal(st.export_buckets(), ref)
        s = st.stream()
        s.submit_pairs(pairs)  # XOR the same stream again: all zero
        s.flush()
        assert not st.export_buckets().any()


def test_full_size_rmat_sampled_vertices_bit_exact():
    """n = 2^17 with the reference geometry (B=34, C=24): the store after 2^22 R-MAT updates in
    2^20-edge batches is bit-exact against the oracle on a sample of vertices (the oracle
    recomputes just their sketches from the edges incident to them), plus the toggle identity
    over the whole store."""
    n = 1 << 17
    pairs = gs.gen_rmat(17, 1 << 22, 20261203)
    rng = np.random.default_rng(0)
    sample = np.unique(np.concatenate([[0, 1, 2, n - 1], rng.integers(0, n, size=60)]))
    with gs.SketchStore(n, 0xABCDEF) as st:
        assert (st.geom.bkt_per_col, st.geom.num_columns) == (34, 24)
        s = st.stream()
        dpairs = gs.DeviceBuffer(pairs)
        for b in range(4):
            s.submit_pairs_device(dpairs.offset(b * (1 << 20) * 8), 1 << 20)
        s.flush()
        got = np.concatenate([st.export_buckets(int(v), 1) for v in sample])
        # toggle everything back out: the whole 1.7 GB store must be zero again
        for b in range(4):
            s.submit_pairs_device(dpairs.offset(b * (1 << 20) * 8), 1 << 20)
        s.flush()
        again = st.export_buckets(0, 4096)
        dpairs.free()
    assert not again.any()
    g = orc.Geometry(n)
    for i, v in enumerate(sample.tolist()):
        others = np.concatenate([pairs[pairs[:, 0] == v, 1], pairs[pairs[:, 1] == v, 0]])
        others = others[others != v]
        assert np.array_equal(got[i], orc.vertex_sketch(g, 0xABCDEF, v, others)), v


@pytest.mark.parametrize("bucket_bits", [0, 13])
def test_large_vertex_ids_both_entry_forms(bucket_bits):
    """n = 2^19 + 7 (20-bit ids): by default 12 local-id bits + 20 id bits fill a packed 32-bit
    grouping entry exactly; 13 local-id bits need the 8-byte entry. Sampled vertices bit-exact
    against the oracle, then the toggle identity."""
    n = (1 << 19) + 7
    rng = np.random.default_rng(70 + bucket_bits)
    pairs = rng.integers(0, n, size=(1 << 20, 2), dtype=np.uint32)
    pairs[:4096, 0] = n - 1  # a hub at the largest id
    sample = np.unique(np.concatenate([[0, 1, n - 2, n - 1], rng.integers(0, n, size=12)]))
    with gs.SketchStore(n, 0x1D5) as st:
        s = st.stream()
        s.set_tuning(bucket_bits=bucket_bits)
        s.submit_pairs(pairs)
        s.flush()
        got = np.concatenate([st.export_buckets(int(v), 1) for v in sample])
        s.submit_pairs(pairs)
        s.flush()
        again = st.export_buckets(n - 64, 64)
    assert not again.any()
    g = orc.Geometry(n)
    for i, v in enumerate(sample.tolist()):
        others = np.concatenate([pairs[pairs[:, 0] == v, 1], pairs[pairs[:, 1] == v, 0]])
        others = others[others != v]
        assert np.array_equal(got[i], orc.vertex_sketch(g, 0x1D5, v, others)), v


@pytest.mark.parametrize("lds_depths,item_updates,bucket_bits", [(8, 128, 4), (16, 512, 6), (12, 64, 0)])
def test_tuned_launch_parameters_match_oracle(lds_depths, item_updates, bucket_bits):
    pairs = gs.gen_rmat(12, 120000, 21, del_frac=0.05)
    n = 1 << 12
    with gs.SketchStore(n, 99) as st:
        s = st.stream()
        s.set_tuning(lds_depths, item_updates, bucket_bits)
        for b in range(0, len(pairs), 40000):
            s.submit_pairs(pairs[b:b + 40000])
        s.flush()
        got = st.export_buckets()
    assert np.array_equal(got, oracle_store(n, 99, pairs))


def test_autotune_leaves_sketches_unchanged_and_keeps_parity():
    pairs = gs.gen_rmat(13, 300000, 33, del_frac=0.05)
    n = 1 << 13
    with gs.SketchStore(n, 5150) as st:
        s = st.stream()
        s.submit_pairs(pairs[:100000])
        s.flush()
        before = st.export_buckets()
        buf = gs.DeviceBuffer(pairs[100000:200000])
        best = s.autotune(buf.offset(0), 100000)
        s.flush()
        assert np.array_equal(st.export_buckets(), before)  # every candidate applied twice
        assert best.item_updates in (256, 512, 1024) and best.lds_depths in (9, 10, 11, 12)
        t = s.get_tuning()
        assert (t.lds_depths, t.item_updates, t.bucket_bits) == (
            best.lds_depths, best.item_updates, best.bucket_bits)
        s.submit_pairs_device(buf.offset(0), 100000)
        s.submit_pairs(pairs[200000:])
        s.flush()
        buf.free()
        got = st.export_buckets()
    assert np.array_equal(got, oracle_store(n, 5150, pairs))
