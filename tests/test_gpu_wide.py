"""GPU parity of wide sketches: the CubeSketch (L0_SAMPLING) update rule of src/sketch.cpp:66-83
with 7 columns per sample (include/sketch.h:208-211), and several columns per sample under the
default rule (test/sketch_test.cpp:316-330 builds such a sketch). Every bucket equals the CPU
oracle bit for bit; components equal the oracle's Borůvka and a DSU over the net graph."""
import numpy as np
import pytest

import gpusketch_amd as gs
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def net_edges(pairs):
    keys = (pairs.min(1).astype(np.uint64) << 32) | pairs.max(1)
    uniq, cnt = np.unique(keys, return_counts=True)
    live = uniq[cnt % 2 == 1]
    return np.stack([live >> 32, live & 0xFFFFFFFF], 1).astype(np.uint32)


def make_store(n, seed, cps, l0):
    st = gs.SketchStore(n, seed, cols_per_sample=cps, cubesketch=l0)
    g = orc.Geometry(n, 1.0, cps, l0)
    assert (st.geom.bkt_per_col, st.geom.num_samples, st.geom.num_columns) == (
        g.bkt_per_col, g.num_samples, g.num_columns)
    return st, g


def oracle_store(g, seed, pairs, l0):
    ref = g.empty_store()
    orc.apply_edges(ref, g, seed, pairs, l0=l0)
    return ref


# (columns per sample, CubeSketch rule): 7/L0 (>= 126 columns), 4/point rule (> 64 columns: the
# wide update kernel), 3/point rule (<= 64 columns: the narrow update kernel, wide sampling)
WIDE = [(7, True), (4, False), (3, False)]


@pytest.mark.parametrize("cps,l0", WIDE)
@pytest.mark.parametrize("n,m,seed", [(64, 400, 1), (1000, 30000, 2)])
def test_wide_edge_pairs_match_oracle(cps, l0, n, m, seed):
    rng = np.random.default_rng(seed)
    pairs = rng.integers(0, n, size=(m, 2), dtype=np.uint32)
    key = 0xB0B + seed
    st, g = make_store(n, key, cps, l0)
    with st:
        s = st.stream()
        s.submit_pairs(pairs)
        s.flush()
        got = st.export_buckets()
    assert np.array_equal(got, oracle_store(g, key, pairs, l0))


@pytest.mark.parametrize("lds_depths", [2, 9, 16])
def test_cubesketch_rmat_hubs_match_oracle(lds_depths):
    # hub vertices span many work items; a 2-depth window sends most depths past it (the direct
    # HBM path of the rule, every depth up to d)
    pairs = gs.gen_rmat(11, 60000, 7, del_frac=0.05)
    n = 1 << 11
    st, g = make_store(n, 321, 7, True)
    with st:
        s = st.stream()
        s.set_tuning(lds_depths=lds_depths)
        for b in range(0, len(pairs), 20000):
            s.submit_pairs(pairs[b:b + 20000])
        s.flush()
        got = st.export_buckets()
    assert np.bincount(pairs.ravel(), minlength=n).max() > 1000
    assert np.array_equal(got, oracle_store(g, 321, pairs, True))


@pytest.mark.parametrize("min_depth", [0, 2])
def test_cubesketch_vertex_batches_match_oracle(min_depth):
    rng = np.random.default_rng(30 + min_depth)
    n, seed, cap = 900, 4141, 20
    pairs = rng.integers(0, n, size=(20000, 2), dtype=np.uint32)
    pairs = pairs[pairs[:, 0] != pairs[:, 1]]
    st, g = make_store(n, seed, 7, True)
    src, sizes, offs, dsts = orc.group_by_vertex(pairs, n, 700)
    keep = np.ones(len(dsts), dtype=bool)
    if min_depth:
        for b in range(len(src)):
            for i in range(int(offs[b]), int(offs[b] + sizes[b])):
                idx = orc.concat_pair(int(src[b]), int(dsts[i]))
                keep[i] = orc.index_depth(idx, seed, cap) >= min_depth
        assert 0 < keep.sum() < len(dsts)
    fsrc, fsizes, foffs, fdsts = [], [], [], []
    for b in range(len(src)):
        lo, hi = int(offs[b]), int(offs[b] + sizes[b])
        sel = dsts[lo:hi][keep[lo:hi]]
        fsrc.append(src[b]); fsizes.append(len(sel)); foffs.append(len(fdsts)); fdsts.extend(sel)
    with st:
        s = st.stream()
        s.submit_vertex_batches(src, sizes, offs, dsts, min_depth=min_depth, depth_cap=cap)
        s.flush()
        got = st.export_buckets()
    ref = g.empty_store()
    orc.apply_vertex_batches(ref, g, seed, fsrc, fsizes, foffs, np.array(fdsts, dtype=np.uint32),
                             l0=True)
    assert np.array_equal(got, ref)


def test_cubesketch_streams_import_and_toggle_identity():
    rng = np.random.default_rng(44)
    n = 700
    pairs = rng.integers(0, n, size=(25000, 2), dtype=np.uint32)
    st, g = make_store(n, 55, 7, True)
    ref = oracle_store(g, 55, pairs, True)
    with st:
        s1, s2 = st.stream(), st.stream()
        s1.submit_pairs(pairs[:12000])
        s2.submit_pairs(pairs[12000:])
        st.sync()
        assert np.array_equal(st.export_buckets(), ref)
        st.zero()
        st.import_buckets(ref)
        s1.submit_pairs(pairs)  # the same stream again: every bucket back to zero
        s1.flush()
        assert not st.export_buckets().any()


@pytest.mark.parametrize("cps,l0", WIDE)
def test_wide_cc_matches_oracle_boruvka_and_dsu(cps, l0):
    n = 3000
    pairs = gs.gen_erdos_renyi(n, 1.3 / n, 0.1, 5 + cps)
    key = 777 + cps
    st, g = make_store(n, key, cps, l0)
    ref = oracle_store(g, key, pairs, l0)
    rc, olab, orounds, _ = orc.boruvka(ref, g, key)
    assert rc == 0
    with st:
        s = st.stream()
        s.submit_pairs(pairs)
        s.flush()
        labels, ncc, rounds, forest = st.query_cc(want_forest=True)
    net = net_edges(pairs)
    assert np.array_equal(labels, olab)
    assert rounds == orounds
    assert np.array_equal(labels, orc.dsu_components(n, net))
    netset = set(map(tuple, net.tolist()))
    assert all(tuple(e) in netset for e in forest.tolist())
    assert len(forest) == n - ncc


def test_cubesketch_cc_giant_split_and_direct_paths_agree():
    # a giant component: pair-fed stores form its sum from the vertices outside it; vertex-batch
    # fed stores accumulate it directly; both give the DSU partition
    n = 20000
    pairs = gs.gen_erdos_renyi(n, 3.0 / n, 0.05, 91)
    net = net_edges(pairs)
    expect = orc.dsu_components(n, net)
    assert np.bincount(expect).max() > n // 2
    st, g = make_store(n, 1234, 7, True)
    with st:
        s = st.stream()
        s.submit_pairs(pairs)
        s.flush()
        lab_pairs, _, _ = st.query_cc()
    src, sizes, offs, dsts = orc.group_by_vertex(pairs, n, 4096)  # both directions
    st2, _ = make_store(n, 1234, 7, True)
    with st2:
        s = st2.stream()
        s.submit_vertex_batches(src, sizes, offs, dsts)
        s.flush()
        lab_batches, _, _ = st2.query_cc()
    assert np.array_equal(lab_pairs, expect)
    assert np.array_equal(lab_batches, expect)


def test_cubesketch_k_forests_keep_planted_cut():
    n = 300
    pairs = gs.gen_planted_cut(n, 0.2, 3, 0.1, 23)
    net = net_edges(pairs)
    assert orc.stoer_wagner(n, net)[0] == 3
    st, g = make_store(n, 0xCAFE, 7, True)
    with st:
        s = st.stream()
        s.submit_pairs(pairs)
        s.flush()
        before = st.export_buckets()
        forests = st.k_spanning_forests(5)
        assert np.array_equal(st.export_buckets(), before)
    netset = set(map(tuple, net.tolist()))
    seen = set()
    for f in forests:
        fs = set(map(tuple, f.tolist()))
        assert fs <= netset and not (fs & seen)
        seen |= fs
    cert = np.concatenate(forests)
    assert gs.min_cut(n, cert)[0] == 3


def test_wide_store_rejected_by_groups():
    with gs.SketchStore(500, 1) as a, gs.SketchStore(500, 2, cubesketch=True) as b:
        with pytest.raises(gs.GsError):
            gs.StoreGroup([a, b])


def test_cubesketch_full_size_sampled_vertices_bit_exact():
    """n = 2^17 under the CubeSketch rule (B=34, S=30, C=210: a 15 GB store) after 2^22 R-MAT
    updates in 2^20-edge batches: sampled vertices bit-exact against the oracle, then the toggle
    identity on the first vertices."""
    n = 1 << 17
    pairs = gs.gen_rmat(17, 1 << 22, 99)
    rng = np.random.default_rng(3)
    sample = np.unique(np.concatenate([[0, 1, n - 1], rng.integers(0, n, size=24)]))
    st, g = make_store(n, 0x5151, 7, True)
    assert (g.bkt_per_col, g.num_columns) == (34, 210)
    with st:
        s = st.stream()
        dpairs = gs.DeviceBuffer(pairs)
        for b in range(4):
            s.submit_pairs_device(dpairs.offset(b * (1 << 20) * 8), 1 << 20)
        s.flush()
        got = np.concatenate([st.export_buckets(int(v), 1) for v in sample])
        for b in range(4):
            s.submit_pairs_device(dpairs.offset(b * (1 << 20) * 8), 1 << 20)
        s.flush()
        again = st.export_buckets(0, 256)
        dpairs.free()
    assert not again.any()
    for i, v in enumerate(sample.tolist()):
        others = np.concatenate([pairs[pairs[:, 0] == v, 1], pairs[pairs[:, 1] == v, 0]])
        others = others[others != v]
        assert np.array_equal(got[i], orc.vertex_sketch(g, 0x5151, v, others, l0=True)), v
