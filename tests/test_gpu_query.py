"""GPU parity of the query path: device Borůvka components equal the oracle's Borůvka and a DSU
over the net graph; spanning forests are valid; k disjoint forests preserve the planted min cut,
which equals exact Stoer–Wagner."""
import os
import struct

import numpy as np
import pytest

import gpusketch_amd as gs
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def net_edges(pairs):
    keys = (pairs.min(1).astype(np.uint64) << 32) | pairs.max(1)
    uniq, cnt = np.unique(keys, return_counts=True)
    live = uniq[cnt % 2 == 1]
    return np.stack([live >> 32, live & 0xFFFFFFFF], 1).astype(np.uint32)


def edge_keys(e):
    e = np.asarray(e, dtype=np.uint32).reshape(-1, 2)
    lo, hi = e.min(1).astype(np.uint64), e.max(1).astype(np.uint64)
    return (lo << np.uint64(32)) | hi


def check_disjoint_forests(n, forests, net):
    """Forest i must be a spanning forest of the net graph minus forests 0..i-1
    (MCSketchAlg::calc_disjoint_spanning_forests, src/mc_sketch_alg.cpp:889-920): its edges lie
    in that graph, it is acyclic, and its components are that graph's components."""
    remaining = np.unique(edge_keys(net))  # sorted
    for i, f in enumerate(forests):
        fk = np.sort(edge_keys(f))
        assert len(np.unique(fk)) == len(fk), f"forest {i} repeats an edge"
        pos = np.searchsorted(remaining, fk)
        inside = (pos < len(remaining)) & (remaining[np.minimum(pos, len(remaining) - 1)] == fk)
        assert inside.all(), f"forest {i} holds an edge outside G - F_0..F_{i-1}"
        g_i = np.stack([remaining >> np.uint64(32), remaining & np.uint64(0xFFFFFFFF)], 1).astype(np.uint32)
        lab_f = orc.dsu_components(n, f)
        assert len(f) == n - len(np.unique(lab_f)), f"forest {i} has a cycle"
        assert np.array_equal(lab_f, orc.dsu_components(n, g_i)), f"forest {i} does not span"
        remaining = np.delete(remaining, pos)


def check_forest(n, forest, net, labels):
    netset = set(map(tuple, net.tolist()))
    assert all(tuple(e) in netset for e in forest.tolist())
    assert len(forest) == n - len(np.unique(labels))
    assert np.array_equal(orc.dsu_components(n, forest), labels)


def test_fixture_stream_78_components():
    raw = open(os.path.join(GOLD, "multiples_graph_1024_stream.data"), "rb").read()
    n, m = struct.unpack_from("<IQ", raw, 0)
    rec = np.frombuffer(raw, dtype=np.dtype([("t", "u1"), ("s", "<u4"), ("d", "<u4")]),
                        count=m, offset=12)
    pairs = np.stack([rec["s"], rec["d"]], 1).astype(np.uint32)
    with gs.SketchStore(n, 0x5EED) as st:
        s = st.stream()
        s.submit_pairs(pairs)
        s.flush()
        labels, ncc, rounds, forest = st.query_cc(want_forest=True)
    assert ncc == 78  # test/cc_alg_test.cpp:56
    net = net_edges(pairs)
    assert np.array_equal(labels, orc.dsu_components(n, net))
    check_forest(n, forest, net, labels)


@pytest.mark.parametrize("seed", [1, 2])
def test_cc_matches_oracle_boruvka_and_dsu(seed):
    n = 8192 if seed == 1 else 3000
    pairs = gs.gen_erdos_renyi(n, 2.0 / n if seed == 1 else 1.2 / n, 0.1, seed)
    key = 1000 + seed
    g = orc.Geometry(n)
    ref = g.empty_store()
    orc.apply_edges(ref, g, key, pairs)
    rc, olab, orounds, _ = orc.boruvka(ref, g, key)
    assert rc == 0
    with gs.SketchStore(n, key) as st:
        s = st.stream()
        s.submit_pairs(pairs)
        s.flush()
        labels, ncc, rounds, forest = st.query_cc(want_forest=True)
        # a second query on the same sketches gives the same answer
        labels2, _, rounds2 = st.query_cc()
    net = net_edges(pairs)
    assert np.array_equal(labels, olab)
    assert rounds == orounds
    assert np.array_equal(labels, orc.dsu_components(n, net))
    assert np.array_equal(labels2, labels) and rounds2 == rounds
    check_forest(n, forest, net, labels)


def test_cc_giant_split_and_direct_paths_agree():
    # A giant component forms after a few rounds. Pairs-only stores (and whole-store imports
    # whose vertex sketches XOR to zero) form the giant supernode's sum from everything outside
    # it; a store filled by partial imports cannot know its sketches XOR to zero and
    # accumulates every supernode directly. Both equal the oracle's Borůvka, round for round.
    n = 6000
    pairs = gs.gen_erdos_renyi(n, 3.0 / n, 0.05, 77)
    key = 4242
    g = orc.Geometry(n)
    ref = g.empty_store()
    orc.apply_edges(ref, g, key, pairs)
    rc, olab, orounds, _ = orc.boruvka(ref, g, key)
    assert rc == 0
    with gs.SketchStore(n, key) as a, gs.SketchStore(n, key) as b, gs.SketchStore(n, key) as c:
        s = a.stream()
        s.submit_pairs(pairs)
        s.flush()
        lab_a, _, rounds_a = a.query_cc()
        data = a.export_buckets()
        b.import_buckets(data[: n // 2], 0)
        b.import_buckets(data[n // 2:], n // 2)
        lab_b, _, rounds_b = b.query_cc()
        c.import_buckets(data, 0)
        lab_c, _, rounds_c = c.query_cc()
    for lab, rounds in ((lab_a, rounds_a), (lab_b, rounds_b), (lab_c, rounds_c)):
        assert np.array_equal(lab, olab)
        assert rounds == orounds
    assert np.array_equal(lab_a, orc.dsu_components(n, net_edges(pairs)))


def test_cc_config1_shape_reduced():
    # the ER(8192, 0.25) config at n=1024 (same density, seeded order, 10% re-deleted)
    n = 1024
    pairs = gs.gen_erdos_renyi(n, 0.25, 0.1, 20261203)
    with gs.SketchStore(n, 8) as st:
        s = st.stream()
        s.submit_pairs(pairs)
        s.flush()
        got = st.export_buckets()
        labels, ncc, rounds = st.query_cc()
    g = orc.Geometry(n)
    ref = g.empty_store()
    orc.apply_edges(ref, g, 8, pairs)
    assert np.array_equal(got, ref)
    assert ncc == 1
    assert np.array_equal(labels, orc.dsu_components(n, net_edges(pairs)))


def test_cc_config1_full_size():
    # the plain-CC config at full size: ER(8192, 0.25) in seeded order, 10% of the inserted
    # edges re-deleted later (about 9.2 M updates); every bucket bit-exact, components vs DSU
    n = 8192
    pairs = gs.gen_erdos_renyi(n, 0.25, 0.1, 20261204)
    assert len(pairs) > 9_000_000
    with gs.SketchStore(n, 0xC0FFEE) as st:
        s = st.stream()
        for b in range(0, len(pairs), 1 << 20):
            s.submit_pairs(pairs[b:b + (1 << 20)])
        s.flush()
        got = st.export_buckets()
        labels, ncc, rounds = st.query_cc()
    g = orc.Geometry(n)
    assert (g.bkt_per_col, g.num_columns, g.num_buckets) == (26, 18, 469)
    ref = g.empty_store()
    orc.apply_edges(ref, g, 0xC0FFEE, pairs)
    assert np.array_equal(got, ref)
    assert ncc == 1
    assert np.array_equal(labels, orc.dsu_components(n, net_edges(pairs)))


def test_cc_hubs_sampled_past_the_shallow_depths():
    # two hubs of 2^17 leaves each, joined by one edge, leaves also in short paths: a hub's first
    # good bucket in round 0 lies at depth ~17, past the 16 shallow depths of the first read
    # pass, so its sample comes from the deep pass; every forest edge must be a live edge
    half = 1 << 17
    n = 2 * half + 2
    leaves_a = np.arange(2, 2 + half, dtype=np.uint32)
    leaves_b = np.arange(2 + half, n, dtype=np.uint32)
    pairs = np.concatenate([
        np.stack([np.zeros(half, np.uint32), leaves_a], 1),
        np.stack([np.ones(half, np.uint32), leaves_b], 1),
        np.array([[0, 1]], dtype=np.uint32),
        np.stack([leaves_a[:-1:2], leaves_a[1::2]], 1)[:1000],
    ]).astype(np.uint32)
    with gs.SketchStore(n, 0xB16) as st:
        assert st.geom.bkt_per_col > 16
        s = st.stream()
        s.submit_pairs(pairs)
        s.flush()
        labels, ncc, rounds, forest = st.query_cc(want_forest=True)
    assert ncc == 1
    net = net_edges(pairs)
    check_forest(n, forest, net, labels)


def test_out_of_samples_reported():
    n = 64
    path = np.array([[i, i + 1] for i in range(n - 1)], dtype=np.uint32)
    with gs.SketchStore(n, 5, num_samples=1) as st:
        s = st.stream()
        s.submit_pairs(path)
        s.flush()
        with pytest.raises(gs.OutOfSamplesError):
            st.query_cc()


def test_k_forests_and_planted_min_cut():
    n = 400
    pairs = gs.gen_planted_cut(n, 0.15, 3, 0.1, 17)
    net = net_edges(pairs)
    exact, _ = orc.stoer_wagner(n, net)
    assert exact == 3
    k = 5
    with gs.SketchStore(n, 0xFACE) as st:
        s = st.stream()
        s.submit_pairs(pairs)
        s.flush()
        before = st.export_buckets()
        forests = st.k_spanning_forests(k)
        after = st.export_buckets()
    assert np.array_equal(before, after)  # forests toggled back in
    check_disjoint_forests(n, forests, net)
    cert = np.concatenate(forests)
    val, side = gs.min_cut(n, cert)
    assert val == exact
    assert orc.stoer_wagner(n, cert)[0] == exact


def test_cc_rmat17_full_size_matches_dsu():
    # the single-stream config's graph family at its real size: R-MAT n=2^17 (a,b,c,d) =
    # (.57,.19,.19,.05), 2^23 updates (about 5% deletions) in 2^20-pair device batches; the
    # partition must equal a DSU over the net graph and the forest must span it
    n = 1 << 17
    pairs = gs.gen_rmat(17, 1 << 23, 20261205, del_frac=0.05)
    with gs.SketchStore(n, 0x5EEDC0DE) as st:
        s = st.stream()
        buf = gs.DeviceBuffer(pairs)
        for b in range(0, len(pairs), 1 << 20):
            s.submit_pairs_device(buf.offset(b * 8), min(1 << 20, len(pairs) - b))
        s.flush()
        labels, ncc, rounds, forest = st.query_cc(want_forest=True)
        buf.free()
    net = net_edges(pairs)
    assert np.array_equal(labels, orc.dsu_components(n, net))
    assert ncc > 1  # isolated vertices remain at this density
    assert len(forest) == n - ncc
    assert np.isin(edge_keys(forest), edge_keys(net)).all()
    assert np.array_equal(orc.dsu_components(n, forest), labels)


def test_mincut_config_full_shape():
    # the min-cut config as stated: n=2^15, two ER(0.1) halves joined by exactly 5 crossing edges,
    # shuffled, 10% insert/delete churn, k=8 sketch forests at the reference tool's sketch factor
    # (tools/min_cut.cu:95-102,156); the certificate's minimum cut is the planted 5 and its side
    # is one of the planted halves
    n, k, cut = 1 << 15, 8, 5
    pairs = gs.gen_planted_cut(n, 0.1, cut, 0.1, 20261206)
    factor = k / np.log2(n) * 1.5
    with gs.SketchStore(n, 0xFACADE, sketches_factor=factor) as st:
        s = st.stream()
        buf = gs.DeviceBuffer(pairs)
        for b in range(0, len(pairs), 1 << 20):
            s.submit_pairs_device(buf.offset(b * 8), min(1 << 20, len(pairs) - b))
        s.flush()
        buf.free()
        before = st.export_buckets(0, 64)
        forests = st.k_spanning_forests(k)
        assert np.array_equal(st.export_buckets(0, 64), before)  # toggled back in
    net = net_edges(pairs)
    half = np.arange(n) < n // 2
    assert int(np.sum(half[net[:, 0]] != half[net[:, 1]])) == cut
    check_disjoint_forests(n, forests, net)
    cert = np.concatenate(forests)
    val, side = gs.min_cut(n, cert)
    assert val == cut
    side = side.astype(bool)
    assert np.array_equal(side, half) or np.array_equal(side, ~half)
