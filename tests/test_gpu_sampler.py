"""The device samplers against Sketch::sample (src/sketch.cpp:110-132) restated by the oracle:
every vertex sketch's sample (GOOD and its index, ZERO, FAIL) as the query's round-0 scan takes
it, compared directly rather than through the partition a query returns (a wrong scan order still
picks a valid edge, so partitions alone would not show it). Stores are planted bucket by bucket:
empty, garbage (gamma not the checksum of alpha) and good buckets at chosen depths, including
columns whose only good bucket lies past the first pass's 16 depths."""
import numpy as np
import pytest

import gpusketch_amd as gs
from oracle import oracle as orc
from tests.test_oracle import crafted_fail_sketch

pytestmark = pytest.mark.gpu
RES = {orc.GOOD: 0, orc.ZERO: 1, orc.FAIL: 2}


def planted_store(rng, g, seed, n):
    """(n, num_buckets, 2) reference-order buckets: per vertex a kind of deterministic bucket
    (empty / good / garbage) and per column empty, garbage or good buckets at random depths."""
    B, C = g.bkt_per_col, g.num_columns
    st = np.zeros((n, g.num_buckets, 2), dtype=np.uint64)

    def good():
        a = int(rng.integers(1, 2**63))
        return a, orc.xxh3_len8(a, seed)

    for v in range(n):
        kind = v % 5  # 0: det empty (ZERO), 1: det good, 2..4: det garbage, columns decide
        if kind == 1:
            st[v, -1] = good()
        elif kind >= 2:
            st[v, -1] = rng.integers(1, 2**63, size=2, dtype=np.uint64)
        for c in range(C):
            mode = rng.integers(0, 6)
            base = c * B
            if mode == 0:
                continue  # empty column: FAIL unless another bucket decides
            ng = int(rng.integers(0, 3))
            garbage = rng.integers(0, B, size=int(rng.integers(0, 6)))
            for d in garbage:
                st[v, base + d] = rng.integers(1, 2**63, size=2, dtype=np.uint64)
            if mode == 5:  # only good buckets deep in the column: the second pass
                depths = rng.integers(min(16, B - 1), B, size=max(1, ng))
                for d in range(min(16, B)):
                    st[v, base + d] = rng.integers(1, 2**63, size=2, dtype=np.uint64)
            else:
                depths = rng.integers(0, B, size=ng)
            for d in depths:
                st[v, base + d] = good()
    return st


def check_samples(store_data, g, seed, dev, sample_idx):
    res, idx = dev.sample_vertices(sample_idx)
    counts = {0: 0, 1: 0, 2: 0}
    for v in range(store_data.shape[0]):
        oidx, ores = orc.sample(store_data[v], g, seed, sample_idx)
        want = RES[ores]
        assert res[v] == want, (v, sample_idx, int(res[v]), want)
        if want == 0:
            assert int(idx[v]) == oidx, (v, sample_idx)
        counts[want] += 1
    return counts


@pytest.mark.parametrize("n", [700, 3000])
def test_round0_sampler_matches_oracle_on_planted_stores(n):
    rng = np.random.default_rng(n)
    seed = 0x5A3B1E + n
    g = orc.Geometry(n)
    data = planted_store(rng, g, seed, n)
    with gs.SketchStore(n, seed) as st:
        assert (st.geom.bkt_per_col, st.geom.num_columns) == (g.bkt_per_col, g.num_columns)
        st.import_buckets(data)
        total = {0: 0, 1: 0, 2: 0}
        for s in (0, 1, g.num_samples // 2, g.num_samples - 1):
            for k, c in check_samples(data, g, seed, st, s).items():
                total[k] += c
        with pytest.raises(gs.OutOfSamplesError):
            st.sample_vertices(g.num_samples)
    assert min(total.values()) > 0  # every outcome exercised


def test_round0_sampler_on_streamed_sketches():
    # sketches built by the device update path itself (no planting): hubs and leaves
    n = 5000
    pairs = gs.gen_rmat(13, 60000, 3, del_frac=0.05)
    pairs = pairs[np.all(pairs < n, axis=1)]
    seed = 777
    g = orc.Geometry(n)
    with gs.SketchStore(n, seed) as st:
        s = st.stream()
        s.submit_pairs(pairs)
        s.flush()
        data = st.export_buckets()
        for smp in (0, 3, g.num_samples - 1):
            check_samples(data, g, seed, st, smp)


@pytest.mark.parametrize("cube", [False, True])
def test_wide_sampler_matches_oracle(cube):
    # 7 columns per sample: the one-vertex-per-wave scan over depth-major slots
    n = 400
    rng = np.random.default_rng(9 + cube)
    seed = 0xC0B3 + cube
    g = orc.Geometry(n, 1.0, 7, cube)
    data = planted_store(rng, g, seed, n)
    with gs.SketchStore(n, seed, cubesketch=cube, cols_per_sample=7) as st:
        assert st.geom.num_columns == g.num_columns
        st.import_buckets(data)
        for smp in (0, 2, g.num_samples - 1):
            check_samples(data, g, seed, st, smp)


def test_crafted_fail_vector_fails_on_device():
    # test/sketch_test.cpp:14-64: Sketch(.., seed 0, 1 sample, 7 columns) with B = 8 depths, whose
    # vector leaves no good bucket; a 16-vertex store has B = 8
    bk, g, seed = crafted_fail_sketch()
    n = 16
    assert gs.geometry_for(n).bkt_per_col == g.bkt_per_col == 8
    with gs.SketchStore(n, seed, num_samples=1, cols_per_sample=7) as st:
        assert st.geom.num_buckets == g.num_buckets
        data = np.zeros((n, g.num_buckets, 2), dtype=np.uint64)
        data[3] = bk
        st.import_buckets(data)
        res, _ = st.sample_vertices(0)
    assert res[3] == 2  # FAIL
    assert all(res[v] == 1 for v in range(n) if v != 3)  # the empty sketches: ZERO
