"""The allocation-free update (gs_apply_edge_pairs_device_noalloc) against the allocating one
(gs_submit_edge_pairs_device + gs_flush): the same sketches bit for bit, on two calls in a row
(the second smaller, so stale scratch from the first would show), and both equal to the oracle.
Sizes past the fixed scratch are refused before any device work."""
import numpy as np
import pytest

import gpusketch_amd as gs
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def oracle_store(n, seed, pairs):
    g = orc.Geometry(n)
    st = g.empty_store()
    orc.apply_edges(st, g, seed, pairs)
    return st


def test_noalloc_matches_allocating_path_two_calls():
    n, seed = 1 << 15, 0xA110C
    first = gs.gen_rmat(15, gs.NOALLOC_MAX_EDGES, 77, del_frac=0.05)  # the largest call allowed
    rng = np.random.default_rng(5)
    second = rng.integers(0, n, size=(90000, 2), dtype=np.uint32)
    second[::97, 1] = second[::97, 0]  # self loops: no-ops on both paths
    with gs.SketchStore(n, seed) as plain, gs.SketchStore(n, seed) as fixed:
        bufs = [gs.DeviceBuffer(first), gs.DeviceBuffer(second)]
        s = plain.stream()
        for buf, pairs in zip(bufs, (first, second)):
            s.submit_pairs_device(buf.ptr, len(pairs))
            s.flush()
            fixed.apply_pairs_device_noalloc(buf.ptr, len(pairs))
            assert np.array_equal(fixed.export_buckets(), plain.export_buckets())
        got = fixed.export_buckets()
        for b in bufs:
            b.free()
        s.close()
    keep = second[second[:, 0] != second[:, 1]]
    assert np.array_equal(got, oracle_store(n, seed, np.concatenate([first, keep])))


def test_noalloc_refusals_and_dropped_pairs():
    n, seed = 5000, 99
    good = np.random.default_rng(6).integers(0, n, size=(3000, 2), dtype=np.uint32)
    with gs.SketchStore(n, seed) as st:
        bad = good.copy()
        bad[10] = (n, 1)
        buf = gs.DeviceBuffer(bad)
        with pytest.raises(gs.GsError, match="exceed the limit"):  # refused before reading it
            st.apply_pairs_device_noalloc(buf.ptr, gs.NOALLOC_MAX_EDGES + 1)
        assert not st.export_buckets().any()
        with pytest.raises(gs.GsError, match="1 submitted updates"):
            st.apply_pairs_device_noalloc(buf.ptr, len(bad))
        keep = np.delete(bad, 10, axis=0)
        assert np.array_equal(st.export_buckets(), oracle_store(n, seed, keep))
        # a stream of the store holding pending updates: refused, nothing applied
        s = st.stream()
        s.submit_pairs(good)
        with pytest.raises(gs.GsError, match="gs_sync first"):
            st.apply_pairs_device_noalloc(buf.ptr, len(bad))
        st.sync()
        assert np.array_equal(st.export_buckets(), oracle_store(n, seed, np.concatenate([keep, good])))
        s.close()
        buf.free()
