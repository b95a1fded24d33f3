"""Reference-order transfers at sizes that take several 256 MB chunks: gs_export_buckets of a
whole store equals the oracle's store, and gs_import_buckets over ranges that split chunks
rebuilds it bit-exact (and answers the same connectivity query)."""
import numpy as np
import pytest

import gpusketch_amd as gs
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def test_export_import_multi_chunk_ranges():
    n = 50000  # about 11 KB per vertex: three transfer chunks
    seed = 0x7A5F
    pairs = gs.gen_erdos_renyi(n, 8.0 / n, 0.1, 99)
    g = orc.Geometry(n)
    ref = g.empty_store()
    orc.apply_edges(ref, g, seed, pairs)
    with gs.SketchStore(n, seed) as a, gs.SketchStore(n, seed) as b:
        s = a.stream()
        s.submit_pairs(pairs)
        s.flush()
        got = a.export_buckets()
        assert np.array_equal(got, ref)
        lab_a, ncc_a, _ = a.query_cc()
        # ranges that start and end inside chunks, imported out of order
        cuts = [0, 1, 23457, 30001, 49999, n]
        for lo, hi in reversed(list(zip(cuts[:-1], cuts[1:]))):
            b.import_buckets(got[lo:hi], lo)
        assert np.array_equal(b.export_buckets(), ref)
        # a middle range read back on its own
        part = b.export_buckets(23000, 5000)
        assert np.array_equal(part, ref[23000:28000])
        lab_b, ncc_b, _ = b.query_cc()
    assert ncc_a == ncc_b and np.array_equal(lab_a, lab_b)
    # components of the net graph (an edge present an odd number of times)
    keys = (pairs.min(1).astype(np.uint64) << np.uint64(32)) | pairs.max(1)
    uniq, cnt = np.unique(keys, return_counts=True)
    live = uniq[cnt % 2 == 1]
    net = np.stack([live >> np.uint64(32), live & np.uint64(0xFFFFFFFF)], 1).astype(np.uint32)
    assert np.array_equal(lab_a, orc.dsu_components(n, net))
