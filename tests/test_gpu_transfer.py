"""Reference-order transfers at sizes that take several 256 MB chunks: gs_export_buckets of a
whole store equals the oracle's store, and gs_import_buckets over ranges that split chunks
rebuilds it bit-exact (and answers the same connectivity query). Device-order snapshots
(gs_store_snapshot / gs_store_restore, the serving loop's spill format) restore a store
bit-exact, keep its pair-fed (XOR-to-zero) state, and are refused by a store of another
geometry or seed."""
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


def test_snapshot_restore_bit_exact_and_refusals():
    n, seed = 1 << 15, 0x5A17
    pairs = gs.gen_rmat(15, 400000, 5, del_frac=0.05)
    g = orc.Geometry(n)
    ref = g.empty_store()
    orc.apply_edges(ref, g, seed, pairs)
    with gs.SketchStore(n, seed) as a, gs.SketchStore(n, seed) as b:
        s = a.stream()
        s.submit_pairs(pairs)
        s.flush()
        snap = a.snapshot()
        lab_a, ncc_a, rounds_a = a.query_cc()
        # a snapshot is taken after the pending updates and restores everything in one call
        b.restore(snap)
        assert np.array_equal(b.export_buckets(), ref)
        lab_b, ncc_b, rounds_b = b.query_cc()
        assert ncc_a == ncc_b and rounds_a == rounds_b and np.array_equal(lab_a, lab_b)
        # restoring over different contents replaces them; a second snapshot is byte-identical
        s.submit_pairs(pairs[:1000])
        s.flush()
        a.restore(snap)
        assert np.array_equal(a.snapshot(), snap)
        s.close()
    with gs.SketchStore(n, seed + 1) as other_seed, gs.SketchStore(n + 1, seed) as other_n:
        for st in (other_seed, other_n):
            with pytest.raises(gs.GsError):
                st.restore(snap)
    keys = (pairs.min(1).astype(np.uint64) << np.uint64(32)) | pairs.max(1)
    uniq, cnt = np.unique(keys, return_counts=True)
    live = uniq[cnt % 2 == 1]
    net = np.stack([live >> np.uint64(32), live & np.uint64(0xFFFFFFFF)], 1).astype(np.uint32)
    assert np.array_equal(lab_a, orc.dsu_components(n, net))
