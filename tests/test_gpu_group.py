"""Batched streams: a group of stores of different sizes and seeds ingested by one launch
sequence and queried by one batched Boruvka. Every member must end bit-exact with the oracle and
with its own single-store run, and every member's partition must equal its single-store query."""
import numpy as np
import pytest

import gpusketch_amd as gs
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

SIZES = [100, 1000, 4096, 777, 20000, 65]


def member_pairs(rng, n, m):
    p = rng.integers(0, n, size=(m, 2), dtype=np.uint32)
    # some toggles twice (a delete) and a few self loops, which the device drops
    p = np.concatenate([p, p[: m // 10], np.array([[1, 1]], dtype=np.uint32)])
    rng.shuffle(p)
    return p


def net_edges(pairs):
    """Edges present an odd number of times (toggle semantics), self loops dropped."""
    p = pairs[pairs[:, 0] != pairs[:, 1]]
    lo, hi = np.minimum(p[:, 0], p[:, 1]), np.maximum(p[:, 0], p[:, 1])
    key = lo.astype(np.uint64) << np.uint64(32) | hi.astype(np.uint64)
    uniq, cnt = np.unique(key, return_counts=True)
    odd = uniq[cnt % 2 == 1]
    return np.stack([(odd >> np.uint64(32)).astype(np.uint32),
                     (odd & np.uint64(0xFFFFFFFF)).astype(np.uint32)], axis=1)


def oracle_store(n, seed, pairs):
    g = orc.Geometry(n)
    st = g.empty_store()
    orc.apply_edges(st, g, seed, pairs[pairs[:, 0] != pairs[:, 1]])
    return st


def test_group_ingest_and_query_match_single_stores():
    rng = np.random.default_rng(41)
    seeds = [1000 + 17 * i for i in range(len(SIZES))]
    stores = [gs.SketchStore(n, s) for n, s in zip(SIZES, seeds)]
    singles = [gs.SketchStore(n, s) for n, s in zip(SIZES, seeds)]
    pairs = [member_pairs(rng, n, 6 * n) for n in SIZES]
    pairs[4] = gs.gen_rmat(14, 150000, 9, del_frac=0.05)  # hubs: the shared-vertex flush path
    pairs[4] = pairs[4][np.all(pairs[4] < SIZES[4], axis=1)]
    try:
        with gs.StoreGroup(stores) as grp:
            half = [p[: len(p) // 2] for p in pairs]
            rest = [p[len(p) // 2:] for p in pairs]
            grp.submit_pairs(half)
            grp.submit_pairs([None, rest[1], rest[2], rest[3], rest[4], rest[5]])
            grp.submit_pairs([rest[0], None, None, None, None, None])
            grp.flush()
            for i, (n, s) in enumerate(zip(SIZES, seeds)):
                got = stores[i].export_buckets()
                assert np.array_equal(got, oracle_store(n, s, pairs[i])), f"member {i}"
                st = singles[i].stream()
                st.submit_pairs(pairs[i])
                st.flush()
                st.close()
            labels, ncc, rounds, status = grp.query_cc()
            for i in range(len(SIZES)):
                lab1, ncc1, r1 = singles[i].query_cc()
                assert np.array_equal(labels[i], lab1), f"member {i}"
                assert ncc[i] == ncc1 and rounds[i] == r1 and status[i] == 0
                assert np.array_equal(labels[i], orc.dsu_components(SIZES[i], net_edges(pairs[i])))
            # more updates, then the captured rounds replay for a second query
            more = [member_pairs(rng, n, n) for n in SIZES]
            grp.submit_pairs(more)
            grp.flush()
            for i in range(len(SIZES)):
                st = singles[i].stream()
                st.submit_pairs(more[i])
                st.flush()
                st.close()
                assert np.array_equal(stores[i].export_buckets(), singles[i].export_buckets())
            labels2, ncc2, _, _ = grp.query_cc()
            for i in range(len(SIZES)):
                lab1, ncc1, _ = singles[i].query_cc()
                assert np.array_equal(labels2[i], lab1) and ncc2[i] == ncc1
            # a subset of members
            mask = np.array([1, 0, 1, 0, 0, 1], dtype=np.uint8)
            labels3, _, _, status3 = grp.query_cc(mask)
            for i in range(len(SIZES)):
                if mask[i]:
                    assert np.array_equal(labels3[i], labels2[i])
    finally:
        for s in stores + singles:
            s.close()


def test_group_device_resident_global_pairs():
    rng = np.random.default_rng(5)
    sizes = [3000, 500, 9000]
    seeds = [7, 8, 9]
    stores = [gs.SketchStore(n, s) for n, s in zip(sizes, seeds)]
    pairs = [member_pairs(rng, n, 5 * n) for n in sizes]
    try:
        with gs.StoreGroup(stores) as grp:
            grp.prepare()  # buffers and query scratch allocated up front; twice is harmless
            grp.prepare()
            glob = np.concatenate([p + np.uint32(b) for p, b in zip(pairs, grp.bases)])
            glob = glob[rng.permutation(len(glob))]  # members interleaved in one stream
            buf = gs.DeviceBuffer(glob)
            grp.submit_pairs_device(buf.offset(0), len(glob))
            grp.flush()
            buf.free()
            for i, (n, s) in enumerate(zip(sizes, seeds)):
                assert np.array_equal(stores[i].export_buckets(), oracle_store(n, s, pairs[i]))
            labels, ncc, _, status = grp.query_cc()
            for i, n in enumerate(sizes):
                assert status[i] == 0
                assert np.array_equal(labels[i], orc.dsu_components(n, net_edges(pairs[i])))
    finally:
        for s in stores:
            s.close()


def test_group_member_out_of_samples():
    rng = np.random.default_rng(3)
    n = 2000
    ok = gs.SketchStore(n, 1)
    tiny = gs.SketchStore(n, 2, num_samples=1)  # one column: cannot finish a dense graph
    p = rng.integers(0, n, size=(20 * n, 2), dtype=np.uint32)
    try:
        with gs.StoreGroup([ok, tiny]) as grp:
            grp.submit_pairs([p, p])
            grp.flush()
            _, _, _, status = grp.query_cc(allow_out_of_samples=True)
            assert status[0] == 0 and status[1] == gs.GS_ERR_OUT_OF_SAMPLES
            with pytest.raises(gs.OutOfSamplesError):
                grp.query_cc()
    finally:
        ok.close()
        tiny.close()


def test_group_query_giants_with_pair_and_batch_fed_members():
    # dense random members form a giant component: pair-fed members (sketches XOR to zero) sum
    # it from the rest of the member; a member fed by vertex batches (one direction only, so its
    # sketches do not XOR to zero) must sum it directly; both agree with single-store queries
    # and with a union-find over the net graph
    rng = np.random.default_rng(77)
    sizes = [5000, 3000, 8000]
    seeds = [31, 32, 33]
    pairs = [member_pairs(rng, n, 3 * n) for n in sizes]
    stores = [gs.SketchStore(n, s) for n, s in zip(sizes, seeds)]
    singles = [gs.SketchStore(n, s) for n, s in zip(sizes, seeds)]
    try:
        # member 1: vertex batches holding both directions of every edge (the same sketches as
        # pairs, but the store no longer knows that they XOR to zero)
        clean = pairs[1][pairs[1][:, 0] != pairs[1][:, 1]]
        src, sizes_b, offs, dsts = orc.group_by_vertex(clean, sizes[1], 512)
        for st in (stores[1], singles[1]):
            s = st.stream()
            s.submit_vertex_batches(src, sizes_b, offs, dsts)
            s.flush()
            s.close()
        with gs.StoreGroup(stores) as grp:
            grp.submit_pairs([pairs[0], None, pairs[2]])
            grp.flush()
            for i in (0, 2):
                s = singles[i].stream()
                s.submit_pairs(pairs[i])
                s.flush()
                s.close()
            labels, ncc, rounds, status = grp.query_cc()
            for i, n in enumerate(sizes):
                lab1, ncc1, r1 = singles[i].query_cc()
                net = net_edges(clean if i == 1 else pairs[i])
                expect = orc.dsu_components(n, net)
                assert np.bincount(expect).max() > n // 2, f"member {i} has no giant"
                assert np.array_equal(labels[i], expect), f"member {i}"
                assert np.array_equal(labels[i], lab1) and rounds[i] == r1 and status[i] == 0
    finally:
        for s in stores + singles:
            s.close()


def test_cross_member_pairs_are_dropped_and_reported():
    """A device submission in global ids whose pair joins two members is skipped (it would XOR an
    index that never cancels into both members) and reported by the flush; the other pairs of the
    submission are applied."""
    rng = np.random.default_rng(7)
    sizes, seeds = [300, 500, 200], [11, 12, 13]
    stores = [gs.SketchStore(n, s) for n, s in zip(sizes, seeds)]
    try:
        with gs.StoreGroup(stores) as grp:
            local = [member_pairs(rng, n, 4 * n) for n in sizes]
            glob = np.concatenate([p + np.uint32(b) for p, b in zip(local, grp.bases)])
            cross = np.array([[5, grp.bases[1] + 7], [grp.bases[2] + 1, grp.bases[1]],
                              [grp.bases[1] - 1, grp.bases[1]]], dtype=np.uint32)
            allp = np.concatenate([glob[:1000], cross, glob[1000:]])
            buf = gs.DeviceBuffer(allp)
            grp.submit_pairs_device(buf.ptr, len(allp))
            with pytest.raises(gs.GsError, match="3 submitted updates"):
                grp.flush()
            buf.free()
            for i, (n, sd) in enumerate(zip(sizes, seeds)):
                assert np.array_equal(stores[i].export_buckets(), oracle_store(n, sd, local[i]))
            # host submissions are checked member by member before anything is submitted
            with pytest.raises(gs.GsError, match="out of range"):
                grp.submit_pairs([local[0], np.array([[0, sizes[1]]], dtype=np.uint32), None])
            grp.flush()
            for i, (n, sd) in enumerate(zip(sizes, seeds)):
                assert np.array_equal(stores[i].export_buckets(), oracle_store(n, sd, local[i]))
    finally:
        for st in stores:
            st.close()


def test_member_pairs_rebased_and_checked_on_device():
    """gs_group_submit_member_pairs: a member's pairs in its own ids go to the device as they are
    (pinned: DMA from the caller's buffer; pageable: staged), the device adds the member's base
    and drops pairs with an id past the member's end (one that a host-side rebase would have
    turned into another member's vertex), reported with their count by the next flush. The
    sketches equal the oracle's of the valid pairs. gs_flush_async launches without waiting, and
    an export (like a query) waits for the launches."""
    rng = np.random.default_rng(23)
    sizes, seeds = [300, 500, 200], [31, 32, 33]
    stores = [gs.SketchStore(n, s) for n, s in zip(sizes, seeds)]
    try:
        with gs.StoreGroup(stores) as grp:
            local = [member_pairs(rng, n, 5 * n) for n in sizes]
            bad = np.array([[0, sizes[1]], [sizes[1] + 5, 3], [sizes[1], sizes[1] + 1]],
                           dtype=np.uint32)  # member 1: id n_1 and past it (member 2's range)
            pinned = gs.PinnedBuffer((len(local[0]), 2))
            pinned.array[:] = local[0]
            try:
                half = len(local[1]) // 2
                grp.submit_member_pairs(1, local[1][:half])
                grp.submit_member_pairs(0, pinned.array)  # pinned: no staging copy
                grp.submit_member_pairs(1, np.concatenate([bad, local[1][half:]]))
                grp.submit_member_pairs(2, local[2])
                with pytest.raises(gs.GsError, match="3 submitted updates"):
                    grp.flush()
            finally:
                pinned.free()
            for i, (n, sd) in enumerate(zip(sizes, seeds)):
                assert np.array_equal(stores[i].export_buckets(), oracle_store(n, sd, local[i]))
            more = [member_pairs(rng, n, 2 * n) for n in sizes]
            for i in (2, 0, 1):
                grp.submit_member_pairs(i, more[i])
            grp.stream.flush_async()
            for i, (n, sd) in enumerate(zip(sizes, seeds)):
                both = np.concatenate([local[i], more[i]])
                assert np.array_equal(stores[i].export_buckets(), oracle_store(n, sd, both))
            grp.flush()  # nothing dropped since the last flush
            with pytest.raises(gs.GsError, match="member 3"):
                grp.submit_member_pairs(3, local[0])
    finally:
        for st in stores:
            st.close()


def test_member_pairs_async_dma_after_return():
    """gs_group_submit_member_pairs_async: many pinned payloads enqueued back to back (their DMAs
    may still run when the calls return), one pageable payload staged, launches crossing fill
    buffers (small launch size); after gs_stream_wait_copies the pinned buffers are overwritten
    and the sketches still equal the oracle's of the pairs as submitted."""
    rng = np.random.default_rng(29)
    sizes, seeds = [4000, 2500, 6000], [41, 42, 43]
    stores = [gs.SketchStore(n, s) for n, s in zip(sizes, seeds)]
    pinned = []
    try:
        with gs.StoreGroup(stores) as grp:
            grp.stream.set_tuning(launch_edges=1 << 16)
            sent = [[] for _ in sizes]
            for k in range(24):
                i = k % 3
                p = member_pairs(rng, sizes[i], int(rng.integers(1000, 30000)))
                buf = gs.PinnedBuffer((len(p), 2))
                buf.array[:] = p
                pinned.append(buf)
                grp.submit_member_pairs_async(i, buf.array)
                sent[i].append(p)
            extra = member_pairs(rng, sizes[1], 7000)  # pageable: staged before the call returns
            grp.submit_member_pairs_async(1, extra)
            sent[1].append(extra)
            grp.wait_copies()
            for buf in pinned:  # the DMAs have read them: the device must not see this
                buf.array[:] = 0
            grp.flush()
            for i, (n, sd) in enumerate(zip(sizes, seeds)):
                assert np.array_equal(stores[i].export_buckets(),
                                      oracle_store(n, sd, np.concatenate(sent[i])))
    finally:
        for buf in pinned:
            buf.free()
        for st in stores:
            st.close()


def test_member_batches_in_one_call():
    """gs_group_submit_member_batches_async: 150 payloads of three members in one call (three
    rebase launches of up to 64 batches, launches crossing fill buffers at a small launch size),
    pinned and pageable, one holding ids past its member's end (dropped and reported by the next
    flush). The sketches equal the oracle's of the valid pairs, and a call naming a member past
    the group's end applies nothing."""
    rng = np.random.default_rng(31)
    sizes, seeds = [3000, 700, 5000], [51, 52, 53]
    stores = [gs.SketchStore(n, s) for n, s in zip(sizes, seeds)]
    pinned = []
    try:
        with gs.StoreGroup(stores) as grp:
            grp.stream.set_tuning(launch_edges=1 << 16)
            sent = [[] for _ in sizes]
            members, batches = [], []
            for k in range(150):
                i = int(rng.integers(0, 3))
                p = member_pairs(rng, sizes[i], int(rng.integers(1, 4000)))
                if k % 5 == 0:  # pageable: staged
                    batches.append(p)
                else:
                    buf = gs.PinnedBuffer((len(p), 2))
                    buf.array[:] = p
                    pinned.append(buf)
                    batches.append(buf.array)
                members.append(i)
                sent[i].append(p)
            bad = np.array([[0, sizes[1]], [sizes[1] + 9, 2]], dtype=np.uint32)
            members.insert(70, 1)
            batches.insert(70, bad)
            members.append(0)
            batches.append(np.zeros((0, 2), dtype=np.uint32))  # empty batch
            grp.submit_member_batches_async(members, batches)
            grp.wait_copies()
            for buf in pinned:  # read by the DMAs already
                buf.array[:] = 0
            with pytest.raises(gs.GsError, match="2 submitted updates"):
                grp.flush()
            for i, (n, sd) in enumerate(zip(sizes, seeds)):
                assert np.array_equal(stores[i].export_buckets(),
                                      oracle_store(n, sd, np.concatenate(sent[i])))
            before = [st.export_buckets() for st in stores]
            with pytest.raises(gs.GsError, match="member 3"):
                grp.submit_member_batches_async([0, 3], [sent[0][0], sent[0][0]])
            grp.flush()
            for st, b in zip(stores, before):
                assert np.array_equal(st.export_buckets(), b)
    finally:
        for buf in pinned:
            buf.free()
        for st in stores:
            st.close()


def test_store_destroy_refused_while_in_group():
    a, b = gs.SketchStore(100, 1), gs.SketchStore(200, 2)
    grp = gs.StoreGroup([a, b])
    grp.submit_pairs([np.array([[0, 1]], dtype=np.uint32), np.array([[2, 3]], dtype=np.uint32)])
    with pytest.raises(gs.GsError, match="destroy the group first"):
        a.close()
    assert a.h  # still alive and usable
    grp.close()
    assert np.array_equal(a.export_buckets(), oracle_store(100, 1, np.array([[0, 1]], dtype=np.uint32)))
    a.close()
    b.close()


def batched_config_workload(graphs, seed):
    """The batched-streams config: n = 2^s, s uniform in [10, 16] (log-uniform n in
    [2^10, 2^16]), 4..64 updates per vertex, 5% deletions; even graphs Erdos-Renyi, odd graphs
    R-MAT(0.57, 0.19, 0.19, 0.05)."""
    rng = np.random.default_rng(seed)
    scales = rng.integers(10, 17, size=graphs)
    upv = rng.integers(4, 65, size=graphs)
    out = []
    for g in range(graphs):
        n = 1 << int(scales[g])
        m = int(upv[g]) * n // 2
        if g % 2 == 0:
            p = min(1.0, m / 1.05 / (n * (n - 1) / 2))
            pairs = gs.gen_erdos_renyi(n, p, 0.05, seed + g)
        else:
            pairs = gs.gen_rmat(int(scales[g]), m, seed + g, del_frac=0.05)
        out.append((n, 0xB0000000 + 7919 * g, pairs))
    return out


def test_batched_config_full_shape_256_members():
    # 256 graphs in one group, all updates interleaved through one launch sequence; sampled
    # members bit-exact with the oracle, every member's partition equal to a DSU of its net graph
    work = batched_config_workload(256, 20261207)
    stores = [gs.SketchStore(n, sd) for n, sd, _ in work]
    try:
        with gs.StoreGroup(stores) as grp:
            glob = np.concatenate([p + np.uint32(b) for (_, _, p), b in zip(work, grp.bases)])
            glob = glob[np.random.default_rng(2).permutation(len(glob))]
            buf = gs.DeviceBuffer(glob)
            for b in range(0, len(glob), 1 << 25):
                grp.submit_pairs_device(buf.offset(b * 8), min(1 << 25, len(glob) - b))
            grp.flush()
            buf.free()
            del glob
            labels, ncc, rounds, status = grp.query_cc()
            assert (status == 0).all()
            sizes = np.array([n for n, _, _ in work])
            sample = sorted({int(np.argmin(sizes)), int(np.argmax(sizes)), 1, 2, 127, 200})
            for i in sample:
                n, sd, p = work[i]
                assert np.array_equal(stores[i].export_buckets(), oracle_store(n, sd, p)), f"member {i}"
            for i, (n, _, p) in enumerate(work):
                assert np.array_equal(labels[i], orc.dsu_components(n, net_edges(p))), f"member {i}"
                assert ncc[i] == len(np.unique(labels[i]))
    finally:
        for st in stores:
            st.close()


def test_fenced_query_sees_updates_before_the_fence():
    """gs_group_fence: a masked query after a fence answers for the updates submitted before it,
    while updates of other members submitted after the fence are applied beside it."""
    rng = np.random.default_rng(77)
    seeds = [3000 + 11 * i for i in range(len(SIZES))]
    stores = [gs.SketchStore(n, s) for n, s in zip(SIZES, seeds)]
    try:
        with gs.StoreGroup(stores) as grp:
            grp.prepare()
            first = [member_pairs(rng, n, 4 * n) for n in SIZES]
            later = [member_pairs(rng, n, 8 * n) for n in SIZES]
            mask = np.array([1, 0, 1, 0, 1, 0], dtype=np.uint8)
            for rep in range(3):
                grp.submit_pairs(first)
                grp.fence()
                # the unqueried members' updates go in after the fence, without a flush
                grp.submit_pairs([None if mask[i] else later[i] for i in range(len(SIZES))])
                labels, ncc, _, status = grp.query_cc(mask)
                for i in range(len(SIZES)):
                    if not mask[i]:
                        continue
                    applied = np.concatenate([first[i]] * (rep + 1))
                    want = orc.dsu_components(SIZES[i], net_edges(applied))
                    assert status[i] == 0 and np.array_equal(labels[i], want), (rep, i)
                grp.flush()
                for i in range(len(SIZES)):
                    seq = [first[i]] * (rep + 1) + ([] if mask[i] else [later[i]] * (rep + 1))
                    got = stores[i].export_buckets()
                    assert np.array_equal(got, oracle_store(SIZES[i], seeds[i], np.concatenate(seq))), (rep, i)
    finally:
        for s in stores:
            s.close()
