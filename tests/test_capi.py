"""The C-ABI library loads, exports every symbol include/gpusketch.h declares, and its host-only
functions (sizing, exact min cut, stream generators) agree with the oracle. No GPU calls."""
import itertools
import os
import re

import numpy as np
import pytest

import gpusketch_amd as gs
from oracle import oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "gpusketch.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\**(gs_\w+)\s*\(", text, re.M)))


def test_header_symbols_exported():
    decl = declared_symbols()
    assert len(decl) >= 25
    for name in decl:
        assert hasattr(gs.lib, name), name
    assert sorted(decl) == gs.EXPORTED


def test_library_built_from_these_sources():
    """The loaded libgpusketch.so carries the digest of the sources it was compiled from; it must
    be the digest of the sources in this tree (a stale build fails here)."""
    build = _build_recipe()
    assert gs.source_digest() == build.source_digest()


def _build_recipe():
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "gpusketch_amd_build", os.path.join(ROOT, "gpusketch_amd", "build.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_geometry_matches_oracle():
    for n in (2, 3, 100, 1024, 8192, 1 << 15, 1 << 17, 1 << 18):
        g = gs.geometry_for(n)
        o = orc.Geometry(n)
        assert (g.bkt_per_col, g.num_samples, g.num_columns, g.num_buckets, g.vector_len) == (
            o.bkt_per_col, o.num_samples, o.num_columns, o.num_buckets, o.vector_len)
    g = gs.geometry_for(1 << 15, 2.7)
    assert g.num_columns == 55
    # CubeSketch sizing: 7 columns per sample, num_samples_div = log2(3) - 1
    for n in (300, 8192, 1 << 17):
        g = gs.geometry_for(n, 1.0, 7, cubesketch=True)
        o = orc.Geometry(n, 1.0, 7, True)
        assert (g.bkt_per_col, g.num_samples, g.num_columns, g.num_buckets, g.flags) == (
            o.bkt_per_col, o.num_samples, o.num_columns, o.num_buckets, gs.STORE_CUBESKETCH)


def test_invalid_arguments_report_errors():
    with pytest.raises(gs.GsError):
        gs.geometry_for(1)
    with pytest.raises(gs.GsError):
        gs.gen_rmat(10, 10, 1, a=0.9, b=0.2, c=0.1)
    # group entry points check their handles before any device call
    assert gs.lib.gs_group_submit_member_batches_async(None, 0, None, None, None) == gs.GS_ERR_INVALID
    assert gs.lib.gs_group_submit_member_pairs_async(None, 0, None, 0) == gs.GS_ERR_INVALID
    assert gs.lib.gs_stream_wait_copies(None) == gs.GS_ERR_INVALID


def test_min_cut_matches_stoer_wagner():
    rng = np.random.default_rng(40)
    for trial in range(60):
        n = int(rng.integers(2, 40))
        allp = list(itertools.combinations(range(n), 2))
        k = int(rng.integers(0, len(allp) + 1))
        idx = rng.choice(len(allp), size=k, replace=False) if k else []
        pairs = np.array([allp[i] for i in idx] + [allp[i] for i in idx[: k // 4]],
                         dtype=np.uint32).reshape(-1, 2)
        val, side = gs.min_cut(n, pairs)
        ref, _ = orc.stoer_wagner(n, pairs)
        assert val == ref, (trial, n, len(pairs))
        assert 0 < side.sum() < n
        crossing = int(np.sum(side[pairs[:, 0]] != side[pairs[:, 1]])) if len(pairs) else 0
        assert crossing == val


def test_min_cut_planted_larger_graph():
    pairs = gs.gen_planted_cut(600, 0.05, 4, 0.0, 9)
    val, side = gs.min_cut(600, pairs)
    assert val == orc.stoer_wagner(600, pairs)[0] == 4
    assert side.sum() in (300, 300)


def _check_toggle_semantics(pairs, n):
    live = set()
    for u, v in pairs.tolist():
        assert u != v and u < n and v < n
        k = (min(u, v), max(u, v))
        if k in live:
            live.remove(k)
        else:
            live.add(k)
    return live


def test_erdos_renyi_stream():
    pairs = gs.gen_erdos_renyi(200, 0.25, 0.1, 5)
    live = _check_toggle_semantics(pairs, 200)
    ins = len(pairs) - (len(pairs) - len(live)) // 2 * 1
    m = (len(pairs) + len(live)) // 2  # inserts = live + deleted, total = inserts + deleted
    assert abs(m - 0.25 * 200 * 199 / 2) < 5 * np.sqrt(m)
    assert abs((len(pairs) - m) - round(0.1 * m)) <= 1
    assert ins > 0


def test_rmat_stream_deletes_only_live_edges():
    pairs = gs.gen_rmat(10, 20000, 3)
    assert pairs.shape == (20000, 2)
    live = _check_toggle_semantics(pairs, 1024)
    dels = (len(pairs) - len(live)) // 2
    assert 0.03 * len(pairs) < dels < 0.07 * len(pairs)
    # R-MAT skew: vertex 0 is the hub
    deg = np.bincount(pairs.ravel(), minlength=1024)
    assert deg.argmax() == 0


def test_rmat_stream_same_for_any_thread_count():
    """The generator shards its duplicate check over threads; the stream must not depend on how
    many there are (GPUSKETCH_GEN_THREADS), and a longer request only extends the kept insertions
    of a shorter one (same candidate sequence)."""
    import subprocess
    import sys
    code = ("import sys, hashlib; sys.path.insert(0, %r); import gpusketch_amd as gs; "
            "p = gs.gen_rmat(13, 400000, 77); print(hashlib.sha1(p.tobytes()).hexdigest())" % ROOT)
    digests = set()
    for t in ("1", "3", "8"):
        r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                           env={**os.environ, "GPUSKETCH_GEN_THREADS": t}, timeout=120)
        assert r.returncode == 0, r.stderr
        digests.add(r.stdout.strip())
    assert len(digests) == 1
    pairs = gs.gen_rmat(13, 400000, 77)
    live = _check_toggle_semantics(pairs, 1 << 13)
    assert len(live) > 0


def test_planted_cut_stream_net_graph():
    pairs = gs.gen_planted_cut(200, 0.2, 5, 0.1, 7)
    live = _check_toggle_semantics(pairs, 200)
    net = np.array(sorted(live), dtype=np.uint32)
    cross = np.sum((net[:, 0] < 100) != (net[:, 1] < 100))
    assert cross == 5
    assert orc.stoer_wagner(200, net)[0] == 5
