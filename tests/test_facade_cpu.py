"""Host-side checks of the C++ API (include/gpusketch/*.h): the self-checking program
facade_cpu (Sketch invariants of the reference's sketch_test.cpp, pairing, union-find, binary
streams, gutters + driver delivery, EdgeStore toggling and contraction), and host Sketch::update
bit-exact against the oracle."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROG = os.path.join(ROOT, "gpusketch_amd", "bin", "facade_cpu")


def _prog():
    if not os.path.exists(PROG):
        pytest.fail("facade_cpu is not built; run `python gpusketch_amd/build.py`")
    return PROG


def test_facade_cpu_checks(tmp_path):
    r = subprocess.run([_prog(), "all", str(tmp_path)], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failure(s)" in r.stdout


@pytest.mark.parametrize("n,m,seed", [(50, 400, 3), (3000, 60000, 0x5EED)])
def test_host_sketch_update_matches_oracle(tmp_path, n, m, seed):
    rng = np.random.default_rng(seed)
    pairs = rng.integers(0, n, size=(m, 2), dtype=np.uint32)
    pairs = pairs[pairs[:, 0] != pairs[:, 1]]
    src = tmp_path / "pairs.bin"
    out = tmp_path / "buckets.bin"
    pairs.tofile(src)
    r = subprocess.run([_prog(), "sketch_dump", str(src), str(n), str(seed), str(out)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    g = orc.Geometry(n)
    ref = g.empty_store()
    orc.apply_edges(ref, g, seed, pairs)
    got = np.fromfile(out, dtype=np.uint64).reshape(ref.shape)
    assert np.array_equal(got, ref)
