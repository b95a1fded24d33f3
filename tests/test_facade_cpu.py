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


def _prog(l0=False):
    path = PROG + ("_l0" if l0 else "")
    if not os.path.exists(path):
        pytest.fail(f"{os.path.basename(path)} is not built; run `python gpusketch_amd/build.py`")
    return path


# l0: the same program and C++ API compiled with L0_SAMPLING (the CubeSketch rule)
@pytest.mark.parametrize("l0", [False, True])
def test_facade_cpu_checks(tmp_path, l0):
    r = subprocess.run([_prog(l0), "all", str(tmp_path)], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failure(s)" in r.stdout


@pytest.mark.parametrize("l0", [False, True])
@pytest.mark.parametrize("n,m,seed", [(50, 400, 3), (3000, 60000, 0x5EED)])
def test_host_sketch_update_matches_oracle(tmp_path, n, m, seed, l0):
    rng = np.random.default_rng(seed)
    pairs = rng.integers(0, n, size=(m, 2), dtype=np.uint32)
    pairs = pairs[pairs[:, 0] != pairs[:, 1]]
    src = tmp_path / "pairs.bin"
    out = tmp_path / "buckets.bin"
    pairs.tofile(src)
    r = subprocess.run([_prog(l0), "sketch_dump", str(src), str(n), str(seed), str(out)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    g = orc.Geometry(n, 1.0, 7, True) if l0 else orc.Geometry(n)
    ref = g.empty_store()
    orc.apply_edges(ref, g, seed, pairs, l0=l0)
    got = np.fromfile(out, dtype=np.uint64).reshape(ref.shape)
    assert np.array_equal(got, ref)
