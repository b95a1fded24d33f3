"""The C++ API on the device: CCGPUSketchAlg / SKGPUSketchAlg / MCGPUSketchAlg driven by
GraphSketchDriver over binary streams (program facade_gpu). Sketch contents are compared bit for
bit with host Sketch::update (pinned against the oracle by test_facade_cpu), components with a
union-find over the net graph, the min cut with the exact minimum cut of the net graph."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROG = os.path.join(ROOT, "gpusketch_amd", "bin", "facade_gpu")
FIXTURE = os.path.join(ROOT, "tests", "golden", "multiples_graph_1024_stream.data")


@pytest.mark.parametrize("prog", ["facade_gpu", "facade_gpu_l0"])
def test_facade_gpu_program(tmp_path, prog):
    """facade_gpu_l0: the same program and C++ API compiled with L0_SAMPLING (CubeSketch rule,
    7 columns per sample), against host Sketch::update compiled the same way."""
    path = os.path.join(os.path.dirname(PROG), prog)
    if not os.path.exists(path):
        pytest.fail(f"{prog} is not built; run `python gpusketch_amd/build.py`")
    r = subprocess.run([path, FIXTURE, str(tmp_path)], capture_output=True, text=True,
                       timeout=600)
    sys.stdout.write(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failure(s)" in r.stdout


def test_verify_samples_build_of_the_api(tmp_path):
    """verify_facade is compiled with the reference's -DVERIFY_SAMPLES_F: the driver records
    each stream's exact graph and the queries of CCSketchAlg / CCGPUSketchAlg / SKGPUSketchAlg /
    MCSketchAlg / MCGPUSketchAlg (held through their base classes) check every answer against it;
    the fixture's cumulative text file gives the expected graph (test/cc_alg_test.cpp:41-57)."""
    path = os.path.join(os.path.dirname(PROG), "verify_facade")
    if not os.path.exists(path):
        pytest.fail("verify_facade is not built; run `python gpusketch_amd/build.py`")
    txt = os.path.join(ROOT, "tests", "golden", "multiples_graph_1024.txt")
    r = subprocess.run([path, txt, str(tmp_path)], capture_output=True, text=True, timeout=600)
    sys.stdout.write(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failure(s)" in r.stdout
