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


def test_facade_gpu_program(tmp_path):
    if not os.path.exists(PROG):
        pytest.fail("facade_gpu is not built; run `python gpusketch_amd/build.py`")
    r = subprocess.run([PROG, FIXTURE, str(tmp_path)], capture_output=True, text=True,
                       timeout=600)
    sys.stdout.write(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failure(s)" in r.stdout
