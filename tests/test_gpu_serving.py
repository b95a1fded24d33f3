"""Serving: the same request trace served one request at a time (plain) and with continuous
batching must give identical answers (connectivity partitions, min cuts) and identical final
sketches, and a tenant served alone must answer exactly as in the mix (program serve_bench)."""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROG = os.path.join(ROOT, "gpusketch_amd", "bin", "serve_bench")


@pytest.mark.parametrize("rate", [0, 3000])
def test_served_equals_plain_and_alone(rate):
    if not os.path.exists(PROG):
        pytest.fail("serve_bench is not built; run `python gpusketch_amd/build.py`")
    r = subprocess.run([PROG, "--tenants", "10", "--requests", "300", "--rate", str(rate),
                        "--max-n", "30000", "--max-edges", "20000", "--seed", str(11 + rate),
                        "--alone-check", "1"], capture_output=True, text=True, timeout=600)
    print(r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["answers_identical"] and out["sketches_identical"] and out["alone_equals_mix"]
    assert out["queries"] > 0


def test_eviction_spills_and_restores_bit_exact():
    # a budget that holds a few of the ten tenants: idle tenants are spilled to host memory and
    # restored when their next request comes; plain and served evict at different times, yet
    # every answer and every final sketch must be identical, and equal to a tenant served alone
    if not os.path.exists(PROG):
        pytest.fail("serve_bench is not built; run `python gpusketch_amd/build.py`")
    r = subprocess.run([PROG, "--tenants", "10", "--requests", "300", "--rate", "3000",
                        "--max-n", "30000", "--max-edges", "20000", "--seed", "23",
                        "--budget-gb", "0.5", "--alone-check", "1"],
                       capture_output=True, text=True, timeout=600)
    print(r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["answers_identical"] and out["sketches_identical"] and out["alone_equals_mix"]
    assert out["served"]["evictions"] > 0 and out["plain"]["evictions"] > 0
    assert out["served"]["restores"] > 0
