"""Serving: the same request trace served one request at a time (plain) and with continuous
batching must give identical answers (connectivity partitions, min cuts) and identical final
sketches, a tenant served alone must answer exactly as in the mix, and every answer must equal
the exact one of the tenant's net graph at that point of the trace: a union-find for
connectivity, min(k, exact minimum cut) for min-cut queries (program serve_bench --verify 1)."""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROG = os.path.join(ROOT, "gpusketch_amd", "bin", "serve_bench")


# pinned 1: payloads handed over in pinned memory (DMA straight from them); 0: copied at arrival.
# launch: the group's launch size (0: the server's default); 4096 edges splits payloads across
# fill buffers and a step's submission across many launches.
@pytest.mark.parametrize("rate,pinned,launch", [(0, 1, 0), (3000, 0, 0), (0, 1, 4096)])
def test_served_equals_plain_and_alone(rate, pinned, launch):
    if not os.path.exists(PROG):
        pytest.fail("serve_bench is not built; run `python gpusketch_amd/build.py`")
    extra = ["--launch-edges", str(launch)] if launch else []
    r = subprocess.run([PROG, "--tenants", "10", "--requests", "300", "--rate", str(rate),
                        "--max-n", "30000", "--max-edges", "20000", "--seed", str(11 + rate + launch),
                        "--alone-check", "1", "--verify", "1", "--pinned-payloads", str(pinned),
                        *extra],
                       capture_output=True, text=True, timeout=600)
    print(r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["answers_identical"] and out["sketches_identical"] and out["alone_equals_mix"]
    assert out["queries"] > 0 and out["pinned_payloads"] == bool(pinned)
    assert out["verified"]["mismatches"] == 0 and out["verified"]["cc_answers"] > 0


def test_eviction_spills_and_restores_bit_exact():
    # a budget that holds a few of the ten tenants: idle tenants are spilled to host memory and
    # restored when their next request comes; plain and served evict at different times, yet
    # every answer and every final sketch must be identical, and equal to a tenant served alone
    if not os.path.exists(PROG):
        pytest.fail("serve_bench is not built; run `python gpusketch_amd/build.py`")
    r = subprocess.run([PROG, "--tenants", "10", "--requests", "300", "--rate", "3000",
                        "--max-n", "30000", "--max-edges", "20000", "--seed", "23",
                        "--budget-gb", "0.5", "--alone-check", "1", "--verify", "1"],
                       capture_output=True, text=True, timeout=600)
    print(r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["answers_identical"] and out["sketches_identical"] and out["alone_equals_mix"]
    assert out["served"]["evictions"] > 0 and out["plain"]["evictions"] > 0
    assert out["served"]["restores"] > 0
    assert out["verified"]["mismatches"] == 0


def test_serving_config_trace_answers_exact():
    # the serving config as stated: 64 tenants with n log-uniform in [2^12, 2^18], 1000 Poisson
    # requests (batches of 1k..1M edges, 5% deletions, k=4 min-cut queries on 8 tenants), 200 GB
    # budget, with the bench's 10% CC queries (about 100); every answer of both modes checked
    # against the exact graph
    if not os.path.exists(PROG):
        pytest.fail("serve_bench is not built; run `python gpusketch_amd/build.py`")
    r = subprocess.run([PROG, "--tenants", "64", "--requests", "1000", "--rate", "2000",
                        "--budget-gb", "200", "--k", "4", "--mc-tenants", "8",
                        "--cc-frac", "0.1", "--alone-check", "1", "--verify", "1"],
                       capture_output=True, text=True, timeout=900)
    print(r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["tenants"] == 64 and out["updates"] > 100_000_000
    assert out["answers_identical"] and out["sketches_identical"] and out["alone_equals_mix"]
    v = out["verified"]
    assert v["mismatches"] == 0, v["first"]
    assert v["cc_answers"] >= 80 and v["mincut_answers"] > 0
