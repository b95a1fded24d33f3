"""Benchmark: sketch edge updates/s at n = 2^17 on one MI355X.

Workload (single-stream GPU CC config): R-MAT n=2^17, (a,b,c,d) = (0.57,0.19,0.19,0.05), 5% of
updates delete a previously inserted live edge, no self loops or duplicate live edges, batches of
2^20 uint32 pairs. One step = one batch through the device update path (vertex grouping +
aggregation kernels) with the batch already resident in HBM. Every step uses a distinct batch of
one continuous seeded stream (warmup batches first, then the timed ones).

Prints one JSON line (the contract of the driver) with a `roofline` object for the dominant
kernel (aggregate_kernel, timed with HIP events on its own stream) and a `cpu_baseline` object
(the CPU restatement of the reference update path, timed on this host's cores over a bounded
prefix of the same stream).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

SCALE = 17
N = 1 << SCALE
BATCH = 1 << 20
RMAT = (0.57, 0.19, 0.19)
DEL_FRAC = 0.05
SEED_STREAM = 20261203
SEED_SKETCH = 0x5EEDC0DE
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(pairs: np.ndarray, threads: int) -> dict:
    """Reference CPU update path (per-thread delta sketch + mutex merge) on pre-grouped vertex
    batches of 4 * num_buckets destinations, the reference's batch size."""
    from oracle import oracle as orc
    g = orc.Geometry(N)
    t0 = time.perf_counter()
    src, sizes, offs, dsts = orc.group_by_vertex(pairs, N, 4 * g.num_buckets)
    group_s = time.perf_counter() - t0
    store = g.empty_store()
    secs = orc.cpu_baseline(store, g, SEED_SKETCH, src, sizes, offs, dsts, threads)
    del store
    return {"value": len(pairs) / secs, "unit": "edges/s", "cores": threads, "kind": "port",
            "sample": f"first {len(pairs)} updates of the same stream ({len(pairs) // BATCH} "
                      f"batches), grouped into {len(src)} vertex batches of <= "
                      f"{4 * g.num_buckets} (grouping {group_s:.2f}s, not timed)",
            "seconds": secs}


def copy_calibration(mib: int, reps: int = 5) -> None:
    """Runs gs_device_copy over `mib` MiB `reps` times: each dispatch reads and writes exactly
    mib MiB, the reference for scaling FETCH_SIZE / WRITE_SIZE counter readings."""
    import ctypes

    import gpusketch_amd as gs
    nbytes = mib << 20
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    gs._check(gs.lib.gs_device_alloc(nbytes, ctypes.byref(src)))
    gs._check(gs.lib.gs_device_alloc(nbytes, ctypes.byref(dst)))
    for _ in range(reps):
        gs._check(gs.lib.gs_device_copy(dst, src, nbytes))
    gs._check(gs.lib.gs_device_free(src))
    gs._check(gs.lib.gs_device_free(dst))
    print(json.dumps({"copy_calibration_bytes_per_dispatch": nbytes, "dispatches": reps}))


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--cpu-batches", type=int, default=32, help="stream prefix for the CPU leg")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--d-sample", type=int, default=2, help="batches used for the bytes model")
    ap.add_argument("--pcie-batches", type=int, default=16,
                    help="batches submitted from host memory for the PCIe-inclusive rate")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run the known-size copy kernel (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the sketch engine "
              "runs on a single MI355X in this build", file=sys.stderr)
        sys.exit(2)

    import gpusketch_amd as gs
    from oracle import oracle as orc

    total_batches = args.warmup + args.steps
    t0 = time.perf_counter()
    pairs = gs.gen_rmat(SCALE, total_batches * BATCH, SEED_STREAM, *RMAT, del_frac=DEL_FRAC)
    log(f"generated {len(pairs)} R-MAT updates in {time.perf_counter() - t0:.1f}s")

    store = gs.SketchStore(N, SEED_SKETCH)
    geo = store.geom
    stream = store.stream()
    dbuf = gs.DeviceBuffer(pairs)
    log(f"store {geo.device_bytes / 1e9:.3f} GB, B={geo.bkt_per_col} C={geo.num_columns}; "
        f"input resident in HBM ({dbuf.nbytes / 1e9:.2f} GB)")

    for b in range(args.warmup):
        stream.submit_pairs_device(dbuf.offset(b * BATCH * 8), BATCH)
    stream.flush()
    log(f"warmup {args.warmup} batches done")

    stream.set_profiling(True)
    gs.lib.gs_device_synchronize()
    t0 = time.perf_counter()
    for b in range(args.warmup, total_batches):
        stream.submit_pairs_device(dbuf.offset(b * BATCH * 8), BATCH)
    stream.flush()
    gs.lib.gs_device_synchronize()
    elapsed = time.perf_counter() - t0
    agg_ms, launches, _ = stream.timing()
    stream.set_profiling(False)
    log(f"timed {args.steps} batches in {elapsed:.4f}s; aggregate_kernel {agg_ms / launches:.4f} ms avg")

    # PCIe-inclusive rate: the same batches handed over as host (pageable) arrays, staged by the
    # library into pinned memory and copied to the device, into a second store
    pcie = None
    if args.pcie_batches > 0:
        nb = min(args.pcie_batches, total_batches)
        with gs.SketchStore(N, SEED_SKETCH) as st2:
            s2 = st2.stream()
            s2.submit_pairs(pairs[:BATCH])
            s2.flush()
            gs.lib.gs_device_synchronize()
            p0 = time.perf_counter()
            for b in range(1, nb):
                s2.submit_pairs(pairs[b * BATCH:(b + 1) * BATCH])
            s2.flush()
            gs.lib.gs_device_synchronize()
            pe = time.perf_counter() - p0
            s2.close()
        pcie = {"value": (nb - 1) * BATCH / pe, "unit": "edges/s", "batches": nb - 1,
                "input": "host pageable uint32 pairs"}
        log(f"PCIe-inclusive {pcie['value']:.3e} edges/s over {nb - 1} host batches")

    # CC query latency on the final store (plain single-stream path)
    lat = []
    for _ in range(5):
        q0 = time.perf_counter()
        _, ncc, rounds = store.query_cc()
        lat.append((time.perf_counter() - q0) * 1e3)
    log(f"CC query: {ncc} components, {rounds} rounds, p50 {np.percentile(lat, 50):.2f} ms")

    # bytes model: 8 B per edge + 32 B per distinct (vertex, bucket) slot of the batch
    g = orc.Geometry(N)
    d_vals = []
    for b in range(args.warmup, args.warmup + args.d_sample):
        d_vals.append(orc.distinct_slots(g, SEED_SKETCH, pairs[b * BATCH:(b + 1) * BATCH]))
    d_mean = float(np.mean(d_vals))
    alg_bytes = 8.0 * BATCH + 32.0 * d_mean
    agg_s = agg_ms / launches / 1e3
    achieved = alg_bytes / agg_s / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_aggregate_rmat17.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath)).get("hbm_bytes_per_launch")

    result = {
        "metric": "edge updates/s at n=2^17 (single-stream GPU CC, R-MAT, 2^20-edge batches)",
        "value": args.steps * BATCH / elapsed,
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": "rmat17_single_stream_cc", "n": N, "batch_edges": BATCH,
                   "rmat_abcd": [*RMAT, round(1 - sum(RMAT), 2)], "del_frac": DEL_FRAC,
                   "bkt_per_col": geo.bkt_per_col, "num_columns": geo.num_columns,
                   "buckets_per_vertex": geo.num_buckets, "stream_updates": len(pairs)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "aggregate_kernel", "kernel_ms": agg_ms / launches,
                     "alg_bytes_per_launch": alg_bytes,
                     "alg_bytes_per_edge": alg_bytes / BATCH,
                     "distinct_slots_per_batch": d_mean},
        "cc_query_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                        "components": ncc, "rounds": rounds},
        "pcie_inclusive": pcie,
    }
    stream.close()
    store.close()
    dbuf.free()
    if not args.no_cpu:
        threads = len(os.sched_getaffinity(0))
        threads = max(1, min(threads, 16))
        nb = min(args.cpu_batches, total_batches)
        cb = cpu_baseline(pairs[: nb * BATCH], threads)
        cb.pop("seconds")
        result["cpu_baseline"] = cb
        log(f"cpu baseline {cb['value']:.3e} edges/s on {threads} threads")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
