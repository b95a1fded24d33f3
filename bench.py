"""Benchmark: sketch edge updates/s at n = 2^17 on one MI355X.

Workload (single-stream GPU CC config): R-MAT n=2^17, (a,b,c,d) = (0.57,0.19,0.19,0.05), about
5% of updates delete an earlier inserted edge, no self loops or duplicate edges, batches of 2^20
uint32 pairs. One step = one launch of the grouping and aggregation kernels over 2^25 edges,
submitted as 32 batches that are already resident in HBM (config.input); the timed region ends
with a flush, so every timed update is applied. Every step takes the next 32 batches of one
continuous seeded stream (warmup steps first); the defaults (5 + 27 steps) cover a 2^30-update
stream. The rate from pinned host memory, PCIe copies included, is reported beside it
(pcie_inclusive), never as the value.

Prints one JSON line (the contract of the driver) with a `roofline` object for the dominant
kernel (aggregate_kernel, timed with HIP events on its own stream; its bytes counted by the
kernel itself; its VALU issue rate from a committed counter profile of the same launch size) and a
`cpu_baseline` object (the CPU restatement of the reference update path, timed on every usable
host core over a bounded prefix of the same stream), plus the CPU Boruvka query on the final
store beside the GPU query latency.
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


def cpu_baseline(pairs: np.ndarray, threads: int, cubesketch: bool = False) -> dict:
    """Reference CPU update path (per-thread delta sketch + mutex merge) on pre-grouped vertex
    batches of 4 * num_buckets destinations, the reference's batch size."""
    from oracle import oracle as orc
    g = orc.Geometry(N, 1.0, 7, True) if cubesketch else orc.Geometry(N)
    t0 = time.perf_counter()
    src, sizes, offs, dsts = orc.group_by_vertex(pairs, N, 4 * g.num_buckets)
    group_s = time.perf_counter() - t0
    store = g.empty_store()
    secs = orc.cpu_baseline(store, g, SEED_SKETCH, src, sizes, offs, dsts, threads, cubesketch)
    del store
    return {"value": len(pairs) / secs, "unit": "edges/s", "cores": threads, "kind": "port",
            "sample": f"first {len(pairs)} updates of the same stream ({len(pairs) // BATCH} "
                      f"batches of 2^20), grouped into {len(src)} vertex batches of <= "
                      f"{4 * g.num_buckets} (grouping {group_s:.2f}s, not timed), "
                      f"{secs:.1f}s timed",
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


def batched_workload(graphs: int, seed: int):
    """Independent graphs for the batched-streams workload: n = 2^s with s uniform in [10, 16]
    (log-uniform n in [2^10, 2^16]), 4..64 updates per vertex, 5% deletions of live edges; even
    graphs Erdos-Renyi, odd graphs R-MAT(0.57, 0.19, 0.19, 0.05)."""
    import gpusketch_amd as gs
    rng = np.random.default_rng(seed)
    scales = rng.integers(10, 17, size=graphs)
    upv = rng.integers(4, 65, size=graphs)
    out = []
    for g in range(graphs):
        n = 1 << int(scales[g])
        m = int(upv[g]) * n // 2  # updates: each touches two vertices
        if g % 2 == 0:
            p = min(1.0, m / (1 + DEL_FRAC) / (n * (n - 1) / 2))
            pairs = gs.gen_erdos_renyi(n, p, DEL_FRAC, seed + g)
        else:
            pairs = gs.gen_rmat(int(scales[g]), m, seed + g, *RMAT, del_frac=DEL_FRAC)
        out.append((n, 0xB0000000 + 7919 * g, pairs))
    return out


def bench_batched(args) -> dict:
    """Config: 256 independent graphs. Plain = every graph through its own store stream, one
    after another; batched = one group: one autotuned ingest launch sequence over all graphs'
    interleaved updates and one batched Boruvka (captured rounds) for all connectivity queries.
    Every member must end bit-exact with its plain run and with the same partition."""
    import gpusketch_amd as gs
    t0 = time.perf_counter()
    work = batched_workload(args.graphs, SEED_STREAM)
    total_upd = sum(len(p) for _, _, p in work)
    log(f"generated {len(work)} graphs, {sum(n for n, _, _ in work)} vertices, {total_upd} "
        f"updates in {time.perf_counter() - t0:.1f}s")
    stores = [gs.SketchStore(n, sd) for n, sd, _ in work]
    sketch_gb = sum(s.geom.device_bytes for s in stores) / 1e9
    # plain: graph after graph, each from its own device-resident array, 2^20-edge batches
    bufs = [gs.DeviceBuffer(p) for _, _, p in work]
    streams = [s.stream() for s in stores]
    gs.lib.gs_device_synchronize()
    t0 = time.perf_counter()
    for (n, sd, p), st, buf in zip(work, streams, bufs):
        for b in range(0, len(p), BATCH):
            st.submit_pairs_device(buf.offset(b * 8), min(BATCH, len(p) - b))
        st.flush()
    gs.lib.gs_device_synchronize()
    plain_s = time.perf_counter() - t0
    for b in bufs:
        b.free()
    sample = list(range(0, len(work), max(1, len(work) // 8)))
    ref = {i: stores[i].export_buckets() for i in sample}
    lat_plain = []
    plain_labels = {}
    for _ in range(args.query_reps):
        q0 = time.perf_counter()
        for i, s in enumerate(stores):
            lab, _, _ = s.query_cc()
            plain_labels[i] = lab
        lat_plain.append((time.perf_counter() - q0) * 1e3)
    for st in streams:
        st.close()
    for s in stores:
        s.zero()
    log(f"plain: {total_upd / plain_s:.3e} updates/s, all-graph CC {np.median(lat_plain):.1f} ms")

    with gs.StoreGroup(stores) as grp:
        glob = np.concatenate([p + np.uint32(b) for (_, _, p), b in zip(work, grp.bases)])
        glob = glob[np.random.default_rng(1).permutation(len(glob))]  # graphs interleaved
        gbuf = gs.DeviceBuffer(glob)
        del glob
        # one launch sequence covers 2^25 updates of all graphs at once: a vertex's updates
        # aggregate in LDS across the whole submission
        chunk = 1 << 25
        sample_edges = min(chunk, total_upd)
        best = grp.stream.autotune(gbuf.offset(0), sample_edges)  # leaves the sketches as they are
        log(f"autotune: {best.as_dict()}")
        gs.lib.gs_device_synchronize()
        t0 = time.perf_counter()
        for b in range(0, total_upd, chunk):
            grp.submit_pairs_device(gbuf.offset(b * 8), min(chunk, total_upd - b))
        grp.flush()
        gs.lib.gs_device_synchronize()
        batched_s = time.perf_counter() - t0
        gbuf.free()
        exact = all(np.array_equal(stores[i].export_buckets(), ref[i]) for i in sample)
        lat_batched = []
        same = True
        for _ in range(args.query_reps):
            q0 = time.perf_counter()
            labels, ncc, rounds, status = grp.query_cc()
            lat_batched.append((time.perf_counter() - q0) * 1e3)
            same = same and all(np.array_equal(labels[i], plain_labels[i]) for i in range(len(work)))
    for s in stores:
        s.close()
    log(f"batched: {total_upd / batched_s:.3e} updates/s, all-graph CC {np.median(lat_batched):.1f} ms, "
        f"bit-exact sample {exact}, partitions equal {same}")
    return {
        "metric": "batched sketch updates/s (256 independent graphs, one group)",
        "value": total_upd / batched_s, "unit": "edges/s", "n_gpus": 1,
        "steps": int((total_upd + (1 << 25) - 1) // (1 << 25)), "warmup": 0,
        "ms_per_step": batched_s * 1e3 / max(1, (total_upd + (1 << 25) - 1) // (1 << 25)),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": "batched_streams", "graphs": len(work),
                   "vertices": int(sum(n for n, _, _ in work)), "updates": int(total_upd),
                   "sketch_gb": sketch_gb, "tuning": best.as_dict()},
        "plain": {"value": total_upd / plain_s, "unit": "edges/s",
                  "all_graph_cc_ms_p50": float(np.median(lat_plain))},
        "batched": {"all_graph_cc_ms_p50": float(np.median(lat_batched)),
                    "all_graph_cc_ms_p99": float(np.percentile(lat_batched, 99))},
        "checks": {"members_bit_exact_sample": exact, "sample_members": len(sample),
                   "partitions_equal_all": same},
    }


def bench_mincut(args) -> dict:
    """Min-cut / k-connectivity: n = 2^15, two Erdos-Renyi(0.1) halves joined by exactly 5
    crossing edges, the stream shuffled with 10% insert/delete churn of non-edges; k = 8 edge-
    disjoint spanning forests from the sketches (device), exact minimum cut of their union
    (host), checked against the exact minimum cut of the whole net graph."""
    import gpusketch_amd as gs
    n, k, cut = 1 << 15, 8, 5
    t0 = time.perf_counter()
    pairs = gs.gen_planted_cut(n, 0.1, cut, 0.1, SEED_STREAM)
    log(f"generated {len(pairs)} updates in {time.perf_counter() - t0:.1f}s")
    factor = k / np.log2(n) * 1.5  # the sketch factor of the reference's min-cut tool
    with gs.SketchStore(n, SEED_SKETCH, sketches_factor=factor) as st:
        s = st.stream()
        buf = gs.DeviceBuffer(pairs)
        gs.lib.gs_device_synchronize()
        t0 = time.perf_counter()
        for b in range(0, len(pairs), BATCH):
            s.submit_pairs_device(buf.offset(b * 8), min(BATCH, len(pairs) - b))
        s.flush()
        gs.lib.gs_device_synchronize()
        ingest_s = time.perf_counter() - t0
        buf.free()
        lat = []
        for _ in range(3):
            q0 = time.perf_counter()
            forests = st.k_spanning_forests(k)
            lat.append((time.perf_counter() - q0) * 1e3)
        cert = np.concatenate(forests)
        c0 = time.perf_counter()
        value, _ = gs.min_cut(n, cert)
        cert_ms = (time.perf_counter() - c0) * 1e3
        geo = st.geom
        s.close()
    keys = (pairs.min(1).astype(np.uint64) << np.uint64(32)) | pairs.max(1).astype(np.uint64)
    uniq, cnt = np.unique(keys, return_counts=True)
    live = uniq[cnt % 2 == 1]
    net = np.stack([live >> np.uint64(32), live & np.uint64(0xFFFFFFFF)], 1).astype(np.uint32)
    e0 = time.perf_counter()
    exact, _ = gs.min_cut(n, net)
    exact_s = time.perf_counter() - e0
    log(f"min cut: certificate {value}, exact {exact}; forests p50 {np.median(lat):.1f} ms")
    return {
        "metric": "min-cut stream: sketch updates/s and k-forest query latency (n=2^15, k=8)",
        "value": len(pairs) / ingest_s, "unit": "edges/s", "n_gpus": 1,
        "steps": int((len(pairs) + BATCH - 1) // BATCH), "warmup": 0,
        "ms_per_step": ingest_s * 1e3 / ((len(pairs) + BATCH - 1) // BATCH),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": "planted_cut_n32768", "n": n, "k": k, "planted_cut": cut,
                   "p": 0.1, "churn": 0.1, "updates": int(len(pairs)),
                   "net_edges": int(len(net)), "num_columns": geo.num_columns,
                   "bkt_per_col": geo.bkt_per_col, "sketches_factor": factor},
        "k_forests_ms": {"p50": float(np.median(lat)), "max": float(np.max(lat))},
        "certificate_edges": int(len(cert)), "certificate_cut_ms": cert_ms,
        "min_cut": int(value), "exact_min_cut": int(exact), "exact_host_s": exact_s,
        "equal": int(value) == int(exact) == cut,
    }


def run_serving(rate: float, requests: int = 1000, tenants: int = 64,
                budget_gb: float = 200, cc_frac: float = 0.1) -> dict:
    """The serving trace (C++ program serve_bench): tenants with n log-uniform in
    [2^12, 2^18], Poisson arrivals of update batches of 1k..1M edges (5% deletions), 10% CC
    queries (about 100 per trace, so that p99 is a percentile), 0.5% min-cut queries (k = 4) on 8
    tenants, 200 GB admission budget; served one request at a time (plain) and with continuous
    batching (served). Update payloads sit in pinned host memory from arrival (the trace is
    placed there before the clock starts) and reach the device by DMA without a host copy."""
    import subprocess
    prog = os.path.join(ROOT, "gpusketch_amd", "bin", "serve_bench")
    r = subprocess.run([prog, "--tenants", str(tenants), "--requests", str(requests),
                        "--rate", str(rate), "--budget-gb", str(budget_gb), "--k", "4",
                        "--mc-tenants", "8", "--cc-frac", str(cc_frac)],
                       capture_output=True, text=True, timeout=900)
    for line in r.stderr.splitlines():
        log(line)
    if r.returncode != 0:
        raise RuntimeError(f"serve_bench failed ({r.returncode}): {r.stdout} {r.stderr}")
    return json.loads(r.stdout.strip().splitlines()[-1])


LAUNCH_EDGES = 1 << 25            # edges per launch of the grouping + aggregation kernels = 1 step
STEP_BATCHES = LAUNCH_EDGES // BATCH
CU_COUNT, CLOCK_GHZ = 256, 2.4
# VALU issue ceiling: each CU's 4 SIMDs are 16 lanes wide and take 4 clocks per wave64
# instruction, so a CU issues one wave64 VALU instruction per clock
VALU_PEAK_GINSTR = CU_COUNT * CLOCK_GHZ


def host_cpu() -> dict:
    """CPU model, the host's logical CPUs, those this process may run on, and the cgroup CPU
    quota; `usable` = the smaller of the last two (one worker thread per usable core)."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    affinity = len(os.sched_getaffinity(0))
    quota = None
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()
        if q != "max":
            quota = max(1, int(round(int(q) / int(period))))
    except (OSError, ValueError):
        pass
    usable = min(affinity, quota) if quota else affinity
    return {"model": model, "nproc": os.cpu_count(), "affinity": affinity,
            "cgroup_quota_cores": quota, "usable": usable}


def build_provenance() -> dict:
    """Digest of the sources the loaded libgpusketch.so was compiled from, and whether it is the
    digest of the sources in this tree."""
    import importlib.util

    import gpusketch_amd as gs
    spec = importlib.util.spec_from_file_location(
        "gpusketch_amd_build", os.path.join(ROOT, "gpusketch_amd", "build.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    loaded = gs.source_digest()
    return {"library": "gpusketch_amd/libgpusketch.so", "source_digest": loaded[:16],
            "matches_tree": loaded == mod.source_digest()}


def load_profile(name: str, edges_per_launch: float):
    """A committed counter summary (profiles/<name>), used only when it was measured at the launch
    size this run times."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    prof = json.load(open(path))
    if abs(prof.get("edges_per_launch", 0) - edges_per_launch) > 1e-3 * edges_per_launch:
        return None
    return prof


def bench_single(args) -> dict:
    """Single-stream GPU CC (R-MAT n=2^17). One step = one launch of the grouping and aggregation
    kernels over 2^25 edges, handed to the stream as 32 device-resident batches of 2^20 pairs;
    every step takes the next 32 batches of one continuous seeded stream (warmup steps first)."""
    import gpusketch_amd as gs
    cube = args.mode == "cubesketch"
    if cube:  # 9x the hashing per update: no serving trace or PCIe leg, a short CPU sample
        args.no_serving = True
        args.pcie_batches = 0
        args.cpu_batches = min(args.cpu_batches, 2)
    total_steps = args.warmup + args.steps
    total = total_steps * LAUNCH_EDGES
    t0 = time.perf_counter()
    pairs = gs.gen_rmat(SCALE, total, SEED_STREAM, *RMAT, del_frac=DEL_FRAC)
    gen_s = time.perf_counter() - t0
    log(f"generated {len(pairs)} R-MAT updates in {gen_s:.1f}s")

    store = gs.SketchStore(N, SEED_SKETCH, cubesketch=cube)
    geo = store.geom
    stream = store.stream()
    stream.set_tuning(lds_depths=args.lds_depths, item_updates=args.item_updates,
                      bucket_bits=args.bucket_bits, launch_edges=LAUNCH_EDGES)
    dbuf = gs.DeviceBuffer(pairs)
    log(f"store {geo.device_bytes / 1e9:.3f} GB, B={geo.bkt_per_col} C={geo.num_columns}; "
        f"input resident in HBM ({dbuf.nbytes / 1e9:.2f} GB)")

    def run_steps(first: int, count: int) -> None:
        for st in range(first, first + count):
            for k in range(STEP_BATCHES):
                b = st * STEP_BATCHES + k
                stream.submit_pairs_device(dbuf.offset(b * BATCH * 8), BATCH)

    run_steps(0, args.warmup)
    stream.flush()
    log(f"warmup {args.warmup} steps done")

    stream.set_profiling(True)
    stream.kernel_times()
    stream.traffic()  # reset the counters
    gs.lib.gs_device_synchronize()
    t0 = time.perf_counter()
    run_steps(args.warmup, args.steps)
    stream.flush()
    gs.lib.gs_device_synchronize()
    elapsed = time.perf_counter() - t0
    grp_ms, agg_ms, launches, agg_updates = stream.kernel_times()
    flushed, direct = stream.traffic()
    stream.set_profiling(False)
    kname = "wide_aggregate_kernel" if cube else "aggregate_kernel"
    log(f"timed {args.steps} steps in {elapsed:.4f}s ({launches} launches); per launch: grouping "
        f"{grp_ms / launches:.4f} ms, {kname} {agg_ms / launches:.4f} ms")

    # CC query latency on the final store (the plain single-stream path): the first query of a
    # store allocates its query scratch (reported apart), then args.query_reps timed queries
    q0 = time.perf_counter()
    labels, _, _ = store.query_cc()
    first_ms = (time.perf_counter() - q0) * 1e3
    lat = []
    for _ in range(max(1, args.query_reps)):
        q0 = time.perf_counter()
        _, ncc, rounds = store.query_cc()
        lat.append((time.perf_counter() - q0) * 1e3)
    log(f"CC query: {ncc} components, {rounds} rounds, p50 {np.percentile(lat, 50):.2f} ms, "
        f"p99 {np.percentile(lat, 99):.2f} ms over {len(lat)} (first query {first_ms:.2f} ms)")

    # PCIe-inclusive rate: batches of the same stream handed over from pinned host memory (the
    # DMA runs on the stream's copy stream while the kernels work), into a second store
    pcie = None
    if args.pcie_batches > 0:
        nb = max(2, args.pcie_batches)
        warm = min(8, nb - 1)
        pin = gs.PinnedBuffer((nb * BATCH, 2))
        for b in range(nb):
            src = b % (total_steps * STEP_BATCHES)
            pin.array[b * BATCH:(b + 1) * BATCH] = pairs[src * BATCH:(src + 1) * BATCH]
        with gs.SketchStore(N, SEED_SKETCH) as st2:
            s2 = st2.stream()
            s2.set_tuning(launch_edges=LAUNCH_EDGES)
            for b in range(warm):
                s2.submit_pairs(pin.array[b * BATCH:(b + 1) * BATCH])
            s2.flush()
            gs.lib.gs_device_synchronize()
            p0 = time.perf_counter()
            for b in range(warm, nb):
                s2.submit_pairs(pin.array[b * BATCH:(b + 1) * BATCH])
            s2.flush()
            gs.lib.gs_device_synchronize()
            pe = time.perf_counter() - p0
            s2.close()
        pin.free()
        timed = nb - warm
        pcie = {"value": timed * BATCH / pe, "unit": "edges/s", "batches": timed,
                "h2d_gb_per_s": timed * BATCH * 8 / pe / 1e9,
                "input": "pinned-host", "batch": "2^20 uint32 pairs per submission"}
        log(f"PCIe-inclusive {pcie['value']:.3e} edges/s ({pcie['h2d_gb_per_s']:.1f} GB/s of "
            f"pairs) over {timed} pinned host batches")

    cpu = host_cpu()
    cpu_query = None
    if not args.no_cpu and not cube:
        # the CPU Boruvka (the oracle's serial restatement of CCSketchAlg::boruvka_emulation) on
        # this very store, exported in reference order; its partition must equal the device's
        from oracle import oracle as orc
        host_store = store.export_buckets()
        g = orc.Geometry(N)
        c0 = time.perf_counter()
        rc, olab, orounds, _ = orc.boruvka(host_store, g, SEED_SKETCH)
        cq = time.perf_counter() - c0
        del host_store
        cpu_query = {"ms": cq * 1e3, "threads": 1, "kind": "port",
                     "rounds": orounds, "labels_equal_device": bool(rc == 0 and np.array_equal(olab, labels))}
        log(f"CPU Boruvka query {cq * 1e3:.0f} ms (1 thread), labels equal: "
            f"{cpu_query['labels_equal_device']}")

    # algorithmic bytes of the aggregation kernel: 8 B per undirected edge (its pair, read once
    # by the grouping) + 16 B read + 16 B written per device bucket the launch changed (counted
    # by the kernel: flushed delta slots, deterministic buckets, updates XORed straight to HBM)
    edges_per_launch = agg_updates / 2 / launches
    slots_per_launch = (flushed + direct) / launches
    alg_bytes = 8.0 * edges_per_launch + 32.0 * slots_per_launch
    agg_s = agg_ms / launches / 1e3
    hbm_gbs = alg_bytes / agg_s / 1e9
    prof = load_profile("counters_wide_aggregate_rmat17.json" if cube
                        else "counters_aggregate_rmat17.json", edges_per_launch)
    hbm = {"achieved": hbm_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": hbm_gbs / HBM_PEAK_GBS}
    roof = {"bound": "hbm", **hbm, "traffic": None}
    if prof:
        roof["traffic"] = prof["hbm_bytes_per_dispatch"]
        valu_g = prof["sq_insts_valu_per_dispatch"] / agg_s / 1e9
        valu = {"achieved": valu_g, "peak": VALU_PEAK_GINSTR, "unit": "G wave64 VALU instr/s",
                "frac": valu_g / VALU_PEAK_GINSTR, "valu_busy_pct": prof.get("valu_busy_pct"),
                "source": "SQ_INSTS_VALU per dispatch (profiles, same launch size) / this run's "
                          "kernel time"}
        if valu["frac"] > hbm["frac"]:  # the binding ceiling first
            roof = {"bound": "valu", **valu, "traffic": roof["traffic"], "hbm": hbm}
        else:
            roof["valu"] = valu
    roof.update({"kernel": kname, "kernel_ms": agg_ms / launches,
                 "grouping_ms": grp_ms / launches, "edges_per_launch": edges_per_launch,
                 "alg_bytes_per_launch": alg_bytes, "alg_bytes_per_edge": alg_bytes / edges_per_launch,
                 "buckets_changed_per_launch": slots_per_launch})

    result = {
        "metric": ("edge updates/s at n=2^17, CubeSketch rule (7 columns per sample)" if cube else
                   "edge updates/s at n=2^17 + p50/p99 CC query latency, served vs plain"),
        "value": args.steps * LAUNCH_EDGES / elapsed,
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
        "config": {"workload": "rmat17_cubesketch_cc" if cube else "rmat17_single_stream_cc",
                   "n": N, "batch_edges": BATCH, "batches_per_step": STEP_BATCHES,
                   "input": "device-resident",
                   "rmat_abcd": [*RMAT, round(1 - sum(RMAT), 2)], "del_frac": DEL_FRAC,
                   "bkt_per_col": geo.bkt_per_col, "num_columns": geo.num_columns,
                   "buckets_per_vertex": geo.num_buckets, "stream_updates": len(pairs),
                   "stream_generation_s": round(gen_s, 1)},
        "roofline": roof,
        "cc_query_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                        "queries": len(lat), "first_query_with_scratch_setup": first_ms,
                        "components": ncc, "rounds": rounds, "cpu_boruvka": cpu_query},
        "pcie_inclusive": pcie,
        "build": build_provenance(),
    }
    stream.close()
    store.close()
    dbuf.free()
    if not args.no_serving:
        sv = run_serving(2000)
        result["serving"] = {
            "workload": f"{sv['tenants']} tenants, {sv['requests']} Poisson requests at "
                        f"{sv['rate_req_per_s']:g}/s, {sv['updates']} updates, {sv['queries']} queries",
            "plain": sv["plain"], "served": sv["served"],
            "answers_identical": sv["answers_identical"],
            "sketches_identical": sv["sketches_identical"],
            "alone_equals_mix": sv["alone_equals_mix"]}
    if not args.no_cpu:
        nb = min(args.cpu_batches, total_steps * STEP_BATCHES)
        cb = cpu_baseline(pairs[: nb * BATCH], cpu["usable"], cube)
        cb.pop("seconds")
        cb["host"] = cpu
        result["cpu_baseline"] = cb
        log(f"cpu baseline {cb['value']:.3e} edges/s on {cpu['usable']} threads ({cpu['model']})")
    return result



def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=27,
                    help="timed steps; a step is one 2^25-edge launch (32 batches of 2^20 pairs)")
    ap.add_argument("--warmup", type=int, default=5,
                    help="untimed steps first (defaults: 5 + 27 steps = the 2^30-update stream)")
    ap.add_argument("--cpu-batches", type=int, default=512,
                    help="stream prefix (2^20-pair batches) timed on the CPU leg")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--pcie-batches", type=int, default=264,
                    help="batches submitted from pinned host memory for the PCIe-inclusive rate "
                         "(8 warmup, then 256 timed: eight 2^25-edge launches whose copies overlap "
                         "the kernels as in a long stream)")
    ap.add_argument("--mode", choices=["single", "cubesketch", "batched", "serving", "mincut"],
                    default="single",
                    help="cubesketch: the single-stream workload under the L0_SAMPLING rule")
    ap.add_argument("--no-serving", action="store_true",
                    help="single mode: skip the served-vs-plain trace")
    ap.add_argument("--graphs", type=int, default=256, help="batched mode: independent graphs")
    ap.add_argument("--query-reps", type=int, default=0,
                    help="query repetitions (default: 3 in batched mode, 100 in single mode)")
    ap.add_argument("--lds-depths", type=int, default=0,
                    help="single mode: LDS depth window of the aggregation kernel (0: the library's)")
    ap.add_argument("--item-updates", type=int, default=0,
                    help="single mode: destinations per work item (0: the library's)")
    ap.add_argument("--bucket-bits", type=int, default=0,
                    help="single mode: vertex-id bits resolved inside one grouping bucket (0: the library's)")
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

    if not args.query_reps:
        args.query_reps = 3 if args.mode == "batched" else 100
    if args.mode == "batched":
        print(json.dumps(bench_batched(args)), flush=True)
        return
    if args.mode == "mincut":
        print(json.dumps(bench_mincut(args)), flush=True)
        return
    if args.mode == "serving":
        lat = run_serving(2000)
        cap = run_serving(0)  # every request queued at once: capacity
        # a budget of about half the tenants' memory: idle tenants spilled and restored
        evict = run_serving(2000, budget_gb=30)
        print(json.dumps({
            "metric": "served vs plain: updates/s and p50/p99 query latency (64 tenants)",
            "value": cap["served"]["updates_per_s"], "unit": "edges/s", "n_gpus": 1,
            "steps": cap["served"]["steps"], "warmup": 0,
            "ms_per_step": cap["served"]["makespan_s"] * 1e3 / max(1, cap["served"]["steps"]),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
            "data": "synthetic", "config": {"workload": "serving_64_tenants"},
            "poisson_2000_req_per_s": lat, "saturation": cap,
            "eviction_30gb_budget": evict}), flush=True)
        return

    print(json.dumps(bench_single(args)), flush=True)


if __name__ == "__main__":
    main()
