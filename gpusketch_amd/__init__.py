"""Python view of libgpusketch's C ABI (ctypes), used by the tests and bench.py.

The library is the product: every call goes through ``libgpusketch.so``'s ``gs_*`` entry points
(declared in include/gpusketch.h). There is no Python or CPU fallback -- importing this module
raises if the shared library is missing, and every failing call raises ``GsError`` with the
library's message.

Bucket arrays returned by ``export_buckets`` use the reference layout (src/sketch.cpp:8-37):
shape (vertices, num_buckets, 2) of uint64 {alpha, gamma}, bucket index
``column * bkt_per_col + depth``, deterministic bucket last.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgpusketch.so")

GS_OK, GS_ERR_INVALID, GS_ERR_OOM, GS_ERR_HIP, GS_ERR_OUT_OF_SAMPLES, GS_ERR_ADMISSION, \
    GS_ERR_UNSUPPORTED = range(7)


class GsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"gs error {code}: {msg}")
        self.code = code


class OutOfSamplesError(GsError):
    pass


class Geometry(ctypes.Structure):
    _fields_ = [("num_vertices", ctypes.c_uint32), ("bkt_per_col", ctypes.c_uint32),
                ("num_samples", ctypes.c_uint32), ("cols_per_sample", ctypes.c_uint32),
                ("num_columns", ctypes.c_uint32), ("num_buckets", ctypes.c_uint32),
                ("seed", ctypes.c_uint64), ("vector_len", ctypes.c_uint64),
                ("device_bytes", ctypes.c_uint64), ("flags", ctypes.c_uint32)]


class Tuning(ctypes.Structure):
    _fields_ = [("lds_depths", ctypes.c_uint32), ("item_updates", ctypes.c_uint32),
                ("bucket_bits", ctypes.c_uint32), ("launch_edges", ctypes.c_uint32),
                ("kernel_ms", ctypes.c_float)]

    def as_dict(self):
        return {"lds_depths": self.lds_depths, "item_updates": self.item_updates,
                "bucket_bits": self.bucket_bits, "launch_edges": self.launch_edges,
                "kernel_ms": self.kernel_ms}


if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is not built; run `python gpusketch_amd/build.py` (hipcc, gfx950)")

lib = ctypes.CDLL(LIB_PATH)
_u32, _u64, _i32, _vp, _dbl, _sz = (ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int,
                                    ctypes.c_void_p, ctypes.c_double, ctypes.c_size_t)
_sigs = {
    "gs_last_error": ([], ctypes.c_char_p),
    "gs_source_digest": ([_vp, _sz], _i32),
    "gs_geometry_for": ([_u32, _dbl, _u32, _vp], _i32),
    "gs_geometry_for_sketch": ([_u32, _dbl, _u32, _u32, _vp], _i32),
    "gs_hash_known_answer": ([_vp, _vp, _u64, _vp], _i32),
    "gs_store_create": ([_u32, _u64, _u32, _u32, _u32, _vp], _i32),
    "gs_store_destroy": ([_vp], _i32),
    "gs_store_geometry": ([_vp, _vp], _i32),
    "gs_store_zero": ([_vp], _i32),
    "gs_stream_create": ([_vp, _vp], _i32),
    "gs_stream_destroy": ([_vp], _i32),
    "gs_submit_vertex_batches": ([_vp, _vp, _vp, _vp, _u32, _vp], _i32),
    "gs_submit_vertex_batches_filtered": ([_vp, _vp, _vp, _vp, _u32, _vp, _u32, _u32], _i32),
    "gs_submit_edge_pairs": ([_vp, _vp, _u64], _i32),
    "gs_submit_edge_pairs_device": ([_vp, _vp, _u64], _i32),
    "gs_submit_edge_pairs_filtered": ([_vp, _vp, _u64, _u32, _u32], _i32),
    "gs_flush": ([_vp], _i32),
    "gs_flush_async": ([_vp], _i32),
    "gs_sync": ([_vp], _i32),
    "gs_export_buckets": ([_vp, _u32, _u32, _vp], _i32),
    "gs_import_buckets": ([_vp, _u32, _u32, _vp], _i32),
    "gs_store_snapshot_bytes": ([_vp, _vp], _i32),
    "gs_store_snapshot": ([_vp, _vp], _i32),
    "gs_store_restore": ([_vp, _vp], _i32),
    "gs_query_cc": ([_vp, _vp, _vp, _vp, _vp, _u64, _vp], _i32),
    "gs_k_spanning_forests": ([_vp, _u32, _vp, _u64, _vp], _i32),
    "gs_sample_vertices": ([_vp, _u32, _vp, _vp], _i32),
    "gs_min_cut": ([_u32, _vp, _u64, _vp, _vp], _i32),
    "gs_stream_set_profiling": ([_vp, _i32], _i32),
    "gs_stream_timing": ([_vp, _vp, _vp, _vp], _i32),
    "gs_stream_kernel_times": ([_vp, _vp, _vp, _vp, _vp], _i32),
    "gs_stream_traffic": ([_vp, _vp, _vp], _i32),
    "gs_stream_set_tuning": ([_vp, _vp], _i32),
    "gs_stream_get_tuning": ([_vp, _vp], _i32),
    "gs_stream_autotune": ([_vp, _vp, _u64, _vp], _i32),
    "gs_group_create": ([_vp, _u32, _vp], _i32),
    "gs_group_destroy": ([_vp], _i32),
    "gs_group_info": ([_vp, _u32, _vp, _vp], _i32),
    "gs_group_stream": ([_vp, _vp], _i32),
    "gs_group_submit_pairs": ([_vp, _vp, _vp], _i32),
    "gs_group_submit_pairs_device": ([_vp, _vp, _u64], _i32),
    "gs_group_submit_member_pairs": ([_vp, _u32, _vp, _u64], _i32),
    "gs_group_submit_member_pairs_async": ([_vp, _u32, _vp, _u64], _i32),
    "gs_group_submit_member_batches_async": ([_vp, _u32, _vp, _vp, _vp], _i32),
    "gs_stream_wait_copies": ([_vp], _i32),
    "gs_group_query_cc": ([_vp, _vp, _vp, _vp, _vp, _vp], _i32),
    "gs_group_fence": ([_vp], _i32),
    "gs_group_prepare": ([_vp], _i32),
    "gs_apply_edge_pairs_device_noalloc": ([_vp, _vp, _u64], _i32),
    "gs_host_alloc": ([_sz, _vp], _i32),
    "gs_host_free": ([_vp], _i32),
    "gs_device_alloc": ([_sz, _vp], _i32),
    "gs_device_free": ([_vp], _i32),
    "gs_memcpy_h2d": ([_vp, _vp, _sz], _i32),
    "gs_memcpy_d2h": ([_vp, _vp, _sz], _i32),
    "gs_device_synchronize": ([], _i32),
    "gs_device_copy": ([_vp, _vp, _sz], _i32),
    "gs_gen_erdos_renyi": ([_u32, _dbl, _dbl, _u64, _vp, _vp], _i32),
    "gs_gen_rmat": ([_u32, _dbl, _dbl, _dbl, _u64, _dbl, _u64, _vp], _i32),
    "gs_gen_planted_cut": ([_u32, _dbl, _u32, _dbl, _u64, _vp, _vp], _i32),
}
EXPORTED = sorted(_sigs)
for _name, (_args, _res) in _sigs.items():
    _f = getattr(lib, _name)
    _f.argtypes = _args
    _f.restype = _res


def _check(rc: int) -> None:
    if rc != GS_OK:
        msg = lib.gs_last_error().decode(errors="replace")
        if rc == GS_ERR_OUT_OF_SAMPLES:
            raise OutOfSamplesError(rc, msg)
        raise GsError(rc, msg)


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _pairs(pairs) -> np.ndarray:
    return np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1, 2)


# most pairs per gs_apply_edge_pairs_device_noalloc call (GS_NOALLOC_MAX_EDGES)
NOALLOC_MAX_EDGES = 1 << 20
# gs_store_create flag: the CubeSketch (L0_SAMPLING) update rule
STORE_CUBESKETCH = 1
# columns per sample of the CubeSketch configuration (include/sketch.h:208-211)
CUBESKETCH_COLS_PER_SAMPLE = 7


def source_digest() -> str:
    """SHA-256 of the sources the loaded library was built from (gs_source_digest)."""
    buf = ctypes.create_string_buffer(80)
    _check(lib.gs_source_digest(buf, len(buf)))
    return buf.value.decode()


def geometry_for(n: int, sketches_factor: float = 1.0, cols_per_sample: int = 1,
                 cubesketch: bool = False) -> Geometry:
    g = Geometry()
    _check(lib.gs_geometry_for_sketch(n, sketches_factor, cols_per_sample,
                                      STORE_CUBESKETCH if cubesketch else 0, ctypes.byref(g)))
    return g


def hash_known_answer(keys, seeds) -> np.ndarray:
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
    out = np.zeros_like(keys)
    _check(lib.gs_hash_known_answer(_p(keys), _p(seeds), keys.size, _p(out)))
    return out


def min_cut(n: int, pairs):
    pairs = _pairs(pairs)
    val = ctypes.c_uint64(0)
    side = np.zeros(n, dtype=np.uint8)
    _check(lib.gs_min_cut(n, _p(pairs), pairs.shape[0], ctypes.byref(val), _p(side)))
    return int(val.value), side


def gen_erdos_renyi(n: int, p: float, del_frac: float, seed: int) -> np.ndarray:
    cnt = ctypes.c_uint64(0)
    _check(lib.gs_gen_erdos_renyi(n, p, del_frac, seed, None, ctypes.byref(cnt)))
    out = np.zeros((cnt.value, 2), dtype=np.uint32)
    _check(lib.gs_gen_erdos_renyi(n, p, del_frac, seed, _p(out), ctypes.byref(cnt)))
    return out[: cnt.value]


def gen_rmat(scale: int, count: int, seed: int, a=0.57, b=0.19, c=0.19,
             del_frac: float = 0.05) -> np.ndarray:
    out = np.zeros((count, 2), dtype=np.uint32)
    _check(lib.gs_gen_rmat(scale, a, b, c, count, del_frac, seed, _p(out)))
    return out


def gen_planted_cut(n: int, p: float, cut: int, churn_frac: float, seed: int) -> np.ndarray:
    cnt = ctypes.c_uint64(0)
    _check(lib.gs_gen_planted_cut(n, p, cut, churn_frac, seed, None, ctypes.byref(cnt)))
    out = np.zeros((cnt.value, 2), dtype=np.uint32)
    _check(lib.gs_gen_planted_cut(n, p, cut, churn_frac, seed, _p(out), ctypes.byref(cnt)))
    return out[: cnt.value]


class PinnedBuffer:
    """Page-locked host memory (gs_host_alloc) with a numpy view: submissions from it are copied
    to the device by DMA without staging."""

    def __init__(self, shape, dtype=np.uint32):
        dt = np.dtype(dtype)
        self.nbytes = int(np.prod(shape)) * dt.itemsize
        self.ptr = ctypes.c_void_p()
        _check(lib.gs_host_alloc(max(1, self.nbytes), ctypes.byref(self.ptr)))
        raw = (ctypes.c_char * max(1, self.nbytes)).from_address(self.ptr.value)
        self.array = np.frombuffer(raw, dtype=dt, count=int(np.prod(shape))).reshape(shape)

    def free(self) -> None:
        if self.ptr:
            self.array = None
            _check(lib.gs_host_free(self.ptr))
            self.ptr = ctypes.c_void_p()


class DeviceBuffer:
    """Device memory owned by the caller (inputs kept resident in HBM)."""

    def __init__(self, host: np.ndarray):
        host = np.ascontiguousarray(host)
        self.nbytes = host.nbytes
        self.ptr = ctypes.c_void_p()
        _check(lib.gs_device_alloc(max(self.nbytes, 1), ctypes.byref(self.ptr)))
        if self.nbytes:
            _check(lib.gs_memcpy_h2d(self.ptr, _p(host), self.nbytes))

    def offset(self, nbytes: int) -> ctypes.c_void_p:
        return ctypes.c_void_p(self.ptr.value + nbytes)

    def free(self):
        if self.ptr:
            _check(lib.gs_device_free(self.ptr))
            self.ptr = ctypes.c_void_p()


class Stream:
    def __init__(self, store: "SketchStore"):
        self.store = store
        self.h = ctypes.c_void_p()
        _check(lib.gs_stream_create(store.h, ctypes.byref(self.h)))

    def submit_pairs(self, pairs) -> None:
        pairs = _pairs(pairs)
        _check(lib.gs_submit_edge_pairs(self.h, _p(pairs), pairs.shape[0]))

    def submit_pairs_filtered(self, pairs, min_depth: int, depth_cap: int) -> None:
        pairs = _pairs(pairs)
        _check(lib.gs_submit_edge_pairs_filtered(self.h, _p(pairs), pairs.shape[0], min_depth,
                                                 depth_cap))

    def submit_pairs_device(self, dptr, nedges: int) -> None:
        _check(lib.gs_submit_edge_pairs_device(self.h, dptr, nedges))

    def submit_vertex_batches(self, src, sizes, offsets, dsts, min_depth: int = 0,
                              depth_cap: int = 0) -> None:
        src = np.ascontiguousarray(src, dtype=np.uint32)
        sizes = np.ascontiguousarray(sizes, dtype=np.uint64)
        offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        dsts = np.ascontiguousarray(dsts, dtype=np.uint32)
        if min_depth:
            _check(lib.gs_submit_vertex_batches_filtered(self.h, _p(src), _p(sizes), _p(offsets),
                                                         len(src), _p(dsts), min_depth, depth_cap))
        else:
            _check(lib.gs_submit_vertex_batches(self.h, _p(src), _p(sizes), _p(offsets), len(src),
                                                _p(dsts)))

    def flush(self) -> None:
        _check(lib.gs_flush(self.h))

    def flush_async(self) -> None:
        """Launches the pending updates without waiting (a later flush/sync/query waits)."""
        _check(lib.gs_flush_async(self.h))

    def set_profiling(self, on: bool) -> None:
        _check(lib.gs_stream_set_profiling(self.h, int(on)))

    def timing(self):
        ms, launches, upd = ctypes.c_double(), ctypes.c_uint64(), ctypes.c_uint64()
        _check(lib.gs_stream_timing(self.h, ctypes.byref(ms), ctypes.byref(launches),
                                    ctypes.byref(upd)))
        return ms.value, launches.value, upd.value

    def kernel_times(self):
        """(grouping ms, aggregation ms, launches, directed updates) of the profiled launches
        since the last call."""
        g, ms, launches, upd = ctypes.c_double(), ctypes.c_double(), ctypes.c_uint64(), ctypes.c_uint64()
        _check(lib.gs_stream_kernel_times(self.h, ctypes.byref(g), ctypes.byref(ms),
                                          ctypes.byref(launches), ctypes.byref(upd)))
        return g.value, ms.value, launches.value, upd.value

    def traffic(self):
        """(device buckets read-modify-written by flushes, updates XORed straight into HBM) of
        the profiled launches since the last call."""
        f, d = ctypes.c_uint64(), ctypes.c_uint64()
        _check(lib.gs_stream_traffic(self.h, ctypes.byref(f), ctypes.byref(d)))
        return f.value, d.value

    def set_tuning(self, lds_depths: int = 0, item_updates: int = 0, bucket_bits: int = 0,
                   launch_edges: int = 0):
        t = Tuning(lds_depths, item_updates, bucket_bits, launch_edges, 0.0)
        _check(lib.gs_stream_set_tuning(self.h, ctypes.byref(t)))

    def get_tuning(self) -> Tuning:
        t = Tuning()
        _check(lib.gs_stream_get_tuning(self.h, ctypes.byref(t)))
        return t

    def autotune(self, dptr, nedges: int) -> Tuning:
        t = Tuning()
        _check(lib.gs_stream_autotune(self.h, dptr, nedges, ctypes.byref(t)))
        return t

    def close(self) -> None:
        if self.h:
            _check(lib.gs_stream_destroy(self.h))
            self.h = ctypes.c_void_p()


class SketchStore:
    """n vertex sketches resident in HBM (one seed, reference geometry).

    cubesketch=True selects the L0_SAMPLING update rule (src/sketch.cpp:66-83) and its sizing;
    cols_per_sample defaults to 7 then (include/sketch.h:208-211), to 1 otherwise."""

    def __init__(self, n: int, seed: int, num_samples: int | None = None,
                 sketches_factor: float = 1.0, cols_per_sample: int | None = None,
                 cubesketch: bool = False):
        if cols_per_sample is None:
            cols_per_sample = CUBESKETCH_COLS_PER_SAMPLE if cubesketch else 1
        if num_samples is None:
            num_samples = geometry_for(n, sketches_factor, cols_per_sample, cubesketch).num_samples
        self.h = ctypes.c_void_p()
        _check(lib.gs_store_create(n, seed & (2**64 - 1), num_samples, cols_per_sample,
                                   STORE_CUBESKETCH if cubesketch else 0, ctypes.byref(self.h)))
        self.geom = Geometry()
        _check(lib.gs_store_geometry(self.h, ctypes.byref(self.geom)))
        self.n = n

    def stream(self) -> Stream:
        return Stream(self)

    def sync(self) -> None:
        _check(lib.gs_sync(self.h))

    def zero(self) -> None:
        _check(lib.gs_store_zero(self.h))

    def apply_pairs_device_noalloc(self, dptr, nedges: int) -> None:
        """gs_apply_edge_pairs_device_noalloc: device pairs applied at once, no device allocation
        (at most NOALLOC_MAX_EDGES per call)."""
        _check(lib.gs_apply_edge_pairs_device_noalloc(self.h, dptr, nedges))

    def export_buckets(self, v0: int = 0, count: int | None = None) -> np.ndarray:
        count = self.n - v0 if count is None else count
        out = np.zeros((count, self.geom.num_buckets, 2), dtype=np.uint64)
        _check(lib.gs_export_buckets(self.h, v0, count, _p(out)))
        return out

    def import_buckets(self, data: np.ndarray, v0: int = 0) -> None:
        data = np.ascontiguousarray(data, dtype=np.uint64)
        _check(lib.gs_import_buckets(self.h, v0, data.shape[0], _p(data)))

    def snapshot(self) -> np.ndarray:
        """gs_store_snapshot: the store's device-order snapshot (bytes)."""
        nbytes = ctypes.c_uint64()
        _check(lib.gs_store_snapshot_bytes(self.h, ctypes.byref(nbytes)))
        out = np.zeros(nbytes.value, dtype=np.uint8)
        _check(lib.gs_store_snapshot(self.h, _p(out)))
        return out

    def restore(self, snap: np.ndarray) -> None:
        """gs_store_restore from a snapshot of a store of the same geometry and seed."""
        snap = np.ascontiguousarray(snap, dtype=np.uint8)
        _check(lib.gs_store_restore(self.h, _p(snap)))

    def query_cc(self, want_forest: bool = False):
        labels = np.zeros(self.n, dtype=np.uint32)
        ncc, rounds, flen = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint64()
        forest = np.zeros((self.n, 2), dtype=np.uint32) if want_forest else None
        _check(lib.gs_query_cc(self.h, _p(labels), ctypes.byref(ncc), ctypes.byref(rounds),
                               _p(forest) if want_forest else None,
                               self.n if want_forest else 0, ctypes.byref(flen)))
        if want_forest:
            return labels, int(ncc.value), int(rounds.value), forest[: flen.value].copy()
        return labels, int(ncc.value), int(rounds.value)

    def sample_vertices(self, sample_idx: int):
        """Every vertex sketch's sample at sample_idx as the device query's round 0 computes it:
        (result codes 0 GOOD / 1 ZERO / 2 FAIL, sampled indices)."""
        res = np.zeros(self.n, dtype=np.int32)
        idx = np.zeros(self.n, dtype=np.uint64)
        _check(lib.gs_sample_vertices(self.h, sample_idx, _p(res), _p(idx)))
        return res, idx

    def k_spanning_forests(self, k: int):
        cap = k * self.n
        edges = np.zeros((cap, 2), dtype=np.uint32)
        counts = np.zeros(k, dtype=np.uint64)
        _check(lib.gs_k_spanning_forests(self.h, k, _p(edges), cap, _p(counts)))
        out, off = [], 0
        for c in counts:
            out.append(edges[off:off + int(c)].copy())
            off += int(c)
        return out

    def close(self) -> None:
        """Frees the store (the library refuses while a group holds it: close the group first)."""
        if self.h:
            _check(lib.gs_store_destroy(self.h))
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class _GroupStream(Stream):
    """The stream a group owns (flush / profiling / timing; destroyed with the group)."""

    def __init__(self, h):
        self.store = None
        self.h = h

    def close(self) -> None:
        self.h = ctypes.c_void_p()


class StoreGroup:
    """Stores ingested and queried together (batched streams): one launch sequence for all
    members' edge batches, one batched Boruvka for all members' connectivity queries."""

    def __init__(self, stores):
        self.stores = list(stores)
        arr = (ctypes.c_void_p * len(self.stores))(*[s.h.value for s in self.stores])
        self.h = ctypes.c_void_p()
        _check(lib.gs_group_create(arr, len(self.stores), ctypes.byref(self.h)))
        self.bases = []
        total = ctypes.c_uint32()
        for i in range(len(self.stores)):
            b = ctypes.c_uint32()
            _check(lib.gs_group_info(self.h, i, ctypes.byref(b), ctypes.byref(total)))
            self.bases.append(int(b.value))
        self.total = int(total.value)
        sh = ctypes.c_void_p()
        _check(lib.gs_group_stream(self.h, ctypes.byref(sh)))
        self.stream = _GroupStream(sh)

    def submit_pairs(self, pairs_per_member) -> None:
        """pairs_per_member: one (m_i, 2) uint32 array (member-local ids) or None per member."""
        keep = [None if p is None else _pairs(p) for p in pairs_per_member]
        ptrs = (ctypes.c_void_p * len(keep))(*[None if p is None else p.ctypes.data for p in keep])
        counts = np.array([0 if p is None else p.shape[0] for p in keep], dtype=np.uint64)
        _check(lib.gs_group_submit_pairs(self.h, ptrs, _p(counts)))

    def submit_pairs_device(self, dptr, nedges: int) -> None:
        _check(lib.gs_group_submit_pairs_device(self.h, dptr, nedges))

    def submit_member_pairs(self, member: int, pairs) -> None:
        """One member's (m, 2) uint32 pairs in its own ids, rebased and range-checked on the
        device (out-of-range pairs are dropped and reported by the next flush)."""
        pairs = _pairs(pairs)
        _check(lib.gs_group_submit_member_pairs(self.h, member, _p(pairs), pairs.shape[0]))

    def submit_member_pairs_async(self, member: int, pairs) -> None:
        """gs_group_submit_member_pairs_async: as submit_member_pairs, but a pinned array is read
        by a DMA that may still run after the call returns; keep it unchanged until
        wait_copies (or a flush) has returned."""
        pairs = _pairs(pairs)
        _check(lib.gs_group_submit_member_pairs_async(self.h, member, _p(pairs), pairs.shape[0]))

    def submit_member_batches_async(self, members, batches) -> None:
        """gs_group_submit_member_batches_async: batch i of `batches` (an (m, 2) uint32 array in
        member members[i]'s ids) as submit_member_pairs_async, all in one call. Keep pinned
        arrays unchanged until wait_copies (or a flush) has returned."""
        batches = [_pairs(b) for b in batches]
        mem = np.ascontiguousarray(members, dtype=np.uint32)
        if mem.shape != (len(batches),):
            raise ValueError("one member per batch")
        ptrs = np.array([b.ctypes.data for b in batches], dtype=np.uint64)
        counts = np.array([b.shape[0] for b in batches], dtype=np.uint64)
        _check(lib.gs_group_submit_member_batches_async(self.h, len(batches), _p(mem), _p(ptrs),
                                                        _p(counts)))

    def wait_copies(self) -> None:
        """gs_stream_wait_copies on the group's stream: every host batch has been read."""
        _check(lib.gs_stream_wait_copies(self.stream.h))

    def fence(self) -> None:
        """gs_group_fence: the next query waits for the updates submitted so far only."""
        _check(lib.gs_group_fence(self.h))

    def prepare(self) -> None:
        """Allocates the group's ingest buffers and query scratch now rather than on first use."""
        _check(lib.gs_group_prepare(self.h))

    def flush(self) -> None:
        self.stream.flush()

    def query_cc(self, mask=None, allow_out_of_samples: bool = False):
        """Returns (labels per member, num_cc, rounds, status) for the masked members."""
        g = len(self.stores)
        labels = np.zeros(self.total, dtype=np.uint32)
        ncc = np.zeros(g, dtype=np.uint32)
        rounds = np.zeros(g, dtype=np.uint32)
        status = np.zeros(g, dtype=np.int32)
        m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
        rc = lib.gs_group_query_cc(self.h, None if m is None else _p(m), _p(labels), _p(ncc),
                                   _p(rounds), _p(status))
        if rc != GS_OK and not (allow_out_of_samples and rc == GS_ERR_OUT_OF_SAMPLES):
            _check(rc)
        # members outside the mask: the library leaves their entries alone; report identity labels
        per = [labels[b:b + s.n].copy() if m is None or m[i] else np.arange(s.n, dtype=np.uint32)
               for i, (b, s) in enumerate(zip(self.bases, self.stores))]
        return per, ncc, rounds, status

    def close(self) -> None:
        if self.h:
            _check(lib.gs_group_destroy(self.h))
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
