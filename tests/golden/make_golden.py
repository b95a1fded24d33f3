"""Writes the golden fixtures under tests/golden/.

* xxh3_len8.json -- XXH3_64bits_withSeed of 8-byte little-endian keys, produced by the xxhash
  library (python ``xxhash`` 3.8.1, libxxhash 0.8.2). XXH3 output is frozen since xxHash 0.8.0.
  The reference binds both of its hashes to this function (include/types.h:8-9).
* geometry.json -- sketch sizing from the formulas of include/sketch.h:63-77,198-216, evaluated
  with Python's IEEE-754 doubles (the same arithmetic as the C++ code).
* multiples_graph_1024.txt / multiples_graph_1024_stream.data -- data fixtures copied from the
  reference's test/res (the CC test expects 78 components, test/cc_alg_test.cpp:41-57).

Run from the repository root: ``python tests/golden/make_golden.py``. The reference fixtures are
read from /root/reference only when that directory exists.
"""
import json
import math
import os
import random
import shutil
import struct

import xxhash

HERE = os.path.dirname(os.path.abspath(__file__))
M64 = (1 << 64) - 1


def xxh3(key, seed):
    return xxhash.xxh3_64_intdigest(struct.pack("<Q", key & M64), seed=seed & M64)


def geometry(n, factor=1.0, cps=1, l0=False):
    vlen = math.ceil(float(n) * ((n - 1) & 0xFFFFFFFF) / 2)
    bpc = int(math.ceil(math.log2(vlen)) + 1)
    if bpc % 16 == 0:
        bpc += 1
    div = (math.log2(3) - 1) if l0 else (1 - 0.263034406)
    s = max(18, int(math.ceil(factor * math.log2(n) / div)))
    cols = s * cps
    return {"n": n, "factor": factor, "cols_per_sample": cps, "l0": l0, "vector_len": vlen,
            "bkt_per_col": bpc, "num_samples": s, "num_columns": cols,
            "num_buckets": cols * bpc + 1}


def main():
    rng = random.Random(20261203)
    vec = []
    special_keys = [0, 1, 2, M64, 1 << 32, (1 << 32) - 1, 0x0000000100000002, 0x1234567890ABCDEF]
    special_seeds = [0, 1, 5, 120, M64, 1 << 63, 0xFFFFFFFF, 0x9E3779B97F4A7C15]
    for k in special_keys:
        for s in special_seeds:
            vec.append([k, s, xxh3(k, s)])
    for _ in range(400):
        k, s = rng.getrandbits(64), rng.getrandbits(64)
        vec.append([k, s, xxh3(k, s)])
    # edge indices (min<<32|max) with column seeds seed+5c, the shape the sketch uses
    for _ in range(100):
        u, v = sorted(rng.sample(range(1 << 17), 2))
        seed = rng.getrandbits(64)
        for c in (0, 1, 23):
            s = (seed + 5 * c) & M64
            vec.append([(u << 32) | v, s, xxh3((u << 32) | v, s)])
    with open(os.path.join(HERE, "xxh3_len8.json"), "w") as f:
        json.dump({"source": "python xxhash %s (libxxhash %s)" % (xxhash.VERSION, xxhash.XXHASH_VERSION),
                   "vectors": [[str(a), str(b), str(c)] for a, b, c in vec]}, f, indent=0)
    geo = [geometry(n) for n in (2, 3, 10, 40, 100, 1024, 4096, 8192, 1 << 15, 1 << 16, 1 << 17,
                                 1 << 18, 1000003)]
    geo.append(geometry(1 << 15, factor=2.7))
    geo.append(geometry(1024, cps=7, l0=True))
    geo.append(geometry(100, cps=7))
    with open(os.path.join(HERE, "geometry.json"), "w") as f:
        json.dump(geo, f, indent=1)
    ref = "/root/reference/test/res"
    if os.path.isdir(ref):
        for name in ("multiples_graph_1024.txt", "multiples_graph_1024_stream.data"):
            shutil.copyfile(os.path.join(ref, name), os.path.join(HERE, name))


if __name__ == "__main__":
    main()
