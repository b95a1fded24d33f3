"""Independent pure-Python restatement of the reference sketch arithmetic, used by the tests to
pin the C oracle on small cases (it shares no code with oracle/oracle.c).

xxh3_len8 follows the published XXH3 short-input path (len 4..8 + rrmxmx); the sketch update
follows src/sketch.cpp:85-99 (CubeSketch rule :66-83) and include/bucket.h:54-73; sample follows
src/sketch.cpp:110-132.
"""
M64 = (1 << 64) - 1
SECRET8 = 0x1CAD21F72C81017C
SECRET16 = 0xDB979083E96DD4DE


def rotl(x, r):
    return ((x << r) | (x >> (64 - r))) & M64


def xxh3_len8(key, seed):
    s32 = seed & 0xFFFFFFFF
    s = seed ^ (int.from_bytes(s32.to_bytes(4, "little"), "big") << 32)
    bitflip = ((SECRET8 ^ SECRET16) - s) & M64
    x = (((key >> 32) | (key << 32)) & M64) ^ bitflip
    x ^= rotl(x, 49) ^ rotl(x, 24)
    x = (x * 0x9FB21C651E98DF25) & M64
    x ^= (x >> 35) + 8
    x = (x * 0x9FB21C651E98DF25) & M64
    x ^= x >> 28
    return x


def depth(idx, seed_col, bpc):
    h = xxh3_len8(idx, seed_col) | (1 << bpc)
    return (h & -h).bit_length() - 1


def concat(i, j):
    i, j = min(i, j), max(i, j)
    return (i << 32) | j


def update(bk, bpc, cols, seed, idx, l0=False):
    """bk: list of [alpha, gamma] of length cols*bpc+1 (reference order). l0: the CubeSketch
    rule of src/sketch.cpp:66-83 (every depth 0..d of the column), else :85-99 (depth d only)."""
    chk = xxh3_len8(idx, seed)
    bk[-1][0] ^= idx
    bk[-1][1] ^= chk
    for c in range(cols):
        d = depth(idx, (seed + 5 * c) & M64, bpc)
        if d < bpc:
            for j in (range(d + 1) if l0 else (d,)):
                bk[c * bpc + j][0] ^= idx
                bk[c * bpc + j][1] ^= chk


def sample(bk, bpc, cps, seed, sample_idx):
    det = bk[-1]
    if det[0] == 0 and det[1] == 0:
        return 0, "ZERO"
    if det[1] == xxh3_len8(det[0], seed):
        return det[0], "GOOD"
    for i in range(cps):
        for j in range(bpc):
            b = bk[(i + sample_idx * cps) * bpc + j]
            if b[1] == xxh3_len8(b[0], seed):
                return b[0], "GOOD"
    return 0, "FAIL"
