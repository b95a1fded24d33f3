"""Build recipe for libgpusketch.so (gfx950) and the C++ host programs.

Compiles every HIP/C++ source under gpusketch_amd/csrc with hipcc into one shared library next
to this file, and the C++ facade programs under gpusketch_amd/host into gpusketch_amd/bin.
Run ``python gpusketch_amd/build.py`` from the repository root (or call ``build()``).
"""
from __future__ import annotations

import glob
import hashlib
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgpusketch.so")
BIN = os.path.join(HERE, "bin")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -Xclang -target-feature -Xclang -packed-fp32-ops keeps the compiler from forming packed FP32
# instructions (none of this code is floating point, the flag is kept uniform for all HIP code).
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
    "-I", os.path.join(ROOT, "include"),
]


def library_sources() -> list[str]:
    """Every file that goes into libgpusketch.so: the sources under csrc and the headers they
    include (include/gpusketch.h, include/gpusketch/*.h), sorted."""
    pats = [os.path.join(CSRC, "*"), os.path.join(ROOT, "include", "gpusketch.h"),
            os.path.join(ROOT, "include", "gpusketch", "*.h")]
    files = []
    for p in pats:
        files += [f for f in glob.glob(p) if os.path.isfile(f)]
    return sorted(files)


def source_digest() -> str:
    """SHA-256 over the library's sources (relative path and bytes of each) and its compile
    flags. Compiled into the library (gs_source_digest), so a loaded library can be checked
    against the tree it runs from."""
    h = hashlib.sha256()
    h.update(" ".join(HIP_FLAGS[:-2]).encode())
    for f in library_sources():
        h.update(os.path.relpath(f, ROOT).encode() + b"\0")
        with open(f, "rb") as fh:
            h.update(fh.read())
        h.update(b"\0")
    return h.hexdigest()


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"compile failed: {cmd[-1] if cmd else ''}")
    return r


def build_library(jobs: int = 8) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))
    objs = []
    cmds = []
    digest = [f"-DGS_SOURCE_DIGEST=\"{source_digest()}\""]
    for s in srcs:
        o = os.path.join(OBJ, os.path.basename(s) + ".o")
        objs.append(o)
        lang = ["-x", "hip"] if s.endswith(".hip") else []
        extra = digest if os.path.basename(s) == "api.hip" else []
        cmds.append([HIPCC, *HIP_FLAGS, *extra, *lang, "-c", s, "-o", o])
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, LIB)
    return LIB


HOST_LIB = os.path.join(HERE, "libgpusketch_host.so")
# the same C++ API compiled with L0_SAMPLING (the reference's CubeSketch build option)
HOST_LIB_L0 = os.path.join(HERE, "libgpusketch_host_l0.so")
HOST_FLAGS = ["-O2", "-std=c++17", "-pthread", "-Wall", "-Wextra", "-Wno-unused-parameter",
              "-I", os.path.join(ROOT, "include")]
# programs built with a define of the reference's build: verify_facade as its VERIFY_SAMPLES_F
# test builds are (the driver records the exact graph and the queries check against it; the
# library itself does not depend on the define)
PROG_DEFINES = {"verify_facade": ["-DVERIFY_SAMPLES_F"]}


def build_host_programs() -> list[str]:
    """The C++ API (include/gpusketch/*.h, gpusketch_amd/host/*.cpp) as libgpusketch_host.so over
    libgpusketch.so, and the C++ test programs host/prog_*.cpp linked against both; the API and
    the two facade programs once more with L0_SAMPLING defined (libgpusketch_host_l0.so,
    bin/facade_cpu_l0, bin/facade_gpu_l0)."""
    os.makedirs(BIN, exist_ok=True)
    lib_srcs = sorted(p for p in glob.glob(os.path.join(HOST, "*.cpp"))
                      if not os.path.basename(p).startswith("prog_"))
    libs = [(HOST_LIB, []), (HOST_LIB_L0, ["-DL0_SAMPLING"])]
    cmds = [["g++", *HOST_FLAGS, *defs, "-fPIC", "-shared", *lib_srcs, "-o", lib + ".tmp",
             f"-L{HERE}", "-lgpusketch", "-Wl,-rpath,$ORIGIN"] for lib, defs in libs]
    with ThreadPoolExecutor(max_workers=2) as ex:
        list(ex.map(_run, cmds))
    for lib, _ in libs:
        os.replace(lib + ".tmp", lib)
    outs = []
    cmds = []
    for p in sorted(glob.glob(os.path.join(HOST, "prog_*.cpp"))):
        name = os.path.basename(p)[len("prog_"):-4]
        out = os.path.join(BIN, name)
        outs.append(out)
        defs = PROG_DEFINES.get(name, [])
        cmds.append(["g++", *HOST_FLAGS, *defs, p, "-o", out, f"-L{HERE}", "-lgpusketch_host",
                     "-lgpusketch", "-Wl,-rpath,$ORIGIN/.."])
    for prog in ("facade_cpu", "facade_gpu"):
        p = os.path.join(HOST, f"prog_{prog}.cpp")
        outs.append(os.path.join(BIN, f"{prog}_l0"))
        cmds.append(["g++", *HOST_FLAGS, "-DL0_SAMPLING", p, "-o", outs[-1], f"-L{HERE}",
                     "-lgpusketch_host_l0", "-lgpusketch", "-Wl,-rpath,$ORIGIN/.."])
    with ThreadPoolExecutor(max_workers=8) as ex:
        list(ex.map(_run, cmds))
    return outs


def build() -> None:
    build_library()
    build_host_programs()


if __name__ == "__main__":
    build()
    print(LIB)
