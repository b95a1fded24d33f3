"""The threaded C++ layer (gutters, GraphSketchDriver workers, EdgeStore contractions, min-cut
routing under back-to-back contractions, verifier) under ThreadSanitizer and under
AddressSanitizer + UndefinedBehaviorSanitizer: the host library and the facade_cpu program
compiled for the host with g++ (no device code is built with a sanitizer; the device library is
linked as built), then facade_cpu's checks run and must report no finding."""
import glob
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "gpusketch_amd", "host")
PKG = os.path.join(ROOT, "gpusketch_amd")

SANITIZERS = {
    "thread": (["-fsanitize=thread"],
               {"TSAN_OPTIONS": "halt_on_error=1 second_deadlock_stack=1"}),
    "address,undefined": (["-fsanitize=address,undefined", "-fno-omit-frame-pointer",
                           "-fno-sanitize-recover=undefined"],
                          {"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=1",
                           "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1"}),
}


def _build(out_dir, flags):
    if not os.path.exists(os.path.join(PKG, "libgpusketch.so")):
        pytest.fail("libgpusketch.so is not built; run `python gpusketch_amd/build.py`")
    common = ["g++", "-O1", "-g", "-std=c++17", "-pthread", "-fPIC", *flags,
              "-I", os.path.join(ROOT, "include")]
    srcs = sorted(p for p in glob.glob(os.path.join(HOST, "*.cpp"))
                  if not os.path.basename(p).startswith("prog_"))

    def compile_one(src):
        obj = os.path.join(out_dir, os.path.basename(src) + ".o")
        r = subprocess.run([*common, "-c", src, "-o", obj], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        return obj

    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(compile_one, srcs))
    prog = os.path.join(out_dir, "facade_cpu")
    r = subprocess.run([*common, os.path.join(HOST, "prog_facade_cpu.cpp"), *objs, "-o", prog,
                        f"-L{PKG}", "-lgpusketch", f"-Wl,-rpath,{PKG}"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return prog


@pytest.mark.parametrize("sanitizer", sorted(SANITIZERS))
def test_facade_cpu_under_sanitizer(tmp_path, sanitizer):
    flags, env_add = SANITIZERS[sanitizer]
    build_dir = tmp_path / "build"
    build_dir.mkdir()
    prog = _build(str(build_dir), flags)
    work = tmp_path / "work"
    work.mkdir()
    env = dict(os.environ, **env_add)
    r = subprocess.run([prog, "all", str(work)], capture_output=True, text=True, timeout=600,
                       env=env)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "0 failure(s)" in r.stdout
    for marker in ("WARNING: ThreadSanitizer", "ERROR: AddressSanitizer", "runtime error:",
                   "ERROR: LeakSanitizer"):
        assert marker not in out, out[-4000:]
