"""Provenance of the library a GPU run loads: built from the sources of this tree (the library
holds gfx950 code only, so every other GPU test fails on another GPU)."""
import os

import pytest

import gpusketch_amd as gs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.gpu
def test_gpu_run_loads_library_of_these_sources():
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "gpusketch_amd_build", os.path.join(ROOT, "gpusketch_amd", "build.py"))
    build = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(build)
    loaded = gs.source_digest()
    print(f"libgpusketch.so source digest {loaded}")
    assert loaded == build.source_digest()
    assert gs.lib.gs_device_synchronize() == gs.GS_OK
