"""The maintainer's drop-in recipe, run verbatim.

The copy, patch and build blocks are taken from the integration guide (its
marked shell blocks), not retyped here, and run in a temporary copy of the reference tree (its src/ and
DrivenCavity.cpp).  The result must be an executable that links the kernel
library and binds to the reference's own abstract headers.  CPU only:
compiling and linking need no device.  Skipped where the reference sources
or the MPI toolchain are absent.
"""
import pathlib
import shutil
import subprocess

import pytest

from oracle import dropin_build

REFERENCE = pathlib.Path("/root/reference")


def test_recipe_blocks_are_documented():
    b = dropin_build.recipe_blocks()
    assert "src/Specifications/HIP" in b["copy"]
    assert "+#if ENABLE_GPU == 2" in b["patch"] and "+#elif ENABLE_GPU" in b["patch"]
    assert "-lgip_hip" in b["build"] and "-DENABLE_GPU=2" in b["build"]
    assert "<repo>" not in b["build"]


def test_documented_recipe_builds_the_driver(tmp_path):
    if not (REFERENCE / "src" / "Algorithms" / "GIP_DNS.h").exists():
        pytest.skip("reference sources absent")
    if not pathlib.Path("/opt/conda/bin/mpic++").exists() or not shutil.which("patch"):
        pytest.skip("MPI wrapper or patch absent")
    if not (dropin_build.REPO / "gip_amd" / "libgip_hip.so").exists():
        pytest.skip("libgip_hip.so not built")
    exe = dropin_build.build(REFERENCE, tmp_path / "DrivenCavityHIP", work=tmp_path)
    assert exe.exists()
    syms = subprocess.run(["nm", "-C", str(exe)], capture_output=True, text=True, check=True).stdout
    # the reference's template instantiated with the HIP specification
    assert "GIP_DNS<GIP_MatrixHIP, GIP_VectorHIP, GIP_ArchHIP, GIP_NonLinearHIP>" in syms
    deps = subprocess.run(["ldd", str(exe)], capture_output=True, text=True, check=True).stdout
    assert "libgip_hip.so" in deps and "not found" not in deps
