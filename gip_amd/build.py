"""Build recipes for the native parts of gip_amd (all in-tree).

    libgip_cavity.so   operator generator + file formats (g++, OpenMP)
    bin/gip_cavity_tool
    libgip_hip.so      gfx950 kernels + C ABI (hipcc)
    libgip_dns.so      HIP specification classes + templated integrator (hipcc, host C++)
    bin/DrivenCavityHIP

and, for the test side, the oracle (oracle/_build) and the reference CPU
build (oracle/_ref, only where the reference sources are present).
"""
from __future__ import annotations

import os
import pathlib
import shutil
import subprocess

PKG = pathlib.Path(__file__).resolve().parent
ROOT = PKG.parent
INC = ROOT / "include"
CSRC = PKG / "csrc"
SPEC = PKG / "spec"
GEN = PKG / "gen"
BIN = PKG / "bin"
ORACLE = ROOT / "oracle"
DATA = ROOT / "data"
REFERENCE = pathlib.Path(os.environ.get("GIP_REFERENCE", "/root/reference"))

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
ROCM_LIB = pathlib.Path(os.environ.get("ROCM_PATH", "/opt/rocm")) / "lib"
# no multiply-add contraction anywhere: products and sums round separately,
# as in the x86-64 CPU specification; packed FP32 instructions disabled.
HIP_FLAGS = [
    "-O3", "--offload-arch=gfx950", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-value", "-Wno-unused-result",
]


def _run(cmd, **kw):
    res = subprocess.run([str(c) for c in cmd], capture_output=True, text=True, **kw)
    if res.returncode != 0:
        raise RuntimeError(f"build step failed: {' '.join(map(str, cmd))}\n{res.stdout}\n{res.stderr}")
    return res


def build_cavity():
    BIN.mkdir(exist_ok=True)
    _run(["g++", "-O2", "-fPIC", "-fopenmp", "-shared", f"-I{INC}", GEN / "gip_cavity.cpp",
          "-o", PKG / "libgip_cavity.so"])
    _run(["g++", "-O2", f"-I{INC}", GEN / "gip_cavity_tool.cpp", f"-L{PKG}", "-lgip_cavity",
          "-Wl,-rpath,$ORIGIN/..", "-o", BIN / "gip_cavity_tool"])


HIP_SOURCES = ["gip_runtime.hip", "gip_blas1.hip", "gip_spmv.hip", "gip_pcg.hip", "gip_batch.hip", "gip_adjoint.hip",
               "gip_stream.hip"]


def build_hip():
    _run([HIPCC, *HIP_FLAGS, "-shared", f"-I{INC}", f"-I{CSRC}", *[CSRC / s for s in HIP_SOURCES],
          "-o", PKG / "libgip_hip.so"])


SPEC_SOURCES = [
    "BasicInterfaces/GIP_Topo.cpp", "BasicInterfaces/GIP_Parameters.cpp",
    "AbstractInterfaces/GIP_Abstract.cpp",
    "Specifications/HIP/GIP_ArchHIP.cpp", "Specifications/HIP/GIP_VectorHIP.cpp",
    "Specifications/HIP/GIP_MatrixHIP.cpp", "Specifications/HIP/GIP_NonLinearHIP.cpp",
    "Specifications/HIP/GIP_TimerHIP.cpp", "Specifications/HIP/GIP_PartitionHIP.cpp",
]


def spec_includes():
    return [f"-I{INC}", f"-I{SPEC}", f"-I{SPEC}/BasicInterfaces", f"-I{SPEC}/AbstractInterfaces",
            f"-I{SPEC}/Specifications/HIP", f"-I{SPEC}/Algorithms"]


def build_dns():
    """Host C++ of the HIP specification; it only calls libgip_hip's C ABI."""
    if not (SPEC / "gip_dns_capi.cpp").exists():
        return
    BIN.mkdir(exist_ok=True)
    srcs = [SPEC / s for s in SPEC_SOURCES] + [SPEC / "gip_dns_capi.cpp"]
    flags = ["-O2", "-fPIC", "-std=c++17", "-fopenmp", *spec_includes()]
    flags += ["-isystem", str(ROCM_LIB.parent / "include")]  # roctx ranges (gip_markers.h)
    roctx = [f"-L{ROCM_LIB}", "-lrocprofiler-sdk-roctx", f"-Wl,-rpath,{ROCM_LIB}"]
    _run(["g++", *flags, "-shared", *srcs, f"-L{PKG}", "-lgip_hip", "-lgip_cavity", *roctx,
          "-Wl,-rpath,$ORIGIN", "-o", PKG / "libgip_dns.so"])
    _run(["g++", *flags, SPEC / "DrivenCavity.cpp", f"-L{PKG}", "-lgip_dns", "-lgip_hip", "-lgip_cavity", *roctx,
          "-Wl,-rpath,$ORIGIN/..", "-o", BIN / "DrivenCavityHIP"])


def build_oracle():
    out = ORACLE / "_build"
    out.mkdir(exist_ok=True)
    # -O2 without -march: no FMA contraction on x86-64, same roundings as the reference build
    _run(["gcc", "-O2", "-std=c99", "-fopenmp", "-fPIC", "-shared", ORACLE / "gip_oracle.c", "-lm", "-o",
          out / "liboracle.so"])
    _run(["gcc", "-O2", "-std=c99", "-fopenmp", ORACLE / "orc_cavity.c", ORACLE / "gip_oracle.c", "-lm", "-o",
          out / "orc_cavity"])


def build_ref():
    """Reference CPU build (test infrastructure); skipped where the sources are absent."""
    if not (REFERENCE / "src" / "Algorithms" / "GIP_DNS.h").exists():
        return False
    _run(["sh", ORACLE / "build_ref.sh"])
    return True


def build_data():
    """C100K under data/: the reference's four-partition files (the ones its
    driver reads with four ranks, for timing the reference CPU path) and the
    one-partition case merged from them (for the GPU)."""
    if not (REFERENCE / "InputFiles").exists():
        return False
    src = REFERENCE / "InputFiles"
    for d in src.rglob("C100K_4p"):
        if d.is_dir():
            dst = DATA / d.relative_to(REFERENCE)
            dst.mkdir(parents=True, exist_ok=True)
            for f in d.glob("*_4p-*"):
                if not (dst / f.name).exists():
                    shutil.copyfile(f, dst / f.name)
    marker = DATA / "InputFiles" / "Topos" / "C100K" / "C100K_1p" / "topoColsC100K_1p-0.topo"
    if marker.exists():
        return True
    _run([BIN / "gip_cavity_tool", "merge", REFERENCE, "C100K", "4", DATA, "C100K"])
    return True


def build_all():
    build_cavity()
    build_hip()
    build_dns()
    build_oracle()
    build_ref()
    build_data()


if __name__ == "__main__":
    build_all()
