"""The HIP specification's MPI ranks (GIP_RanksMPI.cpp) on the CPU: one
process per partition under mpiexec, with the device entry points replaced
by host versions in tests/mpi_ranks_check.cpp.  A generated cavity whose w
field holds every cell's global number is written in N partitions; after
the exchange every halo slot must hold the number the writer stored there
(the halo pairing of GIP_Topo.cpp:124-181), the all-reductions must be
MPI's SUM, MAX and MIN, and each process must sit on device local rank mod
the device count.  A case written for another number of ranks is refused
with a message."""
import os
import pathlib
import subprocess

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
SPEC = ROOT / "gip_amd" / "spec"
MPI_BIN = pathlib.Path("/opt/conda/bin")

pytestmark = pytest.mark.skipif(not (MPI_BIN / "mpic++").exists() or not (MPI_BIN / "mpiexec").exists(),
                                reason="MPICH (mpic++, mpiexec) not installed")


def _env():
    env = dict(os.environ)
    env["PATH"] = f"{MPI_BIN}:{env.get('PATH', '')}"
    env["MPICH_CXX"] = "g++"
    return env


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if not (ROOT / "gip_amd" / "libgip_cavity.so").exists():
        pytest.skip("libgip_cavity.so not built")
    out = tmp_path_factory.mktemp("mpi_ranks_check")
    srcs = [ROOT / "tests" / "mpi_ranks_check.cpp", SPEC / "Specifications/HIP/GIP_RanksMPI.cpp",
            SPEC / "Specifications/HIP/GIP_ArchHIP.cpp", SPEC / "Specifications/HIP/GIP_PartitionHIP.cpp",
            SPEC / "BasicInterfaces/GIP_Topo.cpp"]
    objs = []
    for s in srcs:
        o = out / (s.stem + ".o")
        subprocess.run(["mpic++", "-std=c++17", "-O1", "-Wall", "-Werror", f"-I{ROOT / 'include'}",
                        f"-I{SPEC / 'Specifications/HIP'}", f"-I{SPEC / 'BasicInterfaces'}", "-c", str(s), "-o",
                        str(o)], check=True, env=_env(), timeout=300)
        objs.append(str(o))
    exe = out / "check"
    # libmpi named by path and the system's runtime first on the search path,
    # as the drop-in build links (conda's libstdc++ is older than g++'s)
    subprocess.run(["g++", *objs, "-o", str(exe), str(MPI_BIN.parent / "lib" / "libmpi.so.12"),
                    f"-L{ROOT / 'gip_amd'}", "-lgip_cavity", "-pthread",
                    f"-Wl,-rpath,/usr/lib/x86_64-linux-gnu:{MPI_BIN.parent / 'lib'}:{ROOT / 'gip_amd'}"],
                   check=True, timeout=300)
    return exe


def _labelled_case(root, nparts):
    from gip_amd import cavity
    c = cavity.generate(12, 10, 9, gamma=1e-2)
    c.field("w")[:] = np.arange(c.n_all, dtype=np.float64)
    if nparts == 1:
        c.write(root, "c")
    else:
        c.write_partitions(root, "c", nparts)
    return (f"{root}/InputFiles/Topos/c/c_{nparts}p/topoColsc_{nparts}p-%d.topo",
            f"{root}/InputFiles/Scalarfields/c/c_{nparts}p/wc_{nparts}p-%d.vec")


def _run(exe, nprocs, patterns):
    return subprocess.run([str(MPI_BIN / "mpiexec"), "-n", str(nprocs), str(exe), *patterns],
                          capture_output=True, text=True, env=_env(), timeout=120)


@pytest.mark.parametrize("nparts", [2, 3, 4])
def test_ranks_exchange_and_reduce(checker, tmp_path, nparts):
    r = _run(checker, nparts, _labelled_case(tmp_path, nparts))
    assert r.returncode == 0, r.stdout + r.stderr
    lines = sorted(ln for ln in r.stdout.splitlines() if ln.startswith("rank "))
    assert len(lines) == nparts
    for q, ln in enumerate(lines):
        assert ln.startswith(f"rank {q} of {nparts}: ")
        assert int(ln.split(": ")[1].split()[0]) > 0  # halo slots refreshed


def test_lone_process_has_no_communicator(checker, tmp_path):
    r = _run(checker, 1, _labelled_case(tmp_path, 1))
    assert r.returncode == 0, r.stdout + r.stderr
    assert "rank 0 of 1: no communicator" in r.stdout
    assert "rank 0 of 1: 0 halo slots refreshed" in r.stdout


def test_case_for_another_rank_count_is_refused(checker, tmp_path):
    r = _run(checker, 2, _labelled_case(tmp_path, 3))
    assert r.returncode != 0
    assert "topology partitioned for another number of ranks" in r.stderr

