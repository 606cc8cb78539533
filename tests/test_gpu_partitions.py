"""Several partitions of one case resident on one device (include/gip_dns.h,
gip_dns_open_parts): the reference's own 4-partition C100K files run
unmerged, with the halo refreshed by device pack/unpack kernels
(GIP_Topo.cpp:124-181) and the dot products, norms and CFL bound combined
over the partitions, against the reference binary's 4-rank run
(tests/golden/ref_c100k_4p.json, made with mpirun -np 4) and the README
lines; a generated cavity written in 4 partitions against the oracle on the
whole case; and the drop-in build of the reference's own driver run as the
reference runs its shipped case, one MPI process per partition.

Bar: per-step time, dt and massr within 1e-12 of the 4-rank reference run
(only the order in which partial sums are combined differs), the README
lines to their printed digits, the merged fields' norms and samples within
1e-12; the generated case within 1e-10 of the oracle; the MPI run's printed
lines at their six printed digits.
"""
import json
import pathlib
import subprocess

import numpy as np
import pytest

from gip_amd import cavity, dns
from oracle import pyoracle

from .conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return abs(a - b) / abs(b) if b else abs(a)


def test_c100k_four_partitions_unmerged_match_the_reference_run(c100k_root):
    ref = json.loads((GOLDEN / "ref_c100k_4p.json").read_text())
    readme = pyoracle.parse_lines((GOLDEN / "readme_c100k_np4.txt").read_text())
    sim = dns.PartitionedSimulation(root=c100k_root, name="C100K", nparts=4, gamma=1e-3)
    print(f"[C100K 4p] partitions {sim.sizes}", flush=True)
    worst = 0.0
    for w, line in zip(ref["steps_out"], readme):
        g = sim.step()
        for key in ("time", "dt", "massr"):
            err = _rel(g[key], w[key])
            worst = max(worst, err)
            assert err <= 1e-12, (w["step"], key, g[key], w[key])
            assert f"{g[key]:.6e}" == f"{line[key]:.6e}", (key, g[key], line[key])
    print(f"[C100K 4p] 10 steps, worst relative difference from the 4-rank reference run {worst:.2e}", flush=True)
    for f in ("u", "v", "w", "p"):
        vals = sim.merged_field(f)
        summ = ref["fields"][f]
        assert vals.size == summ["n"]
        l2 = np.linalg.norm(vals)
        print(f"[C100K 4p field {f}] l2 {l2:.15e} reference {summ['l2']:.15e}", flush=True)
        assert abs(l2 - summ["l2"]) <= 1e-12 * summ["l2"]
        np.testing.assert_allclose(vals[summ["sample_index"]], summ["sample"], rtol=0,
                                   atol=1e-12 * summ["max_abs"] * 10)
    sim.close()


def test_generated_cavity_in_four_partitions_matches_the_oracle(tmp_path):
    c = cavity.generate(46, gamma=1e-2)
    c.write_partitions(tmp_path, "cav46", 4)
    merged = cavity.merge(tmp_path, "cav46", 4)  # the same cells, numbered as merged_field returns them
    sim = dns.PartitionedSimulation(root=tmp_path, name="cav46", nparts=4, gamma=1e-2)
    orc = pyoracle.Dns(cavity=merged, gamma=1e-2)
    for k in range(5):
        g, w = sim.step(), orc.step()
        print(f"[46^3 4p step {k + 1}] massr {g['massr']:.12e} oracle {w['massr']:.12e}", flush=True)
        for key in ("time", "dt", "massr"):
            assert _rel(g[key], w[key]) <= 1e-10, (k, key, g[key], w[key])
    for f in ("u", "v", "w", "p"):
        a, r = sim.merged_field(f), orc.field(f)
        err = np.linalg.norm(a - r) / np.linalg.norm(r)
        print(f"[46^3 4p field {f}] rel L2 vs oracle {err:.3e}", flush=True)
        assert err <= 1e-10, (f, err)
    sim.close()
    orc.close()


def test_partitions_need_the_reference_solver_sequence(tmp_path, c100k_root):
    with pytest.raises(dns.DnsError):
        dns.PartitionedSimulation(root=tmp_path, name="absent", nparts=2)
    cfg = dns.Config(1e-3, 1.0, 100, 0.0, dns.MODES["fused"], 0.1, 0, 0)
    h = dns.c_void_p()
    rc = dns.lib().gip_dns_open_parts(str(c100k_root).encode(), b"C100K", 4, dns.byref(cfg), dns.byref(h))
    assert rc != 0 and b"reference" in dns.lib().gip_dns_last_error()


def test_reference_driver_with_four_mpi_ranks(c100k_root):
    """The drop-in build (the reference's unmodified DrivenCavity.cpp and
    templates over the HIP specification, oracle/_ref/DrivenCavityHIP_reftemplates)
    under mpiexec -n 4 on the shipped 4-partition C100K files, as the
    reference runs it (README.md:62-68): four processes share the device,
    the halos are staged through the host between them and the reductions
    are MPI_Allreduce (Specifications/HIP/GIP_RanksMPI.cpp).  Rank 0 prints
    the ten lines; each value is the six-digit rounding of the 4-rank
    reference run's and equals the README line."""
    exe = ROOT / "oracle" / "_ref" / "DrivenCavityHIP_reftemplates"
    mpiexec = pathlib.Path("/opt/conda/bin/mpiexec")
    if not exe.exists() or not mpiexec.exists():
        pytest.skip("drop-in build or mpiexec absent")
    ref = json.loads((GOLDEN / "ref_c100k_4p.json").read_text())
    readme = pyoracle.parse_lines((GOLDEN / "readme_c100k_np4.txt").read_text())
    print("[drop-in, mpiexec -n 4] C100K_4p", flush=True)
    out = subprocess.run([str(mpiexec), "-launcher", "fork", "-n", "4", str(exe), "C100K"], cwd=c100k_root,
                         capture_output=True, text=True, timeout=180)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    got = pyoracle.parse_lines(out.stdout)
    assert len(got) == 10, out.stdout
    for g, w, line in zip(got, ref["steps_out"], readme):
        for key in ("time", "dt", "massr"):
            assert abs(g[key] - w[key]) <= 5.01e-6 * abs(w[key]), (w["step"], key, g[key], w[key])
            assert f"{g[key]:.6e}" == f"{line[key]:.6e}", (key, g[key], line[key])
    print(f"[drop-in, mpiexec -n 4] last line {got[-1]}", flush=True)


def test_reference_driver_with_three_mpi_ranks_on_a_generated_cavity(tmp_path):
    """The same drop-in build under mpiexec -n 3 on a generated 46^3 cavity
    written in 3 partitions (an odd count, so some partitions have two
    neighbours and some one): rank 0's lines equal the oracle's on the
    undivided case to their printed digits."""
    exe = ROOT / "oracle" / "_ref" / "DrivenCavityHIP_reftemplates"
    mpiexec = pathlib.Path("/opt/conda/bin/mpiexec")
    if not exe.exists() or not mpiexec.exists():
        pytest.skip("drop-in build or mpiexec absent")
    c = cavity.generate(46, gamma=1e-2)
    c.write(tmp_path, "cav46")
    c.write_partitions(tmp_path, "cav46", 3)
    print("[drop-in, mpiexec -n 3] generated 46^3 in 3 partitions", flush=True)
    out = subprocess.run([str(mpiexec), "-launcher", "fork", "-n", "3", str(exe), "cav46"], cwd=tmp_path,
                         capture_output=True, text=True, timeout=180)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    got = pyoracle.parse_lines(out.stdout)
    # the reference driver's hardcoded Re=1000, on the one-partition files
    want = pyoracle.run_cli(tmp_path, "cav46", 10, 1e-3)
    assert len(got) == 10, out.stdout
    for g, w in zip(got, want):
        for key in ("time", "dt", "massr"):
            assert _rel(g[key], w[key]) <= 1e-6, (w["step"], key, g[key], w[key])
    print(f"[drop-in, mpiexec -n 3] last line {got[-1]}", flush=True)
