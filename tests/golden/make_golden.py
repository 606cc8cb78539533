"""Regenerate the golden fixtures of tests/golden from the reference CPU build.

Runs the reference's own code (oracle/_ref, built from the reference sources
by oracle/build_ref.sh) on:
  c100k_4p   the shipped C100K case, 4 MPI ranks (the reference's own run)
  c100k_1p   the same case merged into one partition (data/, gip_cavity_tool merge)
  cav46_1p   the generated 46^3 cavity, Re=100 (gamma 0.01), one partition
and writes per-step time/dt/massr at full precision plus field summaries
(SHA-256 of the raw FP64 bytes, L2 norms, strided samples) as JSON.
Needs the reference tree and MPICH (/opt/conda); run from the repo root:
    python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os
import pathlib
import subprocess
import sys
import tempfile

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from oracle import pyoracle  # noqa: E402

HERE = pathlib.Path(__file__).resolve().parent
REF = pathlib.Path(os.environ.get("GIP_REFERENCE", "/root/reference"))
HARNESS = ROOT / "oracle" / "_ref" / "gip_ref_harness"
MPIRUN = "/opt/conda/bin/mpirun"


def field_summary(vals: np.ndarray) -> dict:
    idx = np.linspace(0, vals.size - 1, 64).astype(np.int64)
    return {
        "n": int(vals.size),
        "sha256": hashlib.sha256(vals.astype("<f8").tobytes()).hexdigest(),
        "l2": float(np.linalg.norm(vals)),
        "max_abs": float(np.abs(vals).max()),
        "sample_index": idx.tolist(),
        "sample": [float(v) for v in vals[idx]],
    }


def merge_rank_fields(dump: pathlib.Path, field: str, nparts: int, topo_dir: pathlib.Path, name: str) -> np.ndarray:
    """Reassemble per-rank dumps in the merged one-partition order: every
    rank's inner cells, then every rank's boundary ghosts (halo dropped)."""
    inner, ghosts = [], []
    for r in range(nparts):
        sizes = open(topo_dir / f"topoCols{name}_{nparts}p-{r}.topo").read().split()[:9]
        owned, ninner = int(sizes[3]), int(sizes[5]) + int(sizes[6])
        v = pyoracle.read_vec(dump / f"{field}-{r}.vec")
        inner.append(v[:ninner])
        ghosts.append(v[ninner:owned])
    return np.concatenate(inner + ghosts)


def run(cmd, cwd):
    env = dict(os.environ, PATH="/opt/conda/bin:" + os.environ.get("PATH", ""))
    out = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True, check=True, env=env)
    return pyoracle.parse_lines(out.stdout)


def case(label, cwd, name, steps, gamma, nparts, topo_dir):
    with tempfile.TemporaryDirectory() as td:
        dump = pathlib.Path(td)
        cmd = [str(HARNESS), "dns", name, str(steps), repr(gamma), "100", str(dump)]
        if nparts > 1:
            cmd = [MPIRUN, "-np", str(nparts)] + cmd
        rows = run(cmd, cwd)
        fields = {}
        for f in ("u", "v", "w", "p"):
            if nparts > 1:
                vals = merge_rank_fields(dump, f, nparts, topo_dir, name)
            else:
                vals = pyoracle.read_vec(dump / f"{f}-0.vec")
            fields[f] = field_summary(vals)
    doc = {"case": label, "name": name, "steps": steps, "gamma": gamma, "pcg_maxit": 100, "nparts": nparts,
           "producer": "reference Multicore specification + GIP_DNS template (oracle/_ref/gip_ref_harness)",
           "steps_out": [{k: r[k] for k in ("time", "dt", "massr", "step")} for r in rows], "fields": fields}
    (HERE / f"ref_{label}.json").write_text(json.dumps(doc, indent=1))
    print(f"wrote ref_{label}.json ({len(rows)} steps)")


def main():
    if not HARNESS.exists():
        sys.exit("oracle/_ref/gip_ref_harness not built (oracle/build_ref.sh)")
    with tempfile.TemporaryDirectory() as td:
        os.symlink(REF / "InputFiles", pathlib.Path(td) / "InputFiles")
        case("c100k_4p", td, "C100K", 10, 1e-3, 4, REF / "InputFiles/Topos/C100K/C100K_4p")
    case("c100k_1p", ROOT / "data", "C100K", 10, 1e-3, 1, None)
    with tempfile.TemporaryDirectory() as td:
        subprocess.run([str(ROOT / "gip_amd/bin/gip_cavity_tool"), "gen", td, "cav46", "46", "--gamma", "0.01"],
                       check=True, capture_output=True)
        case("cav46_1p", td, "cav46", 10, 1e-2, 1, None)


if __name__ == "__main__":
    main()
