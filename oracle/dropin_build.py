"""TEST INFRASTRUCTURE ONLY — the reference driver with the HIP specification
plugged in, built by the maintainer's drop-in recipe.

The three recipe blocks (copy, patch, build; fenced shell blocks marked
`<!-- recipe:copy -->` and so on in the integration guide) are read from
the guide and run verbatim, with <repo> replaced by this repository's path, in a temporary
copy of the reference tree (its src/ and DrivenCavity.cpp; the reference
itself is never written).  The resulting executable is copied to `out`.

    python oracle/dropin_build.py <reference> <out>
"""
from __future__ import annotations

import pathlib
import re
import shutil
import subprocess
import sys
import tempfile

REPO = pathlib.Path(__file__).resolve().parents[1]


def recipe_blocks(doc: pathlib.Path = REPO / "INTEGRATION.md") -> dict:
    """{'copy': str, 'patch': str, 'build': str}: the fenced block after each
    <!-- recipe:NAME --> marker, with <repo> substituted."""
    text = doc.read_text()
    out = {}
    for m in re.finditer(r"<!-- recipe:(\w+) -->\n```\w*\n(.*?)```", text, re.S):
        out[m.group(1)] = m.group(2).replace("<repo>", str(REPO))
    missing = {"copy", "patch", "build"} - set(out)
    if missing:
        raise RuntimeError(f"{doc.name} lacks recipe blocks {sorted(missing)}")
    return out


def build(reference: pathlib.Path, out: pathlib.Path, work: pathlib.Path | None = None) -> pathlib.Path:
    blocks = recipe_blocks()
    tmp = pathlib.Path(tempfile.mkdtemp(prefix="gip_dropin_", dir=work))
    try:
        shutil.copytree(reference / "src", tmp / "src")
        shutil.copy2(reference / "DrivenCavity.cpp", tmp / "DrivenCavity.cpp")

        def sh(script, stdin=None):
            r = subprocess.run(["bash", "-e", "-c", script], cwd=tmp, input=stdin, capture_output=True, text=True)
            if r.returncode != 0:
                raise RuntimeError(f"recipe step failed:\n{script}\n{r.stdout}\n{r.stderr}")

        sh(blocks["copy"])
        sh("patch -p1 --no-backup-if-mismatch", stdin=blocks["patch"])
        sh(blocks["build"])
        out.parent.mkdir(parents=True, exist_ok=True)
        shutil.copy2(tmp / "DrivenCavityHIP", out)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return out


if __name__ == "__main__":
    if len(sys.argv) != 3:
        sys.exit(__doc__)
    build(pathlib.Path(sys.argv[1]), pathlib.Path(sys.argv[2]))
    print(f"dropin_build: built {sys.argv[2]}")
