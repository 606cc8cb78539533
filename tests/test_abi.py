"""The C-ABI libraries load without a GPU and export every entry point the
headers in include/ declare (no compute calls here)."""
import ctypes
import pathlib
import re

import pytest

from gip_amd import hip

ROOT = pathlib.Path(__file__).resolve().parents[1]
HEADERS = {"gip_hip.h": "libgip_hip.so", "gip_cavity.h": "libgip_cavity.so", "gip_dns.h": "libgip_dns.so",
           "gip_batch.h": "libgip_hip.so", "gip_adjoint.h": "libgip_hip.so"}


def declared(header):
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gip_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header", sorted(HEADERS))
def test_library_exports_every_declared_symbol(header):
    lib = ctypes.CDLL(str(ROOT / "gip_amd" / HEADERS[header]))
    names = declared(header)
    assert names, header
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_signature_table_covers_gip_hip():
    assert set(declared("gip_hip.h")) == set(hip.SIGNATURES)


def test_headers_compile_as_c(tmp_path):
    import subprocess
    src = tmp_path / "t.c"
    src.write_text("".join(f'#include "{h}"\n' for h in HEADERS) + "int main(void){return 0;}\n")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", f"-I{ROOT / 'include'}", str(src), "-o",
                    str(tmp_path / "t")], check=True)


def test_oracle_is_not_linked_by_the_product():
    import subprocess
    for lib in ("libgip_hip.so", "libgip_dns.so", "libgip_cavity.so", "bin/DrivenCavityHIP"):
        out = subprocess.run(["ldd", str(ROOT / "gip_amd" / lib)], capture_output=True, text=True).stdout
        assert "oracle" not in out, lib
