"""The benchmark's profile summarizers on the CPU, on synthetic profiler
output: the PMC traffic summary scales each kernel's FETCH_SIZE and
WRITE_SIZE (KiB per dispatch) by what the same counters read for the copy
kernel of known size, the roofline reads the committed summary back, and the
kernel summary totals a rocprofv3 results database."""
import csv
import json
import pathlib
import sqlite3
import subprocess
import sys

ROOT = pathlib.Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))

import bench  # noqa: E402

COPY_BYTES = 8 * (1 << 27)


def _pmc_dir(tmp, counter, rows):
    d = tmp / counter
    d.mkdir()
    with open(d / "run_counter_collection.csv", "w", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=["Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writeheader()
        for name, value in rows:
            w.writerow({"Kernel_Name": name, "Counter_Name": counter, "Counter_Value": value})
    return d


def test_traffic_is_scaled_by_the_copy_kernel(tmp_path):
    # FETCH_SIZE reads half a copy's bytes, WRITE_SIZE all of them (as on the device)
    copy_kib = COPY_BYTES / 1024
    fetch = _pmc_dir(tmp_path, "FETCH_SIZE", [
        ("gip::k_copy(int, double const*, double*)", copy_kib / 2),
        ("void gip::k_iter_spmv<-7>(gip::EllView, gip::PcgDev, double const*)", 1000.0),
        ("void gip::k_iter_spmv<-7>(gip::EllView, gip::PcgDev, double const*)", 3000.0),  # median 2000
        ("gip::k_iter_dir(gip::PcgDev, double*, int)", 10.0),
        ("gip::k_iter_dir_rows(gip::PcgDev, double*, int)", 20.0),
        ("void gip::k_iter_spmv_sym<-7>(gip::EllView, gip::PcgDev, double const*)", 400.0),
    ])
    write = _pmc_dir(tmp_path, "WRITE_SIZE", [
        ("gip::k_copy(int, double const*, double*)", copy_kib),
        ("void gip::k_iter_spmv<-7>(gip::EllView, gip::PcgDev, double const*)", 500.0),
        ("gip::k_iter_dir(gip::PcgDev, double*, int)", 1.0),
        ("gip::k_iter_dir_rows(gip::PcgDev, double*, int)", 2.0),
        ("void gip::k_iter_spmv_sym<-7>(gip::EllView, gip::PcgDev, double const*)", 50.0),
    ])
    res, fs, ws = bench.summarize_pmc(fetch, write, bench.PCG_KERNELS)
    assert fs == 2.0 and ws == 1.0
    # "k_iter_spmv<" is the full-layout product only, not the symmetric one
    assert res["k_iter_spmv<"]["bytes"] == 1024 * (2000.0 * 2 + 500.0)
    assert res["k_iter_spmv_sym"]["bytes"] == 1024 * (400.0 * 2 + 50.0)
    # "k_iter_dir(" is the grid-stride sweep only, not k_iter_dir_rows
    assert res["k_iter_dir("]["bytes"] == 1024 * (10.0 * 2 + 1.0)
    assert res["k_iter_dir_rows"]["bytes"] == 1024 * (20.0 * 2 + 2.0)
    assert "k_m_spmv" not in res  # not profiled: left out


def test_traffic_summary_round_trips_into_the_roofline(tmp_path, monkeypatch):
    copy_kib = COPY_BYTES / 1024
    fetch = _pmc_dir(tmp_path, "FETCH_SIZE", [("k_copy", copy_kib / 2), ("k_iter_spmv<-7>", 700.0)])
    write = _pmc_dir(tmp_path, "WRITE_SIZE", [("k_copy", copy_kib), ("k_iter_spmv<-7>", 100.0)])
    out = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--summarize", "traffic", "--fetch", str(fetch),
                          "--write", str(write), "--n", "12"], capture_output=True, text=True, check=True).stdout
    prof = tmp_path / "profiles"
    prof.mkdir()
    (prof / "pcg_kernels_12_traffic.txt").write_text(out)
    monkeypatch.setattr(bench, "ROOT", tmp_path)
    assert bench.measured_traffic(12, "k_iter_spmv<") == int(1024 * (700.0 * 2 + 100.0))
    assert bench.measured_traffic(12, "k_m_spmv") is None
    assert bench.measured_traffic(13, "k_iter_spmv") is None


def test_stream_ceiling_summary_is_read_back(tmp_path, monkeypatch):
    line = {"probe": "stream",
            "all": [{"kind": "copy", "bytes_per_access": 16, "nontemporal": True, "grid": 0, "ms": 1.0,
                     "gbs": 6000.0}],
            "best": {"copy": {"gbs": 6000.0}, "read": {"gbs": 7000.5}}}
    j = tmp_path / "stream.json"
    j.write_text("some log line\n" + json.dumps(line) + "\n")
    out = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--summarize", "stream", "--json", str(j)],
                         capture_output=True, text=True, check=True).stdout
    (tmp_path / "profiles").mkdir()
    (tmp_path / "profiles" / "stream_ceiling.txt").write_text(out)
    monkeypatch.setattr(bench, "ROOT", tmp_path)
    assert bench.stream_ceiling() == {"copy": 6000.0, "read": 7000.5}


def test_kernel_summary_totals_a_results_database(tmp_path):
    db = tmp_path / "run_results.db"
    con = sqlite3.connect(db)
    con.execute("create table kernels (name text, duration integer)")
    con.executemany("insert into kernels values (?, ?)",
                    [("k_a", 3000), ("k_a", 5000), ("k_b", 2000)])  # nanoseconds
    con.commit()
    con.close()
    out = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--summarize", "kernels", "--db", str(db),
                          "--title", "synthetic"], capture_output=True, text=True, check=True).stdout
    rows = [ln.split() for ln in out.splitlines() if not ln.startswith("#")]
    assert rows[0][-1] == "k_a" and rows[0][1] == "2" and rows[0][0] == "80.00%"
    assert float(rows[0][3]) == 4.0  # average microseconds
    assert rows[1][-1] == "k_b" and rows[1][0] == "20.00%"
