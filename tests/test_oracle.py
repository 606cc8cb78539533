"""The oracle (C restatement of the reference CPU path) against the
reference's own golden vectors and against closed forms / library routines.

Pins:
  * README.md:73-82 (the reference's only golden output) to 7 digits;
  * the reference binary's full-precision runs (tests/golden/ref_*.json,
    tests/golden/make_golden.py): bitwise on one partition, 1e-12 relative
    against the 4-rank run (only the reduction order differs);
  * SpMV against scipy's CSR product, the empty-row rule, PCG against a
    direct solve and its residual invariant, CFL against its closed form.
"""
import json

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import pyoracle

from .conftest import GOLDEN


def _readme_lines():
    return pyoracle.parse_lines((GOLDEN / "readme_c100k_np4.txt").read_text())


def test_oracle_reproduces_readme_golden_lines(c100k_root):
    got = pyoracle.run_cli(c100k_root, "C100K", 10, 1e-3)
    want = _readme_lines()
    assert len(got) == len(want) == 10
    for g, w in zip(got, want):
        for k in ("time", "dt", "massr"):
            assert f"{g[k]:.6e}" == f"{w[k]:.6e}", (k, g, w)


@pytest.mark.parametrize("label,exact", [("c100k_1p", True), ("c100k_4p", False)])
def test_oracle_matches_reference_binary_c100k(c100k_root, tmp_path, label, exact):
    ref = json.loads((GOLDEN / f"ref_{label}.json").read_text())
    got = pyoracle.run_cli(c100k_root, "C100K", 10, 1e-3, dump=tmp_path)
    for g, w in zip(got, ref["steps_out"]):
        for k in ("time", "dt", "massr"):
            if exact:
                assert g[k] == w[k], (k, g, w)
            else:
                assert abs(g[k] - w[k]) <= 1e-12 * abs(w[k]), (k, g, w)
    for f in ("u", "v", "w", "p"):
        vals = pyoracle.read_vec(tmp_path / f"{f}.vec")
        summ = ref["fields"][f]
        assert vals.size == summ["n"]
        if exact:
            import hashlib
            assert hashlib.sha256(vals.astype("<f8").tobytes()).hexdigest() == summ["sha256"], f
        else:
            assert abs(np.linalg.norm(vals) - summ["l2"]) <= 1e-12 * summ["l2"], f
            np.testing.assert_allclose(vals[summ["sample_index"]], summ["sample"], rtol=1e-9, atol=1e-12 * summ["max_abs"])


def test_oracle_matches_reference_binary_cav46(cav46_dir, tmp_path):
    import hashlib
    ref = json.loads((GOLDEN / "ref_cav46_1p.json").read_text())
    got = pyoracle.run_cli(cav46_dir, "cav46", 10, 1e-2, dump=tmp_path)
    assert [(g["time"], g["dt"], g["massr"]) for g in got] == \
        [(w["time"], w["dt"], w["massr"]) for w in ref["steps_out"]]
    for f in ("u", "v", "w", "p"):
        vals = pyoracle.read_vec(tmp_path / f"{f}.vec")
        assert hashlib.sha256(vals.astype("<f8").tobytes()).hexdigest() == ref["fields"][f]["sha256"], f


def _random_csr(n, m, density, seed, empty_rows=()):
    rng = np.random.default_rng(seed)
    A = sp.random(n, m, density=density, random_state=rng, format="csr")
    A.data = rng.standard_normal(A.nnz)
    A = A.tolil()
    for r in empty_rows:
        A.rows[r] = []
        A.data[r] = []
    A = A.tocsr()
    A.sort_indices()
    return A


def test_oracle_spmv_matches_scipy_and_keeps_empty_rows():
    A = _random_csr(300, 420, 0.03, 1, empty_rows=(0, 17, 299))
    x = np.random.default_rng(2).standard_normal(420)
    y = np.full(300, 7.25)
    C = pyoracle.Csr(A.indptr, A.indices, A.data)
    pyoracle.spmv(C, x, y)
    ref = A @ x
    nonempty = np.diff(A.indptr) > 0
    np.testing.assert_allclose(y[nonempty], ref[nonempty], rtol=1e-13, atol=1e-13)
    assert np.all(y[~nonempty] == 7.25)
    y2 = np.full(300, 2.0)
    pyoracle.spmv_ab(C, x, y2, -0.5, 3.0)
    np.testing.assert_allclose(y2[nonempty], -0.5 * ref[nonempty] + 6.0, rtol=1e-13, atol=1e-13)
    assert np.all(y2[~nonempty] == 2.0)
    # the fourth variant writes into another operator's value array
    bv = np.linspace(-1.0, 1.0, 300)
    bw = bv.copy()
    pyoracle.spmv_vals_ab(C, x, bw, 1.5, -0.25)
    np.testing.assert_allclose(bw[nonempty], 1.5 * ref[nonempty] - 0.25 * bv[nonempty], rtol=1e-13, atol=1e-13)
    assert np.array_equal(bw[~nonempty], bv[~nonempty])


def test_oracle_pcg_solves_spd_system_and_keeps_residual_invariant():
    from gip_amd import cavity
    cav = cavity.generate(8, gamma=1e-2, ops=cavity.OP_BITS["L"])
    L = cav.op("L")
    n, nall = cav.n_inner, cav.n_all
    b = np.zeros(nall)
    b[:n] = cavity.rhs_normal(n, 7)
    x0 = np.zeros(nall)
    x, st, it, rel = pyoracle.pcg(pyoracle.Csr.of(L), b, x0, n, 400, tol=1e-12)
    assert st == 0 and it < 400 and rel <= 1e-12
    A = L.to_scipy(nall)[:, :n]
    xd = spla.spsolve(A.tocsc(), b[:n])
    np.testing.assert_allclose(x[:n], xd, rtol=1e-8, atol=1e-9 * np.abs(xd).max())
    # fixed-iteration mode runs exactly maxit iterations
    x5, st, it, rel5 = pyoracle.pcg(pyoracle.Csr.of(L), b, x0, n, 5)
    assert it == 5
    r = b[:n] - A @ x5[:n]
    assert abs(np.linalg.norm(r) / np.linalg.norm(b[:n]) - rel5) <= 1e-10


def test_oracle_pcg_breakdown_on_zero_rhs():
    from gip_amd import cavity
    cav = cavity.generate(4, gamma=1e-2, ops=cavity.OP_BITS["L"])
    L = cav.op("L")
    z = np.zeros(cav.n_all)
    x, st, it, rel = pyoracle.pcg(pyoracle.Csr.of(L), z, z, cav.n_inner, 10)
    assert st == -2 and it == 0 and np.all(x == 0)
    x, st, it, rel = pyoracle.pcg(pyoracle.Csr.of(L), z, z, cav.n_inner, 10, tol=1e-8)
    assert st == 0 and it == 0


def test_oracle_cfl_closed_form():
    rng = np.random.default_rng(3)
    n = 500
    u, v, w = rng.standard_normal((3, n))
    dxs = rng.uniform(0.01, 0.02, n)
    gamma, rho = 1e-2, 1.0
    got = pyoracle.cfl(u, v, w, dxs, n, gamma, rho)
    nrm = np.maximum(np.maximum(np.abs(u), np.abs(v)), np.abs(w))
    dt = min(1e20, np.min(0.25 * dxs / nrm), np.min(0.2 * dxs * dxs * rho / gamma))
    assert got == pytest.approx(min(0.8 * dt, 0.1), rel=1e-15)
    # zero velocity: only the diffusive limit applies; large dx: capped at 0.1
    z = np.zeros(n)
    assert pyoracle.cfl(z, z, z, dxs, n, gamma, rho) == pytest.approx(0.8 * 0.2 * dxs.min() ** 2 / gamma, rel=1e-15)
    assert pyoracle.cfl(z, z, z, np.ones(n), n, gamma, rho) == 0.1


def test_oracle_norm_and_dot():
    a = np.array([1.0, -3.5, np.nan, 2.0])
    assert pyoracle.norm_inf(a) == 3.5  # NaN never replaces the running maximum
    x = np.arange(1.0, 6.0)
    assert pyoracle.dot(x, x) == 55.0


def test_reference_binary_live_if_present(c100k_root):
    """When the reference build exists here, its np=1 run on the merged case
    equals the oracle bitwise (refreshes the fixture's claim)."""
    harness = pyoracle.REF / "gip_ref_harness"
    if not harness.exists():
        pytest.skip("reference build absent (oracle/_ref is built only where the reference sources are)")
    import subprocess
    out = subprocess.run([str(harness), "dns", "C100K", "3", "0.001", "100"], cwd=c100k_root, capture_output=True,
                         text=True, check=True).stdout
    ref = pyoracle.parse_lines(out)
    got = pyoracle.run_cli(c100k_root, "C100K", 3, 1e-3)
    assert [(r["time"], r["dt"], r["massr"]) for r in ref] == [(g["time"], g["dt"], g["massr"]) for g in got]


def test_oracle_threads_and_memory_case_are_bitwise_the_serial_file_case(tmp_path):
    """The oracle's threaded row loops (products, axpy) and its in-memory
    case give bitwise the serial, file-read results: each row is computed
    by one thread with the serial expression, reductions stay serial."""
    from gip_amd import cavity
    cav = cavity.generate(48, beta=1.5, gamma=1e-3)  # 110,592 rows: above the threading threshold
    cav.write(tmp_path, "t48")
    runs = []
    for threads, mem in ((1, False), (4, False), (4, True)):
        pyoracle.set_threads(threads)
        d = pyoracle.Dns(cavity=cav, gamma=1e-3) if mem else pyoracle.Dns(tmp_path, "t48", gamma=1e-3)
        rows = [d.step() for _ in range(2)]
        runs.append((rows, [d.field(f) for f in ("u", "v", "w", "p", "mf")]))
        d.close()
    pyoracle.set_threads(1)
    for rows, fields in runs[1:]:
        assert rows == runs[0][0]
        for a, b in zip(fields, runs[0][1]):
            assert np.array_equal(a, b)


def test_oracle_blocked_dot_is_the_serial_sum_up_to_rounding():
    rng = np.random.default_rng(5)
    x, y = rng.standard_normal(300_001), rng.standard_normal(300_001)
    serial = pyoracle.dot(x, y)
    pyoracle.set_blocked_dot(True)
    try:
        blocked = pyoracle.dot(x, y)
        pyoracle.set_threads(3)
        again = pyoracle.dot(x, y)
    finally:
        pyoracle.set_blocked_dot(False)
        pyoracle.set_threads(1)
    exact = float(np.sum(x * y, dtype=np.longdouble))
    assert blocked == again  # independent of the thread count
    assert abs(blocked - exact) <= 1e-12 * np.abs(x * y).sum()
    assert abs(serial - exact) <= 1e-12 * np.abs(x * y).sum()
