"""The device path at the launch shapes of the large configurations, where
the grid and partial-sum sizes differ from the small cases (the sweeping
kernels at their 2048-workgroup cap, products of 8192+ workgroups whose
partials are folded), against the oracle on the same inputs.

Bar: FP64 PCG iterates within 1e-10 of the oracle's (only dot-product
summation orders differ), tolerance stops within one iteration of the
oracle's; the mixed-precision solve reaches a 1e-10 FP64 true residual
(checked with scipy's FP64 product); the integrator's per-step time, dt,
massr and fields within 1e-10 of the oracle's.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from gip_amd import cavity, dns, hip
from oracle import pyoracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = hip.Device(0)
    yield d
    d.close()


@pytest.fixture(scope="module")
def lap128():
    c = cavity.generate(128, gamma=1e-3, ops=cavity.OP_BITS["L"])
    L = c.op("L")
    b = np.zeros(c.n_all)
    b[:c.n_inner] = cavity.rhs_normal(c.n_inner, 7)
    return c, L, b


def test_pcg_fp64_128_fixed_iterations_vs_oracle(dev, lap128):
    c, L, b = lap128
    n, nall = c.n_inner, c.n_all
    x0 = np.zeros(nall)
    A = hip.Matrix.from_csr(dev, L)
    assert A.info()["format"] == hip.FMT_STENCIL
    solver = hip.PCG(A, nall)
    dx = dev.array(nall)
    r = solver.solve(dev.to_device(b), dx, dev.to_device(x0), maxit=50)
    want, st, it, rel = pyoracle.pcg(pyoracle.Csr.of(L), b, x0, n, 50)
    got = dx.download()
    print(f"[128^3 fp64] iters {r.iters} relres {r.relres:.3e} oracle {rel:.3e}", flush=True)
    assert r.iters == it == 50 and r.status == 0
    assert np.linalg.norm(got - want) <= 1e-10 * np.linalg.norm(want)
    assert abs(r.relres - rel) <= 1e-8 * rel


def test_pcg_fp64_128_tolerance_stop_vs_oracle(dev, lap128):
    """A solve stopped by its tolerance: the direction kernel that detects
    convergence applies the last x update on every row (2048 workgroups)."""
    c, L, b = lap128
    n, nall = c.n_inner, c.n_all
    x0 = np.zeros(nall)
    A = hip.Matrix.from_csr(dev, L)
    solver = hip.PCG(A, nall)
    dx = dev.array(nall)
    tol = 1e-5
    r = solver.solve(dev.to_device(b), dx, dev.to_device(x0), maxit=5000, tol=tol)
    want, st, it, rel = pyoracle.pcg(pyoracle.Csr.of(L), b, x0, n, 5000, tol=tol)
    got = dx.download()
    res = b[:n] - (L.to_scipy(nall) @ got)[:n]
    true_rel = np.linalg.norm(res) / np.linalg.norm(b[:n])
    print(f"[128^3 fp64 tol {tol:g}] iters {r.iters} (oracle {it}) relres {r.relres:.3e} true {true_rel:.3e}",
          flush=True)
    assert r.status == 0 and st == 0 and abs(r.iters - it) <= 1
    assert r.relres <= tol and true_rel <= 1.01 * tol
    if r.iters == it:
        assert np.linalg.norm(got - want) <= 1e-10 * np.linalg.norm(want)


def test_pcg_mixed_128_true_residual(dev, lap128):
    c, L, b = lap128
    n, nall = c.n_inner, c.n_all
    A = hip.Matrix.from_csr(dev, L, fp32=True)
    solver = hip.PCG(A, nall)
    dx = dev.array(nall)
    r = solver.solve(dev.to_device(b), dx, dev.to_device(np.zeros(nall)), maxit=20000, tol=1e-10,
                     precision=hip.PCG_MIXED, poll=32)
    x = dx.download()
    Ls = L.to_scipy(nall)
    res = b[:n] - (Ls @ x)[:n]
    true_rel = np.linalg.norm(res) / np.linalg.norm(b[:n])
    print(f"[128^3 mixed] iters {r.iters} corrections {r.corrections} true {true_rel:.3e}", flush=True)
    assert r.status == 0 and true_rel <= 1.01e-10
    assert np.array_equal(x[n:], np.zeros(nall - n))  # ghost entries carried over from x0
    # the inner iteration deflated by the constant vector (A1 taken from its
    # one-entry list on this long operator) needs 832 iterations here, 992
    # without the deflation: a wrong A1 or deflation coefficient loses that
    assert r.iters <= 900, r.iters


def test_dns_96_reference_sequence_vs_oracle(tmp_path):
    cavity.generate(96, gamma=1e-3).write(tmp_path, "cav96")
    sim = dns.Simulation(root=tmp_path, name="cav96", gamma=1e-3, pcg_mode="fused")
    orc = pyoracle.Dns(tmp_path, "cav96", gamma=1e-3)
    for k in range(3):
        g, w = sim.step(), orc.step()
        print(f"[96^3 step {k + 1}] gpu massr {g['massr']:.9e} dt {g['dt']:.9e} | cpu massr {w['massr']:.9e}",
              flush=True)
        assert g["pcg_iters"] == w["pcg_iters"] == 100
        for key in ("time", "dt", "massr"):
            assert abs(g[key] - w[key]) <= 1e-10 * abs(w[key]), (k, key, g[key], w[key])
    for f in ("u", "v", "w", "p", "mf"):
        a, r = sim.field(f), orc.field(f)
        err = np.linalg.norm(a - r) / np.linalg.norm(r)
        print(f"[96^3 field {f}] rel L2 {err:.3e}", flush=True)
        assert err <= 1e-10, (f, err)


@pytest.mark.parametrize("n", [2 ** 19, 2 ** 19 + 1, 2 ** 23, 2 ** 23 + 1])
def test_pcg_either_side_of_the_iteration_switches(dev, n):
    """2^19 rows: 2048 product chunks, read directly by the next kernels (two
    kernels per FP64 iteration); one row more: the chunk partials go through
    the fold kernel and the sweeps re-reduce them in their prologues (four
    kernels); 2^23 + 1 rows: the row-per-thread sweeps with one-workgroup
    alpha and beta kernels (seven).  A shifted 1-D Laplacian: 40 fixed FP64
    iterations and a tolerance stop against the oracle, and a
    mixed-precision solve to a 1e-10 true residual."""
    T = sp.diags([-1.0, 2.05, -1.0], [-1, 0, 1], shape=(n, n)).tocsr()
    T.sort_indices()
    b = np.random.default_rng(n).standard_normal(n)
    x0 = np.zeros(n)
    orc = pyoracle.Csr(T.indptr, T.indices, T.data, n)
    M = hip.Matrix(dev, T.indptr, T.indices, T.data, ncols=n, fp32=True)
    solver = hip.PCG(M, n)
    dx = dev.array(n)
    r = solver.solve(dev.to_device(b), dx, dev.to_device(x0), maxit=40)
    want, st, it, rel = pyoracle.pcg(orc, b, x0, n, 40)
    assert r.iters == it == 40 and r.status == 0
    assert np.linalg.norm(dx.download() - want) <= 1e-10 * np.linalg.norm(want)
    assert abs(r.relres - rel) <= 1e-8 * rel
    r = solver.solve(dev.to_device(b), dx, dev.to_device(x0), maxit=1000, tol=1e-9)
    want, st, it, rel = pyoracle.pcg(orc, b, x0, n, 1000, tol=1e-9)
    print(f"[n={n}] tolerance stop: iters {r.iters} (oracle {it})", flush=True)
    assert r.status == 0 and st == 0 and abs(r.iters - it) <= 1
    if r.iters == it:
        assert np.linalg.norm(dx.download() - want) <= 1e-10 * np.linalg.norm(want)
    r = solver.solve(dev.to_device(b), dx, dev.to_device(x0), maxit=1000, tol=1e-10, precision=hip.PCG_MIXED)
    x = dx.download()
    assert r.status == 0 and np.linalg.norm(b - T @ x) <= 1.01e-10 * np.linalg.norm(b)


@pytest.mark.parametrize("precision", [hip.PCG_FP64, hip.PCG_MIXED])
def test_pcg_long_path_layouts_give_identical_iterates(dev, precision):
    """81^3 cavity (531,441 rows: the three-kernel iteration with the fold):
    stencil, ELL and CSR operators sum each row in stored order, so the
    iterates agree bitwise whichever layout holds L; the FP64 ones also
    against the oracle to 1e-10."""
    c = cavity.generate(81, gamma=1e-3, ops=cavity.OP_BITS["L"])
    L = c.op("L")
    n, nall = c.n_inner, c.n_all
    assert n > 2 ** 19
    b = np.zeros(nall)
    b[:n] = cavity.rhs_normal(n, 3)
    z = dev.to_device(np.zeros(nall))
    outs = []
    for fmt in (hip.FMT_STENCIL, hip.FMT_ELL, hip.FMT_CSR):
        A = hip.Matrix.from_csr(dev, L, fmt=fmt, fp32=True)
        assert A.info()["format"] == fmt
        solver = hip.PCG(A, nall)
        dx = dev.array(nall)
        r = solver.solve(dev.to_device(b), dx, z, maxit=60, precision=precision)
        outs.append((dx.download(), r.iters, r.corrections))
        solver.destroy()
        A.destroy()
    for x, it, corr in outs[1:]:
        assert it == outs[0][1] and corr == outs[0][2]
        assert np.array_equal(x, outs[0][0])
    if precision == hip.PCG_FP64:
        want, st, it, rel = pyoracle.pcg(pyoracle.Csr.of(L), b, np.zeros(nall), n, 60)
        assert outs[0][1] == it == 60
        assert np.linalg.norm(outs[0][0] - want) <= 1e-10 * np.linalg.norm(want)
