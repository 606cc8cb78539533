"""gfx950 kernels through the C ABI against the oracle (C restatement of the
reference CPU path) on the same inputs.

Bar: SpMV (every layout), axpy, axpby, infinity norm and CFL are bitwise
equal to the oracle (same summation order, no contraction); dot products
differ only by summation order (1e-13 relative); the fused FP64 PCG agrees
with the oracle's PCG to 1e-10 relative and its tolerance-mode iteration
counts to +-1; the mixed-precision PCG reaches a 1e-10 FP64 true residual.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from gip_amd import cavity, hip
from oracle import pyoracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = hip.Device(0)
    yield d
    d.close()


@pytest.fixture(scope="module")
def cav46():
    return cavity.generate(46, gamma=1e-2)


@pytest.fixture(scope="module")
def c100k(c100k_root):
    return cavity.read(c100k_root, "C100K")


def _ncols(c, A):
    return max(c.n_all, c.n_faces, int(A.col.max()) + 1 if A.nnz else 1)


def _check_op(dev, c, name, fmt=hip.FMT_AUTO, seed=0, slice=64):
    A = c.op(name)
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(_ncols(c, A))
    y0 = rng.standard_normal(A.nrows)
    M = hip.Matrix.from_csr(dev, A, fmt=fmt, slice=slice)
    dx, dy = dev.to_device(x), dev.to_device(y0)
    M.spmv(dx, dy)
    got = dy.download()
    want = y0.copy()
    pyoracle.spmv(pyoracle.Csr.of(A), x, want)
    assert np.array_equal(got, want), (name, fmt, np.abs(got - want).max())
    dy.upload(y0)
    M.spmv_ab(dx, dy, 0.7, -1.3)
    got = dy.download()
    want = y0.copy()
    pyoracle.spmv_ab(pyoracle.Csr.of(A), x, want, 0.7, -1.3)
    assert np.array_equal(got, want), (name, fmt, "ab")
    return M.info()


@pytest.mark.parametrize("name", cavity.OPS)
def test_spmv_bitwise_cav46(dev, cav46, name):
    print(f"[spmv cav46 {name}]", flush=True)
    _check_op(dev, cav46, name)


@pytest.mark.parametrize("name", cavity.OPS)
def test_spmv_bitwise_c100k(dev, c100k, name):
    print(f"[spmv c100k {name}]", flush=True)
    _check_op(dev, c100k, name)


@pytest.mark.parametrize("fmt", [hip.FMT_CSR, hip.FMT_ELL, hip.FMT_STENCIL])
def test_spmv_forced_layouts(dev, cav46, fmt):
    # Ec, Gx and Lb have empty or irregular rows: a forced stencil layout that
    # does not fit falls back to the automatic choice
    fits = {"L": True, "SO": True, "Mres": True, "Ec": False, "Gx": True, "Lb": False}
    for name in fits:
        info = _check_op(dev, cav46, name, fmt=fmt)
        if fmt != hip.FMT_STENCIL or fits[name]:
            assert info["format"] == fmt, (name, info)
        else:
            assert info["format"] != hip.FMT_STENCIL, (name, info)


@pytest.mark.parametrize("slice", [32, 128])
def test_ell_slice_heights_bitwise(dev, cav46, slice):
    """Sliced ELL with 32 or 128 rows per slice (64 is the default): every
    product variant bitwise against the oracle, the FP32 product equal to the
    64-row layout's, and the PCG iterates on L equal to the stencil
    layout's (the solver reads ELL through the same row sum)."""
    for name in ("L", "Ec", "Gx", "Lb", "SO"):
        info = _check_op(dev, cav46, name, fmt=hip.FMT_ELL, slice=slice)
        assert info["format"] == hip.FMT_ELL, (name, info)
    A = cav46.op("Ec")
    x = np.random.default_rng(7).standard_normal(_ncols(cav46, A)).astype(np.float32)
    outs = []
    for sl in (64, slice):
        M = hip.Matrix.from_csr(dev, A, fmt=hip.FMT_ELL, fp32=True, slice=sl)
        y32 = dev.array(A.nrows, np.float32)
        M.spmv32(dev.to_device(x), y32)
        outs.append(y32.download())
    assert np.array_equal(outs[0], outs[1])
    L = cav46.op("L")
    n, nall = cav46.n_inner, cav46.n_all
    b = np.zeros(nall)
    b[:n] = cavity.rhs_normal(n, 5)
    z = dev.to_device(np.zeros(nall))
    xs = []
    for fmt, sl in ((hip.FMT_STENCIL, 64), (hip.FMT_ELL, slice)):
        M = hip.Matrix.from_csr(dev, L, fmt=fmt, slice=sl)
        assert M.info()["format"] == fmt
        solver = hip.PCG(M, nall)
        dx = dev.array(nall)
        r = solver.solve(dev.to_device(b), dx, z, maxit=30)
        assert r.iters == 30 and r.status == 0
        xs.append(dx.download())
        solver.destroy()
    assert np.array_equal(xs[0], xs[1])


def _banded(n, offsets, seed, extra=0.02, empty=(), shuffle_rows=0):
    """Banded operator with random values, some rows carrying one extra
    off-band entry at a random place in their stored order, some empty rows
    and some rows stored out of column order."""
    rng = np.random.default_rng(seed)
    rowptr, cols, vals = [0], [], []
    shuffled = set(rng.choice(n, size=shuffle_rows, replace=False).tolist()) if shuffle_rows else set()
    for r in range(n):
        c = [r + d for d in offsets if 0 <= r + d < n]
        if r in empty:
            c = []
        elif rng.random() < extra:
            c.insert(int(rng.integers(0, len(c) + 1)), int(rng.integers(0, n + 100)))
        if r in shuffled:
            rng.shuffle(c)
        cols += c
        vals += list(rng.standard_normal(len(c)))
        rowptr.append(len(cols))
    return np.array(rowptr, np.int32), np.array(cols, np.int32), np.array(vals)


@pytest.mark.parametrize("seed", [0, 1])
def test_stencil_layout_bitwise_with_tails_and_empty_rows(dev, seed):
    """Off-band entries, empty rows and rows stored out of column order go
    to the tail; the product keeps every row's stored order."""
    n = 5000
    rp, cl, vl = _banded(n, (-70, -1, 0, 1, 70), seed, empty={0, 64, 65, 4999}, shuffle_rows=20)
    ncx = int(cl.max()) + 1
    rng = np.random.default_rng(seed + 10)
    x, y0 = rng.standard_normal(ncx), rng.standard_normal(n)
    for fp32 in (False, True):
        M = hip.Matrix(dev, rp, cl, vl, ncols=ncx, fmt=hip.FMT_STENCIL, fp32=fp32)
        assert M.info()["format"] == hip.FMT_STENCIL
        dy = dev.to_device(y0)
        M.spmv_ab(dev.to_device(x), dy, 0.5, -2.0)
        want = y0.copy()
        pyoracle.spmv_ab(pyoracle.Csr(rp, cl, vl), x, want, 0.5, -2.0)
        assert np.array_equal(dy.download(), want), fp32
        assert np.array_equal(M.values(), vl)
        if fp32:  # the FP32 product, stored order in FP32 arithmetic
            x32, v32 = x.astype(np.float32), vl.astype(np.float32)
            y32 = dev.array(n, np.float32)
            M.spmv32(dev.to_device(x32), y32)
            want32 = np.zeros(n, np.float32)
            for r in range(n):
                acc = np.float32(0.0)
                for j in range(rp[r], rp[r + 1]):
                    acc = np.float32(acc + np.float32(x32[cl[j]] * v32[j]))
                want32[r] = acc
            assert np.array_equal(y32.download(), want32)


@pytest.mark.parametrize("lanes", [2, 4, 8])
def test_spmv_lanes_per_row_bitwise(dev, cav46, lanes):
    """CSR with several lanes per row keeps the stored summation order."""
    for name in ("L", "Ec", "SO", "Mres"):
        A = cav46.op(name)
        rng = np.random.default_rng(lanes)
        x = rng.standard_normal(_ncols(cav46, A))
        y0 = rng.standard_normal(A.nrows)
        M = hip.Matrix.from_csr(dev, A, fmt=hip.FMT_CSR)
        M.set_lanes(lanes)
        dy = dev.to_device(y0)
        M.spmv_ab(dev.to_device(x), dy, 0.25, 2.0)
        want = y0.copy()
        pyoracle.spmv_ab(pyoracle.Csr.of(A), x, want, 0.25, 2.0)
        assert np.array_equal(dy.download(), want), (name, lanes)


def test_autotuner_picks_a_timed_candidate(dev, cav46):
    for name in ("L", "Ec"):
        A = cav46.op(name)
        M, rep = hip.Matrix.autotune(dev, A, reps=3)
        timed = [u for u in rep["us"] if u > 0]
        # candidate order: ELL 32/64/128 rows per slice, CSR 1/2/4/8 lanes, stencil
        if rep["format"] == hip.FMT_ELL:
            idx = {32: 0, 64: 1, 128: 2}[rep["slice"]]
        elif rep["format"] == hip.FMT_STENCIL:
            idx = 7
        else:
            idx = {1: 3, 2: 4, 4: 5, 8: 6}[rep["lanes"]]
        assert timed and min(timed) == pytest.approx(rep["us"][idx], rel=1e-6)
        x = np.random.default_rng(0).standard_normal(_ncols(cav46, A))
        dy = dev.to_device(np.zeros(A.nrows))
        M.spmv(dev.to_device(x), dy)
        want = np.zeros(A.nrows)
        pyoracle.spmv(pyoracle.Csr.of(A), x, want)
        assert np.array_equal(dy.download(), want)
        print(f"[autotune {name}] {rep}")


def test_layout_choice(dev, cav46):
    # the 7-point Laplacian: per-slice offsets shared by the rows (stencil)
    assert hip.Matrix.from_csr(dev, cav46.op("L")).info()["format"] == hip.FMT_STENCIL
    # convection from faces (empty rows, irregular): ELL or CSR
    assert hip.Matrix.from_csr(dev, cav46.op("Ec")).info()["format"] in (hip.FMT_ELL, hip.FMT_CSR)
    A = cav46.op("L")
    n = A.nrows
    D = hip.Matrix(dev, np.arange(n + 1), np.arange(n), np.ones(n))
    assert D.info()["format"] == hip.FMT_DIAG


def test_spmv_random_with_empty_rows(dev):
    rng = np.random.default_rng(4)
    A = sp.random(1000, 777, density=0.01, random_state=rng, format="lil")
    for r in (0, 5, 63, 64, 999):
        A.rows[r], A.data[r] = [], []
    A = A.tocsr()
    A.sort_indices()
    A.data = rng.standard_normal(A.nnz)
    x = rng.standard_normal(777)
    y0 = rng.standard_normal(1000)
    for fmt, sl in ((hip.FMT_AUTO, 64), (hip.FMT_CSR, 64), (hip.FMT_ELL, 32), (hip.FMT_ELL, 64), (hip.FMT_ELL, 128)):
        M = hip.Matrix(dev, A.indptr, A.indices, A.data, ncols=777, fmt=fmt, slice=sl)
        dy = dev.to_device(y0)
        M.spmv(dev.to_device(x), dy)
        want = y0.copy()
        pyoracle.spmv(pyoracle.Csr(A.indptr, A.indices, A.data), x, want)
        assert np.array_equal(dy.download(), want), (fmt, sl)


def test_spmv_into_values_rebuilds_convection(dev, cav46):
    """Ec.mf written into C's values (GIP_MatrixMulticore.cpp:105-123)."""
    Ec, D = cav46.op("Ec"), cav46.op("D")
    mf = np.random.default_rng(1).standard_normal(cav46.n_faces)
    E = hip.Matrix.from_csr(dev, Ec)
    C = hip.Matrix.from_csr(dev, D)
    E.spmv_vals(dev.to_device(mf), C)
    want = D.val.copy()
    pyoracle.spmv(pyoracle.Csr.of(Ec), mf, want)
    assert np.array_equal(C.values(), want)


@pytest.mark.parametrize("fmt", [None, "ELL", "CSR"])
def test_spmv_into_values_alpha_beta(dev, cav46, fmt):
    """alpha Ec.mf + beta C's values, written into C's values on the device
    (GIP_MatrixMulticore.cpp:125-143): bitwise the oracle's, in every layout
    of C; rows of the product operator that are empty (every 7th row of Ec
    emptied here) leave their value."""
    Ec0, D = cav46.op("Ec"), cav46.op("D")
    S = Ec0.to_scipy(cav46.n_faces).tolil()
    for r in range(0, S.shape[0], 7):
        S.rows[r] = []
        S.data[r] = []
    S = S.tocsr()
    S.sort_indices()
    Ec = cavity.CSR(S.indptr.astype(np.int32), S.indices.astype(np.int32), S.data, S.shape[0], cav46.n_faces)
    assert (np.diff(Ec.rowptr) == 0).any()
    mf = np.random.default_rng(4).standard_normal(cav46.n_faces)
    E = hip.Matrix.from_csr(dev, Ec)
    kw = {} if fmt is None else {"fmt": getattr(hip, "FMT_" + fmt)}
    C = hip.Matrix.from_csr(dev, D, **kw)
    E.spmv_vals_ab(dev.to_device(mf), C, 0.75, -1.5)
    want = D.val.copy()
    pyoracle.spmv_vals_ab(pyoracle.Csr.of(Ec), mf, want, 0.75, -1.5)
    assert np.array_equal(C.values(), want)
    with pytest.raises(ValueError):
        C.spmv_vals_ab(dev.to_device(mf), C, 1.0, 0.0)  # rows(C) != nnz(C)


def _inv_diag(A, vals):
    """1/a_ii from CSR values, the row's last diagonal entry."""
    d = np.zeros(A.nrows)
    for i in range(A.nrows):
        for j in range(A.rowptr[i], A.rowptr[i + 1]):
            if A.col[j] == i:
                d[i] = vals[j]
    with np.errstate(divide="ignore"):  # a zero diagonal gives inf on both sides
        return 1.0 / d


@pytest.mark.parametrize("fmt", [None, "ELL", "CSR"])
def test_jacobi_diagonal_follows_value_changes(dev, cav46, fmt):
    """gip_mat_diag_inv after gip_mat_set_values and after a product into
    the values (gip_dcsrmv_vals) uses the new diagonal, in every layout."""
    kw = {} if fmt is None else {"fmt": getattr(hip, "FMT_" + fmt)}
    L = cav46.op("L")
    A = hip.Matrix.from_csr(dev, L, **kw)
    v2 = L.val * (1.0 + np.random.default_rng(5).random(L.val.size))
    A.set_values(v2)
    out = dev.array(L.nrows)
    A.diag_inv(out)
    assert np.array_equal(out.download(), _inv_diag(L, v2))
    Ec, D = cav46.op("Ec"), cav46.op("D")
    mf = np.random.default_rng(6).standard_normal(cav46.n_faces)
    E = hip.Matrix.from_csr(dev, Ec)
    C = hip.Matrix.from_csr(dev, D, **kw)
    E.spmv_vals(dev.to_device(mf), C)
    want = D.val.copy()
    pyoracle.spmv(pyoracle.Csr.of(Ec), mf, want)
    if all(any(D.col[j] == i for j in range(D.rowptr[i], D.rowptr[i + 1])) for i in range(D.nrows)):
        out = dev.array(D.nrows)
        C.diag_inv(out)
        assert np.array_equal(out.download(), _inv_diag(D, want))


def test_transposed_spmv(dev, cav46):
    for name in ("Mx", "SO", "Ec"):
        A = cav46.op(name)
        nc = _ncols(cav46, A)
        x = np.random.default_rng(2).standard_normal(A.nrows)
        M = hip.Matrix.from_csr(dev, A)
        dy = dev.array(nc)
        M.spmv_t(dev.to_device(x), dy, nc)
        want = A.to_scipy(nc).T @ x
        np.testing.assert_allclose(dy.download(), want, rtol=1e-13, atol=1e-13 * np.abs(want).max())


@pytest.mark.parametrize("n", [0, 1, 63, 64, 65, 1000, 1 << 20])
def test_vector_ops(dev, n):
    rng = np.random.default_rng(n)
    x, y = rng.standard_normal(n), rng.standard_normal(n)
    dx, dy = dev.to_device(x), dev.to_device(y)
    hip.axpy(0.37, dx, dy)
    assert np.array_equal(dy.download(), y + 0.37 * x)
    dy.upload(y)
    hip.axpby(-1.25, dx, 0.5, dy)
    assert np.array_equal(dy.download(), -1.25 * x + 0.5 * y)
    d = hip.dot(dx, dev.to_device(y))
    want = pyoracle.dot(x, y)
    assert abs(d - want) <= 1e-13 * max(1.0, np.abs(x * y).sum())
    assert hip.nrm_inf(dx) == (np.abs(x).max() if n else 0.0)
    hip.fill(dy, 3.5)
    assert np.all(dy.download() == 3.5)
    if n:
        hip.copy(dx, dy, n - 1)  # copyTo over a prefix leaves the tail alone
        got = dy.download()
        assert np.array_equal(got[:n - 1], x[:n - 1]) and got[n - 1] == 3.5


def test_norm_ignores_nan_like_the_cpu(dev):
    a = np.array([1.0, np.nan, -4.0, 2.0])
    assert hip.nrm_inf(dev.to_device(a)) == pyoracle.norm_inf(a) == 4.0


def test_cfl_bitwise(dev, cav46):
    rng = np.random.default_rng(6)
    n = cav46.n_inner
    u, v, w = (rng.standard_normal(cav46.n_all) for _ in range(3))
    u[:100] = 0.0
    v[:100] = 0.0
    w[:100] = 0.0
    dxs = cav46.field("dxs").copy()
    d = [dev.to_device(a) for a in (u, v, w, dxs)]
    for gamma in (1e-2, 1e-3, 1e3):
        got = hip.cfl(*d, n, gamma, 1.0)
        assert got == pyoracle.cfl(u, v, w, dxs, n, gamma, 1.0)


def _pcg_setup(dev, c, seed=11):
    L = c.op("L")
    n, nall = c.n_inner, c.n_all
    b = np.zeros(nall)
    b[:n] = cavity.rhs_normal(n, seed)
    A = hip.Matrix.from_csr(dev, L, fp32=True)
    return L, n, nall, b, A


def test_pcg_fixed_iterations_vs_oracle(dev, cav46):
    L, n, nall, b, A = _pcg_setup(dev, cav46)
    x0 = np.zeros(nall)
    x0[:n] = 0.01 * np.random.default_rng(3).standard_normal(n)
    solver = hip.PCG(A, nall)
    dx = dev.array(nall)
    r = solver.solve(dev.to_device(b), dx, dev.to_device(x0), maxit=100)
    want, st, it, rel = pyoracle.pcg(pyoracle.Csr.of(L), b, x0, n, 100)
    got = dx.download()
    assert r.iters == it == 100 and r.status == 0
    assert np.linalg.norm(got - want) <= 1e-10 * np.linalg.norm(want)
    assert np.array_equal(got[n:], x0[n:])
    assert abs(r.relres - rel) <= 1e-8 * rel


@pytest.mark.parametrize("tol", [1e-6, 1e-8, 1e-10])
def test_pcg_tolerance_iteration_count_within_one(dev, cav46, tol):
    L, n, nall, b, A = _pcg_setup(dev, cav46)
    x0 = np.zeros(nall)
    solver = hip.PCG(A, nall)
    dx = dev.array(nall)
    r = solver.solve(dev.to_device(b), dx, dev.to_device(x0), maxit=2000, tol=tol)
    want, st, it, rel = pyoracle.pcg(pyoracle.Csr.of(L), b, x0, n, 2000, tol=tol)
    assert st == 0 and r.status == 0
    assert abs(r.iters - it) <= 1, (r.iters, it)
    assert r.relres <= tol
    # the same stopping iteration gives the same iterate (every row's last
    # x update applied); one iteration apart, the iterates differ by a step
    bound = 1e-10 if r.iters == it else 1e-6
    assert np.linalg.norm(dx.download() - want) <= bound * np.linalg.norm(want)


@pytest.mark.parametrize("precision", [hip.PCG_FP64, hip.PCG_MIXED])
def test_pcg_product_timing_changes_nothing(dev, cav46, precision):
    """time_products records events around one product launch per host
    poll: the iterates and counts are bitwise those of an untimed solve, and
    the sampled durations are positive."""
    L, n, nall, b, A = _pcg_setup(dev, cav46)
    solver = hip.PCG(A, nall)
    x0 = dev.to_device(np.zeros(nall))
    outs = []
    for timed in (False, True):
        dx = dev.array(nall)
        r = solver.solve(dev.to_device(b), dx, x0, maxit=5000, tol=1e-9, precision=precision, poll=8,
                         time_products=timed)
        outs.append((r, dx.download()))
    (r0, x_plain), (r1, x_timed) = outs
    print(f"[timed products] {r1.product_samples} samples, median {1e3 * r1.product_ms:.2f} us", flush=True)
    assert r0.status == r1.status == 0 and r0.iters == r1.iters and r0.corrections == r1.corrections
    assert np.array_equal(x_plain, x_timed)
    assert r0.product_samples == 0 and r0.product_ms == 0.0
    assert r1.product_samples >= r1.iters // 8 - 1 and r1.product_ms > 0.0


def test_pcg_mixed_precision_reaches_fp64_residual(dev, cav46):
    L, n, nall, b, A = _pcg_setup(dev, cav46)
    x0 = np.zeros(nall)
    solver = hip.PCG(A, nall)
    dx = dev.array(nall)
    r = solver.solve(dev.to_device(b), dx, dev.to_device(x0), maxit=5000, tol=1e-10, precision=hip.PCG_MIXED)
    x = dx.download()
    Ax = np.zeros(nall)
    pyoracle.spmv(pyoracle.Csr.of(L), x, Ax)
    true_rel = np.linalg.norm(b[:n] - Ax[:n]) / np.linalg.norm(b[:n])
    assert r.status == 0 and r.corrections >= 1
    assert true_rel <= 1e-10 * 1.01, true_rel
    want, st, it, rel = pyoracle.pcg(pyoracle.Csr.of(L), b, x0, n, 5000, tol=1e-12)
    assert np.linalg.norm(x - want) <= 1e-7 * np.linalg.norm(want)
    r64 = solver.solve(dev.to_device(b), dx, dev.to_device(x0), maxit=5000, tol=1e-10)
    print(f"[mixed] iters {r.iters} corrections {r.corrections} true {true_rel:.3e} "
          f"fp64 iters {r64.iters} (oracle to 1e-12: {it})")
    # the deflated inner iteration (the small operators' dense A1 read)
    # needs fewer iterations than FP64 PCG to the same residual; without the
    # deflation it needed more (608 against 484 on the 46^3 benchmark case)
    assert r64.status == 0 and r.iters <= 0.85 * r64.iters, (r.iters, r64.iters)


@pytest.mark.parametrize("coarse", [True, False])
def test_pcg_mixed_on_a_general_spd_operator(dev, coarse):
    """A shifted 2-D Laplacian (its constant vector is no near-null mode):
    the mixed solver reaches a 1e-10 FP64 true residual with the coarse step
    on or off, and agrees with a direct solve."""
    m = 80
    n = m * m
    T = sp.diags([-1.0, 2.0, -1.0], [-1, 0, 1], shape=(m, m))
    A = (sp.kron(sp.identity(m), T) + sp.kron(T, sp.identity(m)) + 0.05 * sp.identity(n)).tocsr()
    A.sort_indices()
    b = np.random.default_rng(9).standard_normal(n)
    M = hip.Matrix(dev, A.indptr, A.indices, A.data, ncols=n, fp32=True)
    solver = hip.PCG(M, n)
    dx = dev.array(n)
    r = solver.solve(dev.to_device(b), dx, dev.to_device(np.zeros(n)), maxit=5000, tol=1e-10,
                     precision=hip.PCG_MIXED, coarse=coarse)
    x = dx.download()
    assert r.status == 0
    assert np.linalg.norm(b - A @ x) <= 1.01e-10 * np.linalg.norm(b)
    want = spla.spsolve(A.tocsc(), b)
    assert np.linalg.norm(x - want) <= 1e-8 * np.linalg.norm(want)


@pytest.mark.parametrize("precision", [hip.PCG_FP64, hip.PCG_MIXED])
def test_pcg_layouts_give_identical_iterates(dev, cav46, precision):
    """Stencil, ELL and CSR operators all sum each row in stored order, so
    the PCG iterates agree bitwise whichever layout holds L."""
    L = cav46.op("L")
    n, nall = cav46.n_inner, cav46.n_all
    b = np.zeros(nall)
    b[:n] = cavity.rhs_normal(n, 5)
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
    for x, it, corr in outs[1:]:
        assert it == outs[0][1] and corr == outs[0][2]
        assert np.array_equal(x, outs[0][0])


def test_pcg_breakdown_is_reported(dev, cav46):
    L, n, nall, b, A = _pcg_setup(dev, cav46)
    z = dev.array(nall)
    hip.fill(z, 0.0)
    solver = hip.PCG(A, nall)
    dx = dev.array(nall)
    with pytest.raises(hip.HipError) as e:
        solver.solve(z, dx, z, maxit=10)
    assert e.value.code == hip.EBREAKDOWN
    r = solver.solve(z, dx, z, maxit=10, tol=1e-8)
    assert r.iters == 0 and r.status == 0


def test_pcg_argument_errors_and_statuses(dev, cav46):
    """The PCG's C ABI rejects bad arguments with GIP_EINVAL (nothing is
    launched) and reports a solve that stops at maxit above its tolerance as
    GIP_ENOCONV in its status."""
    from ctypes import byref, c_void_p
    L = cav46.op("L")
    n, nall = cav46.n_inner, cav46.n_all
    b = np.zeros(nall)
    b[:n] = cavity.rhs_normal(n, 2)
    lib = hip.lib()
    A64 = hip.Matrix.from_csr(dev, L)  # no FP32 copy
    solver = hip.PCG(A64, nall)
    db, dx, z = dev.to_device(b), dev.array(nall), dev.to_device(np.zeros(nall))
    st = hip.PcgStats()
    good = hip.PcgOpts(10, 0.0, 1, hip.PCG_FP64, 0.1, 16, 0, 0)
    assert lib.gip_pcg_solve(solver.ptr, None, dx.ptr, z.ptr, byref(good), byref(st)) == hip.EINVAL
    assert lib.gip_pcg_solve(solver.ptr, db.ptr, dx.ptr, z.ptr, None, byref(st)) == hip.EINVAL
    neg = hip.PcgOpts(-1, 0.0, 1, hip.PCG_FP64, 0.1, 16, 0, 0)
    assert lib.gip_pcg_solve(solver.ptr, db.ptr, dx.ptr, z.ptr, byref(neg), byref(st)) == hip.EINVAL
    mixed = hip.PcgOpts(10, 1e-8, 1, hip.PCG_MIXED, 0.1, 16, 0, 0)
    assert lib.gip_pcg_solve(solver.ptr, db.ptr, dx.ptr, z.ptr, byref(mixed), byref(st)) == hip.EINVAL
    assert b"FP32" in lib.gip_last_error()
    out = c_void_p()
    assert lib.gip_pcg_create(dev.ctx, A64.ptr, n - 1, byref(out)) == hip.EINVAL  # vectors shorter than the rows
    other = hip.Device(0)  # a second context (stream) on the same device
    assert lib.gip_pcg_create(other.ctx, A64.ptr, nall, byref(out)) == hip.EINVAL
    assert b"another context" in lib.gip_last_error()
    other.close()
    # a tolerance not reached within maxit: status ENOCONV, the call itself succeeds
    r = solver.solve(db, dx, z, maxit=5, tol=1e-12)
    assert r.status == hip.ENOCONV and r.iters == 5
    # Jacobi needs every diagonal entry: drop row 0's diagonal
    rp, col, val = L.rowptr, L.col, L.val
    rows = np.repeat(np.arange(rp.size - 1), np.diff(rp))
    keep = ~((rows == 0) & (col == 0))
    rp2 = np.concatenate([[0], np.cumsum(np.bincount(rows[keep], minlength=rp.size - 1))]).astype(np.int32)
    A_nodiag = hip.Matrix(dev, rp2, col[keep], val[keep], ncols=nall)
    s2 = hip.PCG(A_nodiag, nall)
    with pytest.raises(hip.HipError) as e:
        s2.solve(db, dx, z, maxit=5)
    assert e.value.code == hip.EINVAL
    r = s2.solve(db, dx, z, maxit=5, precond=False)  # unpreconditioned CG runs
    assert r.iters == 5


def test_invalid_arguments_are_rejected(dev):
    with pytest.raises(hip.HipError) as e:
        hip.Matrix(dev, np.array([0, 2]), np.array([0, -1]), np.array([1.0, 2.0]))
    assert e.value.code == hip.EINVAL
    with pytest.raises(hip.HipError):
        hip.Matrix(dev, np.array([0, 3]), np.array([0, 1]), np.array([1.0, 2.0]))
