"""The symmetric form of a stencil operator's slot values in the FP64 PCG's
product (include/gip_hip.h, GIP_MAT_NO_SYM; gip_layout.h, plan_symmetric):
a slot of offset -d reads row r - d's slot of offset +d, so the product
must give the full layout's bits.  Checked against the same operator built
without it (GIP_MAT_NO_SYM): FP64 iterates bitwise on both long iteration
forms (more and fewer than 2^23 rows) and for 7- and 8-slot layouts (128^3
and 96^3, whose boundary slices pick an eighth offset), a tolerance stop,
the clustered mesh, the mixed solve untouched (its FP32 product keeps all
slots); after a change of
values the form is made again (scaled values) or refused (one value off
by one unit in the last place), and the solve still equals the full one.
"""
import numpy as np
import pytest

from gip_amd import cavity, hip

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = hip.Device(0)
    yield d
    d.close()


def _case(nx, ny=None, nz=None, beta=0.0):
    c = cavity.generate(nx, ny, nz, beta=beta, gamma=1e-3, ops=cavity.OP_BITS["L"])
    b = np.zeros(c.n_all)
    b[:c.n_inner] = cavity.rhs_normal(c.n_inner, 11)
    return c, c.op("L"), b


def _solve(dev, A, c, b, **kw):
    solver = hip.PCG(A, c.n_all)
    dx = dev.array(c.n_all)
    r = solver.solve(dev.to_device(b), dx, dev.to_device(np.zeros(c.n_all)), **kw)
    x = dx.download()
    solver.destroy()
    return r, x


def _pair(dev, L, **kw):
    A = hip.Matrix.from_csr(dev, L, **kw)
    F = hip.Matrix.from_csr(dev, L, sym=False, **kw)
    assert A.info()["format"] == hip.FMT_STENCIL and F.info()["format"] == hip.FMT_STENCIL
    return A, F


@pytest.mark.parametrize("shape", [(96, 96, 96), (128, 128, 128), (256, 256, 129)],
                         ids=["96^3", "128^3", "256x256x129"])
def test_fp64_fixed_iterations_bitwise(dev, shape):
    c, L, b = _case(*shape)
    print(f"[sym fp64 {shape}] rows {c.n_inner}", flush=True)
    A, F = _pair(dev, L)
    assert A.sym_state()[0] == 2 and F.sym_state()[0] == 0
    ra, xa = _solve(dev, A, c, b, maxit=40)
    rf, xf = _solve(dev, F, c, b, maxit=40)
    st, b64 = A.sym_state()
    print(f"[sym fp64 {shape}] state {st} bytes {b64} (full layout {F.product_bytes()}) relres {ra.relres:.6e}",
          flush=True)
    assert st == 1 and F.sym_state()[0] == 0
    assert 0 < b64 < 0.8 * F.product_bytes()
    assert ra.iters == rf.iters == 40 and ra.relres == rf.relres
    assert np.array_equal(xa, xf)
    A.destroy()
    F.destroy()


def test_fp64_tolerance_stop_bitwise_and_mixed_untouched(dev):
    """A tolerance-stopped FP64 solve equals the full one bitwise; the mixed
    solve (whose FP32 product keeps all slots) neither makes nor needs the
    form and equals the full operator's mixed solve bitwise."""
    c, L, b = _case(128)
    A, F = _pair(dev, L, fp32=True)
    kw = dict(maxit=20000, tol=1e-10, precision=hip.PCG_MIXED, poll=32)
    ra, xa = _solve(dev, A, c, b, **kw)
    rf, xf = _solve(dev, F, c, b, **kw)
    print(f"[sym mixed 128^3] iters {ra.iters}/{rf.iters} corrections {ra.corrections}/{rf.corrections}", flush=True)
    assert A.sym_state()[0] == 2
    assert ra.status == rf.status == 0
    assert (ra.iters, ra.corrections) == (rf.iters, rf.corrections)
    assert np.array_equal(xa, xf)
    kw = dict(maxit=5000, tol=1e-6)
    ra, xa = _solve(dev, A, c, b, **kw)
    rf, xf = _solve(dev, F, c, b, **kw)
    print(f"[sym fp64 128^3 tol 1e-6] iters {ra.iters}/{rf.iters}", flush=True)
    assert A.sym_state()[0] == 1
    assert ra.status == rf.status == 0 and ra.iters == rf.iters
    assert np.array_equal(xa, xf)
    A.destroy()
    F.destroy()


def test_clustered_mesh_bitwise(dev):
    c, L, b = _case(128, beta=1.5)
    A, F = _pair(dev, L)
    ra, xa = _solve(dev, A, c, b, maxit=30)
    rf, xf = _solve(dev, F, c, b, maxit=30)
    assert A.sym_state()[0] == 1
    assert np.array_equal(xa, xf)
    A.destroy()
    F.destroy()


def test_changed_values_remake_or_refuse_the_form(dev):
    c, L, b = _case(96, 96, 80)
    A = hip.Matrix.from_csr(dev, L)
    _solve(dev, A, c, b, maxit=5)
    assert A.sym_state()[0] == 1
    # scaled values stay symmetric: the form is made again at the next solve
    v2 = L.val * 2.0
    A.set_values(v2)
    assert A.sym_state()[0] == 2
    F = hip.Matrix(dev, L.rowptr, L.col, v2, nrows=L.nrows, sym=False)
    ra, xa = _solve(dev, A, c, b, maxit=25)
    rf, xf = _solve(dev, F, c, b, maxit=25)
    assert A.sym_state()[0] == 1
    assert np.array_equal(xa, xf)
    F.destroy()
    # one off-diagonal value of an interior row one unit in the last place
    # lower: no longer symmetric, the solve falls back to the full values
    v3 = v2.copy()
    r0 = 96 * 96 * 40 + 96 * 48 + 48
    j = [j for j in range(L.rowptr[r0], L.rowptr[r0 + 1]) if L.col[j] == r0 + 96][0]
    v3[j] = np.nextafter(v3[j], 0.0)
    A.set_values(v3)
    F = hip.Matrix(dev, L.rowptr, L.col, v3, nrows=L.nrows, sym=False)
    ra, xa = _solve(dev, A, c, b, maxit=25)
    rf, xf = _solve(dev, F, c, b, maxit=25)
    print(f"[sym values changed] state {A.sym_state()[0]}", flush=True)
    assert A.sym_state()[0] == -1
    assert np.array_equal(xa, xf)
    A.destroy()
    F.destroy()
