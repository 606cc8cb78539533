"""The allocation-free PCG entry point (gip_pcg_solve_noalloc) against the
solver entry point it mirrors (gip_pcg_create + gip_pcg_solve) on the same
operator and right-hand side: the same kernels, so the iterates, iteration
counts and residuals agree bitwise.  Two calls in a row on the shared
per-device workspace, the second on a smaller operator (what the first call
left in the workspace must not leak into the second), in the reference's
fixed-iteration mode and with a tolerance stop; sizes beyond the stated
limit, mixed precision and product timing are rejected."""
import numpy as np
import pytest

from gip_amd import cavity, hip

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = hip.Device(0)
    yield d
    d.close()


def _case(dev, n, seed):
    c = cavity.generate(n, gamma=1e-3, ops=cavity.OP_BITS["L"])
    b = np.zeros(c.n_all)
    b[:c.n_inner] = cavity.rhs_normal(c.n_inner, seed)
    x0 = np.zeros(c.n_all)
    x0[c.n_inner:] = 0.25  # ghost entries are carried over from x0
    return c, hip.Matrix.from_csr(dev, c.op("L")), b, x0


@pytest.mark.parametrize("tol,maxit", [(0.0, 60), (1e-8, 3000)])
def test_no_alloc_solve_equals_the_solver_entry_point(dev, tol, maxit):
    # 64^3 (262,144 rows, 286,720 entries: the largest the scratch holds), then 24^3
    for n, seed in ((64, 1), (24, 2)):
        c, A, b, x0 = _case(dev, n, seed)
        db, dx0 = dev.to_device(b), dev.to_device(x0)
        want_x = dev.array(c.n_all)
        solver = hip.PCG(A, c.n_all)
        want = solver.solve(db, want_x, dx0, maxit=maxit, tol=tol)
        got_x = dev.array(c.n_all)
        got = hip.pcg_solve_noalloc(A, c.n_all, db, got_x, dx0, maxit=maxit, tol=tol)
        print(f"[no-alloc n={n} tol={tol}] iters {got.iters}/{want.iters} relres {got.relres:.6e}/{want.relres:.6e}",
              flush=True)
        assert (got.iters, got.status, got.relres) == (want.iters, want.status, want.relres)
        assert np.array_equal(got_x.download(), want_x.download())
        solver.destroy()


def test_no_alloc_solve_rejects_what_it_cannot_hold(dev):
    c, A, b, x0 = _case(dev, 65, 3)  # 274,625 rows > 262,144
    db, dx = dev.to_device(b), dev.to_device(x0)
    with pytest.raises(hip.HipError) as e:
        hip.pcg_solve_noalloc(A, c.n_all, db, dx, dx, maxit=10)
    assert e.value.code == hip.EINVAL
    c, A, b, x0 = _case(dev, 16, 3)
    db, dx = dev.to_device(b), dev.to_device(x0)
    o = hip.PcgOpts(10, 0.0, 1, hip.PCG_MIXED, 0.1, 16, 0, 0)
    st = hip.PcgStats()
    rc = hip.lib().gip_pcg_solve_noalloc(dev.ctx, A.ptr, c.n_all, db.ptr, dx.ptr, dx.ptr, hip.byref(o),
                                         hip.byref(st))
    assert rc == hip.EINVAL
