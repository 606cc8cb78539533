"""Hand-written adjoint of the time integration (include/gip_adjoint.h)
against central finite differences of its own forward map (dt frozen).

Bar: adjoint directional derivatives match central differences
(eps = 1e-6) to 1e-6 relative on random directions in the space of the lid
profile and the initial velocity field.  The forward map itself is checked
against the oracle's integrator (same operators, same dt sequence)."""
import numpy as np
import pytest

from gip_amd import adjoint, cavity, hip
from oracle import pyoracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = hip.Device(0)
    yield d
    d.close()


def _setup(n, re, steps, ckpt, seed=1):
    rng = np.random.default_rng(seed)
    lid = 1.0 + 0.05 * rng.standard_normal(n * n)
    cav = cavity.generate(n, gamma=1.0 / re, lid_profile=lid)
    return cav, lid


def test_adjoint_forward_matches_oracle(dev, tmp_path):
    cav, lid = _setup(10, 400, 6, 2)
    cav.write(tmp_path, "adj10")
    a = adjoint.Adjoint(dev, cav, steps=6, checkpoint_every=2, gamma=1 / 400, pcg_tol=1e-12)
    z = np.zeros(cav.n_inner)
    J = a.forward(lid, z, z, z)
    orc = pyoracle.Dns(tmp_path, "adj10", gamma=1 / 400, pcg_maxit=20000, pcg_tol=1e-12)
    for dt in a.dts():
        s = orc.step()
        assert abs(s["dt"] - dt) <= 1e-12 * dt
    u, v, w = orc.field("u"), orc.field("v"), orc.field("w")
    vol = cav.field("vol")
    n = cav.n_inner
    Jo = 0.5 * float(np.sum(vol * (u[:n] ** 2 + v[:n] ** 2 + w[:n] ** 2)))
    assert abs(J - Jo) <= 1e-9 * Jo, (J, Jo)


@pytest.mark.parametrize("n,steps,ckpt", [(8, 5, 2), (12, 12, 4)])
def test_adjoint_gradient_matches_central_differences(dev, n, steps, ckpt):
    cav, lid = _setup(n, 400, steps, ckpt)
    a = adjoint.Adjoint(dev, cav, steps=steps, checkpoint_every=ckpt, gamma=1 / 400, pcg_tol=1e-13)
    rng = np.random.default_rng(7)
    u0, v0, w0 = (0.1 * rng.standard_normal(cav.n_inner) for _ in range(3))
    J0, ad, fd, rel = adjoint.fd_check(a, lid, u0, v0, w0, ndir=4, eps=1e-6)
    print(f"[adjoint n={n} steps={steps}] J={J0:.6e} adjoint={ad} fd={fd} rel={rel}")
    assert np.all(rel <= 1e-6), rel


def test_adjoint_checkpoint_interval_does_not_change_the_gradient(dev):
    cav, lid = _setup(8, 400, 6, 1)
    z = 0.05 * np.ones(cav.n_inner)
    grads = []
    for ck in (1, 3, 6):
        a = adjoint.Adjoint(dev, cav, steps=6, checkpoint_every=ck, gamma=1 / 400, pcg_tol=1e-13)
        a.forward(lid, z, z, z)
        grads.append(a.backward())
        a.close()
    for g in grads[1:]:
        for x, y in zip(grads[0], g):
            assert np.array_equal(x, y)
