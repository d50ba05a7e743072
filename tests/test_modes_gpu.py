"""GPU parity of ceviche_amd.modes (shift-invert Arnoldi on the dense LU
kernels) against the oracle (scipy eigs on the restated operator), on the
ridge waveguide of ceviche/modes.py's own example (:129-160: lambda/100 grid,
10 lambda wide, eps 4 core one lambda wide, npml 10) and on insert_mode.
Eigenvalues to 1e-9 relative; eigenvectors (unit norm, defined up to a phase)
after aligning the phase, to 1e-7."""
import numpy as np
import pytest

from oracle import modes_ref as M
from oracle.fdfd_ref import C_0

pytestmark = pytest.mark.gpu


def _ridge():
    lam = 1.550e-6
    dL = lam / 100
    omega = 2 * np.pi * C_0 / lam
    Nx = int(lam * 10 / dL)
    eps = np.ones((Nx,))
    pts = np.arange(Nx // 2 - int(lam / dL / 2), Nx // 2 + int(lam / dL / 2))
    eps[pts] = 4.0
    return omega, dL, eps


def _align(v, ref):
    return v * (np.vdot(v, ref) / abs(np.vdot(v, ref)))


@pytest.mark.parametrize('m', [1, 4, 10])
def test_get_modes_matches_oracle(m):
    from ceviche_amd import modes
    omega, dL, eps = _ridge()
    vals, vecs = modes.get_modes(eps, omega, dL, npml=10, m=m)
    rv, rvec = M.get_modes(eps, omega, dL, npml=10, m=m)
    assert vals.shape == rv.shape and vecs.shape == rvec.shape
    assert np.allclose(vals, rv, rtol=1e-9, atol=1e-12)
    A = M.mode_operator(eps, omega, dL, 10)
    for i in range(vals.size):
        assert np.linalg.norm(A @ vecs[:, i] - vals[i] * vecs[:, i]) <= 1e-8
        gap = np.min(np.abs(np.delete(rv, i) - rv[i])) if rv.size > 1 else 1.0
        if gap > 1e-6:   # non-degenerate: the vector is fixed up to a phase
            assert np.linalg.norm(_align(vecs[:, i], rvec[:, i]) - rvec[:, i]) <= 1e-7


def test_insert_mode_matches_oracle():
    from ceviche_amd import modes
    lam = 1.55e-6
    dL = lam / 60
    omega = 2 * np.pi * C_0 / lam
    epsr = np.ones((80, 200))
    epsr[:, 90:110] = 12.25
    for m in (1, 2):
        t = modes.insert_mode(omega, dL, 20, slice(20, 180), epsr, m=m)
        r = M.insert_mode(omega, dL, 20, slice(20, 180), epsr, m=m)
        assert np.count_nonzero(t[:20]) == 0 and np.count_nonzero(t[21:]) == 0
        assert np.linalg.norm(_align(t[20], r[20]) - r[20]) <= 1e-7
