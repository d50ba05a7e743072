"""Pins the modes oracle (oracle/modes_ref.py, a restatement of
ceviche/modes.py) against the closed-form spectrum of the periodic discrete
Laplacian and against a dense eigen-decomposition of the same matrix."""
import numpy as np

from oracle import modes_ref as M
from oracle.fdfd_ref import C_0


def test_periodic_homogeneous_closed_form():
    lam, N, e0 = 1.55e-6, 64, 2.25
    dL = lam / 40
    omega = 2 * np.pi * C_0 / lam
    k0 = omega / C_0
    vals, vecs = M.get_modes(np.full(N, e0), omega, dL, npml=0, m=3)
    q = np.array([0, 1, 1])
    ref = e0 + (2 * np.cos(2 * np.pi * q / N) - 2) / (k0 * dL) ** 2
    assert np.allclose(np.sort(vals.real), np.sort(ref), rtol=1e-12, atol=1e-12)
    assert np.allclose(vals.imag, 0, atol=1e-12)
    # the q = 0 mode is flat
    v0 = vecs[:, np.argmin(np.abs(vals - e0))]
    assert np.allclose(np.abs(v0), 1 / np.sqrt(N), rtol=1e-9)


def test_pml_waveguide_against_dense_eig():
    lam = 1.55e-6
    dL = lam / 50
    omega = 2 * np.pi * C_0 / lam
    N = 160
    eps = np.ones(N)
    eps[70:90] = 6.0
    A = M.mode_operator(eps, omega, dL, npml=12).toarray()
    w = np.linalg.eigvals(A)
    sigma = eps.max()
    ref = w[np.argsort(np.abs(w - sigma))[:4]]
    vals, vecs = M.get_modes(eps, omega, dL, npml=12, m=4, filtering=False)
    assert np.allclose(vals, ref, rtol=1e-10, atol=1e-12)
    for i in range(4):
        assert np.linalg.norm(A @ vecs[:, i] - vals[i] * vecs[:, i]) <= 1e-9
        assert abs(np.linalg.norm(vecs[:, i]) - 1) <= 1e-12


def test_insert_mode_places_profile():
    lam = 1.55e-6
    dL = lam / 50
    omega = 2 * np.pi * C_0 / lam
    epsr = np.ones((40, 120))
    epsr[:, 50:70] = 6.0
    t = M.insert_mode(omega, dL, 10, slice(0, 120), epsr, m=1)
    assert t.shape == epsr.shape and np.count_nonzero(t[11:]) == 0 and np.count_nonzero(t[:10]) == 0
    _, v = M.get_modes(epsr[10, :], omega, dL, 0, m=1, filtering=False)
    assert np.allclose(np.abs(t[10]), np.abs(v[:, 0]))
