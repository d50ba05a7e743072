"""Pins the multi-frequency FDFD oracle (oracle/fdfd_mf_ref.py).  The
reference's only test of fdfd_mf_ez is a gradient self-consistency check
(tests/test_gradients_fdfd_mf.py), so the operator is fixed here by reduction
to the pinned single-frequency oracle and by a spectral solve of the coupled
system that shares no assembly code with the oracle."""
import numpy as np
import pytest

from oracle import fdfd_mf_ref as MF
from oracle import fdfd_ref as R

OMEGA = 2 * np.pi * 200e12
OMEGA_MOD = 2 * np.pi * 2e12


def _rng_fields(rng, M, shape):
    return rng.standard_normal((M,) + shape) + 1j * rng.standard_normal((M,) + shape)


def test_uncoupled_blocks_are_single_frequency_solves():
    rng = np.random.default_rng(1)
    shape, dL, Nsb = (14, 11), 0.2e-6, 2
    eps = 1 + rng.random(shape)
    src = _rng_fields(rng, 2 * Nsb + 1, shape)
    zero = np.zeros((1,) + shape)
    hx, hy, ez = MF.solve_mf(OMEGA, dL, eps, 7 * OMEGA_MOD, zero, zero, Nsb, [0, 0], src)
    W = MF.sidebands(OMEGA, 7 * OMEGA_MOD, Nsb)
    for m in range(2 * Nsb + 1):
        # npml = 0: the operator's only frequency dependence is eps0 W^2
        rhx, rhy, rez = R.solve_ez(W[m], dL, eps, [0, 0], src[m])
        np.testing.assert_allclose(ez[m], rez, rtol=1e-9, atol=1e-9 * np.abs(rez).max())
        np.testing.assert_allclose(hx[m], rhx, rtol=1e-9, atol=1e-9 * np.abs(rhx).max())
        np.testing.assert_allclose(hy[m], rhy, rtol=1e-9, atol=1e-9 * np.abs(rhy).max())


def test_uncoupled_blocks_with_pml_use_the_base_frequency_profiles():
    rng = np.random.default_rng(2)
    shape, dL, Nsb = (24, 20), 0.1e-6, 1
    eps = 1 + 3 * rng.random(shape)
    zero = np.zeros((1,) + shape)
    A = MF.make_A_mf(OMEGA, dL, eps, OMEGA_MOD, zero, zero, Nsb, [5, 4])
    N = eps.size
    W = MF.sidebands(OMEGA, OMEGA_MOD, Nsb)
    A0 = R.make_A_ez(OMEGA, dL, eps, [5, 4])
    for m in range(3):
        blk = A[m * N:(m + 1) * N, m * N:(m + 1) * N]
        ref = A0 + R.EPSILON_0 * (OMEGA ** 2 - W[m] ** 2) * np.diag(eps.ravel())
        np.testing.assert_allclose(blk.toarray(), ref.toarray() if hasattr(ref, 'toarray') else ref,
                                   rtol=1e-12, atol=1e-12 * abs(A0).max())
        for q in range(3):
            if q != m:
                assert A[m * N:(m + 1) * N, q * N:(q + 1) * N].nnz == 0


def _spectral_solve(eps, dL, omega, omega_mod, deltas, phis, Nsb, src):
    """Uniform eps / delta_h / phi_h on a periodic grid: per Fourier mode
    (p, q) the sidebands obey one M x M system (the discrete Laplacian symbol
    4 sin^2(pi p / n) / dL^2 per axis on the diagonal)."""
    M = 2 * Nsb + 1
    nx, ny = src.shape[1:]
    W = omega + np.arange(-Nsb, Nsb + 1) * omega_mod
    S = np.fft.fft2(src, axes=(1, 2))
    sx = 4 * np.sin(np.pi * np.arange(nx) / nx) ** 2 / dL ** 2
    sy = 4 * np.sin(np.pi * np.arange(ny) / ny) ** 2 / dL ** 2
    X = np.zeros_like(S)
    for p in range(nx):
        for q in range(ny):
            T = np.zeros((M, M), dtype=complex)
            for m in range(M):
                T[m, m] = (sx[p] + sy[q]) / R.MU_0 - R.EPSILON_0 * W[m] ** 2 * eps
                for h, (d, ph) in enumerate(zip(deltas, phis)):
                    if m + h + 1 < M:
                        T[m, m + h + 1] = -R.EPSILON_0 / 2 * W[m] ** 2 * d * np.exp(1j * ph)
                    if m - h - 1 >= 0:
                        T[m, m - h - 1] = -R.EPSILON_0 / 2 * W[m] ** 2 * d * np.exp(-1j * ph)
            X[:, p, q] = np.linalg.solve(T, 1j * W * S[:, p, q])
    return np.fft.ifft2(X, axes=(1, 2))


@pytest.mark.parametrize('deltas,phis,Nsb', [([0.4], [0.7], 1), ([0.3, 0.25], [1.1, -2.0], 2),
                                             ([0.5, 0.2, 0.9], [0.3, 0.6, 0.9], 1)])
def test_coupled_system_matches_the_spectral_solve(deltas, phis, Nsb):
    rng = np.random.default_rng(3)
    shape, dL, eps = (12, 10), 0.15e-6, 2.0 + 0.2j
    M = 2 * Nsb + 1
    src = _rng_fields(rng, M, shape)
    delta = np.stack([np.full(shape, d) for d in deltas])
    phi = np.stack([np.full(shape, p) for p in phis])
    _, _, ez = MF.solve_mf(OMEGA, dL, np.full(shape, eps), 5 * OMEGA_MOD, delta, phi, Nsb, [0, 0], src)
    ref = _spectral_solve(eps, dL, OMEGA, 5 * OMEGA_MOD, deltas, phis, Nsb, src)
    np.testing.assert_allclose(ez, ref, rtol=1e-9, atol=1e-9 * np.abs(ref).max())
    # swapping the phase sign of the couplings (a transposed coupling) must be visible
    bad = _spectral_solve(eps, dL, OMEGA, 5 * OMEGA_MOD, deltas, [-p for p in phis], Nsb, src)
    assert np.abs(bad - ez).max() > 1e-3 * np.abs(ez).max()
