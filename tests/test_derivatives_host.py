"""The reference's sparse derivative-matrix and curl helpers
(ceviche/derivatives.py:16-216) as exported by ceviche_amd.derivatives:
against the oracle's restatement, a plane-wave eigenvalue and the discrete
div-curl identity.  Host-only (no GPU)."""
import numpy as np
import pytest
import torch

from ceviche_amd import derivatives as D
from oracle import fdfd_ref as R

OMEGA = 2 * np.pi * 299792458.0 / 1.55e-6


@pytest.mark.parametrize('shape,npml,bloch', [((12, 9), (3, 2), (0.0, 0.0)), ((7, 11), (2, 4), (0.4, -0.7)),
                                              ((1, 10), (0, 3), (0.0, 0.0)), ((10, 1), (3, 0), (0.2, 0.0))])
def test_derivative_matrices_match_the_oracle(shape, npml, bloch):
    ours = D.compute_derivative_matrices(OMEGA, shape, npml, 50e-9, *bloch)
    ref = R.derivative_matrices(OMEGA, shape, npml, 50e-9, *bloch)
    for a, b in zip(ours, ref):
        assert a.shape == b.shape
        assert abs(a - b).max() <= 1e-15 * abs(b).max()


def test_forward_difference_plane_wave_eigenvalue():
    """Without PML, with the Bloch phase of one period, exp(i k x) is an
    eigenvector of the forward x difference with eigenvalue (e^{i k dL} - 1) / dL,
    and of the backward one with (1 - e^{-i k dL}) / dL."""
    Nx, Ny, dL = 16, 5, 0.1
    k = 2 * np.pi * 3.3 / (Nx * dL)   # not a lattice wave number: the phase closes the period
    x = np.arange(Nx) * dL
    u = np.outer(np.exp(1j * k * x), np.ones(Ny)).ravel()
    bloch = k * Nx * dL
    Dxf = D.createDws('x', 'f', (Nx, Ny), dL, bloch_x=bloch)
    Dxb = D.createDws('x', 'b', (Nx, Ny), dL, bloch_x=bloch)
    assert np.allclose(Dxf @ u, (np.exp(1j * k * dL) - 1) / dL * u, rtol=0, atol=1e-12)
    assert np.allclose(Dxb @ u, (1 - np.exp(-1j * k * dL)) / dL * u, rtol=0, atol=1e-12)
    with pytest.raises(ValueError):
        D.createDws('z', 'f', (Nx, Ny), dL)


def test_sfactor_profiles_and_s_matrices():
    for d, f in (('f', D.create_sfactor_f), ('b', D.create_sfactor_b)):
        assert np.array_equal(D.create_sfactor(d, OMEGA, 50e-9, 30, 6), f(OMEGA, 50e-9, 30, 6, 6 * 50e-9))
    Sxf, Sxb, Syf, Syb = D.create_S_matrices(OMEGA, (5, 4), (2, 1), 50e-9)
    sx = D.create_sfactor('f', OMEGA, 50e-9, 5, 2)
    sy = D.create_sfactor('b', OMEGA, 50e-9, 4, 1)
    assert np.allclose(Sxf.diagonal().reshape(5, 4), (1 / sx)[:, None] * np.ones((1, 4)))
    assert np.allclose(Syb.diagonal().reshape(5, 4), np.ones((5, 1)) * (1 / sy)[None, :])


def test_curl_identities_numpy_and_torch():
    """The forward-difference divergence of the forward-difference curl of E
    vanishes on the periodic grid (forward differences along different axes
    commute), and the backward divergence of the backward curl of H; numpy
    and torch inputs give the same fields."""
    rng = np.random.default_rng(9)
    dL = 0.2
    E = [rng.standard_normal((6, 7, 5)) for _ in range(3)]
    H = [rng.standard_normal((6, 7, 5)) for _ in range(3)]
    cE = [D.curl_E(a, *E, dL) for a in range(3)]
    cH = [D.curl_H(a, *H, dL) for a in range(3)]
    div_f = sum((np.roll(cE[a], -1, axis=a) - cE[a]) / dL for a in range(3))
    div_b = sum((cH[a] - np.roll(cH[a], 1, axis=a)) / dL for a in range(3))
    assert np.abs(div_f).max() < 1e-11 and np.abs(div_b).max() < 1e-11
    # and the curl itself is not trivially zero
    assert np.abs(cE[0]).max() > 1.0
    Et = [torch.from_numpy(e) for e in E]
    for a in range(3):
        assert np.allclose(D.curl_E(a, *Et, dL).numpy(), cE[a], rtol=0, atol=1e-14)
    assert D.curl_E(3, *E, dL) is None
