"""Pins the FDFD oracle (oracle/fdfd_ref.py) against closed forms, invariants
and finite differences.  The reference ships no golden field values
(its field tests only plot), so these checks are what fix the oracle."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import fdfd_ref as R
from oracle.numdiff import der_num


def test_constants_are_the_truncated_reference_values():
    # ceviche/constants.py:7-11
    assert R.EPSILON_0 == 8.85418782e-12
    assert R.MU_0 == 1.25663706e-6
    assert abs(R.C_0 - 299792458.13099605) < 1e-6
    assert abs(R.ETA_0 - 376.7303131959) < 1e-9   # sqrt(mu0/eps0) of the truncated constants


def test_sfactor_profile_shape_and_limits():
    omega = 2 * np.pi * R.C_0 / 1.55e-6
    sf = R.sfactor('f', omega, 50e-9, 120, 20)
    sb = R.sfactor('b', omega, 50e-9, 120, 20)
    # interior untouched; the asymmetric index ranges of derivatives.py:193-206
    assert np.all(sf[21:100] == 1) and np.all(sb[21:100] == 1)
    assert np.all(sf[:21] != 1) and np.all(sf[101:] != 1) and sf[100] == 1
    # cubic grading: -Im(s) ~ (l/dw)^3, largest at the outer edge
    assert np.all(np.diff(-sf[:21].imag) < 0) and np.all(np.diff(-sf[101:].imag) > 0)
    assert np.all(sf.real == 1)
    # i = 0 sits at l = (npml + 0.5) dL > dw (quirk B1), so |s-1| exceeds sigma_max's
    dw = 20 * 50e-9
    smax = 60 / (R.ETA_0 * dw) / (omega * R.EPSILON_0)
    assert -sf[0].imag == pytest.approx(smax * (20.5 / 20) ** 3, rel=1e-12)
    assert -sb[0].imag == pytest.approx(smax * (21 / 20) ** 3, rel=1e-12)
    assert np.all(R.sfactor('f', omega, 50e-9, 30, 0) == 1)


def test_periodic_vacuum_operator_has_the_dft_symbol():
    # eps = 1, npml = 0: A is circulant; e^{i(kx x + ky y)} are eigenvectors with
    # eigenvalue (1/mu0)(4 sin^2(kx dL/2) + 4 sin^2(ky dL/2))/dL^2 - eps0 w^2.
    nx, ny, dL = 12, 10, 1e-7
    omega = 2 * np.pi * 2e14
    A = R.make_A_ez(omega, dL, np.ones((nx, ny)), [0, 0])
    for mx, my in [(0, 0), (1, 3), (5, 7), (11, 9)]:
        kx, ky = 2 * np.pi * mx / nx, 2 * np.pi * my / ny
        v = np.exp(1j * (kx * np.arange(nx)[:, None] + ky * np.arange(ny)[None, :])).ravel()
        lam = (4 * np.sin(kx / 2) ** 2 + 4 * np.sin(ky / 2) ** 2) / dL ** 2 / R.MU_0 - R.EPSILON_0 * omega ** 2
        assert np.linalg.norm(A @ v - lam * v) <= 1e-12 * np.linalg.norm(lam * v) + 1e-6


def test_bloch_operator_eigenvector():
    # With Bloch phase theta_x, e^{i (2 pi m + theta) x / nx} are eigenvectors.
    nx, ny, dL, th = 9, 7, 1e-7, 0.7
    omega = 2 * np.pi * 2e14
    A = R.make_A_ez(omega, dL, np.ones((nx, ny)), [0, 0], bloch_phases=(th, 0.0))
    kx = (2 * np.pi * 2 + th) / nx
    v = np.exp(1j * kx * np.arange(nx))[:, None] * np.ones(ny)[None, :]
    lam = 4 * np.sin(kx / 2) ** 2 / dL ** 2 / R.MU_0 - R.EPSILON_0 * omega ** 2
    assert np.linalg.norm(A @ v.ravel() - lam * v.ravel()) <= 1e-12 * np.linalg.norm(lam * v)


@pytest.mark.parametrize('maker', [R.make_A_ez, R.make_A_hz])
def test_scaled_operator_is_complex_symmetric(maker):
    # W A with W = diag(s_xf s_yf) (Ez) / diag(s_xb s_yb) (Hz) is complex symmetric
    # when the Bloch phases vanish.
    omega = 2 * np.pi * R.C_0 / 1.55e-6
    rng = np.random.default_rng(3)
    eps = rng.uniform(1, 5, (24, 20))
    A = maker(omega, 50e-9, eps, [5, 4])
    d = 'f' if maker is R.make_A_ez else 'b'
    w = np.outer(R.sfactor(d, omega, 50e-9, 24, 5), R.sfactor(d, omega, 50e-9, 20, 4)).ravel()
    S = sp.diags(w) @ A
    assert abs(S - S.T).max() <= 1e-12 * abs(S).max()
    # and not symmetric without the scaling (a dropped PML factor would hide here)
    assert abs(A - A.T).max() > 1e-6 * abs(A).max()


def test_operator_nnz_structure():
    A = R.make_A_ez(1e15, 50e-9, np.ones((10, 12)), [2, 2])
    assert A.nnz == 5 * 120
    assert np.all(A.diagonal() != 0)


def test_point_source_matches_hankel_green_function():
    lam0, dL, N = 1.55e-6, 20e-9, 200
    omega = 2 * np.pi * R.C_0 / lam0
    J = np.zeros((N, N))
    c = N // 2
    J[c, c] = 1.0
    _, _, ez = R.solve_ez(omega, dL, np.ones((N, N)), [30, 30], J)
    for d in (20, 40, 60):
        ana = R.hankel_point_source_ez(omega, dL, d * dL)
        assert abs(ez[c + d, c] - ana) / abs(ana) < 3e-3
        assert abs(ez[c, c + d] - ana) / abs(ana) < 3e-3


def test_plane_wave_wavelength():
    # ceviche tests/test_plane_wave.py:14-27: lambda/dL = 100 cells expected
    lam0, dL = 1.5e-6, 1.5e-8
    omega = 2 * np.pi * R.C_0 / lam0
    nx, ny = 400, 10
    src = np.zeros((nx, ny))
    src[nx // 2, :] = 1e-8
    for solver in (R.solve_ez, R.solve_hz):
        f = solver(omega, dL, np.ones((nx, ny)), [20, 0], src)[2]
        line = f[nx // 2 + 10:nx - 40, 5]
        k = np.polyfit(np.arange(line.size), np.unwrap(np.angle(line)), 1)[0]
        assert 2 * np.pi / abs(k) == pytest.approx(100.0, rel=1e-3)
        # PML absorbs: the travelling wave has a flat envelope (no standing wave)
        a = np.abs(line)
        assert a.max() / a.min() < 1.001


def test_field_recovery_is_faraday_law():
    # Hx, Hy from Ez satisfy the discrete curl: i w mu0 H = curl E (fdfd.py:138-142)
    omega = 2 * np.pi * R.C_0 / 1.55e-6
    rng = np.random.default_rng(1)
    eps = rng.uniform(1, 4, (30, 26))
    src = np.zeros((30, 26))
    src[15, 13] = 1
    hx, hy, ez = R.solve_ez(omega, 50e-9, eps, [6, 6], src)
    Dxf, Dxb, Dyf, Dyb = R.derivative_matrices(omega, (30, 26), [6, 6], 50e-9)
    # Faraday (field recovery): Dxf Hy - Dyf Hx = (1/(i w mu0)) (Dxf Dxb + Dyf Dyb) Ez
    lhs = Dxf @ hy.ravel() - Dyf @ hx.ravel()
    rhs = (1 / (1j * omega * R.MU_0)) * ((Dxf @ Dxb + Dyf @ Dyb) @ ez.ravel())
    assert np.linalg.norm(lhs - rhs) <= 1e-10 * np.linalg.norm(rhs)
    # Ampere from A Ez = i w Jz, A = -(1/mu0) DD - eps0 w^2 eps:  curl H = i w eps0 eps Ez - Jz
    amp = 1j * omega * R.EPSILON_0 * eps.ravel() * ez.ravel() - src.ravel()
    assert np.linalg.norm(lhs - amp) <= 1e-8 * np.linalg.norm(amp)


def test_ez_adjoint_gradient_matches_central_differences():
    # the set-up of ceviche tests/test_gradients_fdfd.py:34-59 (30x30, dL = 1 um,
    # omega = 2 pi 200 THz, pml 10, eps in [1, 2)), objective |Ez[p]|^2
    rng = np.random.default_rng(7)
    omega, dL, npml = 2 * np.pi * 200e12, 1e-6, [10, 10]
    eps = rng.random((30, 30)) + 1
    src = np.zeros((30, 30))
    src[15, 15] = 1
    p = (20, 11)

    def obj(e):
        return abs(R.solve_ez(omega, dL, e, npml, src)[2][p]) ** 2

    _, _, ez = R.solve_ez(omega, dL, eps, npml, src)
    g = np.zeros((30, 30), dtype=complex)
    g[p] = 2 * ez[p]  # torch-convention cotangent of |Ez_p|^2
    grad, _ = R.grad_eps_ez(omega, dL, eps, npml, src, g)
    idx = [(20, 11), (15, 15), (3, 4), (25, 27), (12, 18)]
    num = np.array([der_num(lambda e: obj(e), eps.copy(), i, 1e-4) for i in idx])
    ana = np.array([grad[i] for i in idx])
    assert np.linalg.norm(num - ana) <= 1e-4 * np.linalg.norm(num)
