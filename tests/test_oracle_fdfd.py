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


def test_hz_grid_average_hand_values():
    # ceviche/fdfd.py:200-209: eps_xx averages each cell with its neighbour at
    # y - 1 (axis 1), eps_yy with its neighbour at x - 1 (axis 0), periodically.
    eps = np.array([[1.0, 2.0, 3.0],
                    [4.0, 5.0, 6.0]])
    exx, eyy = R.grid_average_hz(eps)
    # by hand: exx[0, 0] = (1 + 3) / 2 (wraps to y = 2), exx[0, 1] = (2 + 1) / 2, ...
    assert np.array_equal(exx, np.array([[2.0, 1.5, 2.5],
                                         [5.0, 4.5, 5.5]]))
    # eyy[0, j] = (eps[0, j] + eps[1, j]) / 2 (wraps to x = 1), eyy[1, j] the same pair
    assert np.array_equal(eyy, np.array([[2.5, 3.5, 4.5],
                                         [2.5, 3.5, 4.5]]))


def _hz_layered_1d(omega, dL, e):
    """Hz operator of a medium layered along one axis, for fields constant
    along the other, written out entry by entry: the forward difference
    h[i+1] - h[i] is divided by the average of cells i and i-1 (ceviche's
    index convention, fdfd.py:200-215: the average stored at index i is
    formed with the neighbour at i-1), then differenced backward; periodic,
    no PML."""
    n = e.size
    A = np.zeros((n, n), dtype=complex)
    avg = np.array([(e[i] + e[i - 1]) / 2 + 1e-5 for i in range(n)])
    c = 1 / R.EPSILON_0 / dL ** 2
    for i in range(n):
        ip, im = (i + 1) % n, (i - 1) % n
        # forward difference at i, scaled by 1/avg[i]; backward difference of that flux
        A[i, ip] += c / avg[i]
        A[i, i] -= c / avg[i]
        A[i, i] -= c / avg[im]
        A[i, im] += c / avg[im]
        A[i, i] += R.MU_0 * omega ** 2
    return A


@pytest.mark.parametrize('axis', [0, 1])
def test_hz_operator_pairs_each_average_with_the_right_derivative(axis):
    # eps layered along x (axis 0) only acts through eps_yy and Dx; layered
    # along y only through eps_xx and Dy.  A y-uniform (x-uniform) field sees
    # exactly the 1D operator; a swapped pairing or averaging axis fails.
    omega, dL = 2 * np.pi * R.C_0 / 1.55e-6, 50e-9
    rng = np.random.default_rng(3)
    n, m = 11, 7
    prof = rng.uniform(1, 12.25, n)
    eps = np.repeat(prof[:, None], m, axis=1) if axis == 0 else np.repeat(prof[None, :], m, axis=0)
    A = R.make_A_hz(omega, dL, eps, [0, 0])
    u = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    field = np.repeat(u[:, None], m, axis=1) if axis == 0 else np.repeat(u[None, :], m, axis=0)
    got = (A @ field.ravel()).reshape(eps.shape)
    want = _hz_layered_1d(omega, dL, prof) @ u
    want = np.repeat(want[:, None], m, axis=1) if axis == 0 else np.repeat(want[None, :], m, axis=0)
    assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want)


def test_hz_regulariser_closed_form():
    # eps = 0 everywhere: 1 / (eps + 1e-5) = 1e5 exactly, so a DFT mode of the
    # periodic grid (no PML) has the symbol
    #   -(1e5 / eps0) (4 / dL^2) (sin^2(kx dL / 2) + sin^2(ky dL / 2)) + mu0 w^2.
    omega, dL = 2 * np.pi * R.C_0 / 1.55e-6, 50e-9
    nx, ny = 8, 6
    A = R.make_A_hz(omega, dL, np.zeros((nx, ny)), [0, 0])
    for p, q in ((1, 2), (3, 0), (0, 5)):
        kx, ky = 2 * np.pi * p / (nx * dL), 2 * np.pi * q / (ny * dL)
        X, Y = np.meshgrid(np.arange(nx) * dL, np.arange(ny) * dL, indexing='ij')
        v = np.exp(1j * (kx * X + ky * Y)).ravel()
        lam = -(1e5 / R.EPSILON_0) * (4 / dL ** 2) * (np.sin(kx * dL / 2) ** 2 + np.sin(ky * dL / 2) ** 2) \
            + R.MU_0 * omega ** 2
        assert np.linalg.norm(A @ v - lam * v) <= 1e-10 * abs(lam) * np.linalg.norm(v)


def test_hz_field_recovery_is_faraday_law():
    # Ex, Ey from Hz (fdfd.py:150-162) with A Hz = i w Mz satisfy Faraday's law
    # Dxb Ey - Dyb Ex = -i w mu0 Hz - Mz, and Ampere's law per component:
    # i w eps0 (eps_xx + 1e-5) Ex = Dyf Hz, i w eps0 (eps_yy + 1e-5) Ey = -Dxf Hz.
    omega = 2 * np.pi * R.C_0 / 1.55e-6
    rng = np.random.default_rng(4)
    shape = (28, 23)
    eps = rng.uniform(1, 6, shape)
    src = np.zeros(shape)
    src[14, 11] = 1
    ex, ey, hz = R.solve_hz(omega, 50e-9, eps, [6, 6], src)
    Dxf, Dxb, Dyf, Dyb = R.derivative_matrices(omega, shape, [6, 6], 50e-9)
    lhs = Dxb @ ey.ravel() - Dyb @ ex.ravel()
    rhs = -1j * omega * R.MU_0 * hz.ravel() - src.ravel()
    assert np.linalg.norm(lhs - rhs) <= 1e-9 * np.linalg.norm(rhs)
    exx = 0.5 * (eps + np.roll(eps, 1, axis=1))
    eyy = 0.5 * (eps + np.roll(eps, 1, axis=0))
    assert np.allclose(1j * omega * R.EPSILON_0 * (exx.ravel() + 1e-5) * ex.ravel(), Dyf @ hz.ravel(),
                       rtol=1e-12, atol=1e-12 * np.abs(Dyf @ hz.ravel()).max())
    assert np.allclose(1j * omega * R.EPSILON_0 * (eyy.ravel() + 1e-5) * ey.ravel(), -(Dxf @ hz.ravel()),
                       rtol=1e-12, atol=1e-12 * np.abs(Dxf @ hz.ravel()).max())
