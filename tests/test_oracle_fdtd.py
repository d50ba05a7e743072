"""Pins the FDTD oracle (oracle/fdtd_ref.py) with closed forms and invariants."""
import numpy as np
import pytest

from oracle import fdtd_ref as T
from oracle.fdfd_ref import C_0


def test_time_step_is_the_3d_courant_limit():
    dL = 20e-9
    assert T.time_step(dL) == pytest.approx(0.5 * dL / (np.sqrt(3) * C_0), rel=1e-15)
    assert T.time_step(20e-9) == pytest.approx(1.925833e-17, rel=1e-6)


def test_sigma_profile_index_map():
    # fdtd.py:235-239: low side indices 2p+1 .. 2 get values for n = 0..2p-1;
    # indices 0, 1 stay zero; high side 2N-2p .. 2N-1.
    dt = T.time_step(50e-9)
    sH, sD = T.sigmas((10, 8, 1), [2, 3, 0], dt)
    S = 0.5 * T.EPSILON_0 / dt
    # sigma_Hx at i: 2x index 2i+1; sigma_Dx at i: 2x index 2i
    hx = sH[0][:, 0, 0]
    dx = sD[0][:, 0, 0]
    n2 = np.zeros(20)
    for n in range(4):
        n2[2 * 2 - n + 1] = S * (n / 4) ** 3
        n2[20 - 4 + n] = S * (n / 4) ** 3
    assert np.allclose(hx, n2[1::2], rtol=0, atol=0)
    assert np.allclose(dx, n2[0::2], rtol=0, atol=0)
    # quirk B6: 2x indices 0 and 1 stay zero, so the outermost cell has no loss
    assert dx[0] == 0 and hx[0] == 0 and dx[1] != 0 and hx[1] != 0
    assert np.all(sH[2] == 0) and np.all(sD[2] == 0)


def test_yee_dispersion_recurrence():
    # Periodic vacuum (npml = 0): a standing mode cos(kx i) cos(ky j) is an
    # eigenmode; its time samples obey x[n+1] + x[n-1] = 2 cos(theta) x[n] with
    # sin^2(theta/2) = (c dt / dL)^2 (sin^2(kx dL/2) + sin^2(ky dL/2)) / dL^2 * dL^2.
    nx, ny, dL = 16, 12, 50e-9
    F = T.FDTD(np.ones((nx, ny)), dL, [0, 0, 0])
    kx, ky = 2 * np.pi * 3 / nx, 2 * np.pi * 2 / ny
    mode = np.cos(kx * np.arange(nx))[:, None] * np.cos(ky * np.arange(ny))[None, :]
    F.D[2] = mode[:, :, None].copy()
    F.E[2] = F.mE[2] * F.D[2]
    xs = []
    for _ in range(40):
        xs.append(F.forward()['Ez'][0, 0, 0])
    xs = np.array(xs)
    r = F.dt * C_0 / dL
    s2 = r ** 2 * (np.sin(kx / 2) ** 2 + np.sin(ky / 2) ** 2)
    cos_t = 1 - 2 * s2
    resid = xs[2:] + xs[:-2] - 2 * cos_t * xs[1:-1]
    assert np.max(np.abs(resid)) <= 1e-12 * np.max(np.abs(xs))
    # and the TE / z-derivative components stay exactly zero
    f = F.fields()
    for k in ('Ex', 'Ey', 'Hz', 'Dx', 'Dy'):
        assert np.all(f[k] == 0)


def _pulse(t, t0=40, s=10):
    return np.exp(-(t - t0) ** 2 / (2 * s ** 2))


def _run(eps, J_at, steps, npml=(4, 4, 0), comp='Jz', scale=1.0):
    F = T.FDTD(eps, 50e-9, list(npml))
    J = np.zeros(F.shape)
    out = []
    for t in range(steps):
        J[J_at] = scale * _pulse(t)
        out.append(F.forward(**{comp: J.copy()}))
    return F, out


def test_linearity_in_the_source():
    rng = np.random.default_rng(0)
    eps = rng.uniform(1, 4, (14, 12, 1))
    _, a = _run(eps, (7, 6, 0), 60, scale=1.0)
    _, b = _run(eps, (7, 6, 0), 60, scale=-2.5)
    for k in ('Ez', 'Hx', 'Hy', 'Dz'):
        assert np.allclose(b[-1][k], -2.5 * a[-1][k], rtol=1e-12, atol=1e-30)


def test_mirror_symmetry():
    rng = np.random.default_rng(1)
    half = rng.uniform(1, 4, (9, 16, 1))
    eps = np.concatenate([half, half[::-1]], axis=0)  # symmetric about the x-midplane (18 cells)
    # eps_xx averages along x with roll(+1): use a y-line source so that the
    # symmetry is exact for Ez under x -> 17 - x (cells), H_y odd
    F = T.FDTD(eps, 50e-9, [0, 4, 0])
    J = np.zeros(F.shape)
    for t in range(50):
        J[8, 8, 0] = J[9, 8, 0] = _pulse(t)
        f = F.forward(Jz=J.copy())
    ez = f['Ez'][:, :, 0]
    assert np.allclose(ez, ez[::-1, :], rtol=1e-11, atol=1e-12 * np.abs(ez).max())


def test_tm_drive_keeps_te_components_zero():
    rng = np.random.default_rng(2)
    eps = rng.uniform(1, 4, (10, 10, 1))
    F, out = _run(eps, (5, 5, 0), 30)
    for k in ('Ex', 'Ey', 'Hz', 'Dx', 'Dy'):
        assert np.all(out[-1][k] == 0)
    assert np.abs(out[-1]['Ez']).max() > 0


def test_pml_absorbs_the_pulse():
    eps = np.ones((60, 60, 1))
    F, out = _run(eps, (30, 30, 0), 700, npml=(15, 15, 0))
    peak = max(np.abs(o['Ez']).max() for o in out[:150])
    assert np.abs(out[-1]['Ez']).max() < 2e-2 * peak
    assert np.all(np.isfinite(out[-1]['Ez']))


def test_complex_step_matches_central_differences_on_reference_setup():
    # ceviche tests/test_gradients_fdtd.py:28-88 (8x8x1, dL = 50 nm, pml [2,2,0],
    # Gaussian t0 = 300, sigma = 20; here 400 steps), objective sum_t sum(Ex+Ey+Ez)
    rng = np.random.default_rng(3)
    eps = rng.random((8, 8, 1)) + 1

    def obj(e, dtype=np.float64):
        F = T.FDTD(e, 5e-8, [2, 2, 0], dtype=dtype)
        J = np.zeros((8, 8, 1), dtype=dtype)
        S = 0.0
        for t in range(400):
            J[4, 4, 0] = np.exp(-(t - 300) ** 2 / 2 / 20 ** 2)
            f = F.forward(Jz=J)
            S = S + np.sum(f['Ex'] + f['Ey'] + f['Ez'])
        return S

    h = 1e-20
    for idx in [(4, 4, 0), (1, 6, 0), (7, 0, 0)]:
        d = np.zeros_like(eps)
        d[idx] = 1.0
        cs = np.imag(obj(eps + 1j * h * d, dtype=np.complex128)) / h
        fd = (obj(eps + 1e-6 * d) - obj(eps - 1e-6 * d)) / 2e-6
        assert cs == pytest.approx(fd, rel=1e-5, abs=1e-8 * abs(cs) + 1e-30)


def test_tm_400_golden_fixtures_are_the_oracle_run():
    """The committed 400x400 TM fixtures: the first 300 steps of the trace
    recomputed by the oracle, and the adjoint fixture's probe energy equal to
    the sum of squares of the committed trace."""
    import os
    from ceviche_amd.workloads import fdtd_tm_blobs
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')
    tr = np.load(os.path.join(gold, 'tm_blobs_400_trace.npy'))
    c = fdtd_tm_blobs()
    assert tr.shape == (c['steps'], 4)
    head = T.run_probes(c['eps'], c['dL'], c['npml'], c['src'], lambda t: c['amp'][t], 300,
                        [p for _, p in c['probes']], component='Ez')
    assert np.array_equal(head, tr[:300])
    fx = np.load(os.path.join(gold, 'tm_blobs_400_adjoint.npz'))
    assert float(fx['L']) == pytest.approx(float(np.sum(tr ** 2)), rel=1e-13)
    assert fx['directional_derivatives'].shape == (len(fx['seeds']),)
