"""GPU parity of the probe / spectrum post-processing (measure_fields with
device-side weighted probe sums, weighted probes fused into fdtd.run, the
torch.fft spectrum helpers) against the numpy oracle."""
import numpy as np
import pytest
import torch

from oracle import fdtd_ref as FR
from oracle import spectrum_ref as S

pytestmark = pytest.mark.gpu


def _rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def _setup(shape, seed=0):
    rng = np.random.default_rng(seed)
    eps = 1 + 3 * rng.random(shape)
    line = np.zeros(shape)
    line[shape[0] // 3, 2:shape[1] - 2, 0] = np.sin(np.linspace(0, np.pi, shape[1] - 4))
    mode = np.zeros(shape)
    mode[2 * shape[0] // 3, :, 0] = np.exp(-((np.arange(shape[1]) - shape[1] / 2) / 4.0) ** 2)
    point = np.zeros(shape)
    point[shape[0] // 2, shape[1] // 4, 0] = 1.0
    src_pos = np.zeros(shape)
    src_pos[shape[0] // 4, shape[1] // 2 - 3:shape[1] // 2 + 3, 0] = 1.0
    return eps, [line, mode, point], src_pos


@pytest.mark.parametrize('component', ['Ez', 'Hx', 'Dz'])
def test_measure_fields_matches_oracle_2d(component):
    import ceviche_amd as cv
    shape = (48, 40, 1)
    eps, probes, src_pos = _setup(shape)
    steps = 240
    gauss = lambda t: np.exp(-((t - 60) / 18.0) ** 2) * np.cos(0.3 * t)
    source = lambda t: src_pos * gauss(t)
    F = cv.fdtd(eps, 25e-9, [6, 6, 0])
    m = cv.utils.measure_fields(F, source, steps, probes, component=component)
    ref = S.measure_fields(FR.FDTD(eps, 25e-9, [6, 6, 0]), source, steps, probes, component=component)
    assert isinstance(m, np.ndarray) and m.shape == (steps, 3) and m.dtype == np.float64
    assert _rel(m, ref) <= 1e-11
    # a single probe array (not a list), as the reference accepts
    m1 = cv.utils.measure_fields(F, source, 30, probes[0], component=component)
    assert m1.shape == (30, 1) and _rel(m1[:, 0], ref[:30, 0]) <= 1e-11


def test_measure_fields_3d_and_float32():
    import ceviche_amd as cv
    rng = np.random.default_rng(3)
    shape = (12, 10, 9)
    eps = 1 + rng.random(shape)
    probes = [rng.standard_normal(shape) * (rng.random(shape) < 0.2), np.ones(shape)]
    src_pos = np.zeros(shape)
    src_pos[5, 5, 4] = 1.0
    source = lambda t: src_pos * np.exp(-((t - 15) / 5.0) ** 2)
    ref = S.measure_fields(FR.FDTD(eps, 25e-9, [2, 2, 2]), source, 50, probes, component='Ex')
    m = cv.utils.measure_fields(cv.fdtd(eps, 25e-9, [2, 2, 2]), source, 50, probes, component='Ex')
    assert _rel(m, ref) <= 1e-11
    m32 = cv.utils.measure_fields(cv.fdtd(eps, 25e-9, [2, 2, 2], precision='float32'), source, 50, probes,
                                  component='Ex')
    assert _rel(m32, ref) <= 1e-4


@pytest.mark.parametrize('prec,kernel', [('float64', 'generic'), ('float32', 'auto'), ('float32', 'generic')])
def test_weighted_probes_in_run_match_measure_fields(prec, kernel):
    """fdtd.run with weight-array probes (recorded in-kernel, combined after
    the run) against the oracle's measure_fields for the same point source."""
    import ceviche_amd as cv
    shape = (48, 40, 1)
    eps, probes, _ = _setup(shape, seed=2)
    steps = 300
    amp = np.exp(-((np.arange(steps) - 60) / 18.0) ** 2)
    src = (12, 20, 0)
    src_pos = np.zeros(shape)
    src_pos[src] = 1.0
    F = cv.fdtd(torch.tensor(eps), 25e-9, [6, 6, 0], precision=prec)
    plist = [('Ez', p) for p in probes] + [('Ez', (30, 10, 0)), ('Hy', probes[1])]
    rec = F.run(steps, src, amp, plist, kernel=kernel).double().cpu().numpy()
    R = FR.FDTD(eps, 25e-9, [6, 6, 0])
    ref = S.measure_fields(R, lambda t: src_pos * amp[t], steps, probes, component='Ez')
    refh = S.measure_fields(R, lambda t: src_pos * amp[t], steps, [probes[1]], component='Hy')
    tol = 1e-11 if prec == 'float64' else 1e-4
    assert rec.shape == (steps, 5)
    assert _rel(rec[:, :3], ref) <= tol
    assert _rel(rec[:, 4], refh[:, 0]) <= tol
    point = FR.run_probes(eps, 25e-9, [6, 6, 0], src, lambda t: amp[t], steps, [(30, 10, 0)])
    assert _rel(rec[:, 3], point[:, 0]) <= tol


def _oracle_weighted_energy(eps, dL, npml, src, amp, probes, steps, dtype=np.float64):
    R = FR.FDTD(eps, dL, npml, dtype=dtype)
    J = np.zeros(R.shape, dtype=dtype)
    total = 0.0
    for t in range(steps):
        J[src] = amp[t]
        f = R.forward(Jz=J)
        for comp, w in probes:
            total = total + np.sum(f[comp] * w) ** 2
    return total


def test_weighted_probe_gradients_match_complex_step():
    """Reverse and forward mode through run -> weighted probes, against the
    complex-step derivative of the oracle along a random eps direction."""
    import torch.autograd.forward_ad as fwAD
    import ceviche_amd as cv
    shape = (30, 26, 1)
    eps0, probes, _ = _setup(shape, seed=5)
    steps, dL, npml, src = 120, 25e-9, [4, 4, 0], (8, 13, 0)
    amp = np.exp(-((np.arange(steps) - 40) / 12.0) ** 2)
    plist = [('Ez', probes[0]), ('Hx', probes[1])]
    eps = torch.tensor(eps0, requires_grad=True)
    F = cv.fdtd(eps, dL, npml)
    rec = F.run(steps, src, amp, plist, checkpoint_every=29)
    L = (rec ** 2).sum()
    L.backward()
    g = eps.grad.cpu().numpy()
    d = np.random.default_rng(9).standard_normal(shape)
    h = 1e-30
    Lc = _oracle_weighted_energy(eps0 + 1j * h * d, dL, npml, src, amp.astype(complex), plist, steps,
                                 dtype=np.complex128)
    ref = np.imag(Lc) / h
    assert float(L) == pytest.approx(np.real(Lc), rel=1e-11)
    assert np.sum(g * d) == pytest.approx(ref, rel=1e-8)
    with fwAD.dual_level():
        e = fwAD.make_dual(torch.tensor(eps0, device='cuda'), torch.tensor(d, device='cuda'))
        G = cv.fdtd(e, dL, npml)
        Lf = (G.run(steps, src, amp, plist) ** 2).sum()
        assert fwAD.unpack_dual(Lf).tangent.item() == pytest.approx(ref, rel=1e-8)


def test_reference_fft_example_forward_mode():
    """examples/autograd_fft.py:50-67 pattern: spectral power |my_fft(measured)|^2
    of a weighted probe, forward-mode jacobian with respect to a scale of
    eps, against central differences of the oracle."""
    import ceviche_amd as cv
    shape = (36, 30, 1)
    rng = np.random.default_rng(4)
    eps_r = rng.random(shape) + 1
    steps, dL, npml, src = 160, 25e-9, [5, 5, 0], (10, 15, 0)
    amp = np.exp(-((np.arange(steps) - 40) / 12.0) ** 2) * np.cos(0.4 * np.arange(steps))
    probe = np.zeros(shape)
    probe[-8, 15, 0] = 1.0
    probe[-8, 10:20, 0] += 0.5
    base = torch.tensor(eps_r, device='cuda')

    def objective(c):
        F = cv.fdtd(base * c, dL, npml)
        measured = F.run(steps, src, amp, [('Ez', probe)])[:, 0]
        return torch.square(torch.abs(cv.utils.my_fft(measured)))

    jf = cv.jacobian(objective, mode='forward')(1.0)

    def oracle_power(c):
        R = FR.FDTD(eps_r * c, dL, npml)
        J = np.zeros(R.shape)
        m = []
        for t in range(steps):
            J[src] = amp[t]
            m.append(np.sum(R.forward(Jz=J)['Ez'] * probe))
        return np.square(np.abs(S.my_fft(np.array(m))))

    p0 = objective(torch.tensor(1.0, dtype=torch.float64)).detach().cpu().numpy()
    assert _rel(p0, oracle_power(1.0)) <= 1e-10
    h = 1e-6
    num = ((oracle_power(1 + h) - oracle_power(1 - h)) / (2 * h))[:, None]
    assert jf.shape == num.shape
    assert np.linalg.norm(jf - num) <= 1e-5 * np.linalg.norm(num)


def test_spectrum_helpers_match_numpy():
    import ceviche_amd as cv
    rng = np.random.default_rng(6)
    dt = 2.1e-17
    for series in (rng.standard_normal(500), rng.standard_normal((301, 3)),
                   rng.standard_normal(64) + 1j * rng.standard_normal(64)):
        f, sf = cv.utils.get_spectrum(series, dt)
        rf, rsf = S.get_spectrum(series, dt)
        np.testing.assert_allclose(f, rf, rtol=1e-15)
        np.testing.assert_allclose(sf, rsf, rtol=1e-12, atol=1e-12 * np.abs(rsf).max())
        _, pw = cv.utils.get_spectral_power(series, dt)
        np.testing.assert_allclose(pw, S.get_spectral_power(series, dt)[1], rtol=1e-12,
                                   atol=1e-12 * np.abs(rsf).max() ** 2)
        if np.ndim(series) == 1:
            assert cv.utils.get_max_power_freq(series, dt) == S.get_max_power_freq(series, dt)
    x = rng.standard_normal((4, 50))
    np.testing.assert_allclose(cv.utils.my_fft(x), S.my_fft(x), rtol=1e-12, atol=1e-12)
    t = torch.tensor(x, device='cuda')
    out = cv.utils.my_fft(t)
    assert isinstance(out, torch.Tensor) and out.is_cuda
