"""GPU parity of the FDTD path (Yee kernels + adjoint kernels) against the numpy
oracle of ceviche/fdtd.py."""
import numpy as np
import pytest
import torch

from oracle import fdtd_ref as T

pytestmark = pytest.mark.gpu

FIELDS = ('Ex', 'Ey', 'Ez', 'Dx', 'Dy', 'Dz', 'Hx', 'Hy', 'Hz')


def _pulse(t, t0=30, s=8):
    return np.exp(-(t - t0) ** 2 / (2 * s ** 2))


def _rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


@pytest.mark.parametrize('shape,npml', [((12, 10, 8), [3, 2, 2]), ((9, 1, 7), [2, 0, 2]), ((14, 11, 1), [3, 3, 0]),
                                        ((5, 9, 70), [1, 2, 3]), ((18, 4, 128), [2, 1, 4])])
def test_forward_all_fields_float64(shape, npml):
    import ceviche_amd as cv
    print('fdtd fwd', shape, npml, flush=True)
    rng = np.random.default_rng(0)
    eps = rng.uniform(1, 5, shape)
    pat = [rng.standard_normal(shape) for _ in range(3)]
    F = cv.fdtd(eps, 40e-9, npml)
    R = T.FDTD(eps, 40e-9, npml)
    assert F.dt == R.dt
    for t in range(40):
        J = [p * _pulse(t) for p in pat]
        f = F.forward(Jx=J[0], Jy=J[1], Jz=J[2])
        r = R.forward(Jx=J[0], Jy=J[1], Jz=J[2])
        if t % 13 == 0 or t == 39:
            for k in FIELDS:
                assert isinstance(f[k], np.ndarray) and f[k].shape == shape
                assert _rel(f[k], r[k]) <= 1e-10, (t, k, _rel(f[k], r[k]))
    assert F.t_index == 40


def test_forward_2d_tm_float32_against_float64_oracle():
    import ceviche_amd as cv
    rng = np.random.default_rng(1)
    shape = (64, 48, 1)
    eps = np.where(rng.standard_normal(shape) > 0, 12.25, 1.0)
    F = cv.fdtd(eps, 20e-9, [8, 8, 0], precision='float32')
    R = T.FDTD(eps, 20e-9, [8, 8, 0])
    J = np.zeros(shape)
    for t in range(200):
        J[32, 24, 0] = _pulse(t, 60, 20)
        f = F.forward(Jz=J)
        r = R.forward(Jz=J)
    for k in ('Ez', 'Hx', 'Hy', 'Dz'):
        assert _rel(f[k], r[k]) <= 1e-4
    for k in ('Ex', 'Ey', 'Hz'):
        assert np.all(f[k] == 0)


def test_eps_setter_resets_fields_and_time():
    import ceviche_amd as cv
    eps = np.ones((10, 10, 1))
    F = cv.fdtd(eps, 50e-9, [2, 2, 0])
    J = np.zeros((10, 10, 1))
    J[5, 5, 0] = 1.0
    F.forward(Jz=J)
    assert F.t_index == 1 and np.abs(F.forward(Jz=J)['Ez']).max() > 0
    F.eps_r = 2 * eps
    assert F.t_index == 0
    assert np.all(F.fields['Ez'] == 0)
    with pytest.raises(ValueError):
        F.forward(Jz=np.zeros((10, 10)))
    with pytest.raises(ValueError):
        cv.fdtd(np.ones((2, 2, 2, 2)), 50e-9, [0, 0, 0])


@pytest.mark.parametrize('field_kind', ['E', 'H'])
def test_per_step_reverse_mode_matches_complex_step(field_kind):
    """ceviche tests/test_gradients_fdtd.py:66-140: 8x8x1, dL = 50 nm, pml [2,2,0],
    Gaussian t0 = 300, sigma = 20, 500 steps; objective sum_t sum(Ex+Ey+Ez) with a
    Jz source (E) or sum_t sum(Hx+Hy+Hz) with a Jx source (H); all 64 gradient
    entries against complex-step derivatives of the oracle, ratio <= 1e-4."""
    import ceviche_amd as cv
    rng = np.random.default_rng(5)
    eps0 = rng.random((8, 8, 1)) + 1
    steps = 500
    comp = 'Jz' if field_kind == 'E' else 'Jx'
    keys = ('Ex', 'Ey', 'Ez') if field_kind == 'E' else ('Hx', 'Hy', 'Hz')

    def src(t):
        s = np.zeros((8, 8, 1))
        s[4, 4, 0] = np.exp(-(t - 300) ** 2 / 2 / 20 ** 2)
        return s

    eps = torch.tensor(eps0, requires_grad=True)
    F = cv.fdtd(eps0, 5e-8, [2, 2, 0])
    F.eps_r = eps
    S = 0.0
    for t in range(steps):
        f = F.forward(**{comp: torch.from_numpy(src(t))})
        S = S + sum(f[k].sum() for k in keys)
    S.backward()
    g = eps.grad.numpy()

    def obj_cs(e):
        R = T.FDTD(e, 5e-8, [2, 2, 0], dtype=np.complex128)
        tot = 0.0
        for t in range(steps):
            f = R.forward(**{comp: src(t).astype(np.complex128)})
            tot = tot + sum(np.sum(f[k]) for k in keys)
        return tot

    h = 1e-30
    num = np.zeros((8, 8, 1))
    for idx in np.ndindex(8, 8, 1):
        d = np.zeros((8, 8, 1))
        d[idx] = 1.0
        num[idx] = np.imag(obj_cs(eps0 + 1j * h * d)) / h
    ref_val = np.real(obj_cs(eps0.astype(np.complex128)))
    assert S.item() == pytest.approx(ref_val, rel=1e-10)
    assert np.linalg.norm(num - g) <= 1e-4 * np.linalg.norm(num)


def _oracle_probe_energy(eps, dL, npml, src_pos, amp, probes, dtype=np.float64):
    R = T.FDTD(eps, dL, npml, dtype=dtype)
    J = np.zeros(R.shape, dtype=dtype)
    L = 0.0
    tr = []
    for t in range(len(amp)):
        J[src_pos] = amp[t]
        f = R.forward(Jz=J)
        vals = [f[c][p] for c, p in probes]
        tr.append(vals)
        L = L + sum(v * v for v in vals)
    return L, np.array(tr)


@pytest.mark.parametrize('prec', ['float64', 'float32'])
def test_run_probes_and_checkpointed_adjoint_2d(prec):
    import ceviche_amd as cv
    rng = np.random.default_rng(7)
    shape = (40, 36, 1)
    eps0 = np.where(rng.standard_normal(shape) > 0, 12.25, 1.0)
    steps, dL, npml = 300, 20e-9, [6, 6, 0]
    amp = np.exp(-(np.arange(steps) - 60) ** 2 / (2 * 15 ** 2))
    probes = [('Ez', (10, 18, 0)), ('Ez', (30, 18, 0)), ('Hx', (20, 8, 0)), ('Ez', (20, 28, 0))]
    eps = torch.tensor(eps0, requires_grad=True)
    F = cv.fdtd(eps, dL, npml, precision=prec)
    rec = F.run(steps, (20, 18, 0), amp, probes, checkpoint_every=37)
    L = (rec.double() ** 2).sum()
    L.backward()
    g = eps.grad.detach().cpu().numpy().astype(np.float64)
    Lref, tr = _oracle_probe_energy(eps0, dL, npml, (20, 18, 0), amp, [(c, p) for c, p in probes])
    tol = 1e-10 if prec == 'float64' else 1e-4
    assert _rel(rec.detach().cpu().numpy(), tr) <= tol
    # directional derivatives by complex step of the oracle
    for seed in (0, 1):
        d = np.random.default_rng(100 + seed).standard_normal(shape)
        h = 1e-30
        Lc, _ = _oracle_probe_energy(eps0 + 1j * h * d, dL, npml, (20, 18, 0), amp.astype(complex),
                                     [(c, p) for c, p in probes], dtype=np.complex128)
        dd = np.imag(Lc) / h
        gtol = 1e-8 if prec == 'float64' else 1e-3
        assert np.sum(g * d) == pytest.approx(dd, rel=gtol)


def _oracle_forward_then_run(eps, dL, npml, Jpre, src_pos, amp, probes, dtype=np.float64):
    """Probe energy of len(Jpre) reference steps with dense Jz sources (not
    recorded) followed by a point-source run recording ``probes``."""
    R = T.FDTD(eps, dL, npml, dtype=dtype)
    for Jz in Jpre:
        R.forward(Jz=Jz.astype(dtype))
    J = np.zeros(R.shape, dtype=dtype)
    L = 0.0
    for t in range(len(amp)):
        J[src_pos] = amp[t]
        f = R.forward(Jz=J)
        L = L + sum(f[c][p] ** 2 for c, p in probes)
    return L


@pytest.mark.parametrize('kernel', ['auto', 'generic'])
def test_chained_runs_gradients_reach_eps_and_amplitudes(kernel):
    """Two run() calls in a row: the second starts from the first's final
    state, so eps and the amplitudes of the first run also get the second
    run's cotangents through the state.  Against complex step of one
    continuous oracle run."""
    import ceviche_amd as cv
    rng = np.random.default_rng(21)
    shape = (30, 26, 1)
    eps0 = np.where(rng.standard_normal(shape) > 0, 12.25, 1.0)
    dL, npml, n1, n2 = 20e-9, [5, 5, 0], 70, 90
    amp0 = np.exp(-(np.arange(n1 + n2) - 40) ** 2 / (2 * 12 ** 2))
    probes = [('Ez', (8, 13, 0)), ('Hy', (22, 13, 0)), ('Ez', (15, 20, 0))]
    src = (15, 13, 0)
    eps = torch.tensor(eps0, requires_grad=True)
    amp = torch.tensor(amp0, requires_grad=True)
    F = cv.fdtd(eps, dL, npml)
    r1 = F.run(n1, src, amp[:n1], probes, checkpoint_every=23, kernel=kernel)
    r2 = F.run(n2, src, amp[n1:], probes, checkpoint_every=31, kernel=kernel)
    L = (r1 ** 2).sum() + (r2 ** 2).sum()
    L.backward()
    Lref, tr = _oracle_probe_energy(eps0, dL, npml, src, amp0, probes)
    assert float(L) == pytest.approx(Lref, rel=1e-10)
    h = 1e-30
    d = np.random.default_rng(5).standard_normal(shape)
    Lc, _ = _oracle_probe_energy(eps0 + 1j * h * d, dL, npml, src, amp0.astype(complex), probes, dtype=np.complex128)
    assert np.sum(eps.grad.cpu().numpy() * d) == pytest.approx(np.imag(Lc) / h, rel=1e-8)
    da = np.random.default_rng(6).standard_normal(n1 + n2)
    Lc, _ = _oracle_probe_energy(eps0.astype(complex), dL, npml, src, amp0 + 1j * h * da, probes,
                                 dtype=np.complex128)
    assert amp.grad is not None
    assert np.sum(amp.grad.cpu().numpy() * da) == pytest.approx(np.imag(Lc) / h, rel=1e-8)
    # the first run's amplitudes alone (they act through the carried state in run 2 as well)
    da1 = np.concatenate([da[:n1], np.zeros(n2)])
    Lc, _ = _oracle_probe_energy(eps0.astype(complex), dL, npml, src, amp0 + 1j * h * da1, probes,
                                 dtype=np.complex128)
    assert np.sum(amp.grad.cpu().numpy() * da1) == pytest.approx(np.imag(Lc) / h, rel=1e-8)


def test_forward_steps_then_run_gradient_through_the_state():
    """forward() steps (per-step adjoint) followed by run() (checkpointed
    adjoint): the run's cotangent flows back through the carried state into
    the per-step part.  3D grid, against complex step of the oracle."""
    import ceviche_amd as cv
    rng = np.random.default_rng(22)
    shape = (9, 8, 7)
    eps0 = rng.uniform(1, 4, shape)
    dL, npml = 25e-9, [2, 2, 2]
    Jpre = [rng.standard_normal(shape) * np.exp(-(t - 6) ** 2 / 8.0) for t in range(12)]
    amp0 = np.exp(-(np.arange(50) - 10) ** 2 / (2 * 5 ** 2))
    probes = [('Ex', (3, 4, 3)), ('Hz', (6, 2, 4)), ('Ez', (4, 4, 1))]
    src = (4, 4, 3)
    eps = torch.tensor(eps0, requires_grad=True)
    F = cv.fdtd(eps, dL, npml)
    for Jz in Jpre:
        F.forward(Jz=torch.from_numpy(Jz))
    rec = F.run(50, src, amp0, probes, checkpoint_every=13)
    L = (rec ** 2).sum()
    L.backward()
    Lref = _oracle_forward_then_run(eps0, dL, npml, Jpre, src, amp0, probes)
    assert float(L) == pytest.approx(Lref, rel=1e-10)
    h = 1e-30
    for seed in (0, 1):
        d = np.random.default_rng(40 + seed).standard_normal(shape)
        Lc = _oracle_forward_then_run(eps0 + 1j * h * d, dL, npml, Jpre, src, amp0.astype(complex), probes,
                                      dtype=np.complex128)
        assert np.sum(eps.grad.cpu().numpy() * d) == pytest.approx(np.imag(Lc) / h, rel=1e-8)


def test_3d_checkpointed_adjoint_with_e_and_h_probes():
    import ceviche_amd as cv
    rng = np.random.default_rng(8)
    shape = (10, 9, 8)
    eps0 = rng.uniform(1, 4, shape)
    steps, dL, npml = 80, 25e-9, [2, 2, 2]
    amp = np.exp(-(np.arange(steps) - 25) ** 2 / (2 * 8 ** 2))
    probes = [('Ex', (3, 4, 5)), ('Hy', (6, 2, 3)), ('Ez', (5, 5, 1)), ('Dz', (2, 7, 6)), ('Hz', (8, 1, 4))]
    eps = torch.tensor(eps0, requires_grad=True)
    F = cv.fdtd(eps, dL, npml)
    rec = F.run(steps, (5, 4, 4), amp, probes, checkpoint_every=17)
    L = (rec ** 2).sum()
    L.backward()
    g = eps.grad.detach().cpu().numpy()
    Lref, tr = _oracle_probe_energy(eps0, dL, npml, (5, 4, 4), amp, probes)
    assert _rel(rec.detach().cpu().numpy(), tr) <= 1e-10
    d = np.random.default_rng(3).standard_normal(shape)
    h = 1e-30
    Lc, _ = _oracle_probe_energy(eps0 + 1j * h * d, dL, npml, (5, 4, 4), amp.astype(complex), probes,
                                 dtype=np.complex128)
    assert np.sum(g * d) == pytest.approx(np.imag(Lc) / h, rel=1e-8)


@pytest.mark.parametrize('cfg', range(12))
@pytest.mark.parametrize('shape', [(40, 36, 1), (13, 7, 1), (70, 130, 1)])
def test_tm_blocked_kernel_matches_oracle(cfg, shape):
    """Every tile configuration of the temporally blocked 2D TM kernel, on
    domains smaller than a tile and not multiples of it, with PML corners."""
    import ceviche_amd as cv
    print('tm cfg', cfg, shape, flush=True)
    rng = np.random.default_rng(11)
    eps0 = np.where(rng.standard_normal(shape) > 0, 12.25, 1.0)
    steps, dL = 45, 20e-9
    npml = [min(6, shape[0] // 3), min(6, shape[1] // 3), 0]
    amp = np.exp(-(np.arange(steps) - 15) ** 2 / (2 * 5 ** 2))
    src = (shape[0] // 2, shape[1] // 2, 0)
    probes = [('Ez', (1, 2, 0)), ('Hx', (shape[0] - 1, shape[1] // 3, 0)), ('Hy', (shape[0] // 2, shape[1] - 1, 0)),
              ('Ez', src)]
    F = cv.fdtd(torch.tensor(eps0), dL, npml)
    rec = F.run(steps, src, amp, probes, kernel=cfg).cpu().numpy()
    fields = {k: v.detach().cpu().numpy() for k, v in F.fields.items()}
    G = cv.fdtd(torch.tensor(eps0), dL, npml)
    rec_g = G.run(steps, src, amp, probes, kernel='generic').cpu().numpy()
    _, tr = _oracle_probe_energy(eps0, dL, npml, src, amp, probes)
    assert _rel(rec, tr) <= 1e-11
    assert _rel(rec, rec_g) <= 1e-12
    R = T.FDTD(eps0, dL, npml)
    J = np.zeros(shape)
    for t in range(steps):
        J[src] = amp[t]
        r = R.forward(Jz=J)
    for k in ('Ez', 'Hx', 'Hy', 'Dz'):
        assert _rel(fields[k], r[k]) <= 1e-11, k


def test_tm_blocked_gradient_matches_generic():
    import ceviche_amd as cv
    rng = np.random.default_rng(12)
    shape = (50, 44, 1)
    eps0 = np.where(rng.standard_normal(shape) > 0, 12.25, 1.0)
    amp = np.exp(-(np.arange(120) - 30) ** 2 / (2 * 8 ** 2))
    probes = [('Ez', (10, 20, 0)), ('Ez', (40, 20, 0))]
    grads = []
    for kernel in ('auto', 'generic'):
        eps = torch.tensor(eps0, requires_grad=True)
        F = cv.fdtd(eps, 20e-9, [8, 8, 0])
        rec = F.run(120, (25, 22, 0), amp, probes, checkpoint_every=25, kernel=kernel)
        (rec ** 2).sum().backward()
        grads.append(eps.grad.numpy())
    assert _rel(grads[0], grads[1]) <= 1e-10


@pytest.mark.parametrize('shape,npml,comp', [((12, 10, 9), [3, 2, 2], 'Jz'), ((15, 13, 1), [4, 3, 0], 'Jx'),
                                             ((7, 9, 11), [0, 2, 3], 'Jy'), ((9, 20, 70), [2, 4, 5], 'Jz'),
                                             ((5, 3, 2), [1, 1, 0], 'Jx'), ((12, 33, 130), [3, 3, 3], 'Jy'),
                                             # slab kernel: nz a multiple of 64 up to 256, bands of 4, 2, 1 rows
                                             ((11, 8, 64), [3, 2, 4], 'Jz'), ((9, 6, 128), [2, 2, 5], 'Jx'),
                                             ((10, 7, 64), [0, 3, 3], 'Jy'), ((5, 12, 192), [1, 2, 6], 'Jz'),
                                             ((3, 2, 64), [1, 0, 2], 'Jx'), ((1, 4, 64), [0, 1, 3], 'Jy'),
                                             ((20, 3, 256), [4, 1, 8], 'Jz'),
                                             # interior bands and slices take the z-PML-only path
                                             ((12, 16, 64), [3, 3, 4], 'Jz'), ((14, 12, 128), [2, 2, 0], 'Jy')])
def test_fused_single_pass_kernel_matches_oracle(shape, npml, comp):
    """The single-pass kernels (slab marching when nz is a multiple of 64 up
    to 256, x-marching tiles otherwise in 3D, per-cell in 2D), including
    (y, z) extents that are not multiples of the tiles, grids smaller than one
    tile and x extents of one slice."""
    print('fused', shape, npml, comp, flush=True)
    import ceviche_amd as cv
    rng = np.random.default_rng(13)
    eps0 = rng.uniform(1, 5, shape)
    steps, dL = 37, 30e-9
    amp = np.exp(-(np.arange(steps) - 12) ** 2 / (2 * 4 ** 2))
    src = tuple(s // 2 for s in shape)
    probes = [('Ex', (min(1, shape[0] - 1), min(2, shape[1] - 1), 0)), ('Hy', (shape[0] - 1, 1, shape[2] - 1)), ('Dz', src),
              ('Hz', (min(2, shape[0] - 1), min(3, shape[1] - 1), shape[2] // 2))]
    F = cv.fdtd(torch.tensor(eps0), dL, npml)
    rec = F.run(steps, src, amp, probes, source_component=comp, kernel='fused').cpu().numpy()
    fields = {k: v.detach().cpu().numpy() for k, v in F.fields.items()}
    R = T.FDTD(eps0, dL, npml)
    J = np.zeros(shape)
    tr = []
    for t in range(steps):
        J[src] = amp[t]
        r = R.forward(**{comp: J})
        tr.append([r[c][p] for c, p in probes])
    assert _rel(rec, np.array(tr)) <= 1e-11
    for k in FIELDS:
        assert _rel(fields[k], r[k]) <= 1e-11, k


def test_fused_and_two_phase_runs_agree_at_64_cubed():
    """float32 64^3 with PML on every face, 120 steps: the slab-marching
    single-pass run and the two-phase step kernels give the same probe
    records and final fields (same arithmetic term by term; the float32 bound
    allows for multiply-adds contracted differently, which over 120 steps
    reach 1e-5 relative on the weak Hz of a Jz source; an index or halo
    mistake gives O(1) differences)."""
    print('fused vs two-phase 64^3 float32', flush=True)
    import ceviche_amd as cv
    rng = np.random.default_rng(21)
    shape = (64, 64, 64)
    eps0 = rng.uniform(1, 6, shape).astype(np.float32)
    steps = 120
    amp = np.exp(-(np.arange(steps) - 40) ** 2 / (2 * 10 ** 2))
    probes = [('Ex', (30, 31, 32)), ('Hy', (5, 60, 2)), ('Ez', (63, 0, 63)), ('Hz', (32, 32, 10))]
    out = []
    for kernel in ('fused', 'generic'):
        F = cv.fdtd(torch.tensor(eps0), 25e-9, [8, 8, 8], precision='float32')
        rec = F.run(steps, (32, 32, 20), amp, probes, kernel=kernel).cpu().numpy().astype(np.float64)
        out.append((rec, {k: v.detach().cpu().numpy().astype(np.float64) for k, v in F.fields.items()}))
    assert _rel(out[0][0], out[1][0]) <= 1e-4
    for k in FIELDS:
        assert _rel(out[0][1][k], out[1][1][k]) <= 1e-4, k


def test_forward_mode_run_matches_complex_step():
    """Forward mode through fdtd.run (FdtdRun.jvp -> cv_fdtd_step_tangent) on a
    3D grid with E, D and H probes: the directional derivative of the probe
    energy along a random eps direction equals the complex-step derivative of
    the oracle (1e-8 relative), the tangent of every probe trace equals the
    oracle's, and jacobian(mode='forward') agrees with mode='reverse'."""
    import torch.autograd.forward_ad as fwAD
    import ceviche_amd as cv
    rng = np.random.default_rng(18)
    shape = (9, 8, 7)
    eps0 = rng.uniform(1, 4, shape)
    steps, dL, npml = 60, 25e-9, [2, 2, 2]
    amp = np.exp(-(np.arange(steps) - 20) ** 2 / (2 * 6 ** 2))
    probes = [('Ex', (3, 4, 5)), ('Hy', (6, 2, 3)), ('Dz', (2, 6, 5)), ('Ez', (4, 4, 3))]
    src = (4, 4, 3)
    d = rng.standard_normal(shape)
    with fwAD.dual_level():
        eps = fwAD.make_dual(torch.tensor(eps0, device='cuda'), torch.tensor(d, device='cuda'))
        F = cv.fdtd(eps, dL, npml)
        rec = F.run(steps, src, amp, probes)
        L = (rec ** 2).sum()
        dL_dir = fwAD.unpack_dual(L).tangent.item()
        drec = fwAD.unpack_dual(rec).tangent.cpu().numpy()
    h = 1e-30
    Lc, trc = _oracle_probe_energy(eps0 + 1j * h * d, dL, npml, src, amp.astype(complex), probes,
                                   dtype=np.complex128)
    assert dL_dir == pytest.approx(np.imag(Lc) / h, rel=1e-8)
    assert _rel(drec, np.imag(trc) / h) <= 1e-9

    def fn(e):
        G = cv.fdtd(e, dL, npml)
        return G.run(20, src, amp, probes[:2]).sum(0)

    x0 = eps0[:3, :3, :3].copy()
    base = torch.tensor(eps0, device='cuda')

    def fsub(sub):
        e = base.clone()
        e[:3, :3, :3] = sub
        return fn(e)

    Jf = cv.jacobian(fsub, mode='forward')(x0)
    Jr = cv.jacobian(fsub, mode='reverse')(x0)
    assert Jf.shape == Jr.shape == (2, 27)
    assert np.allclose(Jf, Jr, rtol=1e-9, atol=1e-14 * np.abs(Jr).max())


def test_forward_mode_per_step_api_matches_complex_step():
    """Forward mode through fdtd.forward (FdtdStep.jvp, FdtdEFields.jvp) on the
    reference's gradient set-up (8x8x1, pml [2,2,0], Jz source): the
    directional derivative of sum_t sum(Ex+Ey+Ez) against complex step."""
    import torch.autograd.forward_ad as fwAD
    import ceviche_amd as cv
    rng = np.random.default_rng(6)
    eps0 = rng.random((8, 8, 1)) + 1
    steps = 200
    d = rng.standard_normal((8, 8, 1))

    def src(t):
        s = np.zeros((8, 8, 1))
        s[4, 4, 0] = np.exp(-(t - 80) ** 2 / 2 / 20 ** 2)
        return s

    with fwAD.dual_level():
        eps = fwAD.make_dual(torch.tensor(eps0, device='cuda'), torch.tensor(d, device='cuda'))
        F = cv.fdtd(eps0, 5e-8, [2, 2, 0])
        F.eps_r = eps
        S = 0.0
        for t in range(steps):
            f = F.forward(Jz=torch.from_numpy(src(t)))
            S = S + sum(f[k].sum() for k in ('Ex', 'Ey', 'Ez'))
        dS = fwAD.unpack_dual(S).tangent.item()

    def obj_cs(e):
        R = T.FDTD(e, 5e-8, [2, 2, 0], dtype=np.complex128)
        tot = 0.0
        for t in range(steps):
            f = R.forward(Jz=src(t).astype(np.complex128))
            tot = tot + sum(np.sum(f[k]) for k in ('Ex', 'Ey', 'Ez'))
        return tot

    h = 1e-30
    assert dS == pytest.approx(np.imag(obj_cs(eps0 + 1j * h * d)) / h, rel=1e-8)
