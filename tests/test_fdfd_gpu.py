"""GPU parity of the FDFD path (HIP stencils + BiCGSTAB + adjoints) against the
scipy oracle (ceviche's own algorithm: CSR assembly + SuperLU spsolve)."""
import os

import numpy as np
import pytest
import scipy.sparse.linalg as spla
import torch

from oracle import fdfd_ref as R
from oracle.numdiff import der_num
from tests._cfg import cfg1

pytestmark = pytest.mark.gpu


def _op(pol, shape, npml, bloch=(0.0, 0.0), omega=None, dL=50e-9, dtype=torch.complex128):
    from ceviche_amd._fdfd_op import FdfdOperator
    omega = omega or 2 * np.pi * R.C_0 / 1.55e-6
    return FdfdOperator(pol, [dict(omega=omega, dL=dL, shape=shape, npml=npml, bloch=bloch)], dtype=dtype), omega


CASES = [((12, 9), (3, 2), (0.0, 0.0)), ((7, 11), (2, 3), (0.7, -1.3)), ((2, 5), (0, 1), (0.3, 0.2)),
         ((1, 16), (0, 4), (0.0, 0.0)), ((16, 1), (4, 0), (0.5, 0.0)), ((120, 120), (20, 20), (0.0, 0.0)),
         ((97, 203), (10, 15), (0.2, 0.0))]


@pytest.mark.parametrize('pol', ['ez', 'hz'])
@pytest.mark.parametrize('shape,npml,bloch', CASES)
@pytest.mark.parametrize('dtype', [torch.complex128, torch.complex64])
def test_stencil_matches_assembled_operator(pol, shape, npml, bloch, dtype):
    print('stencil', pol, shape, npml, bloch, dtype, flush=True)
    op, omega = _op(pol, shape, npml, bloch, dtype=dtype)
    rng = np.random.default_rng(11)
    eps = rng.uniform(1, 12.25, shape)
    A = (R.make_A_ez if pol == 'ez' else R.make_A_hz)(omega, 50e-9, eps, npml, bloch)
    x = rng.standard_normal(eps.size) + 1j * rng.standard_normal(eps.size)
    m0, m1 = op.materials(torch.from_numpy(eps.ravel()))
    tol = 1e-14 if dtype == torch.complex128 else 2e-6
    for tr in (False, True):
        y = op.apply(torch.from_numpy(x).cuda(), m0, m1, transpose=tr).cpu().numpy()
        ref = op.sys_scale[0] * ((A.T if tr else A) @ x)
        assert np.linalg.norm(y - ref) <= tol * np.linalg.norm(ref)


@pytest.mark.parametrize('which', range(4))
@pytest.mark.parametrize('mode', range(3))
def test_derivative_operators(which, mode):
    shape, npml, bloch = (13, 10), (3, 2), (0.4, -0.9)
    op, omega = _op('ez', shape, npml, bloch)
    D = R.derivative_matrices(omega, shape, npml, 50e-9, *bloch)[which]
    rng = np.random.default_rng(5)
    u = rng.standard_normal(130) + 1j * rng.standard_normal(130)
    mat = rng.uniform(0.5, 2, 130)
    coef = 0.3 - 1.7j
    out = op.deriv(torch.from_numpy(u).cuda(), which, coef, mat=torch.from_numpy(mat).cuda(), mode=mode).cpu().numpy()
    if mode == 0:
        ref = coef * mat * (D @ u)
    elif mode == 1:
        ref = coef * (D.T @ (mat * u))
    else:
        ref = np.conj(coef) * (D.conj().T @ (mat * u))
    assert np.linalg.norm(out - ref) <= 1e-14 * np.linalg.norm(ref)


@pytest.mark.parametrize('method', ['auto', 'bicgstab', 'cocg'])
def test_cfg1_ez_fields_match_spsolve(method):
    import ceviche_amd as cv
    c = cfg1()
    F = cv.fdfd_ez(c['omega'], c['dL'], c['eps'], c['npml'], solver_options=cv.SolverOptions(method=method))
    Hx, Hy, Ez = F.solve(c['src'])
    info = F.last_solve_info
    print('cfg1 info', method, info, flush=True)
    assert info['method'] == ('direct' if method == 'auto' else method)
    hx, hy, ez = R.solve_ez(c['omega'], c['dL'], c['eps'], c['npml'], c['src'])
    for a, b in ((Ez, ez), (Hx, hx), (Hy, hy)):
        assert isinstance(a, np.ndarray) and a.dtype == np.complex128 and a.shape == (120, 120)
        assert np.linalg.norm(a - b) / np.linalg.norm(b) <= 1e-6
    assert info['rel_residual'] <= 1e-10


@pytest.mark.parametrize('method', ['auto', 'bicgstab', 'cocg'])
def test_cfg1_ez_gradient_matches_oracle_adjoint(method):
    import ceviche_amd as cv
    c = cfg1()
    eps = torch.tensor(c['eps'], requires_grad=True)
    F = cv.fdfd_ez(c['omega'], c['dL'], eps, c['npml'], solver_options=cv.SolverOptions(method=method))
    _, _, Ez = F.solve(torch.from_numpy(c['src']))
    L = Ez[c['probe']].abs() ** 2
    L.backward()
    g = eps.grad.numpy()
    ez = Ez.detach().cpu().numpy()
    cot = np.zeros_like(ez)
    cot[c['probe']] = 2 * ez[c['probe']]
    gref, _ = R.grad_eps_ez(c['omega'], c['dL'], c['eps'], c['npml'], c['src'], cot)
    box = (slice(40, 80), slice(40, 80))
    assert np.linalg.norm(g[box] - gref[box]) <= 1e-4 * np.linalg.norm(gref[box])
    assert np.linalg.norm(g - gref) <= 1e-4 * np.linalg.norm(gref)


def _fd_objective(pol, eps, omega, dL, npml, src):
    f = (R.solve_ez if pol == 'ez' else R.solve_hz)(omega, dL, eps, npml, src)
    return sum(np.sum(np.abs(v) ** 2) for v in f)


@pytest.mark.parametrize('pol', ['ez', 'hz'])
def test_reference_gradient_test_setup(pol):
    """ceviche tests/test_gradients_fdfd.py:34-132: 30x30 (here 20x18 to keep the
    finite differences fast), omega = 2 pi 200 THz, dL = 1 um, eps in [1, 2),
    objective sum |E|^2 + |H|^2 over the three returned components, all entries
    of the gradient against central differences, ratio <= 1e-4."""
    import ceviche_amd as cv
    rng = np.random.default_rng(21)
    shape = (20, 18)
    omega, dL, npml = 2 * np.pi * 200e12, 1e-6, [5, 5]
    eps0 = rng.random(shape) + 1
    src = np.zeros(shape)
    src[10, 9] = 1.0
    cls = cv.fdfd_ez if pol == 'ez' else cv.fdfd_hz
    eps = torch.tensor(eps0, requires_grad=True)
    F = cls(omega, dL, eps, npml)
    out = F.solve(torch.from_numpy(src))
    L = sum((v.abs() ** 2).sum() for v in out)
    L.backward()
    g = eps.grad.numpy()
    scale = _fd_objective(pol, eps0, omega, dL, npml, src)
    num = np.zeros(shape)
    h = 1e-5
    for idx in np.ndindex(shape):
        num[idx] = der_num(lambda e: _fd_objective(pol, e, omega, dL, npml, src), eps0.copy(), idx, h)
    assert L.item() == pytest.approx(scale, rel=1e-8)
    assert np.linalg.norm(num - g) <= 1e-4 * np.linalg.norm(num)


def test_hz_fields_match_spsolve():
    import ceviche_amd as cv
    rng = np.random.default_rng(2)
    omega = 2 * np.pi * R.C_0 / 1.55e-6
    eps = rng.uniform(1, 12.25, (64, 80))
    src = np.zeros((64, 80))
    src[20, 30] = 1.0
    F = cv.fdfd_hz(omega, 50e-9, eps, [10, 12])
    Ex, Ey, Hz = F.solve(src)
    ex, ey, hz = R.solve_hz(omega, 50e-9, eps, [10, 12], src)
    for a, b in ((Hz, hz), (Ex, ex), (Ey, ey)):
        assert np.linalg.norm(a - b) / np.linalg.norm(b) <= 1e-6


def test_bloch_and_complex_eps_ez():
    """Lossy complex eps (Im eps < 0 is loss in ceviche's e^{+i w t} convention)
    with a Bloch phase and no PML along x."""
    import ceviche_amd as cv
    rng = np.random.default_rng(4)
    omega = 2 * np.pi * R.C_0 / 1.55e-6
    eps = rng.uniform(1, 4, (40, 30)) - 1j * rng.uniform(0, 0.5, (40, 30))
    src = np.zeros((40, 30))
    src[20, :] = 1.0
    F = cv.fdfd_ez(omega, 50e-9, eps, [0, 6], bloch_phases=[0.6, 0.0])
    Hx, Hy, Ez = F.solve(src)
    hx, hy, ez = R.solve_ez(omega, 50e-9, eps, [0, 6], src, bloch_phases=(0.6, 0.0))
    assert np.linalg.norm(Ez - ez) / np.linalg.norm(ez) <= 1e-6


@pytest.mark.parametrize('method', ['auto', 'bicgstab'])
def test_complex64_mixed_precision_reaches_fp64_residual(method):
    import ceviche_amd as cv
    c = cfg1()
    F = cv.fdfd_ez(c['omega'], c['dL'], c['eps'], c['npml'], precision='complex64',
                   solver_options=cv.SolverOptions(method=method))
    _, _, Ez = F.solve(c['src'])
    info = F.last_solve_info
    print('mixed info', info, flush=True)
    _, _, ez = R.solve_ez(c['omega'], c['dL'], c['eps'], c['npml'], c['src'])
    assert info['rel_residual'] <= 1e-6
    assert np.linalg.norm(Ez - ez) / np.linalg.norm(ez) <= 1e-4


def test_batched_uneven_instances_match_spsolve():
    """Several instances of different sizes and frequencies in one launch."""
    from ceviche_amd._fdfd_op import FdfdOperator
    rng = np.random.default_rng(9)
    specs, epss, srcs = [], [], []
    for shape, lam in [((40, 30), 1.5e-6), ((64, 64), 1.55e-6), ((33, 50), 1.6e-6)]:
        omega = 2 * np.pi * R.C_0 / lam
        specs.append(dict(omega=omega, dL=40e-9, shape=shape, npml=(6, 6)))
        epss.append(rng.uniform(1, 6, shape))
        s = np.zeros(shape)
        s[shape[0] // 3, :] = 1.0
        srcs.append(s)
    op = FdfdOperator('ez', specs)
    eps = torch.from_numpy(np.concatenate([e.ravel() for e in epss])).cuda()
    b = torch.from_numpy(np.concatenate([1j * sp['omega'] * s.ravel() for sp, s in zip(specs, srcs)])).cuda()
    m0, m1 = op.materials(eps)
    x = op.solve_scaled(b * op.scale_vector(torch.float64), m0, m1).cpu().numpy()
    for k, (sp, e, s) in enumerate(zip(specs, epss, srcs)):
        ref = R.solve_ez(sp['omega'], sp['dL'], e, sp['npml'], s)[2].ravel()
        got = x[op.offsets[k]:op.offsets[k] + e.size]
        assert np.linalg.norm(got - ref) / np.linalg.norm(ref) <= 1e-6
        assert op.last_info[k]['rel_residual'] <= 1e-10


@pytest.mark.parametrize('pol', ['ez', 'hz'])
def test_reference_forward_mode_setup(pol):
    """ceviche tests/test_gradients_fdfd.py:104-132 / :162-190: jacobian(mode='forward')
    of |E|^2 + |H|^2 (all cells) w.r.t. a scalar c with eps = c eps_r and
    source = c eps_r src, against central differences of the oracle."""
    import ceviche_amd as cv
    rng = np.random.default_rng(31)
    shape = (20, 18)
    omega, dL, npml = 2 * np.pi * 200e12, 1e-6, [5, 5]
    eps_r = rng.random(shape) + 1
    src = np.zeros(shape)
    src[10, 9] = 1.0
    cls = cv.fdfd_ez if pol == 'ez' else cv.fdfd_hz
    F = cls(omega, dL, eps_r, npml)

    eps_t, src_t = torch.from_numpy(eps_r), torch.from_numpy(src)

    def J(c):
        c = torch.as_tensor(c, dtype=torch.float64)
        F.eps_r = c * eps_t
        a, b, z = F.solve(c * eps_t * src_t)
        return (z.abs() ** 2 + a.abs() ** 2 + b.abs() ** 2).reshape(-1)

    jf = cv.jacobian(J, mode='forward')(1.0)
    solver = R.solve_ez if pol == 'ez' else R.solve_hz

    def Jr(c):
        a, b, z = solver(omega, dL, c * eps_r, npml, c * eps_r * src)
        return (np.abs(z) ** 2 + np.abs(a) ** 2 + np.abs(b) ** 2).ravel()

    h = 1e-6
    num = ((Jr(1 + h) - Jr(1 - h)) / (2 * h))[:, None]
    assert jf.shape == (shape[0] * shape[1], 1)
    assert np.linalg.norm(jf - num) <= 1e-4 * np.linalg.norm(num)
    jn = cv.jacobian(J, mode='numerical')(1.0)
    assert np.linalg.norm(jn - num) <= 1e-3 * np.linalg.norm(num)


def test_reverse_jacobian_matches_reference_convention():
    """jacobian(mode='reverse') of a real objective w.r.t. real eps (one row)."""
    import ceviche_amd as cv
    c = cfg1()
    F = cv.fdfd_ez(c['omega'], c['dL'], c['eps'], c['npml'])

    def J(eps):
        F.eps_r = eps
        return F.solve(torch.from_numpy(c['src']))[2][c['probe']].abs() ** 2

    jr = cv.jacobian(J, mode='reverse')(c['eps'])
    g = np.load(os.path.join(os.path.dirname(__file__), 'golden', 'waveguide_box_120.npz'))['grad_eps']
    assert jr.shape == (1, 120 * 120)
    assert np.linalg.norm(jr[0] - g.ravel()) <= 1e-4 * np.linalg.norm(g)


@pytest.mark.parametrize('pol', ['ez', 'hz'])
@pytest.mark.parametrize('dtype', [torch.complex128, torch.complex64])
def test_cocg_forward_and_transposed_solves_match_spsolve(pol, dtype):
    """COCG on W A (W = 1/(kxf kyf) for Ez, 1/(kxb kyb) for Hz) for A x = b and
    A^T y = b on a batch of two uneven instances with PML, against scipy."""
    from ceviche_amd._fdfd_op import FdfdOperator, SolverOptions
    rng = np.random.default_rng(12)
    specs, epss = [], []
    for shape, lam, npml in [((70, 56), 1.5e-6, (8, 6)), ((48, 90), 1.6e-6, (6, 10))]:
        specs.append(dict(omega=2 * np.pi * R.C_0 / lam, dL=40e-9, shape=shape, npml=npml))
        epss.append(rng.uniform(1, 8, shape))
    op = FdfdOperator(pol, specs, dtype=dtype)
    assert op.cocg_applies()
    eps = torch.from_numpy(np.concatenate([e.ravel() for e in epss])).cuda()
    m0, m1 = op.materials(eps)
    b = rng.standard_normal(op.total_cells) + 1j * rng.standard_normal(op.total_cells)
    tol = 1e-10 if dtype == torch.complex128 else 1e-5
    for tr in (False, True):
        bs = torch.from_numpy(b).cuda().to(dtype) * op.scale_vector(dtype)
        x = op.solve_scaled(bs, m0, m1, transpose=tr, opts=SolverOptions(tol=tol, method='cocg'))
        x = x.cpu().numpy().astype(np.complex128)
        for k, (sp, e) in enumerate(zip(specs, epss)):
            A = (R.make_A_ez if pol == 'ez' else R.make_A_hz)(sp['omega'], sp['dL'], e, sp['npml'])
            A = (A.T if tr else A).tocsc()
            sl = slice(op.offsets[k], op.offsets[k] + e.size)
            ref = spla.spsolve(A, b[sl])
            assert op.last_info[k]['method'] == 'cocg'
            assert op.last_info[k]['rel_residual'] <= tol
            assert np.linalg.norm(x[sl] - ref) / np.linalg.norm(ref) <= (1e-7 if dtype == torch.complex128 else 5e-3)


def test_cocg_is_refused_with_bloch_phases():
    from ceviche_amd._fdfd_op import SolverOptions
    op, _ = _op('ez', (20, 16), (3, 3), bloch=(0.3, 0.0))
    assert not op.cocg_applies()
    assert op.krylov_method(SolverOptions()) == 'bicgstab'
    assert op.krylov_method(SolverOptions(method='auto')) == 'bicgstab'
    with pytest.raises(ValueError):
        op.krylov_method(SolverOptions(method='cocg'))
