"""GPU parity of the block-tridiagonal direct path (csrc/btd.hip + FGMRES with
FP64 true residual) against the scipy oracle: the factorisation blocks
against a numpy block elimination, the sweep against the block-tridiagonal
part of the assembled operator, and fields / transposed solves / gradients
against spsolve."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl
import torch

from oracle import fdfd_ref as R

pytestmark = pytest.mark.gpu

OMEGA = 2 * np.pi * R.C_0 / 1.55e-6


def _op(pol, shape, npml, bloch=(0.0, 0.0)):
    from ceviche_amd._fdfd_op import FdfdOperator
    return FdfdOperator(pol, [dict(omega=OMEGA, dL=50e-9, shape=shape, npml=npml, bloch=bloch)])


def _scaled_A(op, pol, eps, npml, bloch):
    A = (R.make_A_ez if pol == 'ez' else R.make_A_hz)(OMEGA, 50e-9, eps, npml, bloch)
    return (A * op.sys_scale[0]).tocsr()


def _without_x_wrap(A, nx, ny):
    C = A.tocoo()
    keep = np.abs(C.row // ny - C.col // ny) <= 1
    return sp.csr_matrix((C.data[keep], (C.row[keep], C.col[keep])), shape=A.shape)


CASES = [('ez', (24, 18), (4, 3), (0.0, 0.0)), ('hz', (19, 23), (3, 4), (0.0, 0.0)),
         ('ez', (13, 17), (2, 3), (0.5, -0.8)), ('ez', (1, 20), (0, 4), (0.0, 0.0)),
         ('ez', (20, 1), (4, 0), (0.3, 0.0)), ('hz', (30, 128), (5, 10), (0.0, 0.0)),
         ('ez', (120, 120), (20, 20), (0.0, 0.0))]


CASES_BCR = CASES + [('ez', (2, 9), (0, 2), (0.0, 0.0)), ('hz', (3, 14), (1, 3), (0.0, 0.4)),
                     ('ez', (65, 40), (8, 6), (0.0, 0.0)),
                     # y extents whose products use the 80- and 96-wide tiles
                     ('ez', (40, 80), (6, 8), (0.0, 0.0)), ('hz', (33, 90), (5, 8), (0.2, 0.0)),
                     # the smallest x extent that defaults to reduction; a long one with odd
                     # block counts on most levels
                     ('ez', (8, 30), (2, 4), (0.0, 0.0)), ('hz', (1025, 12), (20, 2), (0.1, 0.0))]


@pytest.mark.parametrize('kind', ['btd', 'bcr'])
@pytest.mark.parametrize('pol,shape,npml,bloch', CASES_BCR)
def test_factor_and_sweep_invert_the_block_tridiagonal_part(pol, shape, npml, bloch, kind):
    from ceviche_amd._fdfd_op import Direct
    print('direct factor', kind, pol, shape, npml, bloch, flush=True)
    op = _op(pol, shape, npml, bloch)
    rng = np.random.default_rng(3)
    eps = rng.uniform(1, 12.25, shape)
    m0, m1 = op.materials(torch.from_numpy(eps.ravel()))
    D = Direct(op, m0, m1, kind=kind)
    nx, ny = shape
    A = _scaled_A(op, pol, eps, npml, bloch)
    if nx > 2:
        Anw = _without_x_wrap(A, nx, ny)
    else:
        Anw = A if nx == 1 else None
    v = rng.standard_normal(nx * ny) + 1j * rng.standard_normal(nx * ny)
    if Anw is not None:
        for tr in (False, True):
            z = D.apply(torch.from_numpy(v).cuda(), transpose=tr).cpu().numpy()
            ref = spl.spsolve((Anw.T if tr else Anw).tocsc(), v)
            assert np.linalg.norm(z - ref) <= 1e-3 * np.linalg.norm(ref), tr
    if kind != 'btd':
        return
    # the first two stored blocks against a float64 block elimination
    S = np.frombuffer(D.factor.cpu().numpy()[:nx * ny * ny * 8].tobytes(), dtype=np.complex64).reshape(nx, ny, ny)
    if Anw is not None:
        D0 = Anw[:ny, :ny].toarray()
        np.testing.assert_allclose(S[0], np.linalg.inv(D0), rtol=0, atol=1e-4 * np.abs(np.linalg.inv(D0)).max())
        if nx > 1:
            lo = Anw[ny:2 * ny, :ny].diagonal()
            up = Anw[:ny, ny:2 * ny].diagonal()
            S1 = np.linalg.inv(Anw[ny:2 * ny, ny:2 * ny].toarray() - (lo[:, None] * np.linalg.inv(D0)) * up[None, :])
            np.testing.assert_allclose(S[1], S1, rtol=0, atol=1e-4 * np.abs(S1).max())


@pytest.mark.parametrize('kind', ['btd', 'bcr'])
@pytest.mark.parametrize('pol,shape,npml,bloch', CASES_BCR)
def test_direct_solves_match_spsolve(pol, shape, npml, bloch, kind):
    """A x = b and A^T x = b to the FP64 residual 1e-10 (the x-wrap restored
    by the outer FGMRES)."""
    from ceviche_amd._fdfd_op import Direct, SolverOptions
    op = _op(pol, shape, npml, bloch)
    rng = np.random.default_rng(4)
    eps = rng.uniform(1, 12.25, shape)
    m0, m1 = op.materials(torch.from_numpy(eps.ravel()))
    D = Direct(op, m0, m1, kind=kind)
    A = _scaled_A(op, pol, eps, npml, bloch)
    b = rng.standard_normal(eps.size) + 1j * rng.standard_normal(eps.size)
    for tr in (False, True):
        x = D.solve(torch.from_numpy(b).cuda(), m0, m1, transpose=tr, opts=SolverOptions(tol=1e-10)).cpu().numpy()
        info = op.last_info[0]
        ref = spl.spsolve((A.T if tr else A).tocsc(), b)
        assert info['status'] == 0 and info['rel_residual'] <= 1e-10, info
        assert np.linalg.norm(x - ref) <= 1e-8 * np.linalg.norm(ref), (tr, info)
        print(kind, pol, shape, 'transpose' if tr else 'forward', info, flush=True)


# tall grids (ny > 128 >= nx): factorised and solved on the transposed grid
CASES_TALL = [('ez', (40, 200), (5, 12), (0.3, -0.5)), ('ez', (9, 150), (2, 10), (0.0, 0.0)),
              ('ez', (128, 129), (10, 10), (0.0, 0.0)), ('hz', (33, 170), (4, 9), (0.0, 0.7)),
              ('hz', (12, 140), (0, 6), (0.2, 0.0))]


@pytest.mark.parametrize('kind', ['btd', 'bcr'])
@pytest.mark.parametrize('pol,shape,npml,bloch', CASES_TALL)
def test_tall_grids_solve_on_the_transposed_grid(pol, shape, npml, bloch, kind):
    """A x = b and A^T x = b for tall grids (the factor and the FGMRES on the
    transposed operator, vectors transposed in and out) vs spsolve of the
    user's operator."""
    from ceviche_amd._fdfd_op import Direct, SolverOptions, direct_applies
    op = _op(pol, shape, npml, bloch)
    assert direct_applies(op)
    rng = np.random.default_rng(5)
    eps = rng.uniform(1, 12.25, shape)
    m0, m1 = op.materials(torch.from_numpy(eps.ravel()))
    D = Direct(op, m0, m1, kind=kind)
    assert D.tr and D.op.shapes[0] == (shape[1], shape[0])
    A = _scaled_A(op, pol, eps, npml, bloch)
    b = rng.standard_normal(eps.size) + 1j * rng.standard_normal(eps.size)
    for tr in (False, True):
        x = D.solve(torch.from_numpy(b).cuda(), m0, m1, transpose=tr, opts=SolverOptions(tol=1e-10)).cpu().numpy()
        info = op.last_info[0]
        ref = spl.spsolve((A.T if tr else A).tocsc(), b)
        assert info['status'] == 0 and info['rel_residual'] <= 1e-10, info
        assert info['factor'] == kind + '-transposed'
        assert np.linalg.norm(x - ref) <= 1e-8 * np.linalg.norm(ref), (tr, info)


def test_tall_grid_fields_and_gradient():
    """fdfd_ez on a 48 x 180 grid takes the transposed direct path by default:
    fields vs spsolve, the eps gradient vs the oracle adjoint."""
    import ceviche_amd as cv
    rng = np.random.default_rng(8)
    shape, npml = (48, 180), [6, 10]
    eps0 = rng.uniform(1, 6, shape)
    src = np.zeros(shape)
    src[:, 30] = 1.0
    eps = torch.tensor(eps0, requires_grad=True)
    F = cv.fdfd_ez(OMEGA, 50e-9, eps, npml)
    _, _, Ez = F.solve(torch.from_numpy(src))
    assert F.last_solve_info['method'] == 'direct' and F.last_solve_info['factor'] == 'bcr-transposed'
    _, _, rez = R.solve_ez(OMEGA, 50e-9, eps0, npml, src)
    ez = Ez.detach().cpu().numpy()
    assert np.linalg.norm(ez - rez) <= 1e-8 * np.linalg.norm(rez)
    (Ez[30, 150].abs() ** 2).backward()
    cot = np.zeros(shape, dtype=complex)
    cot[30, 150] = 2 * ez[30, 150]
    gref, _ = R.grad_eps_ez(OMEGA, 50e-9, eps0, npml, src, cot)
    assert np.linalg.norm(eps.grad.numpy() - gref) <= 1e-6 * np.linalg.norm(gref)


def test_direct_path_fields_gradient_and_forward_mode():
    """fdfd_ez / fdfd_hz through the direct path: fields vs spsolve, the eps
    gradient (one factorisation for the forward and adjoint solves) vs the
    oracle adjoint, forward mode vs central differences."""
    import ceviche_amd as cv
    rng = np.random.default_rng(6)
    shape, npml = (40, 36), [6, 6]
    eps0 = rng.uniform(1, 6, shape)
    src = np.zeros(shape)
    src[10, :] = 1.0
    F = cv.fdfd_ez(OMEGA, 50e-9, eps0, npml)
    _, _, ez = F.solve(src)
    assert F.last_solve_info['method'] == 'direct'
    _, _, rez = R.solve_ez(OMEGA, 50e-9, eps0, npml, src)
    assert np.linalg.norm(ez - rez) <= 1e-8 * np.linalg.norm(rez)
    eps = torch.tensor(eps0, requires_grad=True)
    F = cv.fdfd_ez(OMEGA, 50e-9, eps, npml)
    _, _, Ez = F.solve(torch.from_numpy(src))
    (Ez[30, 18].abs() ** 2).backward()
    cot = np.zeros(shape, dtype=complex)
    cot[30, 18] = 2 * Ez.detach().cpu().numpy()[30, 18]
    gref, _ = R.grad_eps_ez(OMEGA, 50e-9, eps0, npml, src, cot)
    assert np.linalg.norm(eps.grad.numpy() - gref) <= 1e-6 * np.linalg.norm(gref)
    G = cv.fdfd_hz(OMEGA, 50e-9, eps0, npml)

    def J(c):
        c = torch.as_tensor(c, dtype=torch.float64)
        G.eps_r = c * torch.from_numpy(eps0)
        a, b, z = G.solve(c * torch.from_numpy(src))
        return (z.abs() ** 2).reshape(-1)

    jf = cv.jacobian(J, mode='forward')(1.0)
    h = 1e-6
    num = (((np.abs(R.solve_hz(OMEGA, 50e-9, (1 + h) * eps0, npml, (1 + h) * src)[2]) ** 2)
            - (np.abs(R.solve_hz(OMEGA, 50e-9, (1 - h) * eps0, npml, (1 - h) * src)[2]) ** 2)) / (2 * h)).ravel()
    assert np.linalg.norm(jf[:, 0] - num) <= 1e-5 * np.linalg.norm(num)


def _nowrap(A, nx, ny):
    """A without its periodic wrap entries (what the nested-dissection factor inverts)."""
    C = A.tocoo()
    ri, rj, ci, cj = C.row // ny, C.row % ny, C.col // ny, C.col % ny
    keep = (np.abs(ri - ci) + np.abs(rj - cj)) <= 1
    return sp.coo_matrix((C.data[keep], (C.row[keep], C.col[keep])), shape=A.shape).tocsr()


@pytest.mark.parametrize('pol', ['ez', 'hz'])
@pytest.mark.parametrize('shape,npml,tol', [((150, 140), (10, 12), 0.15), ((37, 210), (5, 9), 1e-2),
                                            ((9, 7), (1, 1), 1e-4)])
def test_nested_dissection_factor_inverts_the_unwrapped_operator(pol, shape, npml, tol):
    """cv_nd_apply with the complex64 factor against the scaled operator
    without its periodic wrap, A x = b and A^T x = b, to complex64 accuracy
    as the conditioning amplifies it: the 150 x 140 random medium with a
    Bloch phase reaches 2.5e-3 (Ez) and 6.6e-2 (Hz) in a complex64 CPU
    emulation of the same factorisation, and a wrong factor gives errors of
    order one; FGMRES with the FP64 residual removes the rounding (the
    solves below reach 1e-10)."""
    from ceviche_amd._fdfd_op import NdDirect
    op = _op(pol, shape, npml, bloch=(0.3, 0.0))
    rng = np.random.default_rng(17)
    eps = rng.uniform(1, 12.25, shape)
    m0, m1 = op.materials(torch.from_numpy(eps.ravel()))
    nd = NdDirect(op, m0, m1)
    An = _nowrap(_scaled_A(op, pol, eps, npml, (0.3, 0.0)), *shape)
    x = rng.standard_normal(eps.size) + 1j * rng.standard_normal(eps.size)
    for tr in (False, True):
        b = (An.T if tr else An) @ x
        y = nd.apply(torch.from_numpy(b).cuda(), transpose=tr).cpu().numpy().astype(np.complex128)
        err = np.linalg.norm(y - x) / np.linalg.norm(x)
        print('nd', pol, shape, 'transpose', tr, 'err %.2e' % err, flush=True)
        assert err <= tol


@pytest.mark.parametrize('pol', ['ez', 'hz'])
def test_wide_grids_take_the_nested_dissection_path(pol):
    """A grid wider than 128 in both directions: method 'auto' (and
    'direct') factorises it by nested dissection; fields against spsolve and
    the adjoint gradient against the oracle / finite differences."""
    import ceviche_amd as cv
    rng = np.random.default_rng(23)
    shape, npml = (160, 150), (12, 10)
    eps0 = rng.uniform(1, 6, shape)
    src = np.zeros(shape)
    src[40, 30:120] = 1.0
    cls = cv.fdfd_ez if pol == 'ez' else cv.fdfd_hz
    for method in ('auto', 'direct'):
        F = cls(OMEGA, 50e-9, eps0, npml, solver_options=cv.SolverOptions(method=method))
        out = F.solve(src)
        info = F.last_solve_info
        print('wide', pol, method, info, flush=True)
        assert info['method'] == 'direct' and info['factor'] == 'nd' and info['rel_residual'] <= 1e-10
        ref = (R.solve_ez if pol == 'ez' else R.solve_hz)(OMEGA, 50e-9, eps0, npml, src)
        for a, b in zip(out, ref):
            assert np.linalg.norm(a - b) / np.linalg.norm(b) <= 1e-8
    if pol == 'ez':
        eps = torch.tensor(eps0, requires_grad=True)
        F = cls(OMEGA, 50e-9, eps, npml)
        _, _, Ez = F.solve(torch.from_numpy(src))
        (Ez[100, 75].abs() ** 2).backward()
        ez = Ez.detach().cpu().numpy()
        cot = np.zeros_like(ez)
        cot[100, 75] = 2 * ez[100, 75]
        gref, _ = R.grad_eps_ez(OMEGA, 50e-9, eps0, npml, src, cot)
        assert np.linalg.norm(eps.grad.numpy() - gref) <= 1e-6 * np.linalg.norm(gref)


def test_direct_failure_falls_back_to_bicgstab():
    """The direct path forced to fail from the host (its FGMRES limited to one
    step, far from the 1e-12 tolerance): under method='auto' the solve and its
    gradient fall back to the Jacobi-preconditioned Krylov method (BiCGSTAB)
    with a warning and still match spsolve; under
    method='direct' the failure raises."""
    import warnings
    import ceviche_amd as cv
    from ceviche_amd._lib import NotConvergedError
    from tests._cfg import cfg1
    c = cfg1()
    opts = cv.SolverOptions(tol=1e-12, direct_max_iter=1)
    eps = torch.tensor(c['eps'], requires_grad=True)
    F = cv.fdfd_ez(c['omega'], c['dL'], eps, c['npml'], solver_options=opts)
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter('always')
        _, _, Ez = F.solve(torch.from_numpy(c['src']))
        info = F.last_solve_info
        (Ez[90, 60].abs() ** 2).backward()
    assert any('falling back' in str(x.message) for x in w)
    assert info['method'] == 'bicgstab' and info['rel_residual'] <= 1e-12
    hx, hy, ez = R.solve_ez(c['omega'], c['dL'], c['eps'], c['npml'], c['src'])
    assert np.linalg.norm(Ez.detach().cpu().numpy() - ez) / np.linalg.norm(ez) <= 1e-6
    cot = np.zeros_like(ez)
    cot[90, 60] = 2 * ez[90, 60]
    gref, _ = R.grad_eps_ez(c['omega'], c['dL'], c['eps'], c['npml'], c['src'], cot)
    assert np.linalg.norm(eps.grad.cpu().numpy() - gref) / np.linalg.norm(gref) <= 1e-4
    F2 = cv.fdfd_ez(c['omega'], c['dL'], c['eps'], c['npml'],
                    solver_options=cv.SolverOptions(tol=1e-12, direct_max_iter=1, method='direct'))
    with pytest.raises(NotConvergedError):
        F2.solve(c['src'])


@pytest.mark.parametrize('N', [200, 300, 400, 600])
def test_nested_dissection_pivots_on_blob_designs(N):
    """One blob design per size class of the multi-frequency batch (dL 40 nm,
    npml 20, lambda 1.50 um): their separator fronts need row interchanges in
    the blocked LU of the fronts above 64 eliminated cells, where an
    interchange applied wrongly leaves a zero pivot or an inaccurate factor.
    Under method='direct' (no fallback) the factorisation must succeed and
    FGMRES with it reach 1e-10 in at most 40 steps; at 200^2 the field
    matches spsolve."""
    import ceviche_amd as cv
    from ceviche_amd import workloads
    d = workloads.fdfd_blob_batch(n_inst=1, n_lambda=1, sizes=(N,))[0]
    print('nd blob', N, flush=True)
    F = cv.fdfd_ez(d['omega'], d['dL'], d['eps'], d['npml'],
                   solver_options=cv.SolverOptions(tol=1e-10, method='direct'))
    _, _, Ez = F.solve(torch.from_numpy(d['src']))
    info = F.last_solve_info
    print('  method', info['method'], 'iterations', info['iterations'], 'residual %.2e' % info['rel_residual'],
          flush=True)
    assert info.get('factor') == 'nd' and info['rel_residual'] <= 1e-10 and info['iterations'] <= 40
    if N == 200:
        ez = R.solve_ez(d['omega'], d['dL'], d['eps'], d['npml'], d['src'])[2]
        assert np.linalg.norm(Ez.cpu().numpy() - ez) / np.linalg.norm(ez) <= 1e-8


def test_instance_worker_autotune_keeps_the_fields():
    """autotune_instance_workers times the default solve of an uncoupled batch
    (six instances, uneven shapes and frequencies) per candidate count of
    instances solved at once and keeps the fastest for the device; every
    instance is solved on its own, so the fields with the chosen count match
    one instance at a time (to 1e-12)."""
    import ceviche_amd._fdfd_op as fo
    from ceviche_amd import autotune_instance_workers
    from ceviche_amd._fdfd_op import FdfdOperator, FdfdSolve, SolverOptions
    rng = np.random.default_rng(5)
    shapes = [(40, 56), (64, 48), (52, 52), (36, 70), (60, 60), (48, 40)]
    specs = [dict(omega=OMEGA * (1 + 0.01 * i), dL=50e-9, shape=sh, npml=(6, 6)) for i, sh in enumerate(shapes)]
    op = FdfdOperator('ez', specs)
    eps = torch.from_numpy(np.concatenate([rng.uniform(1, 6, sh).ravel() for sh in shapes])).cuda()
    b = torch.from_numpy(np.concatenate([rng.standard_normal(sh).ravel() + 0j for sh in shapes])).cuda()
    opts = SolverOptions(tol=1e-10)
    try:
        times = autotune_instance_workers(op, eps, b, candidates=(1, 2, 4), opts=opts)
        print('autotune', times, flush=True)
        assert set(times) == {1, 2, 4}
        assert fo.instance_workers(b.device) == min(times, key=times.get)
        m0, m1 = op.materials(eps)
        x_tuned = FdfdSolve.apply(m0, m1, b, op, None, opts)
        assert all(i['method'] == 'direct' for i in op.last_info)
        fo._TUNED_WORKERS[fo._device_key(b.device)] = 1
        x_one = FdfdSolve.apply(m0, m1, b, op, None, opts)
        assert torch.linalg.vector_norm(x_tuned - x_one) <= 1e-12 * torch.linalg.vector_norm(x_one)
    finally:
        fo._TUNED_WORKERS.clear()


def test_nested_dissection_root_above_1024_cells():
    """A 1100 x 1100 grid: the root separator's interior block holds more than
    1024 cells, so its first panels take the 1024-thread panel kernel (two
    rows per thread) and its last panel is 12 columns wide.  Under
    method='direct' the factorisation must succeed and FGMRES reach 1e-10 in
    at most 40 steps."""
    import ceviche_amd as cv
    N = 1100
    rng = np.random.default_rng(21)
    eps = np.ones((N, N))
    c0, c1 = N // 3, 2 * N // 3
    eps[c0:c1, c0:c1] = rng.uniform(1, 4, (c1 - c0, c1 - c0))
    src = np.zeros((N, N))
    src[N // 2, N // 4:N // 4 + 20] = 1.0
    print('nd root 1100', flush=True)
    F = cv.fdfd_ez(OMEGA, 40e-9, eps, [20, 20], solver_options=cv.SolverOptions(tol=1e-10, method='direct'))
    F.solve(torch.from_numpy(src))
    info = F.last_solve_info
    print('  method', info['method'], 'iterations', info['iterations'], 'residual %.2e' % info['rel_residual'],
          flush=True)
    assert info.get('factor') == 'nd' and info['rel_residual'] <= 1e-10 and info['iterations'] <= 40
