"""sp_solve above the dense-LU limit and the solver hook solve_linear /
_solve_cuda (ceviche/primitives.py:103-148, solvers.py:37-79) on the device
CSR BiCGSTAB, against scipy spsolve on the oracle's assembled operators."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch
from scipy.sparse.linalg import spsolve

from oracle import fdfd_ref as R
from tests._cfg import cfg1

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _cfg1_system():
    c = cfg1()
    A = R.make_A_ez(c['omega'], c['dL'], c['eps'], c['npml']).tocoo()
    b = 1j * c['omega'] * c['src'].ravel()
    return A, b


def test_sp_solve_cfg1_operator_matches_spsolve_with_vjp_and_jvp():
    """N = 14 400 > 8192: the reference's own fdfd._make_A -> sp_solve route
    (fdfd.py:172-192) with the operator given as COO entries (diagonal
    entries duplicated, as _make_A builds it)."""
    from ceviche_amd.primitives import sp_solve
    import ceviche_amd.primitives as P
    A, b = _cfg1_system()
    # duplicate the diagonal like _make_A's hstack of D and eps0 w^2 eps terms
    n = A.shape[0]
    d = A.diagonal()
    entries = np.concatenate([A.data, 0.25 * d, -0.25 * d])
    indices = np.hstack([np.vstack((A.row, A.col)), np.vstack((np.arange(n), np.arange(n))),
                         np.vstack((np.arange(n), np.arange(n)))])
    print('sp_solve N', n, 'nnz', entries.size, flush=True)
    x_ref = spsolve(A.tocsc(), b)
    x = sp_solve(entries, indices, b)
    assert isinstance(x, np.ndarray) and x.dtype == np.complex128
    assert _rel(x, x_ref) <= 1e-6
    # reverse mode: L = |x[p]|^2; torch cotangent g = 2 x[p] e_p, lam = A^-H g
    p = 90 * 120 + 60
    e_t = torch.tensor(entries, requires_grad=True)
    b_t = torch.tensor(b, requires_grad=True)
    xt = sp_solve(e_t, torch.from_numpy(indices), b_t)
    (xt[p].abs() ** 2).backward()
    g = np.zeros(n, dtype=complex)
    g[p] = 2 * x_ref[p]
    lam = spsolve(A.tocsc().conj().T.tocsc(), g)
    assert _rel(b_t.grad.numpy(), lam) <= 1e-6
    ge_ref = -lam[indices[0]] * np.conj(x_ref[indices[1]])
    assert _rel(e_t.grad.numpy(), ge_ref) <= 1e-6
    # forward mode: d x = A^-1 (db - dA x)
    import torch.autograd.forward_ad as fwAD
    rng = np.random.default_rng(0)
    db = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    de = 1e-3 * (rng.standard_normal(entries.size) + 1j * rng.standard_normal(entries.size)) * np.abs(entries).max()
    with fwAD.dual_level():
        et = fwAD.make_dual(torch.tensor(entries).cuda(), torch.tensor(de).cuda())
        bt = fwAD.make_dual(torch.tensor(b).cuda(), torch.tensor(db).cuda())
        out = sp_solve(et, torch.from_numpy(indices), bt)
        tang = fwAD.unpack_dual(out).tangent.cpu().numpy()
    dA = sp.coo_matrix((de, (indices[0], indices[1])), shape=(n, n)).tocsc()
    t_ref = spsolve(A.tocsc(), db - dA @ x_ref)
    assert _rel(tang, t_ref) <= 1e-6


def test_sp_solve_general_sparse_matrix_with_duplicates():
    """A random (non-FDFD) complex sparse matrix, N = 9000, random pattern with
    repeated positions plus a dominant diagonal."""
    from ceviche_amd.primitives import sp_solve
    rng = np.random.default_rng(3)
    n, m = 9000, 60000
    rows = rng.integers(0, n, m)
    cols = rng.integers(0, n, m)
    vals = rng.standard_normal(m) + 1j * rng.standard_normal(m)
    rows = np.concatenate([rows, np.arange(n)])
    cols = np.concatenate([cols, np.arange(n)])
    vals = np.concatenate([vals, 20 + 5j + rng.standard_normal(n)])
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    A = sp.coo_matrix((vals, (rows, cols)), shape=(n, n)).tocsc()
    x = sp_solve(vals, np.vstack((rows, cols)), b)
    assert _rel(x, spsolve(A, b)) <= 1e-8


def test_solve_linear_and_solve_cuda_on_a_scipy_csr():
    from ceviche_amd import solvers as S
    A, b = _cfg1_system()
    Acsr = A.tocsr()
    assert Acsr.indices.dtype == np.int32
    x_ref = spsolve(A.tocsc(), b)
    x = S._solve_cuda(Acsr, b)
    assert _rel(x, x_ref) <= 1e-6
    assert S.last_solve_info['rel_residual'] <= 1e-10
    xd = S.solve_linear(Acsr, b)
    assert _rel(xd, x_ref) <= 1e-6
    # the reference's iterative stopping rule: ||r|| <= max(1e-5 ||b||, 1e-8)
    xi = S.solve_linear(Acsr, b, iterative_method='bicgstab')
    assert np.linalg.norm(b - Acsr @ xi) <= 1e-5 * np.linalg.norm(b) * (1 + 1e-9)
    with pytest.raises(ValueError):
        S.solve_linear(Acsr, b, iterative_method='not-a-method')
    # a small system takes the exact dense LU route
    As = sp.random(50, 50, density=0.2, random_state=1) + sp.eye(50) * (4 + 1j)
    bs = np.arange(50) + 1j
    assert _rel(S.solve_linear(As.tocsr(), bs), spsolve(As.tocsc(), bs)) <= 1e-12
