"""Waveguide modes of a 1D cross section: the reference's ceviche/modes.py
(get_modes :10-47, insert_mode :50-66, solver_eigs :69-79, filter_modes
:82-108, normalize_modes :111-119), same names, arguments and results.

The eigenproblem is  A v = n_eff^2 v  with  A = diag(eps) + Dxf Dxb / k0^2
on the cross section (periodic / SC-PML differences of
ceviche/derivatives.py, shape (N, 1), npml on the one axis), solved for the
`m` eigenvalues closest to n_max^2 by shift-invert Arnoldi like scipy's
`eigs(A, k=m, sigma=n_max^2, which='LM')`: (A - sigma I) is factorised once by
the dense complex LU kernels of csrc/sparse.hip and every Arnoldi step is one
triangular solve pair on the GPU; the small Hessenberg eigenproblem is solved
on the host.  Results are ordered like scipy's (closest to sigma first).
Eigenvectors are defined up to a complex phase; here each is scaled to unit
2-norm with its largest-magnitude entry real and positive.

ceviche's Ez_to_H (:121-127) calls a function its module never defines and
cannot run; it is not restated.
"""
import numpy as np
import torch

from . import _lib
from ._lib import check, lib, ptr, stream_ptr
from .constants import C_0
from .derivatives import axis_profiles


def _operator(eps, omega, dL, npml):
    """Dense diag(eps) + Dxf Dxb / k0^2 (complex128, N x N, host)."""
    N = eps.size
    k0 = omega / C_0
    kf, kb = axis_profiles(omega, dL, N, int(npml))
    A = np.zeros((N, N), dtype=np.complex128)
    idx = np.arange(N)
    A[idx, idx] = eps
    if N == 1:
        A[0, 0] += kf[0] * kb[0] / k0 ** 2
        return A
    ip = (idx + 1) % N
    im = (idx - 1) % N
    # (Dxf Dxb u)[i] = kf[i] (kb[i+1] (u[i+1] - u[i]) - kb[i] (u[i] - u[i-1])), periodic
    np.add.at(A, (idx, ip), kf * kb[ip] / k0 ** 2)
    np.add.at(A, (idx, idx), -kf * (kb[ip] + kb) / k0 ** 2)
    np.add.at(A, (idx, im), kf * kb / k0 ** 2)
    return A


class _ShiftInvert:
    """LU of (A - sigma I) on the GPU; apply(v) = (A - sigma I)^-1 v."""

    def __init__(self, A, sigma):
        self.n = A.shape[0]
        dev = _lib.device()
        B = A - sigma * np.eye(self.n)
        # the LU kernels take column-major storage
        self.B = torch.from_numpy(np.ascontiguousarray(B.T)).to(dev)
        self.piv = torch.empty(self.n, dtype=torch.int32, device=dev)
        info = torch.zeros(1, dtype=torch.int32, device=dev)
        check(lib().cv_dense_lu(self.n, ptr(self.B), ptr(self.piv), ptr(info), stream_ptr()), 'dense LU')
        if int(info.item()) != 0:
            raise np.linalg.LinAlgError('A - sigma I is singular: sigma is an eigenvalue')

    def apply(self, v):
        x = v.clone()
        check(lib().cv_dense_lu_solve(self.n, ptr(self.B), ptr(self.piv), ptr(x), 1, 0, stream_ptr()),
              'dense LU solve')
        return x


def _arnoldi(op, n, k, v0):
    V = torch.zeros((k + 1, n), dtype=torch.complex128, device=v0.device)
    H = np.zeros((k + 1, k), dtype=np.complex128)
    V[0] = v0 / torch.linalg.vector_norm(v0)
    for j in range(k):
        w = op.apply(V[j])
        for _ in range(2):      # classical Gram-Schmidt, twice
            h = (V[:j + 1].conj() @ w)
            w = w - V[:j + 1].T @ h
            H[:j + 1, j] += h.cpu().numpy()
        H[j + 1, j] = float(torch.linalg.vector_norm(w))
        if H[j + 1, j] == 0.0:
            return V[:j + 1], H[:j + 1, :j + 1], 0.0, j + 1
        V[j + 1] = w / H[j + 1, j]
    return V[:k], H[:k, :k], H[k, k - 1], k


def solver_eigs(A, Neigs, guess_value=1.0):
    """The Neigs eigenpairs of A closest to guess_value, closest first (as
    scipy.sparse.linalg.eigs(A, k=Neigs, sigma=guess_value, which='LM'))."""
    A = A.toarray() if hasattr(A, 'toarray') else np.asarray(A, dtype=np.complex128)
    n = A.shape[0]
    if not 0 < Neigs < max(n - 1, 2):
        raise ValueError('Neigs must be in [1, N - 2] for an N x N matrix (N = %d)' % n)
    sigma = complex(guess_value)
    op = _ShiftInvert(A, sigma)
    rng = np.random.default_rng(0)
    v0 = torch.from_numpy(rng.standard_normal(n) + 1j * rng.standard_normal(n)).to(op.B.device)
    k = min(n, max(2 * Neigs + 20, 40))
    while True:
        V, H, h_next, kk = _arnoldi(op, n, k, v0)
        theta, Y = np.linalg.eig(H)
        order = np.argsort(-np.abs(theta), kind='stable')[:Neigs]
        # Arnoldi residual of each wanted Ritz pair of the shifted inverse
        res = np.abs(h_next) * np.abs(Y[kk - 1, order]) / np.maximum(np.abs(theta[order]), 1e-300)
        if kk < k or k >= n or np.all(res <= 1e-12):
            break
        k = min(n, 2 * k)
    th = theta[order]
    vals = sigma + 1.0 / th
    X = (V.T @ torch.from_numpy(np.ascontiguousarray(Y[:, order])).to(V.device)).cpu().numpy()
    X = X / np.linalg.norm(X, axis=0)
    big = np.argmax(np.abs(X), axis=0)
    X = X * (np.abs(X[big, np.arange(X.shape[1])]) / X[big, np.arange(X.shape[1])])
    return vals, X


def filter_modes(values, vectors, filters=None):
    """Keeps the modes whose values satisfy every filter (modes.py:82-108)."""
    if filters is None:
        return values, vectors
    keep = np.ones(values.shape, dtype=bool)
    for f in filters:
        keep = np.logical_and(keep, f(values))
    idx = np.where(keep)[0]
    return values[idx], vectors[:, idx]


def normalize_modes(vectors):
    """Each column scaled to sum(|v|^2) = 1 (modes.py:111-119)."""
    return vectors / np.sqrt(np.sum(np.square(np.abs(vectors)), axis=0))


def get_modes(eps_cross, omega, dL, npml, m=1, filtering=True):
    """Effective-index eigenvalues (n_eff^2) and mode profiles of the cross
    section eps_cross (modes.py:10-47): vals (m,), vecs (N, m)."""
    eps = np.asarray(eps_cross).flatten()
    A = _operator(eps.astype(np.complex128), omega, dL, npml)
    n_max = np.sqrt(np.max(eps_cross))
    vals, vecs = solver_eigs(A, m, guess_value=n_max ** 2)
    if filtering:
        vals, vecs = filter_modes(vals, vecs, filters=[lambda v: np.real(v) > 0.0])
    if vals.size == 0:
        raise BaseException("Could not find any eigenmodes for this waveguide")
    return vals, normalize_modes(vecs)


def insert_mode(omega, dx, x, y, epsr, target=None, npml=0, m=1, filtering=False):
    """Mode m of the cross section epsr[x, y] written into target[x, y]
    (a new complex array shaped like epsr if target is None); modes.py:50-66."""
    if target is None:
        target = np.zeros(epsr.shape, dtype=complex)
    _, mode_field = get_modes(epsr[x, y], omega, dx, npml, m=m, filtering=filtering)
    target[x, y] = np.atleast_2d(mode_field)[:, m - 1].squeeze()
    return target
