"""The reference's solver hook (ceviche/solvers.py:37-79) on the GPU.

    x = solve_linear(A, b, iterative_method=False)

A is any scipy.sparse matrix (converted to CSR, complex128) and b a numpy
vector, as in the reference; x comes back as a numpy array.  Every branch runs
on the device:

* direct (``iterative_method`` false): the dense LU kernels for n <= 8192
  (an exact factorisation like spsolve, solvers.py:50-62); above that the
  Jacobi-preconditioned BiCGSTAB of ``cv_csr_solve`` to a relative FP64
  residual of ``DIRECT_TOL`` (1e-10), verified on a freshly computed
  residual.  It raises NotConvergedError rather than return an unconverged
  result.
* iterative (a method name from ITERATIVE_METHODS, or None for the default):
  the same device BiCGSTAB to the reference's stopping rule, scipy's
  ``||b - A x|| <= max(1e-5 ||b||, ATOL)`` (solvers.py:64-75, scipy's default
  rtol with atol = 1e-8), whatever Krylov method is named (names are checked
  like the reference's, ValueError for unknown ones).
* ``_solve_cuda(A, b)`` (solvers.py:77-79): the device BiCGSTAB on
  A.indptr / A.indices / A.data, to DIRECT_TOL.

The FDFD classes do not come through here: they solve their operators
matrix-free.
"""
import ctypes

import numpy as np
import scipy.sparse as sp
import torch

from . import _lib
from ._lib import check, lib, ptr, stream_ptr

DEFAULT_ITERATIVE_METHOD = 'bicg'
ITERATIVE_METHODS = ('bicg', 'bicgstab', 'cg', 'cgs', 'gmres', 'lgmres', 'qmr', 'gcrotmk')
ATOL = 1e-8
DIRECT_TOL = 1e-10
MAX_ITER = 200000

last_solve_info = None


def csr_solve(indptr, indices, data, b, tol=DIRECT_TOL, max_iter=MAX_ITER, x0=None, check_every=32):
    """x with ||b - A x|| <= tol ||b|| (FP64) for the CSR matrix (indptr,
    indices, data) -- device tensors, indices int32 or int64 (the same type for
    both), data and b complex128 -- by Jacobi-preconditioned BiCGSTAB
    (cv_csr_solve).  Returns (x, info dict); raises on non-convergence."""
    global last_solve_info
    n = int(b.numel())
    if indptr.numel() != n + 1:
        raise ValueError('indptr has %d entries for %d rows' % (indptr.numel(), n))
    if indptr.dtype != indices.dtype or indptr.dtype not in (torch.int32, torch.int64):
        raise ValueError('indptr and indices must both be int32 or both int64')
    bits = 32 if indptr.dtype == torch.int32 else 64
    dev = b.device
    data = data.to(dev, torch.complex128).resolve_conj().contiguous()
    b = b.to(torch.complex128).resolve_conj().contiguous()
    x = torch.zeros_like(b) if x0 is None else x0.to(dev, torch.complex128).clone().contiguous()
    ws = torch.empty(int(lib().cv_csr_solve_workspace_bytes(n)), dtype=torch.uint8, device=dev)
    info = _lib.SolveInfo()
    rc = lib().cv_csr_solve(n, bits, ptr(indptr.contiguous()), ptr(indices.contiguous()), ptr(data), ptr(b), ptr(x),
                            float(tol), int(max_iter), int(check_every), ptr(ws), ws.numel(), ctypes.byref(info),
                            stream_ptr())
    last_solve_info = dict(iterations=info.iterations, status=info.status, restarts=info.restarts,
                           rel_residual=info.rel_residual, method='bicgstab-jacobi-csr')
    check(rc, 'csr solve')
    return x, last_solve_info


def _device_csr(A):
    A = sp.csr_matrix(A, dtype=np.complex128)
    A.sort_indices()
    dev = _lib.device()
    it = np.int32 if A.indices.dtype == np.int32 and A.indptr.dtype == np.int32 else np.int64
    indptr = torch.from_numpy(np.ascontiguousarray(A.indptr, dtype=it)).to(dev)
    indices = torch.from_numpy(np.ascontiguousarray(A.indices, dtype=it)).to(dev)
    data = torch.from_numpy(np.ascontiguousarray(A.data)).to(dev)
    return A, indptr, indices, data


def _solve_cuda(A, b, tol=DIRECT_TOL, max_iter=MAX_ITER, **kwargs):
    """x = A^-1 b on the GPU from the CSR arrays of A (solvers.py:77-79)."""
    A, indptr, indices, data = _device_csr(A)
    bt = torch.from_numpy(np.ascontiguousarray(b, dtype=np.complex128).ravel()).to(indptr.device)
    x, _ = csr_solve(indptr, indices, data, bt, tol=tol, max_iter=max_iter)
    return x.cpu().numpy()


def _solve_direct(A, b):
    """Direct solve (solvers.py:50-62): dense LU on the device for n <= 8192,
    else the device BiCGSTAB to DIRECT_TOL."""
    from .primitives import DENSE_MAX_N, sp_solve
    A = sp.coo_matrix(A)
    if A.shape[0] <= DENSE_MAX_N:
        return sp_solve(A.data.astype(np.complex128), np.vstack((A.row, A.col)), np.asarray(b, dtype=np.complex128))
    return _solve_cuda(A, b)


def _solve_iterative(A, b, iterative_method=DEFAULT_ITERATIVE_METHOD):
    """Iterative solve with the reference's stopping rule (solvers.py:64-75)."""
    if iterative_method not in ITERATIVE_METHODS:
        raise ValueError('iterative method {} not found.\n supported methods are:\n {}'.format(
            iterative_method, ITERATIVE_METHODS))
    bn = float(np.linalg.norm(np.asarray(b)))
    tol = max(1e-5, ATOL / bn) if bn > 0 else 1e-5
    return _solve_cuda(A, b, tol=tol)


def solve_linear(A, b, iterative_method=False):
    """Dispatch of solvers.py:37-48."""
    if iterative_method and iterative_method is not None:
        return _solve_iterative(A, b, iterative_method=iterative_method)
    elif iterative_method and iterative_method is None:
        return _solve_iterative(A, b, iterative_method=DEFAULT_ITERATIVE_METHOD)
    else:
        return _solve_direct(A, b)


__all__ = ['solve_linear', 'csr_solve', 'DEFAULT_ITERATIVE_METHOD', 'ITERATIVE_METHODS', 'ATOL']
