"""Differentiable sparse primitives with the reference's interface
(ceviche/primitives.py:56-259):

    sp_mult(entries, indices, x)                      -> A x
    sp_solve(entries, indices, b)                     -> A^-1 b
    spsp_mult(entries_a, indices_a, entries_x, indices_x, N) -> (entries_b, indices_b) of A X

``entries`` (M,) complex, ``indices`` (2, M) integer [rows; cols]; duplicates
are summed (ceviche/utils.py:13-25); matrices are N x N with N = x.size /
b.size.  numpy inputs give numpy outputs; torch inputs give GPU tensors that
carry torch autograd: ``backward`` (VJP) and ``jvp`` (forward mode) rules are
the reference's defvjp/defjvp rules (primitives.py:74-98, 121-148, 178-257)
written in torch's complex convention (cotangent dL/dRe + i dL/dIm, hence
A^H where the reference has A^T).

All numeric work runs in HIP kernels (segmented product sums, dense LU,
CSR BiCGSTAB); the host only groups index sets (the symbolic analysis), once
per index set.
"""
import numpy as np
import torch

from . import _lib
from ._lib import check, lib, ptr, stream_ptr
from .utils import (make_sparse, transpose_indices, make_rand, make_rand_complex, make_rand_indeces,  # noqa: F401
                    make_rand_sparse, der_num, grad_num, get_entries_indices, make_IO_matrices)

DENSE_MAX_N = 8192


def _dev():
    return _lib.device()


def _i64(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.int64)).to(_dev())


try:
    import xxhash

    def _digest(buf):
        return xxhash.xxh3_128_intdigest(buf)
except ImportError:      # pragma: no cover - xxhash ships with this environment
    import hashlib

    def _digest(buf):
        return hashlib.blake2b(buf, digest_size=16).digest()


def _index_np(indices):
    """(2, M) int64 numpy view of an index set (copied only when needed)."""
    a = indices.detach().cpu().numpy() if isinstance(indices, torch.Tensor) else np.asarray(indices)
    return np.ascontiguousarray(a, dtype=np.int64).reshape(2, -1)


def _index_key(indices):
    """Cache key of an index set.  A torch tensor is identified by its memory,
    layout and in-place version counter (no copy to the host); any other array
    by a 128-bit digest of its int64 contents."""
    if isinstance(indices, torch.Tensor):
        return ('t', indices.data_ptr(), tuple(indices.shape), tuple(indices.stride()), str(indices.dtype),
                str(indices.device), indices._version)
    a = _index_np(indices)
    return ('n', a.shape, _digest(memoryview(a).cast('B')))


class _Groups:
    """Entries grouped by a key: ptr (n_groups + 1), order (entry indices)."""

    def __init__(self, keys, n_groups):
        order = np.argsort(keys, kind='stable')
        counts = np.bincount(keys, minlength=n_groups)
        self.n = int(n_groups)
        self.ptr = _i64(np.concatenate([[0], np.cumsum(counts)]))
        self.order_np = order
        self.order = _i64(order)


class Pattern:
    """Symbolic structure of an N x N COO index set."""

    _cache = {}

    def __init__(self, rows, cols, N):
        self.N = int(N)
        self.M = int(rows.size)
        if self.M and (rows.min() < 0 or cols.min() < 0 or rows.max() >= N or cols.max() >= N):
            raise ValueError('indices out of range for an %d x %d matrix' % (N, N))
        self.rows_np, self.cols_np = rows, cols
        self.rows, self.cols = _i64(rows), _i64(cols)
        self.by_row = _Groups(rows, N)
        self.by_col = _Groups(cols, N)
        self.cols_by_row = _i64(cols[self.by_row.order_np])
        self.rows_by_col = _i64(rows[self.by_col.order_np])
        self.unit = _i64(np.arange(self.M + 1))
        key = rows.astype(np.int64) * N + cols
        uniq, inv = np.unique(key, return_inverse=True)
        self.uniq = _Groups(inv.ravel(), uniq.size)
        # column-major dense position of each unique (row, col)
        self.dense_pos = _i64((uniq % N) * N + uniq // N)

    @classmethod
    def get(cls, indices, N):
        k = (int(N),) + _index_key(indices)
        p = cls._cache.get(k)
        if p is None:
            idx = _index_np(indices)
            p = cls(idx[0].copy(), idx[1].copy(), N)
            # a tensor key names its memory: keep the tensor alive with the entry
            p.keep = indices if isinstance(indices, torch.Tensor) else None
            if len(cls._cache) > 64:
                cls._cache.clear()
            cls._cache[k] = p
        return p


def segsum(n_seg, ptr_t, iu, iv, u, v, out, conj_u=False, conj_v=False, alpha=1.0, oidx=None, accumulate=False):
    check(lib().cv_segsum(int(n_seg), ptr(ptr_t), ptr(iu), ptr(iv), ptr(u), ptr(v), int(conj_u), int(conj_v),
                          float(np.real(alpha)), float(np.imag(alpha)), ptr(oidx), int(accumulate), ptr(out),
                          stream_ptr()), 'segsum')
    return out


def _c(t):
    return t.to(_dev(), torch.complex128).contiguous()


# --------------------------------------------------------------- kernels
def _spmv(P, e, x, transpose=False, herm=False):
    """A x (CSR rows), A^T x or A^H x (CSC columns)."""
    y = torch.empty(P.N, dtype=torch.complex128, device=_dev())
    if not (transpose or herm):
        return segsum(P.N, P.by_row.ptr, P.by_row.order, P.cols_by_row, e, x, y)
    return segsum(P.N, P.by_col.ptr, P.by_col.order, P.rows_by_col, e, x, y, conj_u=herm)


def _outer(P, a, b, conj_b=True, alpha=1.0):
    """out[k] = alpha a[row_k] conj?(b[col_k])."""
    out = torch.empty(P.M, dtype=torch.complex128, device=_dev())
    return segsum(P.M, P.unit, P.rows, P.cols, a, b, out, conj_v=conj_b, alpha=alpha)


def _densify(P, e):
    A = torch.zeros(P.N * P.N, dtype=torch.complex128, device=_dev())
    ones = torch.ones(1, dtype=torch.complex128, device=_dev())
    zeros_idx = torch.zeros(P.M, dtype=torch.int64, device=_dev())
    segsum(P.uniq.n, P.uniq.ptr, P.uniq.order, zeros_idx, e, ones, A, oidx=P.dense_pos)
    return A


KRYLOV_TOL = 1e-10
KRYLOV_MAX_ITER = 200000


class _Krylov:
    """sp_solve factor for N > DENSE_MAX_N: no factorisation; each solve is the
    device Jacobi-preconditioned BiCGSTAB (cv_csr_solve) on the CSR of A, or of
    A^T / A^H (the entries gathered in column order) for the reverse mode, to a
    relative FP64 residual of KRYLOV_TOL.  Non-convergence raises."""

    def __init__(self, P, e):
        self.P, self.e = P, e
        self.info = []

    def solve(self, b, trans=0):
        from .solvers import csr_solve
        P = self.P
        if trans == 0:
            indptr, col, val = P.by_row.ptr, P.cols_by_row, self.e[P.by_row.order]
        else:
            indptr, col, val = P.by_col.ptr, P.rows_by_col, self.e[P.by_col.order]
            if trans == 2:
                val = val.conj()
        x, info = csr_solve(indptr, col, val.resolve_conj(), _c(b), tol=KRYLOV_TOL, max_iter=KRYLOV_MAX_ITER)
        self.info.append(info)
        return x


def _factor(P, e):
    return _LU(P, e) if P.N <= DENSE_MAX_N else _Krylov(P, e)


class _LU:
    def __init__(self, P, e):
        self.n = P.N
        self.A = _densify(P, e)
        self.piv = torch.empty(P.N, dtype=torch.int32, device=_dev())
        info = torch.zeros(1, dtype=torch.int32, device=_dev())
        check(lib().cv_dense_lu(P.N, ptr(self.A), ptr(self.piv), ptr(info), stream_ptr()), 'dense LU')
        self.singular = int(info.item()) != 0

    def solve(self, b, trans=0):
        x = _c(b).clone()
        check(lib().cv_dense_lu_solve(self.n, ptr(self.A), ptr(self.piv), ptr(x), 1, trans, stream_ptr()),
              'dense LU solve')
        if self.singular:
            import warnings
            warnings.warn('Matrix is exactly singular', RuntimeWarning)
            x.fill_(complex('nan'))
        return x


# ---------------------------------------------------------------- sp_mult
class SpMult(torch.autograd.Function):

    @staticmethod
    def forward(ctx, entries, x, P):
        e, xx = _c(entries), _c(x)
        ctx.P = P
        ctx.save_for_backward(e, xx)
        ctx.save_for_forward(e, xx)
        return _spmv(P, e, xx)

    @staticmethod
    def backward(ctx, g):
        e, x = ctx.saved_tensors
        g = _c(g)
        ge = _outer(ctx.P, g, x) if ctx.needs_input_grad[0] else None     # g[i] conj(x[j])
        gx = _spmv(ctx.P, e, g, herm=True) if ctx.needs_input_grad[1] else None  # A^H g
        return ge, gx, None

    @staticmethod
    def jvp(ctx, de, dx, _):
        e, x = ctx.saved_tensors
        out = torch.zeros(ctx.P.N, dtype=torch.complex128, device=_dev())
        if de is not None:
            out = out + _spmv(ctx.P, _c(de), x)
        if dx is not None:
            out = out + _spmv(ctx.P, e, _c(dx))
        return out


# --------------------------------------------------------------- sp_solve
class SpSolve(torch.autograd.Function):

    @staticmethod
    def forward(ctx, entries, b, P):
        e = _c(entries)
        lu = _factor(P, e)
        x = lu.solve(b)
        ctx.P, ctx.lu = P, lu
        ctx.save_for_backward(e, x)
        ctx.save_for_forward(e, x)
        return x

    @staticmethod
    def backward(ctx, g):
        e, x = ctx.saved_tensors
        lam = ctx.lu.solve(_c(g), trans=2)  # A^H lam = g (one factorisation for all rules)
        ge = _outer(ctx.P, lam, x, alpha=-1.0) if ctx.needs_input_grad[0] else None
        gb = lam if ctx.needs_input_grad[1] else None
        return ge, gb, None

    @staticmethod
    def jvp(ctx, de, db, _):
        e, x = ctx.saved_tensors
        rhs = torch.zeros(ctx.P.N, dtype=torch.complex128, device=_dev())
        if db is not None:
            rhs = rhs + _c(db)
        if de is not None:
            rhs = rhs - _spmv(ctx.P, _c(de), x)
        return ctx.lu.solve(rhs)


# -------------------------------------------------------------- spsp_mult
class ProductPattern:
    """Symbolic SpGEMM B = A X: the products (ka, kx) that land in each output
    entry p, grouped three ways (by p for B, by ka and by kx for the VJPs)."""

    _cache = {}

    def __init__(self, ia, ix, N):
        ra, ca = ia
        rx, cx = ix
        ox = np.argsort(rx, kind='stable')
        rxs = rx[ox]
        lo = np.searchsorted(rxs, ca, 'left')
        hi = np.searchsorted(rxs, ca, 'right')
        cnt = hi - lo
        ka = np.repeat(np.arange(ra.size), cnt)
        start = np.repeat(lo, cnt)
        within = np.arange(ka.size) - np.repeat(np.cumsum(cnt) - cnt, cnt)
        kx = ox[start + within]
        key = ra[ka].astype(np.int64) * N + cx[kx]
        uniq, p = np.unique(key, return_inverse=True)
        p = p.ravel()
        self.N, self.n_out, self.n_prod = int(N), int(uniq.size), int(ka.size)
        self.indices_b = np.vstack((uniq // N, uniq % N)).astype(np.int64)
        byp = _Groups(p, uniq.size)
        self.p_ptr, self.p_ka, self.p_kx = byp.ptr, _i64(ka[byp.order_np]), _i64(kx[byp.order_np])
        bya = _Groups(ka, ra.size)
        self.a_ptr, self.a_p, self.a_kx = bya.ptr, _i64(p[bya.order_np]), _i64(kx[bya.order_np])
        byx = _Groups(kx, rx.size)
        self.x_ptr, self.x_p, self.x_ka = byx.ptr, _i64(p[byx.order_np]), _i64(ka[byx.order_np])
        self.Ma, self.Mx = int(ra.size), int(rx.size)

    @classmethod
    def get(cls, ia_in, ix_in, N):
        k = (int(N),) + _index_key(ia_in) + _index_key(ix_in)
        pp = cls._cache.get(k)
        if pp is None:
            ia, ix = _index_np(ia_in), _index_np(ix_in)
            for a, name in ((ia, 'A'), (ix, 'X')):
                if a.size and (a.min() < 0 or a.max() >= N):
                    raise ValueError('indices of %s out of range for an %d x %d matrix' % (name, N, N))
            pp = cls(ia, ix, N)
            pp.keep = tuple(a for a in (ia_in, ix_in) if isinstance(a, torch.Tensor))
            if len(cls._cache) > 64:
                cls._cache.clear()
            cls._cache[k] = pp
        return pp

    def numeric(self, ea, ex, out=None, accumulate=False):
        if out is None:
            out = torch.empty(self.n_out, dtype=torch.complex128, device=_dev())
        return segsum(self.n_out, self.p_ptr, self.p_ka, self.p_kx, ea, ex, out, accumulate=accumulate)


class SpspMult(torch.autograd.Function):

    @staticmethod
    def forward(ctx, ea, ex, PP):
        a, x = _c(ea), _c(ex)
        ctx.PP = PP
        ctx.save_for_backward(a, x)
        ctx.save_for_forward(a, x)
        return PP.numeric(a, x)

    @staticmethod
    def backward(ctx, g):
        a, x = ctx.saved_tensors
        PP = ctx.PP
        g = _c(g)
        ga = gx = None
        if ctx.needs_input_grad[0]:  # (V X^H)[ia, ja]: sum g[p] conj(x[kx]) over a's products
            ga = torch.empty(PP.Ma, dtype=torch.complex128, device=_dev())
            segsum(PP.Ma, PP.a_ptr, PP.a_p, PP.a_kx, g, x, ga, conj_v=True)
        if ctx.needs_input_grad[1]:  # (A^H V)[ix, jx]
            gx = torch.empty(PP.Mx, dtype=torch.complex128, device=_dev())
            segsum(PP.Mx, PP.x_ptr, PP.x_ka, PP.x_p, a, g, gx, conj_u=True)
        return ga, gx, None

    @staticmethod
    def jvp(ctx, dea, dex, _):
        a, x = ctx.saved_tensors
        PP = ctx.PP
        out = torch.zeros(PP.n_out, dtype=torch.complex128, device=_dev())
        if dea is not None:
            PP.numeric(_c(dea), x, out=out, accumulate=True)
        if dex is not None:
            PP.numeric(a, _c(dex), out=out, accumulate=True)
        return out


# ------------------------------------------------------------- public API
def _is_t(*a):
    return any(isinstance(v, torch.Tensor) for v in a)


def _to_t(a):
    if isinstance(a, torch.Tensor):
        return a.to(_dev())
    return torch.from_numpy(np.ascontiguousarray(a)).to(_dev())


def _out(t, as_torch):
    return t if as_torch else t.detach().cpu().numpy()


def sp_mult(entries, indices, x):
    """A x for A given by (entries, indices); primitives.py:58-72."""
    as_t = _is_t(entries, x)
    N = int(x.numel() if isinstance(x, torch.Tensor) else np.asarray(x).size)
    P = Pattern.get(indices, N)
    return _out(SpMult.apply(_to_t(entries), _to_t(x).reshape(-1), P), as_t)


def sp_solve(entries, indices, b):
    """A^-1 b; primitives.py:103-119.  N <= DENSE_MAX_N: dense LU with partial
    pivoting on the GPU (exact, like spsolve); larger N: the device
    Jacobi-preconditioned BiCGSTAB on the CSR of A to a relative FP64 residual
    of KRYLOV_TOL (the VJP and JVP solves likewise on A^H and A)."""
    as_t = _is_t(entries, b)
    N = int(b.numel() if isinstance(b, torch.Tensor) else np.asarray(b).size)
    P = Pattern.get(indices, N)
    return _out(SpSolve.apply(_to_t(entries), _to_t(b).reshape(-1), P), as_t)


def spsp_mult(entries_a, indices_a, entries_x, indices_x, N):
    """(entries_b, indices_b) of B = A X (N x N), entries in row-major order of
    the unique output positions; primitives.py:153-176."""
    as_t = _is_t(entries_a, entries_x)
    PP = ProductPattern.get(indices_a, indices_x, N)
    eb = SpspMult.apply(_to_t(entries_a), _to_t(entries_x), PP)
    ib = PP.indices_b
    if as_t:
        return eb, torch.from_numpy(ib).to(_dev())
    return eb.detach().cpu().numpy(), ib


__all__ = ['sp_mult', 'sp_solve', 'spsp_mult',
           # the helpers the reference's primitives module imports from its utils (primitives.py:5-6)
           'make_sparse', 'transpose_indices', 'make_rand', 'make_rand_complex', 'make_rand_indeces',
           'make_rand_sparse', 'der_num', 'grad_num', 'get_entries_indices', 'make_IO_matrices']
