"""Sparse-primitive oracle: sp_mult, sp_solve, spsp_mult and their VJP/JVP rules.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Convention: these functions use the autograd convention of the reference
(cotangent of a complex z is dL/dx - i dL/dy, plain transposes), exactly as
ceviche/primitives.py:58-259 does, so they can be compared with the
reference's formulas directly.  Pinning (tests/test_oracle_primitives.py):
dense brute force at N = 8 and central finite differences in the pattern of
ceviche/tests/test_primitives.py:39-121 (np.random.seed(2), M = N^2 - 20).
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spl


def make_sparse(entries, indices, shape):
    """COO -> CSR, duplicates summed, complex128; ceviche/utils.py:13-25."""
    coo = sp.coo_matrix((entries, (indices[0], indices[1])), shape=shape, dtype=np.complex128)
    return coo.tocsr()


def get_entries_indices(csr):
    """CSR -> (data, [rows; cols]) in CSR storage order; ceviche/utils.py:27-35."""
    coo = csr.tocoo()
    return csr.data, np.vstack((coo.row, coo.col))


def transpose_indices(indices):
    """ceviche/utils.py:37-39."""
    return np.flip(indices, axis=0)


def make_IO_matrices(indices, N):
    """One-hot incidence matrices I (N x M) and O = J^T (M x N); utils.py:50-73."""
    M = indices.shape[1]
    ones = np.ones(M)
    ik, jk = indices
    I = make_sparse(ones, np.vstack((ik, np.arange(M))), (N, M))
    J = make_sparse(ones, np.vstack((jk, np.arange(M))), (N, M))
    return I, J.T


# ------------------------------------------------------------------ sp_mult
def sp_mult(entries, indices, x):
    """b = A x; primitives.py:58-72."""
    N = x.size
    return make_sparse(entries, indices, (N, N)).dot(x)


def sp_mult_vjp_entries(v, entries, indices, x):
    """primitives.py:74-79: v[i] * x[j]."""
    i, j = indices
    return v[i] * x[j]


def sp_mult_vjp_x(v, entries, indices, x):
    """primitives.py:81-86: A^T v."""
    return sp_mult(entries, transpose_indices(indices), v)


def sp_mult_jvp_entries(g, entries, indices, x):
    """primitives.py:90-92."""
    return sp_mult(g, indices, x)


def sp_mult_jvp_x(g, entries, indices, x):
    """primitives.py:94-96."""
    return sp_mult(entries, indices, g)


# ----------------------------------------------------------------- sp_solve
def sp_solve(entries, indices, b):
    """x = A^-1 b with SuperLU (ceviche's default path); primitives.py:103-119,
    solvers.py:50-62."""
    N = b.size
    A = make_sparse(entries, indices, (N, N))
    return spl.spsolve(A.tocsc(), b)


def sp_solve_vjp_entries(v, entries, indices, b):
    """primitives.py:121-128: (A^-T (-v))[i] * x[j]."""
    x = sp_solve(entries, indices, b)
    adj = sp_solve(entries, transpose_indices(indices), -v)
    i, j = indices
    return adj[i] * x[j]


def sp_solve_vjp_b(v, entries, indices, b):
    """primitives.py:130-135: A^-T v."""
    return sp_solve(entries, transpose_indices(indices), v)


def sp_solve_jvp_entries(g, entries, indices, b):
    """primitives.py:139-142: A^-1 (-(G x))."""
    x = sp_solve(entries, indices, b)
    return sp_solve(entries, indices, -sp_mult(g, indices, x))


def sp_solve_jvp_b(g, entries, indices, b):
    """primitives.py:144-146: A^-1 g."""
    return sp_solve(entries, indices, g)


# ---------------------------------------------------------------- spsp_mult
def spsp_mult(entries_a, indices_a, entries_x, indices_x, N):
    """B = A X as (entries, indices) in CSR order; primitives.py:153-176."""
    A = make_sparse(entries_a, indices_a, (N, N))
    X = make_sparse(entries_x, indices_x, (N, N))
    return get_entries_indices(A.dot(X))


def spsp_mult_vjp_a(v_entries, indices_b, entries_a, indices_a, entries_x, indices_x, N):
    """primitives.py:178-204: diag(I_a^T (V X^T) O_a^T) = (V X^T)[ia, ja]."""
    V = make_sparse(v_entries, indices_b, (N, N))
    X = make_sparse(entries_x, indices_x, (N, N))
    VXt = (V.dot(X.T)).toarray()
    ia, ja = indices_a
    return VXt[ia, ja]


def spsp_mult_vjp_x(v_entries, indices_b, entries_a, indices_a, entries_x, indices_x, N):
    """primitives.py:206-221: (A^T V)[ix, jx]."""
    V = make_sparse(v_entries, indices_b, (N, N))
    A = make_sparse(entries_a, indices_a, (N, N))
    AtV = (A.T.dot(V)).toarray()
    ix, jx = indices_x
    return AtV[ix, jx]


def spsp_mult_jvp_a(g, indices_b, entries_a, indices_a, entries_x, indices_x, N):
    """primitives.py:225-244: (G X) sampled at B's pattern."""
    G = make_sparse(g, indices_a, (N, N))
    X = make_sparse(entries_x, indices_x, (N, N))
    GX = (G.dot(X)).toarray()
    ib, jb = indices_b
    return GX[ib, jb]


def spsp_mult_jvp_x(g, indices_b, entries_a, indices_a, entries_x, indices_x, N):
    """primitives.py:246-257: (A G) sampled at B's pattern."""
    A = make_sparse(entries_a, indices_a, (N, N))
    G = make_sparse(g, indices_x, (N, N))
    AG = (A.dot(G)).toarray()
    ib, jb = indices_b
    return AG[ib, jb]


# ------------------------------------------------------------ test matrices
def make_rand(N):
    """ceviche/utils.py:78-80 (values in [-0.5, 0.5))."""
    return np.random.random(N) - 0.5


def make_rand_complex(N):
    """ceviche/utils.py:82-84."""
    return (np.random.random(N) - 0.5) + 1j * (np.random.random(N) - 0.5)


def make_rand_indices(N, M):
    """ceviche/utils.py:86-88."""
    return np.random.randint(low=0, high=N, size=(2, M))
