"""Numerical differentiation helpers of the reference.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

grad_num / der_num follow ceviche/utils.py:108-125 (central differences with
real and imaginary steps: the result is dL/dx - i dL/dy for complex inputs,
which is the autograd cotangent convention).  jacobian_numerical follows
ceviche/jacobians.py:54-73 (one-sided, shape (n_out, m_in)).  Pinned by
exact derivatives of polynomials in tests/test_oracle_primitives.py.
"""
import numpy as np


def der_num(fn, arg, index, delta):
    """Central difference of fn at arg along one entry; utils.py:108-115."""
    fwd = arg.copy()
    bwd = arg.copy()
    fwd[index] += delta / 2
    bwd[index] -= delta / 2
    return (fn(fwd) - fn(bwd)) / delta


def grad_num(fn, arg, delta=1e-6):
    """Complex numerical gradient; utils.py:117-125."""
    arg = np.asarray(arg, dtype=np.complex128)
    g = np.zeros(arg.size, dtype=np.complex128)
    for i in range(arg.size):
        g[i] = der_num(fn, arg, i, delta)
        g[i] += der_num(fn, arg, i, 1j * delta)
    return g


def jacobian_numerical(fn, x, step_size=1e-7):
    """One-sided numerical jacobian (n_out, m_in); jacobians.py:54-73."""
    x = np.asarray(x, dtype=float) if not np.iscomplexobj(x) else np.asarray(x)
    flat = np.atleast_1d(x).ravel()
    out0 = np.atleast_1d(fn(x)).ravel()
    jac = np.zeros((out0.size, flat.size))
    for i in range(flat.size):
        xi = flat.copy()
        xi[i] += step_size
        oi = np.atleast_1d(fn(xi.reshape(np.shape(x)))).ravel()
        jac[:, i] = np.real((oi - out0) / step_size)
    return jac
