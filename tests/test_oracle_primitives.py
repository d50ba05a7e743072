"""Pins the primitive oracle (oracle/primitives_ref.py) by dense brute force and
central differences, in the set-up of ceviche tests/test_primitives.py
(np.random.seed(2), N = 8, M = N^2 - 20, random indices with duplicates)."""
import numpy as np
import pytest

from oracle import primitives_ref as P
from oracle.numdiff import grad_num


def _setup():
    np.random.seed(2)
    N = 8
    M = N ** 2 - 20
    ind = P.make_rand_indices(N, M)
    ent = P.make_rand_complex(M)
    ind2 = P.make_rand_indices(N, M - 2)
    ent2 = P.make_rand_complex(M - 2)
    x = P.make_rand_complex(N)
    b = P.make_rand_complex(N)
    return N, M, ind, ent, ind2, ent2, x, b


def _dense(e, ind, N):
    A = np.zeros((N, N), complex)
    np.add.at(A, (ind[0], ind[1]), e)
    return A


def test_duplicates_present_and_summed():
    N, M, ind, ent, *_ = _setup()
    key = ind[0] * N + ind[1]
    assert np.unique(key).size < M  # the reference's random indices repeat
    assert np.allclose(P.make_sparse(ent, ind, (N, N)).toarray(), _dense(ent, ind, N))


def test_forward_values_against_dense():
    N, M, ind, ent, ind2, ent2, x, b = _setup()
    A, X = _dense(ent, ind, N), _dense(ent2, ind2, N)
    assert np.allclose(P.sp_mult(ent, ind, x), A @ x, rtol=1e-13, atol=1e-13)
    assert np.allclose(A @ P.sp_solve(ent, ind, b), b, rtol=1e-10, atol=1e-10)
    eb, ib = P.spsp_mult(ent, ind, ent2, ind2, N)
    assert np.allclose(_dense(eb, ib, N), A @ X, rtol=1e-13, atol=1e-13)


def test_vjps_against_central_differences():
    # autograd convention: gradient of the real objective |sum(out)| is dL/dx - i dL/dy
    N, M, ind, ent, ind2, ent2, x, b = _setup()
    out_fn = lambda o: np.abs(np.sum(o))

    def v_of(out):  # cotangent of |sum(out)| in autograd convention
        s = np.sum(out)
        return np.conj(s) / abs(s) * np.ones_like(out)

    # sp_mult w.r.t. entries and x
    v = v_of(P.sp_mult(ent, ind, x))
    g = P.sp_mult_vjp_entries(v, ent, ind, x)
    assert np.allclose(g, grad_num(lambda e: out_fn(P.sp_mult(e, ind, x)), ent), atol=1e-6)
    g = P.sp_mult_vjp_x(v, ent, ind, x)
    assert np.allclose(g, grad_num(lambda xx: out_fn(P.sp_mult(ent, ind, xx)), x), atol=1e-6)
    # sp_solve w.r.t. entries and b
    v = v_of(P.sp_solve(ent, ind, b))
    g = P.sp_solve_vjp_entries(v, ent, ind, b)
    assert np.allclose(g, grad_num(lambda e: out_fn(P.sp_solve(e, ind, b)), ent), atol=1e-4, rtol=1e-4)
    g = P.sp_solve_vjp_b(v, ent, ind, b)
    assert np.allclose(g, grad_num(lambda bb: out_fn(P.sp_solve(ent, ind, bb)), b), atol=1e-4, rtol=1e-4)


def test_spsp_vjp_and_jvp_against_dense_products():
    N, M, ind, ent, ind2, ent2, x, b = _setup()
    eb, ib = P.spsp_mult(ent, ind, ent2, ind2, N)
    rng = np.random.default_rng(0)
    v = rng.standard_normal(eb.size) + 1j * rng.standard_normal(eb.size)
    # <v, dB> = <VJP, dA> for a random direction: linearity identity
    dA = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    dB = P.spsp_mult_jvp_a(dA, ib, ent, ind, ent2, ind2, N)
    ga = P.spsp_mult_vjp_a(v, ib, ent, ind, ent2, ind2, N)
    # bilinear pairing (autograd convention uses plain transposes)
    assert np.sum(v * dB) == pytest.approx(np.sum(ga * dA), rel=1e-12)
    dX = rng.standard_normal(M - 2) + 1j * rng.standard_normal(M - 2)
    dB = P.spsp_mult_jvp_x(dX, ib, ent, ind, ent2, ind2, N)
    gx = P.spsp_mult_vjp_x(v, ib, ent, ind, ent2, ind2, N)
    assert np.sum(v * dB) == pytest.approx(np.sum(gx * dX), rel=1e-12)
    # JVP equals the finite difference of the product (exact: bilinear)
    h = 1e-3
    e_plus, _ = P.spsp_mult(ent + h * dA, ind, ent2, ind2, N)
    assert np.allclose((e_plus - eb) / h, P.spsp_mult_jvp_a(dA, ib, ent, ind, ent2, ind2, N), atol=1e-10)
