"""The reference's utility API (ceviche/utils.py:13-125) as exported by
ceviche_amd.utils and ceviche_amd.primitives: sparse-format helpers, random
test data and numerical derivatives.  Host-only (no GPU)."""
import warnings

import numpy as np
import pytest
import scipy.sparse as sp
import torch

from ceviche_amd import utils as U
from oracle import numdiff


def test_make_sparse_sums_duplicates_and_is_complex_csr():
    entries = np.array([1.0, 2.0, 3.0, 4.0])
    indices = np.array([[0, 1, 0, 2], [1, 2, 1, 0]])
    A = U.make_sparse(entries, indices, (3, 3))
    assert sp.isspmatrix_csr(A) and A.dtype == np.complex128 and A.shape == (3, 3)
    dense = np.zeros((3, 3), dtype=complex)
    dense[0, 1] = 4.0        # 1 + 3, the duplicate (0, 1) summed
    dense[1, 2] = 2.0
    dense[2, 0] = 4.0
    assert np.array_equal(A.toarray(), dense)


def test_make_sparse_torch_input_matches_scipy():
    rng = np.random.default_rng(0)
    e = rng.standard_normal(30) + 1j * rng.standard_normal(30)
    idx = rng.integers(0, 6, (2, 30))
    A = U.make_sparse(e, idx, (6, 6)).toarray()
    T = U.make_sparse(torch.from_numpy(e), torch.from_numpy(idx), (6, 6))
    assert T.layout == torch.sparse_csr
    assert np.allclose(T.to_dense().numpy(), A, rtol=0, atol=1e-15)
    te, ti = U.get_entries_indices(T)
    assert np.allclose(U.make_sparse(te.numpy(), ti.numpy(), (6, 6)).toarray(), A, rtol=0, atol=1e-15)


def test_get_entries_indices_round_trip():
    rng = np.random.default_rng(1)
    A = sp.random(9, 9, density=0.3, random_state=3, format='csr') * (1 + 2j)
    e, idx = U.get_entries_indices(A)
    assert idx.shape == (2, A.nnz) and np.array_equal(e, A.data)
    assert (U.make_sparse(e, idx, A.shape) != A).nnz == 0
    assert np.array_equal(U.transpose_indices(idx), idx[::-1])


def test_make_io_matrices_identities():
    # A = I diag(a) O, and for unique positions a = diag(I^T A O^T) (utils.py:50-73)
    rng = np.random.default_rng(2)
    N, M = 7, 15
    flat = rng.choice(N * N, M, replace=False)
    idx = np.vstack((flat // N, flat % N))
    a = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    I, O = U.make_IO_matrices(idx, N)
    assert I.shape == (N, M) and O.shape == (M, N)
    A = U.make_sparse(a, idx, (N, N))
    assert np.allclose((I @ sp.diags(a) @ O).toarray(), A.toarray(), rtol=0, atol=1e-15)
    assert np.allclose((I.T @ A @ O.T).diagonal(), a, rtol=0, atol=1e-15)


def test_random_helpers_shapes_and_ranges():
    np.random.seed(2)
    r = U.make_rand(50)
    assert r.shape == (50,) and r.min() >= -0.5 and r.max() < 0.5
    c = U.make_rand_complex(20)
    assert c.dtype == np.complex128 and np.abs(c.real).max() <= 0.5 and np.abs(c.imag).max() <= 0.5
    idx = U.make_rand_indeces(8, 44)
    assert idx.shape == (2, 44) and idx.min() >= 0 and idx.max() < 8
    e, i = U.make_rand_entries_indices(8, 44)
    assert e.shape == (44,) and i.shape == (2, 44)
    S = U.make_rand_sparse(8, 44)
    assert sp.isspmatrix_csr(S) and S.shape == (8, 8) and S.nnz <= 44
    D = U.make_rand_sparse_density(6, density=0.5)
    assert D.shape == (6, 6) and np.iscomplexobj(D.toarray())


def test_the_reference_seed_2_setup_draws_from_the_global_generator():
    # ceviche tests/test_primitives.py:11-27 draw their matrices this way; the
    # same calls after np.random.seed(2) give the same values here as numpy's
    # own generator calls in the reference's order.
    np.random.seed(2)
    idx = U.make_rand_indeces(8, 44)
    ent = U.make_rand_complex(44)
    np.random.seed(2)
    idx_np = np.random.randint(low=0, high=8, size=(2, 44))
    re = np.random.random(44) - 0.5
    im = np.random.random(44) - 0.5
    assert np.array_equal(idx, idx_np) and np.array_equal(ent, re + 1j * im)


def test_grad_num_exact_on_polynomials_and_matches_the_oracle():
    # f(z) = sum |z_k|^2: in autograd's convention grad = df/dx - i df/dy
    # = 2 (x - i y) = 2 conj(z), and the central difference of a quadratic is exact
    z = np.array([0.3 + 0.2j, -1.1 + 0.5j, 0.7 - 0.4j])
    f = lambda v: np.sum(np.abs(v) ** 2)
    g = U.grad_num(f, z, delta=1e-4)
    assert np.allclose(g, 2 * np.conj(z), rtol=1e-8, atol=1e-9)
    h = lambda v: np.real(np.sum(v ** 3))
    assert np.allclose(U.grad_num(h, z, 1e-5), numdiff.grad_num(h, z, 1e-5), rtol=1e-12, atol=1e-12)
    # der_num: central difference, exact for a quadratic
    x = np.array([1.5, -2.0])
    q = lambda v: 3 * v[0] ** 2 + v[0] * v[1]
    assert U.der_num(q, x, 0, 1e-3) == pytest.approx(6 * 1.5 - 2.0, rel=1e-10)


def test_grad_num_real_argument_keeps_the_reference_behaviour():
    # with a real array the imaginary step is discarded by numpy (ComplexWarning):
    # the gradient is the real central difference
    x = np.array([0.5, 2.0])
    f = lambda v: v[0] ** 2 + 3 * v[1]
    with warnings.catch_warnings():
        warnings.simplefilter('ignore')
        g = U.grad_num(f, x, 1e-4)
    assert np.allclose(g, [1.0, 3.0], rtol=1e-8)


def test_primitives_module_exports_the_reference_helpers():
    import ceviche_amd.primitives as P
    for name in ('make_sparse', 'transpose_indices', 'make_rand', 'make_rand_complex', 'make_rand_indeces',
                 'make_rand_sparse', 'der_num', 'grad_num', 'get_entries_indices', 'make_IO_matrices'):
        assert name in P.__all__ and callable(getattr(P, name))


def test_jac_num_of_a_linear_map_is_its_transpose_layout():
    """jac_num lays the Jacobian out (input entries, output entries)
    (utils.py:127-150): for f(x) = A x it is A^T up to the forward-difference
    rounding; a quadratic's rows are its gradient at x (+ h/2 per entry)."""
    rng = np.random.default_rng(3)
    A = rng.standard_normal((4, 6))
    x = rng.standard_normal(6)
    J = U.jac_num(lambda v: A @ v, x, step_size=1e-6)
    assert J.shape == (6, 4) and np.allclose(J, A.T, atol=1e-8)
    Jq = U.jac_num(lambda v: np.sum(v ** 2), x, step_size=1e-7)
    assert Jq.shape == (6, 1) and np.allclose(Jq[:, 0], 2 * x, atol=1e-6)
    assert U.jac_num(lambda v: 3.0 * v, 2.0).shape == (1, 1)


def test_vjp_maker_num_against_the_closed_form():
    """vjp_maker_num (utils.py:239-276): one maker per listed argument; for
    f(a, b) = a * b (elementwise) the VJP with cotangent v is v * b for a and
    v * a for b, exactly for a bilinear map up to the step's rounding."""
    rng = np.random.default_rng(4)
    a, b, v = rng.standard_normal(5), rng.standard_normal(5), rng.standard_normal(5)
    f = lambda p, q: p * q
    mk_a, mk_b = U.vjp_maker_num(f, [0, 1], [1e-6, 1e-6])
    out = f(a, b)
    assert np.allclose(mk_a(out, a, b)(v), v * b, atol=1e-8)
    assert np.allclose(mk_b(out, a, b)(v), v * a, atol=1e-8)


def test_value_and_shape_helpers():
    """float_2_array, reshape_to_ND, get_value, get_shape (utils.py:200-236)."""
    assert np.array_equal(U.float_2_array(2.5), np.array([2.5]))
    arr = np.zeros((2, 3))
    assert U.float_2_array(arr) is arr
    assert U.reshape_to_ND(np.zeros((4, 5)), 3).shape == (4, 5, 1)
    assert U.reshape_to_ND(np.zeros(4), 3).shape == (4, 1, 1)
    with pytest.raises(ValueError):
        U.reshape_to_ND(np.zeros((2, 2, 2, 2)), 3)
    t = torch.ones(3, requires_grad=True) * 2
    g = U.get_value(t)
    assert not g.requires_grad and torch.equal(g, torch.full((3,), 2.0))
    assert U.get_value(arr) is arr
    assert U.get_shape(1.0) == (1,) and U.get_shape(3) == (1,)
    assert U.get_shape([1, 2, 3]) == (3,) and U.get_shape((1, 2)) == (2,)
    assert U.get_shape(np.zeros((2, 7))) == (2, 7) and U.get_shape(torch.zeros(4, 1)) == (4, 1)
