"""CPU restatement of ceviche/modes.py (test infrastructure only: imported by
tests/ and never by the product).

get_modes (ceviche/modes.py:10-47): A = diag(eps) + Dxf Dxb / k0^2 on the
(N, 1) cross section with npml PML cells (derivatives of fdfd_ref), the m
eigenvalues nearest n_max^2 by scipy.sparse.linalg.eigs in shift-invert mode
(:69-79), real(val) > 0 filter (:82-108), unit-norm columns (:111-119).
insert_mode (:50-66).  Pinned by tests/test_oracle_modes.py against the
closed-form spectrum of the periodic discrete Laplacian and against a dense
numpy.linalg.eig of the same matrix with PML.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spl

from .fdfd_ref import C_0, derivative_matrices


def mode_operator(eps_cross, omega, dL, npml):
    N = eps_cross.size
    k0 = omega / C_0
    Dxf, Dxb, _, _ = derivative_matrices(omega, (N, 1), [npml, 0], dL)
    return sp.spdiags(eps_cross.flatten(), [0], N, N) + Dxf.dot(Dxb) * (1 / k0) ** 2


def get_modes(eps_cross, omega, dL, npml, m=1, filtering=True):
    A = mode_operator(eps_cross, omega, dL, npml)
    n_max = np.sqrt(np.max(eps_cross))
    vals, vecs = spl.eigs(A, k=m, sigma=n_max ** 2, v0=None, which='LM')
    if filtering:
        keep = np.where(np.real(vals) > 0.0)[0]
        vals, vecs = vals[keep], vecs[:, keep]
    if vals.size == 0:
        raise BaseException("Could not find any eigenmodes for this waveguide")
    return vals, vecs / np.sqrt(np.sum(np.abs(vecs) ** 2, axis=0))


def insert_mode(omega, dx, x, y, epsr, target=None, npml=0, m=1, filtering=False):
    if target is None:
        target = np.zeros(epsr.shape, dtype=complex)
    _, mode_field = get_modes(epsr[x, y], omega, dx, npml, m=m, filtering=filtering)
    target[x, y] = np.atleast_2d(mode_field)[:, m - 1].squeeze()
    return target
