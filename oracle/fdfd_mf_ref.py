"""CPU restatement of ceviche's multi-frequency Ez FDFD (fdfd_mf_ez,
ceviche/fdfd.py:247-359).  TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

The system has M = 2 Nsb + 1 blocks of N = Nx Ny cells, sideband frequencies
W_m = omega + (m - Nsb) omega_mod:
    A[m, m]     = -(1/mu0)(Dxf Dxb + Dyf Dyb) - eps0 W_m^2 diag(eps)
                  (derivatives at the base omega, fdfd.py:289-298)
    A[m, m+h+1] = diag(-eps0/2 W_m^2 delta_h exp(+i phi_h))      (fdfd.py:307-313)
    A[m, m-h-1] = diag(-eps0/2 W_m^2 delta_h exp(-i phi_h))      (fdfd.py:314-320)
    b_m = i W_m Jz_m                                               (fdfd.py:326-333)
    Hx_m = -Dyb Ez_m / (i mu0 W_m),  Hy_m = Dxb Ez_m / (i mu0 W_m)  (fdfd.py:337-351)
solved by scipy's spsolve (solvers.py:50-62).  Harmonics with h + 1 >= M
add nothing (their index ranges are empty in the reference).

Pinning (tests/test_oracle_fdfd_mf.py):
  * delta = 0: every block is the single-frequency Ez solve of fdfd_ref at
    W_m (npml = 0, where the PML does not depend on the frequency), and with
    PML the block equals make_A_ez(omega) shifted by eps0 (omega^2 - W_m^2) eps;
  * coupling: on a periodic vacuum-like grid with uniform eps, delta, phi the
    spatial DFT diagonalises every block, leaving one M x M system per
    Fourier mode (the discrete Laplacian symbol on the diagonal, the
    harmonic couplings off it), solved densely and compared with spsolve
    for one and two harmonics.
"""
import numpy as np
import scipy.sparse as sp

from .fdfd_ref import EPSILON_0, MU_0, _bloch, derivative_matrices, spsolve


def sidebands(omega, omega_mod, Nsb):
    return omega + np.arange(-Nsb, Nsb + 1) * omega_mod


def make_A_mf(omega, dL, eps_r, omega_mod, delta, phi, Nsb, npml, bloch_phases=(0.0, 0.0)):
    """The (M N) x (M N) multi-frequency operator as a block matrix."""
    eps = np.asarray(eps_r)
    shape = eps.shape
    N = eps.size
    M = 2 * Nsb + 1
    W = sidebands(omega, omega_mod, Nsb)
    Dxf, Dxb, Dyf, Dyb = derivative_matrices(omega, shape, npml, dL, *_bloch(bloch_phases))
    lap = (-1 / MU_0 * (Dxf.dot(Dxb) + Dyf.dot(Dyb))).tocsr()
    ev = eps.ravel().astype(np.complex128)
    blocks = [[None] * M for _ in range(M)]
    for m in range(M):
        blocks[m][m] = lap - EPSILON_0 * W[m] ** 2 * sp.diags(ev)
    delta = np.asarray(delta).reshape(-1, N)
    phi = np.asarray(phi).reshape(-1, N)
    for h in range(delta.shape[0]):
        dd = h + 1
        up = -0.5 * EPSILON_0 * delta[h] * np.exp(1j * phi[h])
        dn = -0.5 * EPSILON_0 * delta[h] * np.exp(-1j * phi[h])
        for m in range(M):
            if m + dd < M:
                blocks[m][m + dd] = sp.diags(W[m] ** 2 * up)
            if m - dd >= 0:
                blocks[m][m - dd] = sp.diags(W[m] ** 2 * dn)
    for m in range(M):
        for q in range(M):
            if blocks[m][q] is None:
                blocks[m][q] = sp.csr_matrix((N, N), dtype=np.complex128)
    return sp.bmat(blocks, format='csr')


def solve_mf(omega, dL, eps_r, omega_mod, delta, phi, Nsb, npml, source, bloch_phases=(0.0, 0.0)):
    """(Hx, Hy, Ez), each (M, Nx, Ny), for a source of that shape."""
    eps = np.asarray(eps_r)
    shape = eps.shape
    N = eps.size
    M = 2 * Nsb + 1
    W = sidebands(omega, omega_mod, Nsb)
    A = make_A_mf(omega, dL, eps, omega_mod, delta, phi, Nsb, npml, bloch_phases)
    b = 1j * np.repeat(W, N) * np.asarray(source).reshape(-1)
    ez = spsolve(A, b)
    _, Dxb, _, Dyb = derivative_matrices(omega, shape, npml, dL, *_bloch(bloch_phases))
    ezb = ez.reshape(M, N)
    hx = np.stack([-1 / 1j / MU_0 / W[m] * Dyb.dot(ezb[m]) for m in range(M)])
    hy = np.stack([1 / 1j / MU_0 / W[m] * Dxb.dot(ezb[m]) for m in range(M)])
    full = (M,) + shape
    return hx.reshape(full), hy.reshape(full), ez.reshape(full)
