"""1D SC-PML stretch profiles and derivative-profile setup for the FDFD stencils,
and the reference's sparse derivative-matrix and curl helpers.

The reference builds full sparse matrices (ceviche/derivatives.py:34-216); the
matrix-free kernels only need, per axis, the diagonal of S^-1 D0 scaled by the
grid step: k_f[i] = 1/(s_f[i] dL) and k_b[i] = 1/(s_b[i] dL).  These are O(Nx + Ny)
host values computed once per operator, in float64/complex128 like the
reference, then uploaded in the operator's precision.
"""
import numpy as np
import scipy.sparse as sp
import torch

from .constants import EPSILON_0, ETA_0


def sig_w(l, dw, m=3, lnR=-30):
    """Polynomial PML conductivity; ceviche/derivatives.py:209-212."""
    sig_max = -(m + 1) * lnR / (2 * ETA_0 * dw)
    return sig_max * (l / dw) ** m


def s_value(l, dw, omega):
    """Complex stretch factor; ceviche/derivatives.py:214-216."""
    return 1 - 1j * sig_w(l, dw) / (omega * EPSILON_0)


def _sfactor(omega, dL, N, N_pml, dw, shift):
    i = np.arange(N)
    out = np.ones(N, dtype=np.complex128)
    lo = i <= N_pml
    hi = (i > N - N_pml) & ~lo
    out[lo] = s_value(dL * (N_pml - i[lo] + shift), dw, omega)
    out[hi] = s_value(dL * (i[hi] - (N - N_pml) - shift), dw, omega)
    return out


def create_sfactor_f(omega, dL, N, N_pml, dw):
    """S-factors for forward differences (ceviche/derivatives.py:189-197):
    cells i <= N_pml take s(dL (N_pml - i + 0.5)), cells i > N - N_pml take
    s(dL (i - (N - N_pml) - 0.5)), the rest 1."""
    return _sfactor(omega, dL, N, N_pml, dw, 0.5)


def create_sfactor_b(omega, dL, N, N_pml, dw):
    """S-factors for backward differences (ceviche/derivatives.py:199-207): as
    create_sfactor_f with the half-cell shift replaced by a whole cell."""
    return _sfactor(omega, dL, N, N_pml, dw, 1.0)


def create_sfactor(direction, omega, dL, N, N_pml):
    """S-factor cross-section for 'f' or 'b' differences; ceviche/derivatives.py:173-187
    (all ones without PML, ValueError for another direction)."""
    if N_pml == 0:
        return np.ones(N, dtype=np.complex128)
    dw = N_pml * dL
    if direction == 'f':
        return create_sfactor_f(omega, dL, N, N_pml, dw)
    if direction == 'b':
        return create_sfactor_b(omega, dL, N, N_pml, dw)
    raise ValueError("Dir value {} not recognized".format(direction))


def axis_profiles(omega, dL, N, N_pml):
    """(k_f, k_b) for one axis: 1/(s dL), or 1/s when N == 1 (the reference's
    identity 'derivative', ceviche/derivatives.py:76-79)."""
    inv_f = 1 / create_sfactor('f', omega, dL, N, N_pml)
    inv_b = 1 / create_sfactor('b', omega, dL, N, N_pml)
    if N == 1:
        return inv_f, inv_b
    return inv_f * (1 / dL), inv_b * (1 / dL)


# ---- the reference's sparse-matrix and curl helpers (ceviche/derivatives.py:16-171),
# host scipy / numpy like the reference; the solvers here never build these
# matrices (the stencil kernels apply the same rows matrix-free), they serve
# callers that assemble operators of their own.

def _roll(F, shift, axis):
    return torch.roll(F, shifts=shift, dims=axis) if isinstance(F, torch.Tensor) else np.roll(F, shift, axis=axis)


def curl_E(axis, Ex, Ey, Ez, dL):
    """Component `axis` (0, 1, 2) of the curl of E by forward differences on a
    periodic 3D grid (ceviche/derivatives.py:16-22); numpy arrays or torch
    tensors.  Another axis gives None, as in the reference."""
    d = lambda F, ax: (_roll(F, -1, ax) - F) / dL
    if axis == 0:
        return d(Ez, 1) - d(Ey, 2)
    if axis == 1:
        return d(Ex, 2) - d(Ez, 0)
    if axis == 2:
        return d(Ey, 0) - d(Ex, 1)
    return None


def curl_H(axis, Hx, Hy, Hz, dL):
    """Component `axis` of the curl of H by backward differences on a periodic
    3D grid (ceviche/derivatives.py:24-30); numpy arrays or torch tensors."""
    d = lambda F, ax: (F - _roll(F, 1, ax)) / dL
    if axis == 0:
        return d(Hz, 1) - d(Hy, 2)
    if axis == 1:
        return d(Hx, 2) - d(Hz, 0)
    if axis == 2:
        return d(Hy, 0) - d(Hx, 1)
    return None


def make_Dxf(dL, shape, bloch_x=0.0):
    """Forward x difference on the (Nx, Ny) grid, cells flattened x-major:
    (-1 on the diagonal, +1 on the next x, the Bloch phasor e^{i bloch_x}
    closing the period) / dL (ceviche/derivatives.py:94-100)."""
    Nx, Ny = shape
    D = sp.diags([-1, 1, np.exp(1j * bloch_x)], [0, 1, -Nx + 1], shape=(Nx, Nx), dtype=np.complex128)
    return 1 / dL * sp.kron(D, sp.eye(Ny))


def make_Dxb(dL, shape, bloch_x=0.0):
    """Backward x difference, -conj(e^{i bloch_x}) closing the period
    (ceviche/derivatives.py:102-108)."""
    Nx, Ny = shape
    D = sp.diags([1, -1, -np.conj(np.exp(1j * bloch_x))], [0, -1, Nx - 1], shape=(Nx, Nx), dtype=np.complex128)
    return 1 / dL * sp.kron(D, sp.eye(Ny))


def make_Dyf(dL, shape, bloch_y=0.0):
    """Forward y difference (ceviche/derivatives.py:110-116)."""
    Nx, Ny = shape
    D = sp.diags([-1, 1, np.exp(1j * bloch_y)], [0, 1, -Ny + 1], shape=(Ny, Ny))
    return 1 / dL * sp.kron(sp.eye(Nx), D)


def make_Dyb(dL, shape, bloch_y=0.0):
    """Backward y difference (ceviche/derivatives.py:118-124)."""
    Nx, Ny = shape
    D = sp.diags([1, -1, -np.conj(np.exp(1j * bloch_y))], [0, -1, Ny - 1], shape=(Ny, Ny))
    return 1 / dL * sp.kron(sp.eye(Nx), D)


def createDws(component, dir, shape, dL, bloch_x=0.0, bloch_y=0.0):
    """The x or y, 'f' or 'b' difference matrix without PML
    (ceviche/derivatives.py:63-92); an axis of length 1 gives the identity
    over the other axis; ValueError for another component or direction."""
    Nx, Ny = shape
    if component == 'x' and Nx == 1:
        return sp.eye(Ny)
    if component == 'y' and Ny == 1:
        return sp.eye(Nx)
    make = {'xf': lambda: make_Dxf(dL, shape, bloch_x=bloch_x), 'xb': lambda: make_Dxb(dL, shape, bloch_x=bloch_x),
            'yf': lambda: make_Dyf(dL, shape, bloch_y=bloch_y), 'yb': lambda: make_Dyb(dL, shape, bloch_y=bloch_y)}
    key = component + dir
    if key not in make:
        raise ValueError("component and direction {} and {} not recognized".format(component, dir))
    return make[key]()


def create_S_matrices(omega, shape, npml, dL):
    """Diagonal matrices 1/s of the four stretch profiles over the grid,
    x-major (Sx_f, Sx_b, Sy_f, Sy_b) (ceviche/derivatives.py:129-171)."""
    Nx, Ny = shape
    N = Nx * Ny
    Nx_pml, Ny_pml = npml
    sx_f, sx_b = create_sfactor('f', omega, dL, Nx, Nx_pml), create_sfactor('b', omega, dL, Nx, Nx_pml)
    sy_f, sy_b = create_sfactor('f', omega, dL, Ny, Ny_pml), create_sfactor('b', omega, dL, Ny, Ny_pml)
    ones_x, ones_y = np.ones(Nx), np.ones(Ny)
    diag = lambda v: sp.spdiags(v, 0, N, N)
    return (diag(np.outer(1 / sx_f, ones_y).ravel()), diag(np.outer(1 / sx_b, ones_y).ravel()),
            diag(np.outer(ones_x, 1 / sy_f).ravel()), diag(np.outer(ones_x, 1 / sy_b).ravel()))


def compute_derivative_matrices(omega, shape, npml, dL, bloch_x=0.0, bloch_y=0.0):
    """(Dxf, Dxb, Dyf, Dyb) with the PML applied, S D for each
    (ceviche/derivatives.py:34-59)."""
    Dxf = createDws('x', 'f', shape, dL, bloch_x=bloch_x, bloch_y=bloch_y)
    Dxb = createDws('x', 'b', shape, dL, bloch_x=bloch_x, bloch_y=bloch_y)
    Dyf = createDws('y', 'f', shape, dL, bloch_x=bloch_x, bloch_y=bloch_y)
    Dyb = createDws('y', 'b', shape, dL, bloch_x=bloch_x, bloch_y=bloch_y)
    Sxf, Sxb, Syf, Syb = create_S_matrices(omega, shape, npml, dL)
    return Sxf.dot(Dxf), Sxb.dot(Dxb), Syf.dot(Dyf), Syb.dot(Dyb)
