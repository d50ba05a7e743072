"""1D SC-PML stretch profiles and derivative-profile setup for the FDFD stencils.

The reference builds full sparse matrices (ceviche/derivatives.py:34-216); the
matrix-free kernels only need, per axis, the diagonal of S^-1 D0 scaled by the
grid step: k_f[i] = 1/(s_f[i] dL) and k_b[i] = 1/(s_b[i] dL).  These are O(Nx + Ny)
host values computed once per operator, in float64/complex128 like the
reference, then uploaded in the operator's precision.
"""
import numpy as np

from .constants import EPSILON_0, ETA_0


def sig_w(l, dw, m=3, lnR=-30):
    """Polynomial PML conductivity; ceviche/derivatives.py:209-212."""
    sig_max = -(m + 1) * lnR / (2 * ETA_0 * dw)
    return sig_max * (l / dw) ** m


def s_value(l, dw, omega):
    """Complex stretch factor; ceviche/derivatives.py:214-216."""
    return 1 - 1j * sig_w(l, dw) / (omega * EPSILON_0)


def create_sfactor(direction, omega, dL, N, N_pml):
    """S-factor cross-section for 'f' or 'b' differences; ceviche/derivatives.py:173-207.

    Cells i <= N_pml take s(dL (N_pml - i + shift)), cells i > N - N_pml take
    s(dL (i - (N - N_pml) - shift)), with shift 0.5 ('f') or 1 ('b')."""
    if N_pml == 0:
        return np.ones(N, dtype=np.complex128)
    if direction not in ('f', 'b'):
        raise ValueError("Dir value {} not recognized".format(direction))
    shift = 0.5 if direction == 'f' else 1.0
    dw = N_pml * dL
    i = np.arange(N)
    out = np.ones(N, dtype=np.complex128)
    lo = i <= N_pml
    hi = (i > N - N_pml) & ~lo
    out[lo] = s_value(dL * (N_pml - i[lo] + shift), dw, omega)
    out[hi] = s_value(dL * (i[hi] - (N - N_pml) - shift), dw, omega)
    return out


def axis_profiles(omega, dL, N, N_pml):
    """(k_f, k_b) for one axis: 1/(s dL), or 1/s when N == 1 (the reference's
    identity 'derivative', ceviche/derivatives.py:76-79)."""
    inv_f = 1 / create_sfactor('f', omega, dL, N, N_pml)
    inv_b = 1 / create_sfactor('b', omega, dL, N, N_pml)
    if N == 1:
        return inv_f, inv_b
    return inv_f * (1 / dL), inv_b * (1 / dL)
