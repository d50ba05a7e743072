"""FDFD oracle: scipy restatement of ceviche's Ez / Hz frequency-domain path.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Pinning (tests/test_oracle_fdfd.py):
  * operator: complex symmetry of W*A (zero Bloch phase), nnz structure,
    the closed-form DFT symbol of the periodic vacuum operator (npml=0),
    adjoint identity <A x, y> = <x, A^T y>;
  * PML profile + solve: vacuum point source against the 2D Green's function
    (Hankel H0^(2)) far from the source, plane-wave wavelength = lambda/dL
    cells (ceviche tests/test_plane_wave.py:14-27), PML reflection small;
  * adjoint gradient: central finite differences (ceviche
    tests/test_gradients_fdfd.py:61-73 criterion, ratio <= 1e-4).
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spl

# ceviche/constants.py:7-11 (truncated values; parity needs exactly these)
EPSILON_0 = 8.85418782e-12
MU_0 = 1.25663706e-6
C_0 = 1 / np.sqrt(EPSILON_0 * MU_0)
ETA_0 = np.sqrt(MU_0 / EPSILON_0)
Q_E = 1.602176634e-19


# ---------------------------------------------------------------- SC-PML
def _sigma(l, dw, m=3, lnR=-30.0):
    """Polynomial PML conductivity; ceviche/derivatives.py:209-212."""
    sig_max = -(m + 1) * lnR / (2 * ETA_0 * dw)
    return sig_max * (l / dw) ** m


def _stretch(l, dw, omega):
    """Complex coordinate stretch s(l); ceviche/derivatives.py:214-216."""
    return 1 - 1j * _sigma(l, dw) / (omega * EPSILON_0)


def sfactor(direction, omega, dL, n, npml):
    """1D stretch profile for forward ('f') or backward ('b') differences.

    ceviche/derivatives.py:173-207.  Note the asymmetric index ranges
    (i <= npml on the low side, i > n - npml on the high side)."""
    if npml == 0:
        return np.ones(n, dtype=np.complex128)
    dw = npml * dL
    if direction == 'f':
        shift = 0.5
    elif direction == 'b':
        shift = 1.0
    else:
        raise ValueError("Dir value {} not recognized".format(direction))
    out = np.ones(n, dtype=np.complex128)
    i = np.arange(n)
    lo = i <= npml
    hi = (i > n - npml) & ~lo
    out[lo] = _stretch(dL * (npml - i[lo] + shift), dw, omega)
    out[hi] = _stretch(dL * (i[hi] - (n - npml) - shift), dw, omega)
    return out


# ------------------------------------------------------ difference matrices
def _diff_1d(n, dL, forward, phase):
    """Periodic / Bloch first difference on n points; derivatives.py:94-124."""
    ph = np.exp(1j * phase)
    if forward:
        d = sp.diags([-1, 1, ph], [0, 1, -n + 1], shape=(n, n), dtype=np.complex128)
    else:
        d = sp.diags([1, -1, -np.conj(ph)], [0, -1, n - 1], shape=(n, n),
                     dtype=np.complex128)
    return d / dL


def derivative_matrices(omega, shape, npml, dL, bloch_x=0.0, bloch_y=0.0):
    """Dxf, Dxb, Dyf, Dyb with SC-PML applied (S^-1 * D0).

    ceviche/derivatives.py:34-59 (assembly), :63-92 (1D identity special case
    for Nx == 1 / Ny == 1), :129-171 (S matrices hold 1/s)."""
    nx, ny = shape
    n = nx * ny
    if nx == 1:
        d0xf = sp.eye(ny, dtype=np.complex128)
        d0xb = sp.eye(ny, dtype=np.complex128)
    else:
        d0xf = sp.kron(_diff_1d(nx, dL, True, bloch_x), sp.eye(ny))
        d0xb = sp.kron(_diff_1d(nx, dL, False, bloch_x), sp.eye(ny))
    if ny == 1:
        d0yf = sp.eye(nx, dtype=np.complex128)
        d0yb = sp.eye(nx, dtype=np.complex128)
    else:
        d0yf = sp.kron(sp.eye(nx), _diff_1d(ny, dL, True, bloch_y))
        d0yb = sp.kron(sp.eye(nx), _diff_1d(ny, dL, False, bloch_y))
    px, py = npml
    inv = {}
    inv['xf'] = np.repeat(1 / sfactor('f', omega, dL, nx, px), ny)
    inv['xb'] = np.repeat(1 / sfactor('b', omega, dL, nx, px), ny)
    inv['yf'] = np.tile(1 / sfactor('f', omega, dL, ny, py), nx)
    inv['yb'] = np.tile(1 / sfactor('b', omega, dL, ny, py), nx)
    S = {k: sp.spdiags(v, 0, n, n) for k, v in inv.items()}
    return (S['xf'].dot(d0xf).tocsr(), S['xb'].dot(d0xb).tocsr(),
            S['yf'].dot(d0yf).tocsr(), S['yb'].dot(d0yb).tocsr())


# --------------------------------------------------------------- operators
def make_A_ez(omega, dL, eps_r, npml, bloch_phases=(0.0, 0.0)):
    """A = -(1/mu0)(Dxf Dxb + Dyf Dyb) - eps0 w^2 diag(eps); fdfd.py:172-185."""
    Dxf, Dxb, Dyf, Dyb = derivative_matrices(omega, eps_r.shape, npml, dL,
                                             *_bloch(bloch_phases))
    eps = np.asarray(eps_r).ravel()
    C = -1 / MU_0 * Dxf.dot(Dxb) - 1 / MU_0 * Dyf.dot(Dyb)
    return (C - EPSILON_0 * omega ** 2 * sp.diags(eps.astype(np.complex128))).tocsr()


def grid_average_hz(eps_r):
    """eps_xx averages along axis 1, eps_yy along axis 0; fdfd.py:200-209."""
    eps_xx = 0.5 * (eps_r + np.roll(eps_r, 1, axis=1))
    eps_yy = 0.5 * (eps_r + np.roll(eps_r, 1, axis=0))
    return eps_xx, eps_yy


def make_A_hz(omega, dL, eps_r, npml, bloch_phases=(0.0, 0.0)):
    """A = (1/eps0)(Dxb D(1/eps_yy) Dxf + Dyb D(1/eps_xx) Dyf) + mu0 w^2 I.

    fdfd.py:211-233 (the +1e-5 regulariser is at :214-215)."""
    Dxf, Dxb, Dyf, Dyb = derivative_matrices(omega, eps_r.shape, npml, dL,
                                             *_bloch(bloch_phases))
    exx, eyy = grid_average_hz(np.asarray(eps_r))
    ixx = sp.diags(1 / (exx.ravel() + 1e-5))
    iyy = sp.diags(1 / (eyy.ravel() + 1e-5))
    n = exx.size
    A = 1 / EPSILON_0 * (Dxb.dot(iyy).dot(Dxf) + Dyb.dot(ixx).dot(Dyf))
    return (A + MU_0 * omega ** 2 * sp.eye(n)).tocsr()


def _bloch(bloch_phases):
    if bloch_phases is None:
        return 0.0, 0.0
    bx = bloch_phases[0]
    by = bloch_phases[1] if len(bloch_phases) > 1 else 0.0
    return bx, by


# ------------------------------------------------------------------ solves
def spsolve(A, b):
    """ceviche's default direct path (no pyMKL): solvers.py:50-62."""
    return spl.spsolve(sp.csc_matrix(A), b)


def solve_ez(omega, dL, eps_r, npml, source_z, bloch_phases=(0.0, 0.0)):
    """(Hx, Hy, Ez) for a Jz source; fdfd.py:58-76, 135-147, 187-192."""
    shape = eps_r.shape
    A = make_A_ez(omega, dL, eps_r, npml, bloch_phases)
    b = 1j * omega * np.asarray(source_z).ravel()
    ez = spsolve(A, b)
    hx, hy = ez_to_hx_hy(omega, dL, shape, npml, ez, bloch_phases)
    return hx.reshape(shape), hy.reshape(shape), ez.reshape(shape)


def ez_to_hx_hy(omega, dL, shape, npml, ez, bloch_phases=(0.0, 0.0)):
    """Hx = -(1/(i w mu0)) Dyb Ez, Hy = (1/(i w mu0)) Dxb Ez; fdfd.py:138-147."""
    Dxf, Dxb, Dyf, Dyb = derivative_matrices(omega, shape, npml, dL, *_bloch(bloch_phases))
    ez = np.asarray(ez).ravel()
    hx = -1 / 1j / omega / MU_0 * Dyb.dot(ez)
    hy = 1 / 1j / omega / MU_0 * Dxb.dot(ez)
    return hx, hy


def solve_hz(omega, dL, eps_r, npml, source_z, bloch_phases=(0.0, 0.0)):
    """(Ex, Ey, Hz) for an Mz source; fdfd.py:150-162, 235-244."""
    shape = eps_r.shape
    A = make_A_hz(omega, dL, eps_r, npml, bloch_phases)
    b = 1j * omega * np.asarray(source_z).ravel()
    hz = spsolve(A, b)
    ex, ey = hz_to_ex_ey(omega, dL, eps_r, npml, hz, bloch_phases)
    return ex.reshape(shape), ey.reshape(shape), hz.reshape(shape)


def hz_to_ex_ey(omega, dL, eps_r, npml, hz, bloch_phases=(0.0, 0.0)):
    """Ex = (1/(i w eps0)) Dyf Hz / (eps_xx + 1e-5), Ey = -(...) Dxf Hz / (eps_yy + 1e-5).

    fdfd.py:150-162."""
    Dxf, Dxb, Dyf, Dyb = derivative_matrices(omega, eps_r.shape, npml, dL,
                                             *_bloch(bloch_phases))
    exx, eyy = grid_average_hz(np.asarray(eps_r))
    hz = np.asarray(hz).ravel()
    ex = 1 / 1j / omega / EPSILON_0 / (exx.ravel() + 1e-5) * Dyf.dot(hz)
    ey = -1 / 1j / omega / EPSILON_0 / (eyy.ravel() + 1e-5) * Dxf.dot(hz)
    return ex, ey


# ------------------------------------------------------- adjoint gradients
def grad_eps_ez(omega, dL, eps_r, npml, source_z, dL_dEz_conj, bloch_phases=(0.0, 0.0)):
    """Gradient of a real objective L(Ez) with respect to real eps.

    ``dL_dEz_conj`` is the torch-convention cotangent g = dL/dRe(Ez) + i dL/dIm(Ez).
    With x = A^-1 b and A = C - eps0 w^2 diag(eps):
        dL/deps = Re( conj(g) . dx/deps ) = Re( conj(lam) * eps0 w^2 * x ),
    where A^H lam = g.  This is the reference VJP (primitives.py:121-128) with
    the diagonal slice and the -eps0 w^2 factor of fdfd.py:179, rewritten in
    the conjugate convention."""
    A = make_A_ez(omega, dL, eps_r, npml, bloch_phases)
    b = 1j * omega * np.asarray(source_z).ravel()
    x = spsolve(A, b)
    lam = spsolve(A.conj().T.tocsr(), np.asarray(dL_dEz_conj).ravel())
    g = np.real(np.conj(lam) * EPSILON_0 * omega ** 2 * x)
    return g.reshape(eps_r.shape), x.reshape(eps_r.shape)


def hankel_point_source_ez(omega, dL, r, amp=1.0):
    """Continuum 2D field of a one-cell Jz source of amplitude ``amp`` in vacuum.

    A Ez = i w Jz with A ~ -(1/mu0)(lap + k^2) and a one-cell source equal to
    amp * dL^2 * delta(r):  (lap + k^2) Ez = -i w mu0 amp dL^2 delta, so
    Ez = i w mu0 amp dL^2 G with (lap + k^2) G = -delta.  ceviche's stretch
    s = 1 - i sigma/(w eps0) absorbs waves outgoing in the e^{+i w t}
    convention, G = -(i/4) H0^(2)(k r), giving Ez = (w mu0 amp dL^2 / 4) H0^(2)(k r).
    Used only as a closed-form pin far from the source (tolerance ~1e-2)."""
    from scipy.special import hankel2
    k = omega / C_0
    return omega * MU_0 * amp * dL ** 2 / 4 * hankel2(0, k * r)
