"""FDTD oracle: numpy restatement of ceviche's Yee update with UPML integrals.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

The arithmetic follows ceviche/fdtd.py op by op (float64 throughout, B8), so
the oracle is the reference's numbers up to numpy's own rounding.  Passing
``dtype=np.complex128`` with a complex eps gives complex-step derivatives
(the update is analytic in eps: it enters only through 1/eps_aa).

Pinning (tests/test_oracle_fdtd.py): linearity in J, mirror symmetry for a
symmetric eps and a centred source, Ex = Ey = Hz = 0 exactly for a 2D TM Jz
drive, the 1D-vacuum numerical dispersion relation of the Yee scheme, PML
absorption (energy leaves the box), and the reference's own gradient test
set-up (tests/test_gradients_fdtd.py:28-88: complex-step vs central FD,
ratio <= 1e-4).
"""
import numpy as np

from .fdfd_ref import EPSILON_0, C_0

FIELDS = ('Ex', 'Ey', 'Ez', 'Dx', 'Dy', 'Dz', 'Hx', 'Hy', 'Hz')


def reshape_to_3d(arr):
    """ceviche/utils.py:206-214 (raises for > 3 dims)."""
    arr = np.asarray(arr)
    if arr.ndim > 3:
        raise ValueError("array is larger than 3 dimensional, given shape {}".format(arr.shape))
    return arr.reshape(arr.shape + (1,) * (3 - arr.ndim))


def curl_E(axis, Ex, Ey, Ez, dL):
    """Forward-difference curl with periodic wrap; ceviche/derivatives.py:16-22."""
    if axis == 0:
        return (np.roll(Ez, -1, axis=1) - Ez) / dL - (np.roll(Ey, -1, axis=2) - Ey) / dL
    if axis == 1:
        return (np.roll(Ex, -1, axis=2) - Ex) / dL - (np.roll(Ez, -1, axis=0) - Ez) / dL
    return (np.roll(Ey, -1, axis=0) - Ey) / dL - (np.roll(Ex, -1, axis=1) - Ex) / dL


def curl_H(axis, Hx, Hy, Hz, dL):
    """Backward-difference curl with periodic wrap; ceviche/derivatives.py:24-30."""
    if axis == 0:
        return (Hz - np.roll(Hz, 1, axis=1)) / dL - (Hy - np.roll(Hy, 1, axis=2)) / dL
    if axis == 1:
        return (Hx - np.roll(Hx, 1, axis=2)) / dL - (Hz - np.roll(Hz, 1, axis=0)) / dL
    return (Hy - np.roll(Hy, 1, axis=0)) / dL - (Hx - np.roll(Hx, 1, axis=1)) / dL


def time_step(dL, stability_factor=0.5):
    """dt = 0.5 * dL / (sqrt(3) C0), the 3D Courant limit even in 2D; fdtd.py:213-222."""
    dL_sum = 3 / dL ** 2
    dL_avg = 1 / np.sqrt(dL_sum)
    return dL_avg / C_0 * stability_factor


def sigmas(shape, npml, dt):
    """PML conductivities on the doubled grid, sampled per component; fdtd.py:224-263."""
    nx, ny, nz = shape
    sig2 = [np.zeros((2 * nx, 2 * ny, 2 * nz)) for _ in range(3)]
    n2 = (nx, ny, nz)
    for a in range(3):
        p = npml[a]
        for n in range(2 * p):
            i1 = 2 * p - n + 1
            i2 = 2 * n2[a] - 2 * p + n
            val = (0.5 * EPSILON_0 / dt) * (n / 2 / p) ** 3
            idx1 = [slice(None)] * 3
            idx2 = [slice(None)] * 3
            idx1[a] = i1
            idx2[a] = i2
            sig2[a][tuple(idx1)] = val
            sig2[a][tuple(idx2)] = val
    sx2, sy2, sz2 = sig2
    sH = (sx2[1::2, ::2, ::2], sy2[::2, 1::2, ::2], sz2[::2, ::2, 1::2])
    sD = (sx2[::2, 1::2, 1::2], sy2[1::2, ::2, 1::2], sz2[1::2, 1::2, ::2])
    return sH, sD


def grid_center_to_xyz(q):
    """eps_aa = (eps + roll(eps, +1, axis a)) / 2; ceviche/utils.py:153-176."""
    return tuple((q + np.roll(q, 1, axis=a)) / 2 for a in range(3))


def coefficients(sH, sD, dt, mu_r=1.0):
    """The 24 PML update coefficients (mH1..4, mD1..4 per axis); fdtd.py:265-311."""
    e0 = EPSILON_0
    out = {}
    for kind, s, sign in (('H', sH, -1.0), ('D', sD, 1.0)):
        for a, name in enumerate('xyz'):
            b, c = [k for k in range(3) if k != a]
            sa, sb, sc = s[a], s[b], s[c]
            m0 = (1 / dt + (sb + sc) / 2 / e0 + sb * sc * dt / 4 / e0 ** 2)
            m1 = (1 / m0 * (1 / dt - (sb + sc) / 2 / e0 - sb * sc * dt / 4 / e0 ** 2))
            if kind == 'H':
                m2 = (-1 / m0 * C_0 / mu_r)
                m3 = (-1 / m0 * C_0 * dt * sa / e0 / mu_r)
            else:
                m2 = (1 / m0 * C_0)
                m3 = (1 / m0 * C_0 * dt * sa / e0)
            m4 = (-1 / m0 * dt * sb * sc / e0 ** 2)
            out['m%s%s1' % (kind, name)] = m1
            out['m%s%s2' % (kind, name)] = m2
            out['m%s%s3' % (kind, name)] = m3
            out['m%s%s4' % (kind, name)] = m4
    return out


class FDTD:
    """State machine mirroring ceviche.fdtd (fdtd.py:10-147)."""

    def __init__(self, eps_r, dL, npml, dtype=np.float64):
        eps_r = reshape_to_3d(eps_r)
        self.shape = eps_r.shape
        self.dL = dL
        self.npml = list(npml)
        self.dtype = dtype
        self.dt = time_step(dL)
        self.sH, self.sD = sigmas(self.shape, self.npml, self.dt)
        self.m = coefficients(self.sH, self.sD, self.dt)
        self.set_eps(eps_r)

    def set_eps(self, eps_r):
        """Recomputes mE and resets fields and t (the eps_r setter, fdtd.py:63-72)."""
        eps_r = reshape_to_3d(eps_r)
        self.eps_r = eps_r
        exx, eyy, ezz = grid_center_to_xyz(eps_r)
        self.mE = (1 / exx, 1 / eyy, 1 / ezz)
        self.reset()

    def reset(self):
        z = lambda: np.zeros(self.shape, dtype=self.dtype)
        self.t_index = 0
        self.H = [z(), z(), z()]
        self.D = [z(), z(), z()]
        self.E = [z(), z(), z()]
        self.ICE = [z(), z(), z()]
        self.IH = [z(), z(), z()]
        self.ICH = [z(), z(), z()]
        self.ID = [z(), z(), z()]

    def forward(self, Jx=None, Jy=None, Jz=None):
        """One Yee step in the order of fdtd.py:74-144."""
        m = self.m
        dL = self.dL
        self.t_index += 1
        CE = [curl_E(a, *self.E, dL) for a in range(3)]
        self.ICE = [self.ICE[a] + CE[a] for a in range(3)]
        self.IH = [self.IH[a] + self.H[a] for a in range(3)]
        H = []
        for a, n in enumerate('xyz'):
            H.append(m['mH%s1' % n] * self.H[a] + m['mH%s2' % n] * CE[a]
                     + m['mH%s3' % n] * self.ICE[a] + m['mH%s4' % n] * self.IH[a])
        self.H = H
        CH = [curl_H(a, *self.H, dL) for a in range(3)]
        self.ICH = [self.ICH[a] + CH[a] for a in range(3)]
        self.ID = [self.ID[a] + self.D[a] for a in range(3)]
        D = []
        for a, n in enumerate('xyz'):
            D.append(m['mD%s1' % n] * self.D[a] + m['mD%s2' % n] * CH[a]
                     + m['mD%s3' % n] * self.ICH[a] + m['mD%s4' % n] * self.ID[a])
        for a, J in enumerate((Jx, Jy, Jz)):
            if J is not None:
                D[a] = D[a] + J
        self.D = D
        self.E = [self.mE[a] * self.D[a] for a in range(3)]
        return self.fields()

    def fields(self):
        vals = list(self.E) + list(self.D) + list(self.H)
        return dict(zip(FIELDS, vals))


def run_probes(eps_r, dL, npml, src_pos, pulse, steps, probes, component='Ez',
               src_component='Jz', dtype=np.float64):
    """Runs ``steps`` steps with a point source J[src_pos] = pulse(t) for
    t = 0..steps-1 and returns the (steps, n_probes) trace of ``component``."""
    F = FDTD(eps_r, dL, npml, dtype=dtype)
    J = np.zeros(F.shape, dtype=dtype)
    trace = np.zeros((steps, len(probes)), dtype=dtype)
    key = {'Jx': 'Jx', 'Jy': 'Jy', 'Jz': 'Jz'}[src_component]
    for t in range(steps):
        J[src_pos] = pulse(t)
        f = F.forward(**{key: J})
        for k, p in enumerate(probes):
            trace[t, k] = f[component][p]
    return trace
