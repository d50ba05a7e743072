"""FDFD simulation classes with the reference's API (ceviche/fdfd.py:11-244).

    F = fdfd_ez(omega, dL, eps_r, npml, bloch_phases=None)
    Hx, Hy, Ez = F.solve(source_z)
    F = fdfd_hz(omega, dL, eps_r, npml, bloch_phases=None)
    Ex, Ey, Hz = F.solve(source_z)
    F = fdfd_mf_ez(omega, dL, eps_r, omega_mod, delta, phi, Nsb, npml, bloch_phases=None)
    Hx, Hy, Ez = F.solve(source_z)        # fields and source shaped (2 Nsb + 1, Nx, Ny)

Inputs may be numpy arrays (outputs are numpy, complex128, as in the
reference) or torch tensors (outputs are torch tensors on the GPU that carry
autograd history: gradients flow to eps_r and the source through hand-written
adjoint kernels).  All arithmetic runs in HIP kernels: a matrix-free SC-PML
stencil and a device-resident BiCGSTAB replace the COO assembly and the
SuperLU factorisation.

Extra keyword arguments (not in the reference):
    precision: 'complex128' (default; complex128 BiCGSTAB, FP64 reductions) or
               'complex64' (complex64 BiCGSTAB corrections inside complex128
               iterative refinement, FP64 residual).
    solver_options: a SolverOptions (tol = relative FP64 residual).
"""
import numpy as np
import torch

from . import _lib
from ._fdfd_op import FdfdCoupledSolve, FdfdDeriv, FdfdOperator, FdfdSolve, HzMaterials, SolverOptions
from .constants import EPSILON_0, MU_0


def _to_tensor(a, device, dtype=None):
    if isinstance(a, torch.Tensor):
        t = a.to(device)
    else:
        t = torch.from_numpy(np.ascontiguousarray(a)).to(device)
    if dtype is not None and t.dtype != dtype:
        t = t.to(dtype)
    return t


class fdfd:
    """Base class (ceviche/fdfd.py:11-162)."""

    def __init__(self, omega, dL, eps_r, npml, bloch_phases=None, precision='complex128', solver_options=None):
        self.omega = omega
        self.dL = dL
        self.npml = npml
        self._setup_bloch_phases(bloch_phases)
        self.eps_r = eps_r
        if precision not in ('complex128', 'complex64'):
            raise ValueError("precision must be 'complex128' or 'complex64'")
        self.precision = precision
        self.solver_options = solver_options or SolverOptions()
        self._setup_derivatives()

    # eps_r setter saves the shape only; operators are NOT rebuilt (fdfd.py:40-44)
    @property
    def eps_r(self):
        return self._eps_r

    @eps_r.setter
    def eps_r(self, new_eps):
        self._save_shape(new_eps)
        self._eps_r = new_eps

    def _setup_bloch_phases(self, bloch_phases):
        """fdfd.py:101-112."""
        self.bloch_x = 0.0
        self.bloch_y = 0.0
        self.bloch_z = 0.0
        if bloch_phases is not None:
            self.bloch_x = bloch_phases[0]
            if len(bloch_phases) > 1:
                self.bloch_y = bloch_phases[1]
            if len(bloch_phases) > 2:
                self.bloch_z = bloch_phases[2]

    def _save_shape(self, grid):
        """fdfd.py:122-126."""
        self.shape = tuple(grid.shape)
        self.Nx, self.Ny = self.shape
        self.N = self.Nx * self.Ny

    def _setup_derivatives(self):
        """Builds the device operators (PML profiles) once, like fdfd.py:80-99."""
        self._op_shape = self.shape
        spec = dict(omega=self.omega, dL=self.dL, shape=self.shape, npml=self.npml,
                    bloch=(self.bloch_x, self.bloch_y))
        self.op = FdfdOperator(self._pol, [spec], dtype=torch.complex128)
        self.op32 = FdfdOperator(self._pol, [spec], dtype=torch.complex64) if self.precision == 'complex64' else None

    @property
    def last_solve_info(self):
        """Iterations, status and FP64 relative residual of the last solve."""
        op = self.op32 if self.op32 is not None else self.op
        return op.last_info[0] if op.last_info else None

    def _grid_to_vec(self, grid):
        return grid.reshape(-1)

    def _vec_to_grid(self, vec):
        return vec.reshape(self.shape)

    def solve(self, source_z):
        """(Fx, Fy, Fz) for a z source (fdfd.py:58-76)."""
        if self.shape != self._op_shape:
            raise ValueError('eps_r shape changed from %s to %s; the reference does not rebuild its operators '
                             'on eps_r assignment either -- build a new object' % (self._op_shape, self.shape))
        as_torch = isinstance(source_z, torch.Tensor) or isinstance(self.eps_r, torch.Tensor)
        dev = self.op.device
        src = _to_tensor(source_z, dev)
        if tuple(src.shape) != self.shape:
            raise ValueError('source shape %s does not match eps_r shape %s' % (tuple(src.shape), self.shape))
        eps = _to_tensor(self.eps_r, dev)
        if not torch.is_complex(eps):
            eps = eps.to(torch.float64)
        else:
            eps = eps.to(torch.complex128)
        src = src.to(torch.complex128)
        fx, fy, fz = self._solve_fn(self._grid_to_vec(eps), self._grid_to_vec(src))
        out = tuple(self._vec_to_grid(v) for v in (fx, fy, fz))
        if as_torch:
            return out
        return tuple(v.detach().cpu().numpy() for v in out)

    def _solve_linear(self, m0, m1, b):
        return FdfdSolve.apply(m0, m1, b, self.op, self.op32, self.solver_options)


class fdfd_ez(fdfd):
    """Ez polarisation (fdfd.py:166-192)."""
    _pol = 'ez'

    def __init__(self, omega, dL, eps_r, npml, bloch_phases=None, **kw):
        super().__init__(omega, dL, eps_r, npml, bloch_phases=bloch_phases, **kw)

    def _solve_fn(self, eps_vec, Jz_vec):
        m0, m1 = self.op.materials(eps_vec)
        b = 1j * self.omega * Jz_vec
        Ez = self._solve_linear(m0, m1, b)
        Hx = FdfdDeriv.apply(Ez, None, self.op, 3, -1 / 1j / self.omega / MU_0)
        Hy = FdfdDeriv.apply(Ez, None, self.op, 1, 1 / 1j / self.omega / MU_0)
        return Hx, Hy, Ez


class fdfd_hz(fdfd):
    """Hz polarisation (fdfd.py:194-244)."""
    _pol = 'hz'

    def __init__(self, omega, dL, eps_r, npml, bloch_phases=None, **kw):
        super().__init__(omega, dL, eps_r, npml, bloch_phases=bloch_phases, **kw)

    def _complex_ops(self):
        """Operators for complex eps_r: the Hz materials 1/(eps_xx + 1e-5) and
        1/(eps_yy + 1e-5) are then complex (fdfd.py:200-215 on complex
        arrays); built on first use, solved by BiCGSTAB."""
        if getattr(self, '_cops', None) is None:
            spec = dict(omega=self.omega, dL=self.dL, shape=self.shape, npml=self.npml,
                        bloch=(self.bloch_x, self.bloch_y))
            op = FdfdOperator('hz', [spec], dtype=torch.complex128, complex_materials=True)
            op32 = (FdfdOperator('hz', [spec], dtype=torch.complex64, complex_materials=True)
                    if self.precision == 'complex64' else None)
            self._cops = (op, op32)
        return self._cops

    @property
    def last_solve_info(self):
        ops = self._cops if getattr(self, '_last_complex', False) else (self.op, self.op32)
        op = ops[1] if ops[1] is not None else ops[0]
        return op.last_info[0] if op.last_info else None

    def _solve_fn(self, eps_vec, Mz_vec):
        self._last_complex = torch.is_complex(eps_vec)
        op, op32 = self._complex_ops() if self._last_complex else (self.op, self.op32)
        m0, m1 = HzMaterials.apply(eps_vec, op)
        b = 1j * self.omega * Mz_vec
        Hz = FdfdSolve.apply(m0, m1, b, op, op32, self.solver_options)
        Ex = FdfdDeriv.apply(Hz, m0, op, 2, 1 / 1j / self.omega / EPSILON_0)
        Ey = FdfdDeriv.apply(Hz, m1, op, 0, -1 / 1j / self.omega / EPSILON_0)
        return Ex, Ey, Hz


class fdfd_mf_ez(fdfd):
    """Multi-frequency Ez with a time-modulated permittivity
    (fdfd.py:247-359): eps(t) = eps_r + sum_h delta_h cos((h+1) omega_mod t
    + phi_h) couples the 2 Nsb + 1 sidebands W_m = omega + m omega_mod,
    m = -Nsb .. Nsb (Y. Shi, W. Shin, S. Fan, Optica 3(11), 2016).

    Arguments as the reference: delta and phi are (Nfreq, Nx, Ny) arrays of
    modulation depth and phase per harmonic; source_z and the returned fields
    are (2 Nsb + 1, Nx, Ny).  Attributes eps_r, delta, phi, Nsb and omega_mod
    are read at every solve, as the reference reads them in _make_A.

    Block m of the system (all blocks on the base-omega PML grid,
    fdfd.py:289) is  -(1/mu0)(Dxf Dxb + Dyf Dyb) - eps0 W_m^2 eps; harmonic h
    adds -eps0/2 W_m^2 delta_h exp(+i phi_h) on block (m, m+h+1) and
    -eps0/2 W_m^2 delta_h exp(-i phi_h) on block (m, m-h-1) (fdfd.py:303-324);
    b = i W_m Jz_m (fdfd.py:326-335); Hx = -Dyb Ez / (i mu0 W_m),
    Hy = Dxb Ez / (i mu0 W_m) (fdfd.py:337-351).  The whole system is one
    BiCGSTAB solve on the GPU with the coupling fused into the stencil
    kernels; gradients flow to eps_r, delta, phi and the source."""
    _pol = 'ez'

    def __init__(self, omega, dL, eps_r, omega_mod, delta, phi, Nsb, npml, bloch_phases=None, **kw):
        self.omega_mod = omega_mod
        self.delta = delta
        self.phi = phi
        self.Nsb = Nsb
        self._mf_key = None
        super().__init__(omega, dL, eps_r, npml, bloch_phases=bloch_phases, **kw)

    def _setup_derivatives(self):
        # the operators depend on Nsb, omega_mod and the harmonic count: built at solve time
        self._op_shape = self.shape
        self.op = None
        self.op32 = None

    def _mf_ops(self, n_cpl):
        M = 2 * int(self.Nsb) + 1
        key = (M, float(self.omega_mod), n_cpl, self.shape)
        if key != self._mf_key:
            W = self._sidebands()
            specs = []
            for m in range(M):
                specs.append(dict(omega=self.omega, dL=self.dL, shape=self.shape, npml=self.npml,
                                  bloch=(self.bloch_x, self.bloch_y), shift_omega=float(W[m]),
                                  system=(m, M, n_cpl)))
            self.op = FdfdOperator('ez', specs, dtype=torch.complex128)
            self.op32 = (FdfdOperator('ez', specs, dtype=torch.complex64)
                         if self.precision == 'complex64' else None)
            self._mf_key = key
        return self.op, self.op32

    def _sidebands(self):
        return self.omega + np.arange(-int(self.Nsb), int(self.Nsb) + 1) * self.omega_mod

    def _vec_to_grid(self, vec):
        """(2 Nsb + 1, Nx, Ny) grids, fdfd.py:353-357."""
        return vec.reshape((2 * int(self.Nsb) + 1,) + self.shape)

    def solve(self, source_z):
        """(Hx, Hy, Ez), each (2 Nsb + 1, Nx, Ny), for a Jz source of that shape
        (fdfd.py:261-280)."""
        if self.shape != self._op_shape:
            raise ValueError('eps_r shape changed from %s to %s; build a new object' % (self._op_shape, self.shape))
        M = 2 * int(self.Nsb) + 1
        if M < 1:
            raise ValueError('Nsb must be >= 0')
        as_torch = any(isinstance(v, torch.Tensor) for v in (source_z, self.eps_r, self.delta, self.phi))
        dev = _lib.device()
        src = _to_tensor(source_z, dev)
        if tuple(src.shape) != (M,) + self.shape:
            raise ValueError('source shape %s does not match (2 Nsb + 1, Nx, Ny) = %s'
                             % (tuple(src.shape), (M,) + self.shape))
        eps = _to_tensor(self.eps_r, dev)
        eps = eps.to(torch.complex128) if torch.is_complex(eps) else eps.to(torch.float64)
        delta = _to_tensor(self.delta, dev)
        phi = _to_tensor(self.phi, dev)
        if delta.dim() != 3 or tuple(delta.shape[1:]) != self.shape or tuple(phi.shape) != tuple(delta.shape):
            raise ValueError('delta and phi must both be (Nfreq, Nx, Ny) with (Nx, Ny) = %s' % (self.shape,))
        n_cpl = min(int(delta.shape[0]), M - 1)
        op, op32 = self._mf_ops(n_cpl)
        N = self.N
        W = self._sidebands()
        m0, m1 = op.materials(eps.reshape(-1).repeat(M))
        if n_cpl > 0:
            # coupling tables in units of omega^2 * mu0 dL^2 (the system scale);
            # block m multiplies them by (W_m / omega)^2
            kappa = -0.5 * EPSILON_0 * MU_0 * self.dL ** 2 * self.omega ** 2
            d = delta[:n_cpl].reshape(n_cpl, N).to(torch.complex128)
            ph = phi[:n_cpl].reshape(n_cpl, N).to(torch.complex128)
            gp = kappa * d * torch.exp(1j * ph)
            gm = kappa * d * torch.exp(-1j * ph)
            cpl = torch.stack([gp, gm], dim=1).reshape(-1)
        else:
            cpl = torch.zeros(0, dtype=torch.complex128, device=dev)
        w_cells = torch.from_numpy(np.repeat(W, N)).to(dev)
        b = 1j * w_cells * src.reshape(-1).to(torch.complex128)
        Ez = FdfdCoupledSolve.apply(m0, m1, cpl, b, op, op32, self.solver_options)
        Hx = FdfdDeriv.apply(Ez, None, op, 3, -1 / 1j / MU_0) / w_cells
        Hy = FdfdDeriv.apply(Ez, None, op, 1, 1 / 1j / MU_0) / w_cells
        out = tuple(self._vec_to_grid(v) for v in (Hx, Hy, Ez))
        if as_torch:
            return out
        return tuple(v.detach().cpu().numpy() for v in out)
