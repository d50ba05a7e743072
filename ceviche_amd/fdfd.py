"""FDFD simulation classes with the reference's API (ceviche/fdfd.py:11-244).

    F = fdfd_ez(omega, dL, eps_r, npml, bloch_phases=None)
    Hx, Hy, Ez = F.solve(source_z)
    F = fdfd_hz(omega, dL, eps_r, npml, bloch_phases=None)
    Ex, Ey, Hz = F.solve(source_z)

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
from ._fdfd_op import FdfdDeriv, FdfdOperator, FdfdSolve, HzMaterials, SolverOptions
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

    def _solve_fn(self, eps_vec, Mz_vec):
        if torch.is_complex(eps_vec):
            raise NotImplementedError('fdfd_hz with complex eps_r is not supported')
        m0, m1 = HzMaterials.apply(eps_vec, self.op)
        b = 1j * self.omega * Mz_vec
        Hz = self._solve_linear(m0, m1, b)
        Ex = FdfdDeriv.apply(Hz, m0, self.op, 2, 1 / 1j / self.omega / EPSILON_0)
        Ey = FdfdDeriv.apply(Hz, m1, self.op, 0, -1 / 1j / self.omega / EPSILON_0)
        return Ex, Ey, Hz
