"""FDTD simulation class with the reference's API (ceviche/fdtd.py:10-316).

    F = fdtd(eps_r, dL, npml)            # eps_r (Nx, Ny[, Nz]); npml list of 3
    fields = F.forward(Jx=None, Jy=None, Jz=None)   # one Yee step
    fields['Ez'], F.t_index, F.dt

Semantics kept from the reference: dt uses the 3D Courant limit even in 2D
(fdtd.py:213-222); assigning ``eps_r`` recomputes the update coefficients and
resets the fields and t_index (fdtd.py:63-72); J must have the full grid shape
(fdtd.py:125-127); arrays of more than 3 dimensions raise (utils.py:206-214).

Values: numpy in / numpy out (float64, like the reference), or torch in /
torch out on the GPU with autograd through hand-written adjoint kernels.  The
``precision`` keyword ('float64' default, or 'float32') selects the kernel
precision.  ``run`` is the fused many-step runner with device-side probes and a
checkpointed adjoint (no reference counterpart; the reference loops in Python).
"""
import numpy as np
import torch

from . import _lib
from ._fdtd_op import COMPONENTS, FdtdEFields, FdtdGrid, FdtdRun, FdtdStep, _probe_tensors, autotune_tm, time_step
from ._probes import ProbeSums, ProbeWeights

_FIELD_ORDER = ('Ex', 'Ey', 'Ez', 'Dx', 'Dy', 'Dz', 'Hx', 'Hy', 'Hz')


def reshape_to_ND(arr, N):
    """ceviche/utils.py:206-214."""
    ND = len(arr.shape)
    if ND > N:
        raise ValueError("array is larger than {} dimensional, given shape {}".format(N, arr.shape))
    extra_dims = (N - ND) * (1,)
    return arr.reshape(tuple(arr.shape) + extra_dims)


class fdtd:

    def __init__(self, eps_r, dL, npml, precision='float64'):
        if precision not in ('float64', 'float32'):
            raise ValueError("precision must be 'float64' or 'float32'")
        self.precision = precision
        self._tdtype = torch.float64 if precision == 'float64' else torch.float32
        eps_r = reshape_to_ND(eps_r, N=3)
        self.Nx, self.Ny, self.Nz = self.grid_shape = tuple(eps_r.shape)
        self.dL = dL
        self.npml = npml
        self.eps_r = eps_r

    def __repr__(self):
        return "FDTD(eps_r.shape={}, dL={}, NPML={})".format(self.grid_shape, self.dL, self.npml)

    def __str__(self):
        return "FDTD object:\n\tdomain size = {}\n\tdL = {}\n\tNPML = {}".format(self.grid_shape, self.dL, self.npml)

    # ------------------------------------------------------------ properties
    @property
    def dL(self):
        return self.__dL

    @dL.setter
    def dL(self, new_dL):
        """Resets the time step (fdtd.py:41-45)."""
        self.__dL = new_dL
        self.dt = time_step(new_dL)

    @property
    def npml(self):
        return self.__npml

    @npml.setter
    def npml(self, new_npml):
        """Recomputes the PML profiles with the current dt (fdtd.py:52-56)."""
        self.__npml = list(new_npml)
        self._sigma_dt = self.dt

    @property
    def eps_r(self):
        return self.__eps_r

    @eps_r.setter
    def eps_r(self, new_eps):
        """Recomputes coefficients and resets fields and t_index (fdtd.py:63-72)."""
        new_eps = reshape_to_ND(new_eps, N=3)
        self.__eps_r = new_eps
        self.grid_shape = self.Nx, self.Ny, self.Nz = tuple(new_eps.shape)
        self.N = self.Nx * self.Ny * self.Nz
        # The update coefficients are fixed here, with the dt and PML of this moment.
        self._grid = FdtdGrid(self.grid_shape, self.dL, self.npml, self._sigma_dt, dtype=self._tdtype)
        self._grid.desc.dt = self.dt
        self._as_torch = isinstance(new_eps, torch.Tensor)
        if self._as_torch:
            self._eps_dev = new_eps.to(self._grid.device, self._tdtype)
        else:
            self._eps_dev = torch.from_numpy(np.ascontiguousarray(new_eps, dtype=np.float64)).to(
                self._grid.device, self._tdtype)
        self.initialize_fields()

    def initialize_fields(self):
        """Zero fields and integrals, t_index = 0 (fdtd.py:147-211)."""
        self.t_index = 0
        self._state = self._grid.new_state()
        self._te_live = self.Nz != 1  # TE components can only become non-zero through Jx / Jy

    # ---------------------------------------------------------------- steps
    def _prep_J(self, J):
        if J is None:
            return None
        if isinstance(J, torch.Tensor):
            self._as_torch = True
            t = J.to(self._grid.device, self._tdtype)
        else:
            t = torch.from_numpy(np.ascontiguousarray(J, dtype=np.float64)).to(self._grid.device, self._tdtype)
        if tuple(t.shape) != self.grid_shape:
            raise ValueError('J must have the full grid shape %s (got %s)' % (self.grid_shape, tuple(t.shape)))
        return t

    def forward(self, Jx=None, Jy=None, Jz=None):
        """One time step of FDTD (fdtd.py:74-144); returns the 9 fields."""
        self._advance(Jx, Jy, Jz)
        return self.fields

    def _advance(self, Jx=None, Jy=None, Jz=None):
        """forward() without building the returned field dictionary."""
        self._grid.desc.dL = self.dL  # curls use the current dL, like the reference
        Js = [self._prep_J(J) for J in (Jx, Jy, Jz)]
        if Js[0] is not None or Js[1] is not None:
            self._te_live = True
        self.t_index += 1
        self._state = FdtdStep.apply(self._state, self._eps_dev, Js[0], Js[1], Js[2], self._grid)

    def _device_field(self, name):
        """One field component as a device tensor of the grid shape."""
        c = COMPONENTS[name]
        if c < 3:
            return FdtdEFields.apply(self._state, self._eps_dev, self._grid)[c]
        return self._state[c if c < 6 else c - 6]

    @property
    def fields(self):
        E = FdtdEFields.apply(self._state, self._eps_dev, self._grid)
        out = {}
        for name in _FIELD_ORDER:
            c = COMPONENTS[name]
            out[name] = E[c] if c < 3 else self._state[c if c < 6 else c - 6]
        if self._as_torch:
            return out
        return {k: v.detach().cpu().numpy() for k, v in out.items()}

    def __getattr__(self, name):
        # Ex, Hy, ... attributes of the reference object
        if name in COMPONENTS and '_state' in self.__dict__:
            return self.fields[name]
        raise AttributeError(name)

    def run(self, steps, source_pos, amplitudes, probes, source_component='Jz', checkpoint_every=50,
            kernel='auto'):
        """Fused runner: ``steps`` Yee steps with a point source
        J[source_component][source_pos] = amplitudes[t] (t = 0..steps-1, added
        at every step like forward(J)) recording ``probes`` = [(field name,
        position), ...] after each step.  Returns a (steps, n_probes) tensor
        that is differentiable with respect to eps_r (torch) through a
        checkpointed adjoint; advances t_index and the internal state.

        A probe position may also be a weight array of the grid shape (a
        line, plane or mode profile): the probe then records
        sum(field * weights), the quantity measure_fields records
        (ceviche/utils.py:316-332).  The cells of all weighted probes are
        recorded in-kernel and combined by one weighted-sum kernel after the
        run.

        kernel: 'auto' uses the temporally blocked 2D TM kernel when it applies
        (nz == 1, Jz source, TE fields zero, Ez/Hx/Hy probes, <= 256 probes;
        tile shape from the autotuner) and the two-phase per-step kernels
        otherwise (measured faster than the single-pass kernels at 256^3:
        375 us per step against 420 us for the slab-marching kernel);
        'generic' forces the two-phase kernels, 'fused' the single-pass
        kernel (slab-marching when nz is a multiple of 64 up to 256,
        x-marching tiles otherwise in 3D), an int a TM tile configuration."""
        g = self._grid
        g.desc.dL = self.dL
        comp = {'Jx': 0, 'Jy': 1, 'Jz': 2}[source_component]
        pos = tuple(source_pos) + (0,) * (3 - len(source_pos))
        cell = (int(pos[0]) * g.ny + int(pos[1])) * g.nz + int(pos[2])
        if not 0 <= cell < g.n:
            raise ValueError('source position outside the grid')
        amp = amplitudes if isinstance(amplitudes, torch.Tensor) else torch.as_tensor(np.asarray(amplitudes, float))
        if amp.numel() < steps:
            raise ValueError('need one amplitude per step')
        weights = None
        if any(not isinstance(p[1], (tuple, list)) for p in probes):
            weights = ProbeWeights(probes, g.shape, g.device)
            probes = weights.points
        pt = _probe_tensors(g, probes)
        tm_ok = (g.nz == 1 and comp == 2 and not self._te_live and len(probes) <= 256
                 and all(COMPONENTS[p[0]] in (2, 6, 7) for p in probes))
        if isinstance(kernel, int):
            if not tm_ok:
                raise ValueError('the TM kernel does not apply to this run')
            cfg = kernel
        elif kernel in ('generic', 'fused'):
            cfg = kernel
        elif kernel == 'auto':
            cfg = autotune_tm(g, self._eps_dev.detach()) if tm_ok else 'generic'
        else:
            raise ValueError("kernel must be 'auto', 'generic', 'fused' or a TM configuration index")
        rec, state = FdtdRun.apply(self._eps_dev, amp, g, self._state, (comp, cell), pt, int(steps),
                                   int(checkpoint_every), cfg)
        self._state = state
        self.t_index += int(steps)
        if weights is not None:
            rec = ProbeSums.apply(rec, weights)
        return rec


__all__ = ['fdtd']
