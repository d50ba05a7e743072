"""Device-side FDTD grid and its differentiable entry points.

``FdtdGrid`` holds the per-axis PML conductivity profiles (the reference's
sigma tensors vary along one axis each, fdtd.py:224-263) and the time step,
and wraps the FDTD C-ABI of include/ceviche_amd.h.

Two differentiable paths:
  * FdtdStep: one Yee step as a torch.autograd.Function (drop-in for the
    reference's per-step ``forward`` inside an autograd loop); its backward is
    the hand-written adjoint step kernel.
  * FdtdRun: many steps with a point source and probes, reverse mode by
    checkpoint-and-recompute (state saved every ``checkpoint_every`` steps;
    each segment is re-run forward to rebuild its D history, then swept
    backward with the adjoint kernels).
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check, lib, ptr, stream_ptr
from .constants import C_0, EPSILON_0

NSTATE = 18
COMPONENTS = {'Ex': 0, 'Ey': 1, 'Ez': 2, 'Dx': 3, 'Dy': 4, 'Dz': 5, 'Hx': 6, 'Hy': 7, 'Hz': 8}
_CODE = {torch.float32: _lib.CV_C64, torch.float64: _lib.CV_C128}  # CV_F32 == 0, CV_F64 == 1


def time_step(dL, stability_factor=0.5):
    """dt = 0.5 dL / (sqrt(3) C0) -- the 3D Courant limit even in 2D; fdtd.py:213-222."""
    dL_sum = 3 / dL ** 2
    dL_avg = 1 / np.sqrt(dL_sum)
    return dL_avg / C_0 * stability_factor


def sigma_profiles(shape, npml, dt):
    """Per-axis (sigH, sigD) 1D profiles sampled from the doubled grid; fdtd.py:224-263.

    On the 2x grid, index 2p - n + 1 (low side) and 2N - 2p + n (high side) get
    (0.5 eps0/dt) (n / 2p)^3 for n < 2p; H takes odd, D even samples along the
    component's own axis."""
    sH, sD = [], []
    for a in range(3):
        N, p = int(shape[a]), int(npml[a])
        s2 = np.zeros(2 * N)
        for n in range(2 * p):
            val = (0.5 * EPSILON_0 / dt) * (n / 2 / p) ** 3
            s2[2 * p - n + 1] = val
            s2[2 * N - 2 * p + n] = val
        sH.append(s2[1::2].copy())
        sD.append(s2[0::2].copy())
    return sH, sD


class FdtdGrid:
    """Geometry, time step and PML profiles of one FDTD problem on the GPU."""

    def __init__(self, shape, dL, npml, dt, dtype=torch.float64, device=None):
        self.shape = tuple(int(v) for v in shape)
        self.nx, self.ny, self.nz = self.shape
        self.n = self.nx * self.ny * self.nz
        self.dL = float(dL)
        self.dt = float(dt)
        self.npml = list(npml)
        self.dtype = dtype
        self.code = _CODE[dtype]
        self.device = device if device is not None else _lib.device()
        sH, sD = sigma_profiles(self.shape, self.npml, self.dt)
        self.sig_host = np.concatenate(sH + sD)
        self.sig = torch.from_numpy(self.sig_host).to(self.device, dtype)
        self.desc = _lib.FdtdDesc(self.nx, self.ny, self.nz, 0, self.dt, self.dL, C_0, EPSILON_0)
        self._tmp = None

    def new_state(self):
        return torch.zeros((NSTATE,) + self.shape, dtype=self.dtype, device=self.device)

    def _d(self):
        return ctypes.byref(self.desc)

    def step(self, state, eps, J=(None, None, None), jscale=1.0, n_steps=1, src=None, amp=None,
             probes=None, rec=None):
        """Advance ``state`` in place.  src = (comp 0..2, flat cell); amp: host
        float64 array of n_steps; probes = (comp int32 tensor, cell int64 tensor)."""
        comp, cell = (-1, 0) if src is None else (int(src[0]), int(src[1]))
        amp_p = None
        if comp >= 0:
            amp = np.ascontiguousarray(amp, dtype=np.float64)
            assert amp.size >= n_steps
            amp_p = amp.ctypes.data_as(ctypes.c_void_p)
        pc, pl = (None, None) if probes is None else probes
        npb = 0 if probes is None else int(pc.numel())
        Jx, Jy, Jz = (None if j is None else j.to(self.device, self.dtype).contiguous() for j in J)
        check(lib().cv_fdtd_step(self.code, self._d(), ptr(self.sig), ptr(eps), ptr(state), ptr(Jx), ptr(Jy),
                                 ptr(Jz), float(jscale), comp, cell, amp_p, int(n_steps), ptr(pc), ptr(pl), npb,
                                 ptr(rec), stream_ptr()), 'fdtd step')
        return state

    def step_tangent(self, state, tstate, eps, d_eps, n_steps, src=None, amp=None, d_amp=None, probes=None,
                     trec=None, rec=None, J=(None, None, None), dJ=(None, None, None), jscale=1.0):
        """Forward mode: advance ``state`` and its tangent ``tstate`` in place
        along d_eps (tensor like eps), the dense source tangents dJ and d_amp
        (host float64 array or None); tangents of the probes go to trec
        (n_steps, n_probes)."""
        Jx, Jy, Jz = (None if j is None else j.to(self.device, self.dtype).contiguous() for j in J)
        dJx, dJy, dJz = (None if j is None else j.to(self.device, self.dtype).contiguous() for j in dJ)
        comp, cell = (-1, 0) if src is None else (int(src[0]), int(src[1]))
        amp_p = damp_p = None
        if comp >= 0:
            amp = np.ascontiguousarray(amp, dtype=np.float64)
            amp_p = amp.ctypes.data_as(ctypes.c_void_p)
            if d_amp is not None:
                d_amp = np.ascontiguousarray(d_amp, dtype=np.float64)
                damp_p = d_amp.ctypes.data_as(ctypes.c_void_p)
        pc, pl = (None, None) if probes is None else probes
        npb = 0 if probes is None else int(pc.numel())
        check(lib().cv_fdtd_step_tangent(self.code, self._d(), ptr(self.sig), ptr(eps), ptr(state), ptr(tstate),
                                         ptr(d_eps), ptr(Jx), ptr(Jy), ptr(Jz), ptr(dJx), ptr(dJy), ptr(dJz),
                                         float(jscale), comp, cell, amp_p, damp_p, int(n_steps), ptr(pc), ptr(pl),
                                         npb, ptr(rec), ptr(trec), stream_ptr()), 'fdtd step_tangent')

    def efields_tangent(self, state, tstate, eps, d_eps):
        dE = torch.empty((3,) + self.shape, dtype=self.dtype, device=self.device)
        check(lib().cv_fdtd_efields_tangent(self.code, self._d(), ptr(self.sig), ptr(eps), ptr(state), ptr(tstate),
                                            ptr(d_eps), ptr(dE), stream_ptr()), 'fdtd efields_tangent')
        return dE

    def run_tm(self, state, eps, src_cell, amp_dev, t0, n_steps, probes, rec, rec_t0, config):
        """Fused, temporally blocked 2D TM steps (nz == 1, TE fields zero, Jz point source)."""
        pc, pl = (None, None) if probes is None else probes
        npb = 0 if probes is None else int(pc.numel())
        alt = getattr(self, '_tm_alt', None)
        if alt is None or alt.shape != state.shape or alt.dtype != state.dtype:
            alt = self._tm_alt = torch.empty_like(state)
        check(lib().cv_fdtd_run_tm(self.code, self._d(), ptr(self.sig), ptr(eps), ptr(state), ptr(alt),
                                   int(src_cell), ptr(amp_dev), int(t0), int(n_steps), ptr(pc), ptr(pl), npb,
                                   ptr(rec), int(rec_t0), int(config), stream_ptr()), 'fdtd run_tm')
        return state

    def run_fused(self, state, other, eps, comp, src_cell, amp_dev, t0, n_steps, probes, rec, rec_t0):
        """Fused single-pass steps ping-ponging between ``state`` and ``other``;
        returns (current, spare) after the steps."""
        pc, pl = (None, None) if probes is None else probes
        npb = 0 if probes is None else int(pc.numel())
        check(lib().cv_fdtd_run_fused(self.code, self._d(), ptr(self.sig), ptr(eps), ptr(state), ptr(other),
                                      int(comp), int(src_cell), ptr(amp_dev), int(t0), int(n_steps), ptr(pc),
                                      ptr(pl), npb, ptr(rec), int(rec_t0), stream_ptr()), 'fdtd run_fused')
        return (other, state) if n_steps % 2 else (state, other)

    def efields(self, state, eps):
        E = torch.empty((3,) + self.shape, dtype=self.dtype, device=self.device)
        check(lib().cv_fdtd_efields(self.code, self._d(), ptr(self.sig), ptr(eps), ptr(state), ptr(E),
                                    stream_ptr()), 'fdtd efields')
        return E

    def efields_grad(self, state, eps, gE, gD, gmE):
        check(lib().cv_fdtd_efields_grad(self.code, self._d(), ptr(self.sig), ptr(eps), ptr(state),
                                         ptr(gE.contiguous()), ptr(gD), ptr(gmE), stream_ptr()),
              'fdtd efields_grad')

    def tmp(self):
        if self._tmp is None:
            self._tmp = torch.empty((3,) + self.shape, dtype=self.dtype, device=self.device)
        return self._tmp

    def adjoint(self, lam, eps, D_prev, gmE, probes=None, g_probe=None, D_t=None, src=None, g_src=None):
        """Reverse one step in place (lam: after step t -> before step t).
        With src = (comp, cell) and g_src (a one-element device tensor of the
        grid dtype), g_src receives the cotangent of the point-source
        amplitude added at this step."""
        pc, pl = (None, None) if probes is None else probes
        npb = 0 if probes is None else int(pc.numel())
        comp, cell = (-1, 0) if src is None else (int(src[0]), int(src[1]))
        check(lib().cv_fdtd_adjoint_step(self.code, self._d(), ptr(self.sig), ptr(eps), ptr(lam), ptr(self.tmp()),
                                         ptr(D_prev), ptr(gmE), ptr(pc), ptr(pl), npb, ptr(g_probe), ptr(D_t),
                                         comp, cell, ptr(g_src), stream_ptr()), 'fdtd adjoint')

    def eps_grad(self, eps, gmE, geps):
        check(lib().cv_fdtd_eps_grad(self.code, self._d(), ptr(self.sig), ptr(eps), ptr(gmE), ptr(geps),
                                     stream_ptr()), 'fdtd eps_grad')
        return geps


class FdtdStep(torch.autograd.Function):
    """state_t = step(state_{t-1}; eps, J): one reference ``forward`` (fdtd.py:74-144)."""

    @staticmethod
    def forward(ctx, state, eps, Jx, Jy, Jz, grid):
        new = state.detach().clone()
        grid.step(new, eps, J=(Jx, Jy, Jz))
        ctx.grid = grid
        ctx.save_for_backward(state, eps)
        ctx.save_for_forward(state, eps, Jx, Jy, Jz)
        ctx.has_J = (Jx is not None, Jy is not None, Jz is not None)
        return new

    @staticmethod
    def jvp(ctx, d_state, d_eps, dJx, dJy, dJz, _):
        state, eps, Jx, Jy, Jz = ctx.saved_tensors
        grid = ctx.grid
        st = state.detach().clone()
        tst = grid.new_state() if d_state is None else d_state.to(grid.dtype).contiguous().clone()
        de = torch.zeros_like(eps) if d_eps is None else d_eps.to(eps.dtype).contiguous()
        grid.step_tangent(st, tst, eps.detach(), de, 1, J=(Jx, Jy, Jz), dJ=(dJx, dJy, dJz))
        return tst

    @staticmethod
    def backward(ctx, g):
        state, eps = ctx.saved_tensors
        grid = ctx.grid
        lam = g.detach().clone().contiguous()
        gJ = [lam[3 + a].clone() if (ctx.has_J[a] and ctx.needs_input_grad[2 + a]) else None for a in range(3)]
        gmE = torch.zeros((3,) + grid.shape, dtype=grid.dtype, device=grid.device)
        grid.adjoint(lam, eps, state[3:6].contiguous(), gmE)
        geps = None
        if ctx.needs_input_grad[1]:
            geps = grid.eps_grad(eps, gmE, torch.zeros_like(eps))
        return lam, geps, gJ[0], gJ[1], gJ[2], None


class FdtdEFields(torch.autograd.Function):
    """E = D / eps_aa (fdtd.py:134-137)."""

    @staticmethod
    def forward(ctx, state, eps, grid):
        ctx.grid = grid
        ctx.save_for_backward(state, eps)
        ctx.save_for_forward(state, eps)
        return grid.efields(state, eps)

    @staticmethod
    def jvp(ctx, d_state, d_eps, _):
        state, eps = ctx.saved_tensors
        grid = ctx.grid
        tst = grid.new_state() if d_state is None else d_state.to(grid.dtype).contiguous()
        de = torch.zeros_like(eps) if d_eps is None else d_eps.to(eps.dtype).contiguous()
        return grid.efields_tangent(state.detach(), tst, eps.detach(), de)

    @staticmethod
    def backward(ctx, gE):
        state, eps = ctx.saved_tensors
        grid = ctx.grid
        gstate = torch.zeros_like(state)
        gmE = torch.zeros((3,) + grid.shape, dtype=grid.dtype, device=grid.device) if ctx.needs_input_grad[1] else None
        grid.efields_grad(state, eps, gE, gstate[3:6], gmE)
        geps = grid.eps_grad(eps, gmE, torch.zeros_like(eps)) if gmE is not None else None
        return gstate, geps, None


_TUNE_CACHE = {}


def tm_configs():
    """[(waves per workgroup, rows per thread, steps per launch), ...]"""
    n = lib().cv_fdtd_tm_configs(None, None, None, 0)
    rpt = (ctypes.c_int * n)()
    ts = (ctypes.c_int * n)()
    wv = (ctypes.c_int * n)()
    lib().cv_fdtd_tm_configs(rpt, ts, wv, n)
    return [(wv[i], rpt[i], ts[i]) for i in range(n)]


def autotune_tm(grid, eps, steps=64, reps=2):
    """Times every TM tile configuration on this grid (a scratch copy of the
    state) and returns the fastest; cached per (shape, dtype)."""
    key = (grid.shape, str(grid.dtype))
    if key in _TUNE_CACHE:
        return _TUNE_CACHE[key]
    st = grid.new_state()
    amp = torch.zeros(steps, dtype=torch.float64, device=grid.device)
    best, best_t = 0, float('inf')
    for cfg in range(len(tm_configs())):
        grid.run_tm(st, eps, -1, amp, 0, steps, None, None, 0, cfg)  # warm-up
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            grid.run_tm(st, eps, -1, amp, 0, steps, None, None, 0, cfg)
        e1.record()
        e1.synchronize()
        t = e0.elapsed_time(e1)
        if t < best_t:
            best, best_t = cfg, t
    _TUNE_CACHE[key] = best
    return best


def _probe_tensors(grid, probes):
    comps = []
    cells = []
    for name, pos in probes:
        comps.append(COMPONENTS[name] if isinstance(name, str) else int(name))
        pos = tuple(pos) + (0,) * (3 - len(pos))
        i, j, k = (int(v) for v in pos)
        if not (0 <= i < grid.nx and 0 <= j < grid.ny and 0 <= k < grid.nz):
            raise ValueError('probe position %s outside the grid %s' % (pos, grid.shape))
        cells.append((i * grid.ny + j) * grid.nz + k)
    pc = torch.tensor(comps, dtype=torch.int32, device=grid.device)
    pl = torch.tensor(cells, dtype=torch.int64, device=grid.device)
    return pc, pl


class FdtdRun(torch.autograd.Function):
    """rec[t, p] = probe p after step t0+t for a point source amp(t), many steps in
    one call; reverse mode with checkpoint-and-recompute."""

    @staticmethod
    def forward(ctx, eps, amp, grid, state0, src, probes, n_steps, checkpoint_every, tm_config=None):
        # The amplitude series is read by the temporally blocked and fused
        # kernels from device memory (a float64 device tensor is used as it
        # is, anything else is uploaded once) and by the per-step kernels from
        # host memory (copied once, only for those).
        a = amp.detach()
        amp_dev = amp_h = None
        if tm_config != 'generic':
            amp_dev = a.to(grid.device, torch.float64).contiguous()
        if tm_config in ('generic', 'fused'):
            amp_h = a.to('cpu', torch.float64).numpy()
        state = state0.detach().clone()
        spare = torch.empty_like(state) if tm_config == 'fused' else None
        pc, pl = probes
        rec = torch.empty((n_steps, pc.numel()), dtype=grid.dtype, device=grid.device)
        ck = max(1, int(checkpoint_every))
        need_grad = ctx.needs_input_grad[0] or ctx.needs_input_grad[1]
        checkpoints = []
        t = 0
        while t < n_steps:
            m = min(ck, n_steps - t) if need_grad else n_steps - t
            if need_grad:
                checkpoints.append(state.clone())
            if tm_config == 'fused':
                state, spare = grid.run_fused(state, spare, eps, src[0], src[1], amp_dev, t, m, probes, rec, 0)
            elif tm_config == 'generic':
                grid.step(state, eps, n_steps=m, src=src, amp=amp_h[t:t + m], probes=probes, rec=rec[t:t + m])
            else:
                grid.run_tm(state, eps, src[1], amp_dev, t, m, probes, rec, 0, tm_config)
            t += m
        ctx.grid, ctx.src, ctx.probes, ctx.n_steps, ctx.ck = grid, src, probes, n_steps, ck
        ctx.amp_h, ctx.amp_dev, ctx.tm_config = amp_h, amp_dev, tm_config
        ctx.amp_meta = (amp.shape, amp.dtype, amp.device)
        ctx.checkpoints = checkpoints
        ctx.state0 = state0.detach().clone()
        ctx.save_for_backward(eps)
        ctx.save_for_forward(eps)
        return rec, state

    @staticmethod
    def jvp(ctx, d_eps, d_amp, d_grid, d_state0, *_):
        """Forward mode: the run is replayed from the initial state with the
        tangent carried along (cv_fdtd_step_tangent), giving the tangents of
        the records and of the final state."""
        (eps,) = ctx.saved_tensors
        grid, src, probes, n_steps = ctx.grid, ctx.src, ctx.probes, ctx.n_steps
        pc, _ = probes
        state = ctx.state0.clone()
        tstate = grid.new_state() if d_state0 is None else d_state0.to(grid.dtype).contiguous().clone()
        de = torch.zeros_like(eps) if d_eps is None else d_eps.to(eps.dtype).contiguous()
        da = None if d_amp is None else d_amp.detach().to('cpu', torch.float64).numpy()
        amp_h = ctx.amp_h if ctx.amp_h is not None else ctx.amp_dev.cpu().numpy()
        trec = torch.empty((n_steps, pc.numel()), dtype=grid.dtype, device=grid.device)
        grid.step_tangent(state, tstate, eps.detach(), de, n_steps, src=src, amp=amp_h, d_amp=da, probes=probes,
                          trec=trec)
        return trec, tstate

    @staticmethod
    def backward(ctx, grec, gstate_final):
        (eps,) = ctx.saved_tensors
        grid, src, probes, n_steps, ck = ctx.grid, ctx.src, ctx.probes, ctx.n_steps, ctx.ck
        lam = grid.new_state()
        if gstate_final is not None:
            lam += gstate_final
        gmE = torch.zeros((3,) + grid.shape, dtype=grid.dtype, device=grid.device)
        grec = grec.to(grid.dtype).contiguous() if grec is not None else None
        nseg = len(ctx.checkpoints)
        # cotangent of amp[t]: lam(D_src) after step t (source added to D, fdtd.py:122-124)
        gamp = torch.zeros(n_steps, dtype=grid.dtype, device=grid.device) if ctx.needs_input_grad[1] else None
        # D history of one segment: Dhist[s] = D after step (seg_start + s)
        Dhist = torch.empty((ck + 1, 3) + grid.shape, dtype=grid.dtype, device=grid.device)
        for sgi in range(nseg - 1, -1, -1):
            t0 = sgi * ck
            m = min(ck, n_steps - t0)
            st = ctx.checkpoints[sgi].clone()
            Dhist[0].copy_(st[3:6])
            for s in range(m):
                if isinstance(ctx.tm_config, int):
                    grid.run_tm(st, eps, src[1], ctx.amp_dev, t0 + s, 1, None, None, 0, ctx.tm_config)
                else:
                    grid.step(st, eps, n_steps=1, src=src, amp=ctx.amp_h[t0 + s:t0 + s + 1])
                Dhist[s + 1].copy_(st[3:6])
            for s in range(m - 1, -1, -1):
                t = t0 + s
                ga = None if gamp is None else gamp[t:t + 1]
                if grec is not None:
                    gp = grec[t]
                    grid.adjoint(lam, eps, Dhist[s], gmE, probes=probes, g_probe=gp, D_t=Dhist[s + 1],
                                 src=src, g_src=ga)
                else:
                    grid.adjoint(lam, eps, Dhist[s], gmE, src=src, g_src=ga)
            del st
        geps = grid.eps_grad(eps, gmE, torch.zeros_like(eps)) if ctx.needs_input_grad[0] else None
        g_amp = None
        if gamp is not None:
            shape, dtype, dev = ctx.amp_meta
            g_amp = torch.zeros(shape, dtype=dtype, device=dev)
            g_amp.view(-1)[:n_steps] = gamp.to(dev, dtype)
        # lam now holds the cotangent of the state the run started from
        gstate0 = lam if ctx.needs_input_grad[3] else None
        # the checkpoints stay with the graph: a retained graph (one backward
        # per output row of a Jacobian) runs this more than once
        return geps, g_amp, None, gstate0, None, None, None, None, None
