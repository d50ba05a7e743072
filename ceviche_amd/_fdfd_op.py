"""Device-side FDFD operators and their differentiable entry points.

``FdfdOperator`` holds a batch of 2D FDFD instances of one polarisation:
their descriptors (cv_fdfd_desc), the tile table that maps workgroups to
cells, and the 1D SC-PML derivative profiles, all resident on the GPU.  It
wraps the C-ABI calls of include/ceviche_amd.h.

The autograd Functions below are the reference's sparse primitives
specialised to the FDFD operator (ceviche/primitives.py:103-148 for the
solve, :58-98 for the derivative matrices), in torch's complex convention:
the cotangent of a complex z is dL/dRe z + i dL/dIm z, so reverse mode uses
A^H = conj(A)^T where the reference (autograd convention) uses A^T.
"""
import ctypes
import threading
import time
import warnings
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import torch

from . import _lib
from ._lib import CevicheError, check, lib, ptr, stream_ptr
from .constants import EPSILON_0, MU_0
from .derivatives import axis_profiles

_REAL_OF = {torch.complex128: torch.float64, torch.complex64: torch.float32}
_CODE_OF = {torch.complex128: _lib.CV_C128, torch.complex64: _lib.CV_C64}


class SolverOptions:
    """Solver settings.  tol is the relative FP64 residual ||b - A x|| / ||b||.

    method: 'auto' (a direct factorisation + FP64-refined FGMRES for each
    uncoupled instance, one instance after another in a batch: block cyclic
    reduction when ny <= 128, or nx <= 128 on the transposed grid, with real
    materials, nested dissection otherwise; BiCGSTAB for coupled systems or
    when a factor does not fit in half the free device memory), 'direct' (require a
    factorisation), 'bicgstab', or 'cocg' (COCG on the complex-symmetric
    form: zero Bloch phases, uncoupled instances).  restart: FGMRES
    cycle length of the direct path; direct_max_iter: its FGMRES step limit
    (min(max_iter, DIRECT_MAX_ITER) when None: with a factorisation the
    steps number a few to a few tens, and a factor that does not converge
    by then gives way to the fallback).  Under 'auto' a direct solve that fails (singular
    factor, FGMRES not converged, out of memory) falls back to the Krylov
    method with a warning; under 'direct' it raises."""

    def __init__(self, tol=None, max_iter=200000, check_every=32, inner_tol=1e-4, max_sweeps=12, method='auto',
                 restart=30, direct_max_iter=None):
        if method not in ('auto', 'direct', 'bicgstab', 'cocg'):
            raise ValueError("method must be 'auto', 'direct', 'bicgstab' or 'cocg'")
        self.tol = tol
        self.max_iter = max_iter
        self.direct_max_iter = direct_max_iter
        self.check_every = check_every
        self.inner_tol = inner_tol
        self.max_sweeps = max_sweeps
        self.method = method
        self.restart = restart


DEFAULT_TOL = {torch.complex128: 1e-10, torch.complex64: 1e-6}
DIRECT_MAX_ITER = 600


def _direct_max_iter(opts):
    return min(opts.max_iter, DIRECT_MAX_ITER) if opts.direct_max_iter is None else opts.direct_max_iter


class FdfdOperator:
    """A batch of FDFD instances (``specs``: list of dicts with keys omega, dL,
    shape, npml, bloch) of polarisation 'ez' or 'hz' in precision ``dtype``
    (torch.complex128 or torch.complex64).  Vectors of the batch are packed
    instance after instance, each in C order over (nx, ny).

    Frequency-coupled Ez systems (the multi-frequency operator,
    ceviche/fdfd.py:284-324): consecutive specs with
    ``system=(blk, nblk, n_cpl)`` form one linear system of nblk blocks on the
    same grid; block blk uses ``shift_omega`` (its sideband frequency W_blk)
    in the eps0 W^2 eps term and ``omega`` for the PML, and couples to blocks
    blk +- (h+1), h < n_cpl, through a complex table passed to the solve /
    apply calls (see ``prof_with``): for each h, gp_h then gm_h over the
    grid's cells, in units of omega^2 * (system scale), multiplied by
    w2[blk] = (W_blk / omega)^2."""

    def __init__(self, pol, specs, dtype=torch.complex128, device=None, complex_materials=False):
        if pol not in ('ez', 'hz'):
            raise ValueError("pol must be 'ez' or 'hz'")
        # Hz with complex eps: m0 = 1/(eps_xx + 1e-5) and m1 = 1/(eps_yy + 1e-5)
        # are complex arrays (the instances carry CV_FLAG_CPLX_MAT)
        self.cmat = bool(complex_materials) and pol == 'hz'
        if dtype not in _CODE_OF:
            raise ValueError('dtype must be torch.complex128 or torch.complex64')
        self.pol = pol
        self.dtype = dtype
        self.rdtype = _REAL_OF[dtype]
        self.code = _CODE_OF[dtype]
        self.device = device if device is not None else _lib.device()
        self.specs = list(specs)
        n_inst = len(self.specs)
        desc = np.zeros(n_inst, dtype=_lib.DESC_DTYPE)
        tiles = []
        profs = []
        cell_off = 0
        prof_off = 0
        self.sys_scale = np.zeros(n_inst)
        self.shapes = []
        self.offsets = []
        systems = []       # (first instance, nblk, n_cpl) of every coupled system
        w2 = []
        for k, sp in enumerate(self.specs):
            nx, ny = (int(v) for v in sp['shape'])
            omega, dL = float(sp['omega']), float(sp['dL'])
            npml = sp['npml']
            bx, by = sp.get('bloch', (0.0, 0.0))
            kxf, kxb = axis_profiles(omega, dL, nx, int(npml[0]))
            kyf, kyb = axis_profiles(omega, dL, ny, int(npml[1]))
            profs += [kxf, kxb, kyf, kyb]
            n = nx * ny
            ntile = (n + _lib.CV_TILE_CELLS - 1) // _lib.CV_TILE_CELLS
            first_tile = len(tiles)
            tiles += [(k, t * _lib.CV_TILE_CELLS) for t in range(ntile)]
            # The system is scaled by c = mu0 dL^2 (Ez) or eps0 dL^2 (Hz) so the
            # stencil coefficients are O(1) in either precision; the solution is
            # unchanged and the caller's right-hand side is scaled by c as well.
            blk, nblk, n_cpl = sp.get('system', (0, 1, 0))
            w_shift = float(sp.get('shift_omega', omega))
            if nblk > 1 or n_cpl:
                if pol != 'ez':
                    raise ValueError('frequency-coupled systems are Ez only')
                if not (0 <= blk < nblk and n_cpl >= 0):
                    raise ValueError('bad system fields %r' % ((blk, nblk, n_cpl),))
                if blk == 0:
                    systems.append((k, nblk, n_cpl))
                else:
                    r = self.specs[k - blk] if k - blk >= 0 else None
                    if (r is None or tuple(r.get('system', ())) != (0, nblk, n_cpl)
                            or tuple(int(v) for v in r['shape']) != (nx, ny)):
                        raise ValueError('block %d of a coupled system must follow block 0 on the same grid' % k)
                w2.append((w_shift / omega) ** 2)
            if pol == 'ez':
                c = MU_0 * dL ** 2
                scale, dshift = -1 / MU_0 * c, -EPSILON_0 * w_shift ** 2 * c
            else:
                c = EPSILON_0 * dL ** 2
                scale, dshift = 1 / EPSILON_0 * c, MU_0 * omega ** 2 * c
            self.sys_scale[k] = c
            phx, phy = np.exp(1j * bx), np.exp(1j * by)
            flags = (_lib.CV_FLAG_X_IDENTITY if nx == 1 else 0) | (_lib.CV_FLAG_Y_IDENTITY if ny == 1 else 0)
            if self.cmat:
                flags |= _lib.CV_FLAG_CPLX_MAT
            desc[k] = (nx, ny, _lib.CV_POL_EZ if pol == 'ez' else _lib.CV_POL_HZ, flags, cell_off, prof_off,
                       first_tile, ntile, scale, dshift, phx.real, phx.imag, phy.real, phy.imag, omega,
                       n_cpl, blk, nblk, 0, 0, 0)
            self.shapes.append((nx, ny))
            self.offsets.append(cell_off)
            cell_off += n
            prof_off += 2 * (nx + ny)
        # coupled systems: shared tile range, table and w2 offsets (tables after
        # all PML profiles, then every system's w2 entries)
        self.pml_len = prof_off
        cpl_cursor = prof_off
        self.systems = []
        for k0, nblk, n_cpl in systems:
            if k0 + nblk > n_inst:
                raise ValueError('coupled system at instance %d is missing blocks' % k0)
            n = self.shapes[k0][0] * self.shapes[k0][1]
            self.systems.append(dict(first=k0, nblk=nblk, n_cpl=n_cpl, cells=n, cpl_off=cpl_cursor,
                                     cell_off=self.offsets[k0]))
            cpl_cursor += 2 * n_cpl * n
        self.cpl_len = cpl_cursor - prof_off
        w2_cursor = cpl_cursor
        for sy in self.systems:
            k0, nblk = sy['first'], sy['nblk']
            t0 = desc[k0]['tile_first']
            tn = int(desc[k0 + nblk - 1]['tile_first'] + desc[k0 + nblk - 1]['tile_count'] - t0)
            for b in range(nblk):
                d = desc[k0 + b]
                d['tile_first'], d['tile_count'] = t0, tn
                d['cpl_off'], d['w2_off'] = sy['cpl_off'], w2_cursor
            w2_cursor += nblk
        self.total_cells = cell_off
        self.n_inst = n_inst
        self.desc_host = desc
        self.tiles_host = np.array(tiles, dtype=_lib.TILE_DTYPE)
        self.n_tiles = len(tiles)
        dev = self.device
        self.desc = torch.from_numpy(desc.view(np.uint8).copy()).to(dev)
        self.tiles = torch.from_numpy(self.tiles_host.view(np.uint8).copy()).to(dev)
        prof = np.concatenate(profs + [np.zeros(self.cpl_len), np.asarray(w2, dtype=np.float64)]).astype(np.complex128)
        self.prof = torch.from_numpy(prof).to(dev).to(dtype)
        self.prof64 = self.prof if dtype == torch.complex128 else torch.from_numpy(prof).to(dev)
        self._ws = None
        self._ws_mixed = None
        self._scale_vec = None
        self._desc64 = None
        self.last_info = None

    # ------------------------------------------------------------- helpers
    def _args(self):
        return (ptr(self.desc), ptr(self.tiles), self.n_tiles)

    def prof_with(self, cpl, dtype=None):
        """The coefficient array with the coupling tables ``cpl`` (complex,
        cpl_len entries: the systems' tables in order) in place."""
        dtype = dtype or self.dtype
        base = self.prof if self.prof.dtype == dtype else self.prof64.to(dtype)
        if self.cpl_len == 0:
            return base
        if cpl is None or cpl.numel() != self.cpl_len:
            raise ValueError('coupling tables must hold %d entries' % self.cpl_len)
        return torch.cat([base[:self.pml_len], cpl.detach().to(dtype).reshape(-1),
                          base[self.pml_len + self.cpl_len:]]).contiguous()

    def coupling_apply(self, tab, x, y=None, transpose=False):
        """y (+)= C x with the inter-block terms only, tables from ``tab``
        (an array laid out like prof_with's result)."""
        x = x.to(self.dtype).contiguous()
        acc = y is not None
        if y is None:
            y = torch.empty_like(x)
        check(lib().cv_fdfd_coupling_apply(self.code, *self._args(), ptr(tab), ptr(x), ptr(y),
                                           int(bool(transpose)), int(acc), stream_ptr()), 'coupling_apply')
        return y

    def coupling_grad(self, prof, lam, x, scale):
        """scale * lam^T (dA/d table) x for every table entry (cpl_len, complex)."""
        g = torch.zeros_like(prof)
        check(lib().cv_fdfd_coupling_grad(self.code, *self._args(), ptr(prof), ptr(lam.contiguous()),
                                          ptr(x.contiguous()), ptr(g), float(scale), stream_ptr()),
              'coupling_grad')
        return g[self.pml_len:self.pml_len + self.cpl_len]

    def scale_vector(self, dtype=None):
        """Per-cell system scale c (real), for packing right-hand sides."""
        if self._scale_vec is None:
            v = np.concatenate([np.full(s[0] * s[1], c) for s, c in zip(self.shapes, self.sys_scale)])
            self._scale_vec = torch.from_numpy(v).to(self.device)
        return self._scale_vec if dtype is None else self._scale_vec.to(dtype)

    def materials(self, eps):
        """(m0, m1) for a packed eps (real or complex) of total_cells entries."""
        eps = eps.to(self.device)
        if self.pol == 'ez':
            if torch.is_complex(eps):
                return eps.real.to(self.rdtype).contiguous(), eps.imag.to(self.rdtype).contiguous()
            return eps.to(self.rdtype).contiguous(), None
        if torch.is_complex(eps) != self.cmat:
            raise ValueError('this Hz operator was built for %s eps; use complex_materials=%s'
                             % ('complex' if self.cmat else 'real', torch.is_complex(eps)))
        e = eps.to(self.dtype if self.cmat else self.rdtype).contiguous()
        m0 = torch.empty_like(e)
        m1 = torch.empty_like(e)
        check(lib().cv_fdfd_hz_materials(self.code, *self._args(), ptr(e), ptr(m0), ptr(m1), stream_ptr()),
              'hz_materials')
        return m0, m1

    def materials_grad(self, m0, m1, g0, g1):
        """Reverse of materials(): d eps for cotangents on (m0, m1)."""
        out = torch.zeros_like(m0)
        check(lib().cv_fdfd_hz_materials_grad(self.code, *self._args(), ptr(m0), ptr(m1), ptr(g0), ptr(g1),
                                              ptr(out), stream_ptr()), 'hz_materials_grad')
        return out

    def apply(self, x, m0, m1, transpose=False, prof=None):
        """c * A x (or c * A^T x) with the per-instance system scale c
        (``prof``: coefficients with coupling tables, from prof_with)."""
        x = x.to(self.dtype).contiguous()
        y = torch.empty_like(x)
        prof = self.prof if prof is None else prof
        check(lib().cv_fdfd_apply(self.code, *self._args(), ptr(prof), ptr(m0), ptr(m1), ptr(x), ptr(y),
                                  int(bool(transpose)), stream_ptr()), 'apply')
        return y

    def deriv(self, u, which, coef, mat=None, mode=0, out=None):
        """coef * mat .* (D u) (mode 0), coef * D^T(mat .* u) (1), conj(coef) * D^H(mat .* u) (2)."""
        u = u.to(self.dtype).contiguous()
        acc = out is not None
        if out is None:
            out = torch.empty_like(u)
        check(lib().cv_fdfd_deriv(self.code, *self._args(), ptr(self.prof), ptr(mat), ptr(u), ptr(out),
                                  which, mode, float(np.real(coef)), float(np.imag(coef)), int(acc),
                                  stream_ptr()), 'deriv')
        return out

    def deriv_matgrad(self, u, g, which, coef, gmat):
        check(lib().cv_fdfd_deriv_matgrad(self.code, *self._args(), ptr(self.prof), ptr(u), ptr(g), ptr(gmat),
                                          which, float(np.real(coef)), float(np.imag(coef)), stream_ptr()),
              'deriv_matgrad')
        return gmat

    def material_apply(self, dm0, dm1, x, y=None):
        """(dA/dm . dm) x with the scaled system (c A)."""
        x = x.to(self.dtype).contiguous()
        acc = y is not None
        if y is None:
            y = torch.empty_like(x)
        check(lib().cv_fdfd_material_apply(self.code, *self._args(), ptr(self.prof), ptr(dm0), ptr(dm1), ptr(x),
                                           ptr(y), int(acc), stream_ptr()), 'material_apply')
        return y

    def materials_jvp(self, m0, m1, d_eps):
        dm0, dm1 = torch.empty_like(m0), torch.empty_like(m1)
        check(lib().cv_fdfd_hz_materials_jvp(self.code, *self._args(), ptr(m0), ptr(m1),
                                             ptr(d_eps.to(m0.dtype).contiguous()), ptr(dm0), ptr(dm1),
                                             stream_ptr()), 'hz_materials_jvp')
        return dm0, dm1

    def material_grad(self, lam, x, g0, g1, scale):
        check(lib().cv_fdfd_material_grad(self.code, *self._args(), ptr(self.prof), ptr(lam), ptr(x), ptr(g0),
                                          ptr(g1), float(scale), stream_ptr()), 'material_grad')

    # --------------------------------------------------------------- solve
    def _workspace(self, mixed):
        L = lib()
        if mixed:
            need = L.cv_fdfd_solve_mixed_workspace_bytes(self.total_cells, self.n_inst, self.n_tiles)
            if self._ws_mixed is None or self._ws_mixed.numel() < need:
                self._ws_mixed = torch.empty(need, dtype=torch.uint8, device=self.device)
            return self._ws_mixed
        need = L.cv_fdfd_solve_workspace_bytes(self.code, self.total_cells, self.n_inst, self.n_tiles)
        if self._ws is None or self._ws.numel() < need:
            self._ws = torch.empty(need, dtype=torch.uint8, device=self.device)
        return self._ws

    def cocg_applies(self):
        """COCG needs the complex-symmetric form: zero Bloch phases, no
        frequency-coupled blocks."""
        d = self.desc_host
        return (self.cpl_len == 0 and bool(np.all(d['nblk'] == 1)) and bool(np.all(d['phx_re'] == 1.0))
                and bool(np.all(d['phx_im'] == 0.0)) and bool(np.all(d['phy_re'] == 1.0))
                and bool(np.all(d['phy_im'] == 0.0)))

    def krylov_method(self, opts):
        """The Krylov method a solve with ``opts`` runs: 'cocg' only when asked
        for (method='cocg'), else 'bicgstab'.  COCG needs about a third of
        BiCGSTAB's operator applications on small and mid-size blob problems,
        but its residual can swing by orders of magnitude on the way (complex64
        rounding does not survive that, and on the held-out 512^2 and 1024^2
        blob structures it diverges), so it is not the default."""
        method = (opts or SolverOptions()).method
        if method == 'cocg':
            if not self.cocg_applies():
                raise ValueError("method='cocg' needs zero Bloch phases and uncoupled instances")
            return 'cocg'
        return 'bicgstab'

    def solve_scaled(self, bs, m0, m1, transpose=False, x0=None, opts=None, mixed_ops=None, prof=None):
        """Solve (c A) x = bs (bs already multiplied by c).  Returns x.

        For complex64 operators, ``mixed_ops`` = (op64, m0_64, m1_64) selects
        mixed-precision iterative refinement: x and the residual in complex128,
        complex64 BiCGSTAB corrections.  ``prof`` (coupled systems) is
        (prof, prof64) from prof_with for this operator's and op64's
        precision (prof64 unused without mixed_ops)."""
        opts = opts or SolverOptions()
        info = (_lib.SolveInfo * self.n_inst)()
        pr, pr64 = (self.prof, None) if prof is None else prof
        kmethod = self.krylov_method(opts)
        kcode = _lib.CV_KRYLOV_COCG if kmethod == 'cocg' else _lib.CV_KRYLOV_BICGSTAB
        if mixed_ops is not None:
            op64, m0_64, m1_64 = mixed_ops
            pr64 = op64.prof if pr64 is None else pr64
            tol = opts.tol if opts.tol is not None else DEFAULT_TOL[torch.complex64]
            b64 = bs.to(torch.complex128).contiguous()
            x = torch.zeros_like(b64) if x0 is None else x0.to(torch.complex128).clone()
            ws = self._workspace(True)
            rc = lib().cv_fdfd_solve_mixed(
                ptr(op64.desc), self.desc_host.ctypes.data_as(ctypes.c_void_p), self.n_inst, ptr(self.tiles),
                self.n_tiles, ptr(pr64), ptr(m0_64), ptr(m1_64), ptr(pr), ptr(m0), ptr(m1),
                ptr(b64), ptr(x), int(bool(transpose)), kcode, float(tol), float(opts.inner_tol), int(opts.max_iter),
                int(opts.max_sweeps), int(opts.check_every), ptr(ws), ws.numel(), info, stream_ptr())
        else:
            tol = opts.tol if opts.tol is not None else DEFAULT_TOL[self.dtype]
            b = bs.to(self.dtype).contiguous()
            x = torch.zeros_like(b) if x0 is None else x0.to(self.dtype).clone().contiguous()
            ws = self._workspace(False)
            entry = lib().cv_fdfd_cocg if kmethod == 'cocg' else lib().cv_fdfd_solve
            rc = entry(
                self.code, ptr(self.desc), self.desc_host.ctypes.data_as(ctypes.c_void_p), self.n_inst,
                ptr(self.tiles), self.n_tiles, ptr(pr), ptr(m0), ptr(m1), ptr(b), ptr(x),
                int(bool(transpose)), float(tol), int(opts.max_iter), int(opts.check_every), ptr(ws), ws.numel(),
                info, stream_ptr())
        self.last_info = [dict(iterations=i.iterations, status=i.status, restarts=i.restarts,
                               rel_residual=i.rel_residual, method=kmethod) for i in info]
        check(rc, 'fdfd solve')
        return x


BTD_MAX_NY = 128


def direct_applies(op):
    """True when the block-tridiagonal factorisation handles ``op``: one
    uncoupled instance whose y extent fits a workgroup, or whose x extent
    does (factorised on the transposed grid)."""
    if op.n_inst != 1 or op.cpl_len != 0 or op.cmat:
        return False
    nx, ny = op.shapes[0]
    return ny <= BTD_MAX_NY or nx <= BTD_MAX_NY


def direct_factor_bytes(op):
    """Device bytes of the block-cyclic-reduction factor of ``op`` (in the
    orientation the direct path factorises)."""
    nx, ny = op.shapes[0]
    if ny > BTD_MAX_NY:
        nx, ny = ny, nx
    return int(lib().cv_fdfd_bcr_factor_bytes(nx, ny))


DIRECT_MEMORY_FRACTION = 0.5
# Instances of an uncoupled batch solved at once on their own streams: a
# small instance's factorisation and FGMRES are chains of short launches
# that leave most of the GPU idle, so several run side by side.
INSTANCE_WORKERS = 8
_WORKER_STREAMS = {}


_TUNED_WORKERS = {}


def _device_key(device):
    d = torch.device(device)
    return d.index if d.index is not None else torch.cuda.current_device()


def instance_workers(device):
    """Instances of an uncoupled batch solved at once on a device: the count
    autotune_instance_workers chose there, else INSTANCE_WORKERS."""
    return _TUNED_WORKERS.get(_device_key(device), INSTANCE_WORKERS)


def autotune_instance_workers(op, eps, b, candidates=(1, 2, 4, 8, 12), opts=None):
    """Times the default forward solve (each instance by its direct
    factorisation and FGMRES) of an uncoupled batch once per candidate count
    of instances solved at once, after one untimed solve that builds the
    plans and library handles, and keeps the fastest count for this device
    (instance_workers); returns {count: seconds}.  The batch should look like
    the ones to come (sizes, number of instances); its solutions are
    discarded.  eps and b are the batch's packed permittivity and source
    term (as FdfdSolve takes them)."""
    if op.n_inst < 2:
        raise ValueError('autotune_instance_workers needs a batch of at least two instances')
    opts = opts or SolverOptions()
    if opts.method not in ('auto', 'direct'):
        raise ValueError("the instance count only matters for method 'auto' or 'direct'")
    key = _device_key(b.device)
    m0, m1 = op.materials(eps)

    def one():
        torch.cuda.synchronize(b.device)
        t0 = time.perf_counter()
        FdfdSolve.apply(m0, m1, b, op, None, opts)
        torch.cuda.synchronize(b.device)
        return time.perf_counter() - t0

    saved = _TUNED_WORKERS.get(key)
    times = {}
    try:
        one()
        for c in candidates:
            _TUNED_WORKERS[key] = max(1, int(c))
            times[int(c)] = one()
    finally:
        if saved is None:
            _TUNED_WORKERS.pop(key, None)
        else:
            _TUNED_WORKERS[key] = saved
    _TUNED_WORKERS[key] = min(times, key=times.get)
    return times


def _worker_streams(device, n):
    """n long-lived streams of a device (default priority), created once."""
    d = torch.device(device)
    key = d.index if d.index is not None else torch.cuda.current_device()
    lst = _WORKER_STREAMS.setdefault(key, [])
    while len(lst) < n:
        lst.append(torch.cuda.Stream(device=d))
    return lst[:n]


def transposed_operator(op):
    """The same single-instance operator on the transposed grid (x and y
    swapped with their PML widths and Bloch phases): A_T = P A P^T for the
    cell permutation P (i, j) -> (j, i).  Ez: -(1/mu0)(Dxf Dxb + Dyf Dyb) -
    eps0 w^2 eps treats both axes alike (ceviche/fdfd.py:172-185,
    derivatives.py:94-124); Hz: the same with the x- and y-derivative
    materials exchanged, as eps_xx / eps_yy average along y / x
    (fdfd.py:200-244)."""
    if getattr(op, '_transposed', None) is None:
        sp = dict(op.specs[0])
        nx, ny = (int(v) for v in sp['shape'])
        px, py = sp['npml']
        bx, by = sp.get('bloch', (0.0, 0.0))
        sp.update(shape=(ny, nx), npml=(py, px), bloch=(by, bx))
        op._transposed = FdfdOperator(op.pol, [sp], dtype=op.dtype, device=op.device)
    return op._transposed


class Direct:
    """Block-tridiagonal factorisation of one FDFD instance (A without its
    x-wrap, complex64) and FP64-refined FGMRES solves with it in either
    orientation (csrc/btd.hip, cv_fdfd_fgmres_direct).  kind 'bcr' (block
    cyclic reduction: log2(nx) levels of parallel block inversions, the
    default for nx >= 8) or 'btd' (the sequential block elimination along x).
    One factorisation serves the forward, adjoint (A^T) and tangent solves of
    a linear system."""

    def __init__(self, op, m0, m1, kind=None, factorise=True):
        if not direct_applies(op):
            raise ValueError('the block-tridiagonal factorisation needs one uncoupled instance with nx or ny '
                             '<= %d' % BTD_MAX_NY)
        self.user_op = op
        self.shape = tuple(op.shapes[0])
        # a tall grid (ny > 128 >= nx) is factorised and solved on its
        # transpose; vectors are transposed on the way in and out (Hz: the
        # two materials exchanged as well)
        self.tr = self.shape[1] > BTD_MAX_NY
        if self.tr:
            op = transposed_operator(op)
            m0, m1 = self._tm(m0, m1)
        self.op = op
        nx, ny = op.shapes[0]
        if kind is None:
            kind = 'bcr' if nx >= 8 else 'btd'
        if kind not in ('bcr', 'btd'):
            raise ValueError("kind must be 'bcr' or 'btd'")
        self.kind = kind
        self.code = _lib.CV_FACTOR_BCR if kind == 'bcr' else _lib.CV_FACTOR_BTD
        self.prof32 = op.prof.to(torch.complex64).contiguous()
        self.m0 = m0.to(torch.float32).contiguous()
        self.m1 = None if m1 is None else m1.to(torch.float32).contiguous()
        L = lib()
        dh = op.desc_host.ctypes.data_as(ctypes.c_void_p)
        if kind == 'bcr':
            nbytes = L.cv_fdfd_bcr_factor_bytes(nx, ny)
            self.factor = torch.empty(nbytes, dtype=torch.uint8, device=op.device)
            self.factor_bytes = nbytes
            if factorise:       # otherwise the caller factorises several at once
                check(L.cv_fdfd_bcr_factor(ptr(op.desc), dh, ptr(self.prof32), ptr(self.m0), ptr(self.m1),
                                           ptr(self.factor), nbytes, stream_ptr()), 'block cyclic reduction')
        else:
            nbytes = L.cv_fdfd_btd_factor_bytes(nx, ny)
            self.factor = torch.empty(nbytes, dtype=torch.uint8, device=op.device)
            check(L.cv_fdfd_btd_factor(ptr(op.desc), dh, ptr(self.prof32), ptr(self.m0), ptr(self.m1),
                                       ptr(self.factor), nbytes, stream_ptr()), 'block-tridiagonal factorisation')

    def _t(self, v):
        """A vector of the user's (nx, ny) grid in the factorised grid's order."""
        if v is None or not self.tr:
            return v
        return v.reshape(self.shape).transpose(0, 1).contiguous().reshape(-1)

    def _tm(self, m0, m1):
        """The materials (m0, m1) of the user's grid for the factorised grid."""
        if not self.tr:
            return m0, m1
        if self.user_op.pol == 'hz':     # m1 goes with x differences, m0 with y
            return self._t(m1), self._t(m0)
        return self._t(m0), self._t(m1)

    def _tb(self, v):
        """The inverse of _t."""
        if not self.tr:
            return v
        return v.reshape(self.shape[1], self.shape[0]).transpose(0, 1).contiguous().reshape(-1)

    def apply(self, v, transpose=False):
        """M^-1 v (complex64) with the factorisation alone."""
        if self.tr:
            return self._tb(self._apply(self._t(v), transpose))
        return self._apply(v, transpose)

    def _apply(self, v, transpose):
        v = v.to(torch.complex64).contiguous()
        z = torch.empty_like(v)
        op = self.op
        dh = op.desc_host.ctypes.data_as(ctypes.c_void_p)
        if self.kind == 'bcr':
            check(lib().cv_fdfd_bcr_apply(ptr(op.desc), dh, ptr(self.factor), ptr(v), ptr(z), int(bool(transpose)),
                                          stream_ptr()), 'bcr_apply')
        else:
            check(lib().cv_fdfd_btd_apply(ptr(op.desc), dh, ptr(self.prof32), ptr(self.m0), ptr(self.m1),
                                          ptr(self.factor), ptr(v), ptr(z), int(bool(transpose)), stream_ptr()),
                  'btd_apply')
        return z

    def solve(self, bs, m0_64, m1_64, transpose=False, opts=None):
        """(c A) x = bs (or (c A)^T x = bs) to the FP64 tolerance."""
        if self.tr:
            x = self._solve(self._t(bs), *self._tm(m0_64, m1_64), transpose, opts)
            return self._tb(x)
        return self._solve(bs, m0_64, m1_64, transpose, opts)

    def _solve(self, bs, m0_64, m1_64, transpose, opts):
        op = self.op
        opts = opts or SolverOptions()
        tol = opts.tol if opts.tol is not None else DEFAULT_TOL[torch.complex128]
        nx, ny = op.shapes[0]
        L = lib()
        need = L.cv_fdfd_fgmres_workspace_bytes(nx, ny, int(opts.restart), _lib.CV_PRECOND_NONE, 0)
        ws = torch.empty(need, dtype=torch.uint8, device=op.device)
        b = bs.to(torch.complex128).contiguous()
        x = torch.zeros_like(b)
        info = (_lib.SolveInfo * 1)()
        max_iter = _direct_max_iter(opts)
        rc = L.cv_fdfd_fgmres_direct(ptr(op.desc), op.desc_host.ctypes.data_as(ctypes.c_void_p), ptr(op.prof64),
                                     ptr(m0_64), ptr(m1_64), ptr(self.prof32), ptr(self.m0), ptr(self.m1),
                                     self.code, ptr(self.factor), ptr(b), ptr(x), int(bool(transpose)), float(tol),
                                     int(opts.restart), int(max_iter), ptr(ws), need, info, stream_ptr())
        self.user_op.last_info = [dict(iterations=info[0].iterations, status=info[0].status,
                                       restarts=info[0].restarts, rel_residual=info[0].rel_residual,
                                       method='direct', factor=self.kind + ('-transposed' if self.tr else ''))]
        check(rc, 'fdfd direct solve')
        return x


def blockdiag_applies(op):
    """True when ``op`` is one frequency-coupled Ez system whose blocks the
    block factorisation handles (ny <= 128)."""
    return (len(op.systems) == 1 and op.n_inst == op.systems[0]['nblk'] and op.systems[0]['nblk'] >= 1
            and op.shapes[0][1] <= BTD_MAX_NY)


class BlockDirect:
    """Block-diagonal preconditioner of a frequency-coupled system: each
    sideband block's own factorisation (the block as a single instance at its
    W_m, block cyclic reduction), inside FP64-refined FGMRES on the whole
    coupled operator (cv_fdfd_fgmres_blockdiag).  One set of factors serves
    the forward, adjoint and tangent solves."""

    def __init__(self, op, m0, m1):
        if not blockdiag_applies(op):
            raise ValueError('the block-diagonal factorisation needs one coupled system with ny <= %d'
                             % BTD_MAX_NY)
        self.op = op
        nb = op.n_inst
        n0 = op.shapes[0][0] * op.shapes[0][1]
        if getattr(op, '_block_ops', None) is None:
            op._block_ops = [FdfdOperator('ez', [{k: v for k, v in sp.items() if k != 'system'}], dtype=op.dtype,
                                          device=op.device) for sp in op.specs]
        self.blocks = [Direct(bop, m0[m * n0:(m + 1) * n0], None if m1 is None else m1[m * n0:(m + 1) * n0],
                              factorise=False)
                       for m, bop in enumerate(op._block_ops)]
        self.block_desc = np.concatenate([bop.desc_host for bop in op._block_ops])
        self.factor_ptrs = (ctypes.c_void_p * nb)(*[d.factor.data_ptr() for d in self.blocks])
        self.kinds = {d.kind for d in self.blocks}
        if self.kinds != {'bcr'}:
            raise ValueError('block-diagonal preconditioning uses block cyclic reduction (nx >= 8)')
        # every sideband block factorised in the same launches
        arr = lambda vals: (ctypes.c_void_p * nb)(*vals)
        blk = self.blocks
        check(lib().cv_fdfd_bcr_factor_batch(
            nb, arr([bop.desc.data_ptr() for bop in op._block_ops]), self.block_desc.ctypes.data_as(ctypes.c_void_p),
            arr([d.prof32.data_ptr() for d in blk]), arr([d.m0.data_ptr() for d in blk]),
            arr([0 if d.m1 is None else d.m1.data_ptr() for d in blk]), self.factor_ptrs, blk[0].factor_bytes,
            stream_ptr()), 'block cyclic reduction of the sideband blocks')

    def solve(self, bs, m0_64, m1_64, p64, p32, m0_32, m1_32, transpose=False, opts=None):
        op = self.op
        opts = opts or SolverOptions()
        tol = opts.tol if opts.tol is not None else DEFAULT_TOL[torch.complex128]
        nx, ny = op.shapes[0]
        nb = op.n_inst
        L = lib()
        need = L.cv_fdfd_fgmres_workspace_bytes(nx * nb, ny, int(opts.restart), _lib.CV_PRECOND_NONE, 0)
        ws = torch.empty(need, dtype=torch.uint8, device=op.device)
        b = bs.to(torch.complex128).contiguous()
        x = torch.zeros_like(b)
        info = (_lib.SolveInfo * nb)()
        rc = L.cv_fdfd_fgmres_blockdiag(ptr(op.desc), op.desc_host.ctypes.data_as(ctypes.c_void_p), ptr(p64),
                                        ptr(m0_64), ptr(m1_64), ptr(p32), ptr(m0_32), ptr(m1_32),
                                        self.block_desc.ctypes.data_as(ctypes.c_void_p), self.factor_ptrs, ptr(b),
                                        ptr(x), int(bool(transpose)), float(tol), int(opts.restart),
                                        int(opts.max_iter), ptr(ws), need, info, stream_ptr())
        op.last_info = [dict(iterations=i.iterations, status=i.status, restarts=i.restarts,
                             rel_residual=i.rel_residual, method='direct', factor='bcr-blockdiag') for i in info]
        check(rc, 'multi-frequency block-diagonal solve')
        return x


def _use_blockdiag(op, opts):
    method = (opts or SolverOptions()).method
    if method in ('bicgstab', 'cocg'):
        return False
    if method == 'direct' and not (blockdiag_applies(op) and op.shapes[0][0] >= 8):
        raise ValueError("method='direct' needs one coupled system with ny <= %d and nx >= 8" % BTD_MAX_NY)
    return blockdiag_applies(op) and op.shapes[0][0] >= 8


def nd_applies(op):
    """True when the nested-dissection factorisation handles ``op``: one
    uncoupled instance (any shape, real or complex materials)."""
    return op.n_inst == 1 and op.cpl_len == 0 and op.shapes[0][0] * op.shapes[0][1] >= 4


def _direct_kind(op):
    """'bcr' (block cyclic reduction, an extent <= 128 with real materials),
    'nd' (nested dissection) or None."""
    if direct_applies(op):
        return 'bcr'
    if nd_applies(op):
        return 'nd'
    return None


def _factor_bytes(op, kind):
    if kind == 'bcr':
        return direct_factor_bytes(op)
    plan = nd_plan(op.shapes[0], op.device)
    return int(lib().cv_nd_factor_bytes(plan.ptr))


def instance_operators(op):
    """Single-instance operators (same polarisation, precision and specs) of
    the instances of an uncoupled batch, cached on ``op``."""
    if getattr(op, '_instances', None) is None:
        op._instances = [FdfdOperator(op.pol, [sp], dtype=op.dtype, device=op.device, complex_materials=op.cmat)
                         for sp in op.specs]
    return op._instances


def _use_direct(op, opts):
    """The direct factorisation to use ('bcr', 'nd', 'instances' for an
    uncoupled batch factorised instance by instance) or None.  method
    'direct' requires one; 'auto' takes it when it applies and a factor (the
    largest instance's, for a batch) fits in DIRECT_MEMORY_FRACTION of the
    free device memory."""
    method = (opts or SolverOptions()).method
    if method in ('bicgstab', 'cocg'):
        return None
    if op.n_inst > 1 and op.cpl_len == 0:
        subs = instance_operators(op)
        kinds = [_direct_kind(sub) for sub in subs]
        if None in kinds:
            if method == 'direct':
                raise ValueError("method='direct' needs every instance to have a direct factorisation")
            return None
        if method == 'direct':
            return 'instances'
        free, _ = torch.cuda.mem_get_info(op.device)
        need = max(_factor_bytes(sub, kd) for sub, kd in zip(subs, kinds))
        return 'instances' if need <= DIRECT_MEMORY_FRACTION * free else None
    kind = _direct_kind(op)
    if method == 'direct':
        if kind is None:
            raise ValueError("method='direct' needs uncoupled instances")
        return kind
    if kind is None:
        return None
    free, _ = torch.cuda.mem_get_info(op.device)
    return kind if _factor_bytes(op, kind) <= DIRECT_MEMORY_FRACTION * free else None


def make_direct(op, m0, m1, kind):
    if kind == 'instances':
        return InstanceDirect(op, m0, m1)
    return Direct(op, m0, m1) if kind == 'bcr' else NdDirect(op, m0, m1)


class InstanceDirect:
    """Direct solves of an uncoupled batch, instance by instance: each
    instance is factorised by its own kind (block cyclic reduction or nested
    dissection) and solved by FP64-refined FGMRES.  The factors are kept for
    the adjoint and tangent solves when all of them fit in
    DIRECT_MEMORY_FRACTION of the free device memory; otherwise each solve
    factorises its instance again and releases the factor.  Under method
    'auto' an instance whose factorisation or FGMRES fails is solved by
    Jacobi BiCGSTAB alone (and stays on it for the adjoint and tangent
    solves); the other instances keep their factors."""

    kind = 'instances'

    def __init__(self, op, m0, m1):
        self.op = op
        self.subs = instance_operators(op)
        self.kinds = [_direct_kind(sub) for sub in self.subs]
        if None in self.kinds:
            raise ValueError('every instance of the batch needs a direct factorisation')
        self.m0, self.m1 = m0, m1
        self.sizes = [nx * ny for nx, ny in op.shapes]
        free, _ = torch.cuda.mem_get_info(op.device)
        total = sum(_factor_bytes(sub, kd) for sub, kd in zip(self.subs, self.kinds))
        self.keep = total <= DIRECT_MEMORY_FRACTION * free
        self.factors = [None] * op.n_inst
        self.failed = set()   # instances on the Krylov fallback

    def _seg(self, v, k):
        if v is None:
            return None
        o = self.op.offsets[k]
        return v[o:o + self.sizes[k]]

    def _make(self, k):
        return make_direct(self.subs[k], self._seg(self.m0, k), self._seg(self.m1, k), self.kinds[k])

    def _direct_one(self, k, bs, m0_64, m1_64, transpose, opts):
        """Instance k by its factorisation on the current stream; None when it
        is (or now goes) on the Krylov fallback."""
        if k in self.failed:
            return None
        seg = (self._seg(bs, k).contiguous(), self._seg(m0_64, k).contiguous(),
               None if m1_64 is None else self._seg(m1_64, k).contiguous())
        try:
            f = self.factors[k]
            if f is None:
                f = self._make(k)
                if self.keep:   # kept for the adjoint and tangent solves
                    self.factors[k] = f
            return f.solve(*seg, transpose=transpose, opts=opts)
        except (CevicheError, torch.cuda.OutOfMemoryError) as err:
            if not _direct_failed(err, opts):
                raise
            self.failed.add(k)
            self.factors[k] = None
            return None

    def solve(self, bs, m0_64, m1_64, transpose=False, opts=None):
        """Every instance by its direct factorisation, up to
        instance_workers(device) at once on their own streams (largest first); under method 'auto' an
        instance whose factorisation or FGMRES fails (and its later adjoint /
        tangent solves) falls back to Jacobi BiCGSTAB alone, on the caller's
        stream after the others."""
        n = self.op.n_inst
        x = torch.empty(bs.shape, dtype=torch.complex128, device=bs.device)
        outs = [None] * n
        order = sorted(range(n), key=lambda k: -self.sizes[k])
        workers = min(instance_workers(bs.device), n)
        if workers > 1:
            main = torch.cuda.current_stream(bs.device)
            ready = torch.cuda.Event()
            ready.record(main)
            streams = _worker_streams(bs.device, workers)
            it = iter(order)
            lock = threading.Lock()

            def run(ws):
                with torch.cuda.stream(ws):
                    ws.wait_event(ready)
                    while True:
                        with lock:
                            k = next(it, None)
                        if k is None:
                            return
                        out = self._direct_one(k, bs, m0_64, m1_64, transpose, opts)
                        if out is not None:
                            out.record_stream(main)
                        outs[k] = out

            try:
                with ThreadPoolExecutor(workers) as ex:
                    for f in [ex.submit(run, ws) for ws in streams]:
                        f.result()
            finally:
                for ws in streams:
                    main.wait_stream(ws)
        else:
            for k in order:
                outs[k] = self._direct_one(k, bs, m0_64, m1_64, transpose, opts)
        info = []
        for k in range(n):
            sub = self.subs[k]
            out = outs[k]
            if out is None:
                out = sub.solve_scaled(self._seg(bs, k).contiguous(), self._seg(m0_64, k).contiguous(),
                                       None if m1_64 is None else self._seg(m1_64, k).contiguous(),
                                       transpose=transpose, opts=opts)
            x[self.op.offsets[k]:self.op.offsets[k] + self.sizes[k]] = out
            info += sub.last_info
        self.op.last_info = info
        return x


class NdPlan:
    """Host symbolic analysis of an nx x ny grid and its device index arrays."""

    def __init__(self, shape, device):
        nx, ny = (int(v) for v in shape)
        L = lib()
        p = L.cv_nd_plan_create(nx, ny)
        if not p:
            raise CevicheError(_lib.CV_ERR_INVALID, L.cv_last_error().decode(errors='replace'))
        self.ptr = ctypes.c_void_p(p)
        nbytes = int(L.cv_nd_plan_bytes(self.ptr))
        self.dev = torch.empty(nbytes, dtype=torch.uint8, device=device)
        check(L.cv_nd_plan_upload(self.ptr, ptr(self.dev), nbytes, stream_ptr()), 'nd plan upload')

    def __del__(self):
        try:
            if _lib._lib is not None and self.ptr:
                _lib._lib.cv_nd_plan_destroy(self.ptr)
                self.ptr = None
        except Exception:  # interpreter teardown
            pass


_ND_PLANS = {}


def nd_plan(shape, device):
    """The cached NdPlan of a grid shape on a device."""
    key = (tuple(int(v) for v in shape), str(device))
    plan = _ND_PLANS.get(key)
    if plan is None:
        if len(_ND_PLANS) > 16:
            _ND_PLANS.clear()
        plan = _ND_PLANS[key] = NdPlan(shape, device)
    return plan


class NdDirect:
    """Nested-dissection factorisation of one FDFD instance (complex64, the
    periodic wrap dropped) and FP64-refined FGMRES solves with it in either
    orientation (csrc/nd.hip, cv_fdfd_fgmres_nd).  One factorisation serves
    the forward, adjoint (A^T) and tangent solves of a linear system."""

    kind = 'nd'

    def __init__(self, op, m0, m1):
        if not nd_applies(op):
            raise ValueError('the nested-dissection factorisation needs one uncoupled instance')
        self.user_op = self.op = op
        self.plan = nd_plan(op.shapes[0], op.device)
        lo = lambda m: None if m is None else m.to(torch.complex64 if m.is_complex() else torch.float32).contiguous()
        self.prof32 = op.prof.to(torch.complex64).contiguous()
        self.m0, self.m1 = lo(m0), lo(m1)
        L = lib()
        nbytes = int(L.cv_nd_factor_bytes(self.plan.ptr))
        self.factor = torch.empty(nbytes, dtype=torch.uint8, device=op.device)
        check(L.cv_nd_factor(self.plan.ptr, ptr(self.plan.dev), ptr(op.desc),
                             op.desc_host.ctypes.data_as(ctypes.c_void_p), ptr(self.prof32), ptr(self.m0),
                             ptr(self.m1), ptr(self.factor), nbytes, stream_ptr()), 'nested dissection')

    def apply(self, v, transpose=False):
        """M^-1 v (complex64) with the factorisation alone."""
        v = v.to(torch.complex64).contiguous()
        z = torch.empty_like(v)
        check(lib().cv_nd_apply(self.plan.ptr, ptr(self.plan.dev), ptr(self.factor), ptr(v), ptr(z),
                                int(bool(transpose)), stream_ptr()), 'nd_apply')
        return z

    def solve(self, bs, m0_64, m1_64, transpose=False, opts=None):
        """(c A) x = bs (or (c A)^T x = bs) to the FP64 tolerance."""
        op = self.op
        opts = opts or SolverOptions()
        tol = opts.tol if opts.tol is not None else DEFAULT_TOL[torch.complex128]
        nx, ny = op.shapes[0]
        L = lib()
        need = L.cv_fdfd_fgmres_workspace_bytes(nx, ny, int(opts.restart), _lib.CV_PRECOND_NONE, 0)
        ws = torch.empty(need, dtype=torch.uint8, device=op.device)
        b = bs.to(torch.complex128).contiguous()
        x = torch.zeros_like(b)
        info = (_lib.SolveInfo * 1)()
        max_iter = _direct_max_iter(opts)
        rc = L.cv_fdfd_fgmres_nd(ptr(op.desc), op.desc_host.ctypes.data_as(ctypes.c_void_p), ptr(op.prof64),
                                 ptr(m0_64), ptr(m1_64), ptr(self.prof32), ptr(self.m0), ptr(self.m1), self.plan.ptr,
                                 ptr(self.plan.dev), ptr(self.factor), ptr(b), ptr(x), int(bool(transpose)),
                                 float(tol), int(opts.restart), int(max_iter), ptr(ws), need, info, stream_ptr())
        op.last_info = [dict(iterations=info[0].iterations, status=info[0].status, restarts=info[0].restarts,
                             rel_residual=info[0].rel_residual, method='direct', factor='nd')]
        check(rc, 'fdfd direct solve (nested dissection)')
        return x


# status codes after which the direct path gives way to BiCGSTAB under 'auto'
_DIRECT_FALLBACK = (_lib.CV_ERR_SINGULAR, _lib.CV_ERR_NOT_CONVERGED, _lib.CV_ERR_BREAKDOWN, _lib.CV_ERR_NONFINITE)


def _direct_failed(err, opts):
    """True when a direct-path error should fall back to BiCGSTAB (method
    'auto'); warns.  Other errors, and any error under 'direct', propagate."""
    if (opts or SolverOptions()).method == 'direct':
        return False
    if isinstance(err, torch.cuda.OutOfMemoryError) or (isinstance(err, CevicheError)
                                                         and err.status in _DIRECT_FALLBACK):
        warnings.warn('direct FDFD solve failed (%s); falling back to Jacobi-preconditioned Krylov' % err,
                      RuntimeWarning)
        return True
    return False


# ---------------------------------------------------------------- autograd
class FdfdSolve(torch.autograd.Function):
    """x = A(m)^-1 b for the FDFD operator (sp_solve, primitives.py:103-137).

    ``op`` is the complex128 operator; when ``op32`` (complex64, same specs) is
    given, the solve runs complex64 BiCGSTAB corrections inside complex128
    iterative refinement.  Backward: A^H lam = g is solved as A^T mu = conj(g),
    lam = conj(mu) (one adjoint solve shared by both cotangents, unlike the two
    of primitives.py:121-135), grad_b = lam, grad_m = -Re(mu^T dA/dm x)."""

    @staticmethod
    def _krylov(op, op32, opts, m0, m1, bs, transpose):
        if op32 is None:
            return op.solve_scaled(bs, m0, m1, transpose=transpose, opts=opts)
        lo = lambda m: m.to(torch.complex64 if m.is_complex() else torch.float32).contiguous()
        m1s = None if m1 is None else lo(m1)
        return op32.solve_scaled(bs, lo(m0), m1s, transpose=transpose, opts=opts, mixed_ops=(op, m0, m1))

    @staticmethod
    def _solve(op, op32, opts, m0, m1, bs, transpose, direct=None, ctx=None):
        if direct is not None:
            try:
                return direct.solve(bs, m0, m1, transpose=transpose, opts=opts)
            except (CevicheError, torch.cuda.OutOfMemoryError) as err:
                if not _direct_failed(err, opts):
                    raise
                if ctx is not None:
                    ctx.direct = None
        return FdfdSolve._krylov(op, op32, opts, m0, m1, bs, transpose)

    @staticmethod
    def forward(ctx, m0, m1, b, op, op32, opts):
        c = op.scale_vector(torch.float64)
        # one factorisation serves the forward, adjoint and tangent solves
        direct = None
        kind = _use_direct(op, opts)
        if kind is not None:
            try:
                direct = make_direct(op, m0, m1, kind)
            except (CevicheError, torch.cuda.OutOfMemoryError) as err:
                if not _direct_failed(err, opts):
                    raise
        ctx.direct = direct
        x = FdfdSolve._solve(op, op32, opts, m0, m1, b * c, False, direct, ctx)
        direct = ctx.direct
        if direct is not None and op32 is not None:
            op32.last_info = op.last_info
        ctx.op, ctx.op32, ctx.opts = op, op32, opts
        ctx.save_for_backward(m0, m1, x)
        ctx.save_for_forward(m0, m1, x)
        ctx.has_m1 = m1 is not None
        return x

    @staticmethod
    def jvp(ctx, dm0, dm1, db, *_):
        """dx = A^-1 (db - (dA/dm . dm) x)  (primitives.py:137-148)."""
        m0, m1, x = ctx.saved_tensors
        op = ctx.op
        c = op.scale_vector(torch.float64)
        rhs = torch.zeros_like(x) if db is None else (db.to(x.dtype) * c)
        if dm0 is not None or dm1 is not None:
            d0 = torch.zeros_like(m0) if dm0 is None else dm0.to(m0.dtype).contiguous()
            d1 = None
            if ctx.has_m1:
                d1 = torch.zeros_like(m0) if dm1 is None else dm1.to(m0.dtype).contiguous()
            rhs = rhs - op.material_apply(d0, d1, x)
        return FdfdSolve._solve(op, ctx.op32, ctx.opts, m0, m1, rhs.contiguous(), False, ctx.direct, ctx)

    @staticmethod
    def backward(ctx, gx):
        m0, m1, x = ctx.saved_tensors
        op = ctx.op
        mu = FdfdSolve._solve(op, ctx.op32, ctx.opts, m0, m1, torch.conj(gx).resolve_conj().contiguous(), True,
                              ctx.direct, ctx)
        c = op.scale_vector(torch.float64)
        gb = torch.conj(mu).resolve_conj() * c if ctx.needs_input_grad[2] else None
        g0 = g1 = None
        if ctx.needs_input_grad[0] or (ctx.has_m1 and ctx.needs_input_grad[1]):
            g0 = torch.zeros_like(m0)
            g1 = torch.zeros_like(m0) if ctx.has_m1 else None
            op.material_grad(mu.contiguous(), x.contiguous(), g0, g1, -1.0)
        return g0, g1, gb, None, None, None


class FdfdCoupledSolve(torch.autograd.Function):
    """x = A(m, g)^-1 b for frequency-coupled systems, where g are the
    coupling tables (complex, op.cpl_len entries, system-scaled units): the
    sp_solve of the multi-frequency operator (fdfd.py:326-335 ->
    primitives.py:103-148).  Backward as FdfdSolve, plus
    grad_g = conj(-mu^T (dA/dg) x) from the coupling-gradient kernel;
    forward mode subtracts (dA/dm . dm + dA/dg . dg) x."""

    @staticmethod
    def _profs(op, op32, cpl):
        p64 = op.prof_with(cpl, torch.complex128)
        p32 = op32.prof_with(cpl, torch.complex64) if op32 is not None else None
        return p64, p32

    @staticmethod
    def _solve(op, op32, opts, m0, m1, bs, transpose, p64, p32, direct=None):
        if direct is not None:
            pd32 = p64.to(torch.complex64).contiguous()   # coefficients incl. coupling tables, complex64
            x = direct.solve(bs, m0, m1, p64, pd32, direct.m0_32, direct.m1_32, transpose=transpose, opts=opts)
            if op32 is not None:
                op32.last_info = op.last_info
            return x
        if op32 is None:
            return op.solve_scaled(bs, m0, m1, transpose=transpose, opts=opts, prof=(p64, None))
        m0s = m0.to(torch.float32).contiguous()
        m1s = None if m1 is None else m1.to(torch.float32).contiguous()
        return op32.solve_scaled(bs, m0s, m1s, transpose=transpose, opts=opts, mixed_ops=(op, m0, m1),
                                 prof=(p32, p64))

    @staticmethod
    def forward(ctx, m0, m1, cpl, b, op, op32, opts):
        c = op.scale_vector(torch.float64)
        p64, p32 = FdfdCoupledSolve._profs(op, op32, cpl)
        direct = None
        if _use_blockdiag(op, opts):
            direct = BlockDirect(op, m0, m1)
            direct.m0_32 = m0.to(torch.float32).contiguous()
            direct.m1_32 = None if m1 is None else m1.to(torch.float32).contiguous()
        x = FdfdCoupledSolve._solve(op, op32, opts, m0, m1, b * c, False, p64, p32, direct)
        ctx.direct = direct
        ctx.op, ctx.op32, ctx.opts = op, op32, opts
        ctx.p64, ctx.p32 = p64, p32
        ctx.save_for_backward(m0, m1, x)
        ctx.save_for_forward(m0, m1, x)
        ctx.has_m1 = m1 is not None
        return x

    @staticmethod
    def jvp(ctx, dm0, dm1, dcpl, db, *_):
        m0, m1, x = ctx.saved_tensors
        op = ctx.op
        c = op.scale_vector(torch.float64)
        rhs = torch.zeros_like(x) if db is None else (db.to(x.dtype) * c)
        if dm0 is not None or dm1 is not None:
            d0 = torch.zeros_like(m0) if dm0 is None else dm0.to(m0.dtype).contiguous()
            d1 = None
            if ctx.has_m1:
                d1 = torch.zeros_like(m0) if dm1 is None else dm1.to(m0.dtype).contiguous()
            rhs = rhs - op.material_apply(d0, d1, x)
        if dcpl is not None:
            rhs = rhs - op.coupling_apply(op.prof_with(dcpl, torch.complex128), x)
        return FdfdCoupledSolve._solve(op, ctx.op32, ctx.opts, m0, m1, rhs.contiguous(), False, ctx.p64, ctx.p32,
                                       ctx.direct)

    @staticmethod
    def backward(ctx, gx):
        m0, m1, x = ctx.saved_tensors
        op = ctx.op
        mu = FdfdCoupledSolve._solve(op, ctx.op32, ctx.opts, m0, m1, torch.conj(gx).resolve_conj().contiguous(),
                                     True, ctx.p64, ctx.p32, ctx.direct)
        c = op.scale_vector(torch.float64)
        gb = torch.conj(mu).resolve_conj() * c if ctx.needs_input_grad[3] else None
        g0 = g1 = None
        if ctx.needs_input_grad[0] or (ctx.has_m1 and ctx.needs_input_grad[1]):
            g0 = torch.zeros_like(m0)
            g1 = torch.zeros_like(m0) if ctx.has_m1 else None
            op.material_grad(mu.contiguous(), x.contiguous(), g0, g1, -1.0)
        gc = None
        if ctx.needs_input_grad[2]:
            gc = torch.conj(op.coupling_grad(ctx.p64, mu, x, -1.0)).resolve_conj()
        return g0, g1, gc, gb, None, None, None


class FdfdDeriv(torch.autograd.Function):
    """out = coef * mat .* (D u): field recovery (fdfd.py:135-162) through the
    sp_mult of a derivative matrix (primitives.py:58-98)."""

    @staticmethod
    def forward(ctx, u, mat, op, which, coef):
        ctx.op, ctx.which, ctx.coef = op, which, coef
        ctx.save_for_backward(u, mat)
        ctx.save_for_forward(u, mat)
        ctx.has_mat = mat is not None
        return op.deriv(u, which, coef, mat=mat, mode=0)

    @staticmethod
    def jvp(ctx, du, dmat, *_):
        u, mat = ctx.saved_tensors
        op = ctx.op
        out = None
        if du is not None:
            out = op.deriv(du, ctx.which, ctx.coef, mat=mat, mode=0)
        if ctx.has_mat and dmat is not None:
            t = op.deriv(u, ctx.which, ctx.coef, mat=dmat.to(mat.dtype).contiguous(), mode=0)
            out = t if out is None else out + t
        return out if out is not None else torch.zeros_like(op.deriv(u, ctx.which, ctx.coef, mat=mat, mode=0))

    @staticmethod
    def backward(ctx, g):
        u, mat = ctx.saved_tensors
        op = ctx.op
        g = g.to(op.dtype).contiguous()
        gu = op.deriv(g, ctx.which, ctx.coef, mat=mat, mode=2) if ctx.needs_input_grad[0] else None
        gm = None
        if ctx.has_mat and ctx.needs_input_grad[1]:
            gm = torch.zeros_like(mat)
            op.deriv_matgrad(u.to(op.dtype).contiguous(), g, ctx.which, ctx.coef, gm)
        return gu, gm, None, None, None


class HzMaterials(torch.autograd.Function):
    """(m0, m1) = (1/(eps_xx + 1e-5), 1/(eps_yy + 1e-5)); fdfd.py:200-215."""

    @staticmethod
    def forward(ctx, eps, op):
        m0, m1 = op.materials(eps)
        ctx.op = op
        ctx.save_for_backward(m0, m1)
        ctx.save_for_forward(m0, m1)
        return m0, m1

    @staticmethod
    def jvp(ctx, d_eps, _):
        m0, m1 = ctx.saved_tensors
        return ctx.op.materials_jvp(m0, m1, d_eps)

    @staticmethod
    def backward(ctx, g0, g1):
        m0, m1 = ctx.saved_tensors
        g0 = None if g0 is None else g0.contiguous()
        g1 = None if g1 is None else g1.contiguous()
        return ctx.op.materials_grad(m0, m1, g0, g1), None
