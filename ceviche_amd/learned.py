"""Learned preconditioner for the 2D Ez FDFD solve.

A U-Net of 3x3 convolutions with 32 channels (bf16 weights and activations on
the matrix cores, fp32 accumulation; kernels in csrc/cnn.hip) maps a residual
r to an approximate correction e ~ A^-1 r.  It is trained on (residual,
correction) pairs taken from runs of the plain BiCGSTAB solver
(csrc/krylov.hip) on random binary-blob structures, by the project's own
forward and backward kernels, and applied inside flexible GMRES
(csrc/fgmres.hip).  The FGMRES convergence test is the FP64 true residual, so
the network can change the number of iterations but not the accuracy; plain
BiCGSTAB and Jacobi-preconditioned FGMRES remain the baseline and fallback.

The reference has no counterpart: it factorises A (ceviche/solvers.py:37-62,
called from sp_solve, ceviche/primitives.py:103-119).
"""
import ctypes
import math

import numpy as np
import torch

from . import _lib
from ._lib import check, lib, ptr, stream_ptr
from ._fdfd_op import FdfdOperator
from .constants import C_0
from .derivatives import create_sfactor

CH = 32
PACK = _lib.CV_CNN_PACKED_ELEMS
NF = 3


# ------------------------------------------------------------------ inputs
def features(spec, eps):
    """Per-cell network inputs besides the residual, (nx, ny, 3) float32:
    (eps - 1) / 11.25, the local (k0 dL)^2 eps / 0.3, and the PML strength
    |s_x - 1| / max + |s_y - 1| / max of the stretch factors."""
    eps = np.asarray(eps, dtype=np.float64)
    nx, ny = eps.shape
    omega, dL = float(spec['omega']), float(spec['dL'])
    npml = spec['npml']
    f = np.zeros((nx, ny, NF), np.float32)
    f[..., 0] = (eps - 1.0) / 11.25
    f[..., 1] = (omega * dL / C_0) ** 2 * eps / 0.3

    def strength(n, npml_):
        if npml_ == 0:
            return np.zeros(n)
        s = np.abs(create_sfactor('f', omega, dL, n, int(npml_)) - 1.0)
        return s / max(s.max(), 1e-30)

    f[..., 2] = strength(nx, npml[0])[:, None] + strength(ny, npml[1])[None, :]
    return f


# ----------------------------------------------------------------- network
class UNet:
    """Weights, packed bf16 operands and the batched forward / backward pass
    of the network of cv_unet_apply (layer order documented there).

    W: (2L+1, 9, 32, 32) fp32 masters [layer][tap][ci][co]; b: (2L+1, 32)."""

    def __init__(self, levels=6, nf=NF, seed=0, device=None):
        self.L = int(levels)
        self.nf = int(nf)
        self.n_layers = 2 * self.L + 1
        self.device = device if device is not None else _lib.device()
        g = torch.Generator().manual_seed(seed)
        W = torch.randn(self.n_layers, 9, CH, CH, generator=g) * math.sqrt(2.0 / (9 * CH))
        W[0] = torch.randn(9, CH, CH, generator=g) * math.sqrt(2.0 / (9 * (2 + self.nf)))
        W[0, :, 2 + self.nf:, :] = 0.0      # unused input channels
        W[-1] = torch.randn(9, CH, CH, generator=g) * math.sqrt(1.0 / (9 * CH))
        W[-1, :, :, 2:] = 0.0               # only channels 0, 1 are read out
        self.W = W.to(self.device).contiguous()
        self.b = torch.zeros(self.n_layers, CH, device=self.device)
        self.wpack = torch.empty(self.n_layers * PACK, dtype=torch.bfloat16, device=self.device)
        self.wflip = torch.empty_like(self.wpack)
        self._ws = None
        self.pack()

    # -------------------------------------------------------- parameters
    def pack(self):
        L = lib()
        check(L.cv_cnn_pack_weights(ptr(self.W), ptr(self.wpack), 0, self.n_layers, stream_ptr()), 'pack')
        check(L.cv_cnn_pack_weights(ptr(self.W), ptr(self.wflip), 1, self.n_layers, stream_ptr()), 'pack')

    def state(self):
        return dict(W=self.W.detach().cpu().numpy(), b=self.b.detach().cpu().numpy(),
                    levels=np.int64(self.L), nf=np.int64(self.nf))

    def save(self, path, **extra):
        np.savez(path, **self.state(), **extra)

    @classmethod
    def load(cls, path, device=None):
        d = np.load(path, allow_pickle=False)
        net = cls(levels=int(d['levels']), nf=int(d['nf']), device=device)
        net.W.copy_(torch.from_numpy(d['W']))
        net.b.copy_(torch.from_numpy(d['b']))
        net.pack()
        extra = {k: d[k] for k in d.files if k not in ('W', 'b', 'levels', 'nf')}
        return net, extra

    def _wp(self, layer, flip=False):
        t = self.wflip if flip else self.wpack
        return ctypes.c_void_p(t.data_ptr() + layer * PACK * 2)

    # --------------------------------------------------------- primitives
    def _conv(self, layer, x, relu):
        B, H, W, _ = x.shape
        out = torch.empty_like(x)
        check(lib().cv_conv3x3_fwd_bf16(ptr(x), self._wp(layer), ptr(self.b[layer]), ptr(out), B, H, W,
                                        int(relu), stream_ptr()), 'conv3x3')
        return out

    def _dgrad(self, layer, g, mask):
        B, H, W, _ = g.shape
        out = torch.empty_like(g)
        check(lib().cv_conv3x3_bwd_data_bf16(ptr(g), ptr(mask), self._wp(layer, True), ptr(out), B, H, W, 0,
                                             stream_ptr()), 'conv3x3 bwd data')
        return out

    def _wgrad(self, layer, x, g, mask, dW, db):
        B, H, W, _ = x.shape
        need = lib().cv_conv3x3_bwd_weight_workspace_bytes(B, H, W)
        if self._ws is None or self._ws.numel() < need:
            self._ws = torch.empty(need, dtype=torch.uint8, device=self.device)
        check(lib().cv_conv3x3_bwd_weight_bf16(ptr(x), ptr(g), ptr(mask), ptr(dW[layer]), ptr(db[layer]), B, H, W,
                                               0, ptr(self._ws), self._ws.numel(), stream_ptr()),
              'conv3x3 bwd weight')

    @staticmethod
    def _down(x, scale):
        B, H, W, _ = x.shape
        out = torch.empty(B, (H + 1) // 2, (W + 1) // 2, CH, dtype=x.dtype, device=x.device)
        check(lib().cv_cnn_down2(ptr(x), ptr(out), B, H, W, float(scale), stream_ptr()), 'down2')
        return out

    @staticmethod
    def _up(x, H, W, scale, add=None, out=None):
        B = x.shape[0]
        if out is None:
            out = torch.empty(B, H, W, CH, dtype=x.dtype, device=x.device)
        check(lib().cv_cnn_up2(ptr(x), ptr(add), ptr(out), B, H, W, float(scale), stream_ptr()), 'up2')
        return out

    # -------------------------------------------------------------- passes
    def forward(self, x0):
        """y = net(x0) for x0 (B, H, W, 32) bf16; returns (y, cache)."""
        L = self.L
        e, h, u, bb = [None] * L, [None] * L, [None] * L, [None] * L
        e[0] = self._conv(0, x0, True)
        for l in range(L - 1):
            h[l] = self._conv(1 + l, e[l], True)
            e[l + 1] = self._down(h[l], 0.25)
        bb[L - 1] = self._conv(L, e[L - 1], True)
        for l in range(L - 2, -1, -1):
            u[l] = self._up(bb[l + 1], h[l].shape[1], h[l].shape[2], 1.0, add=h[l])
            bb[l] = self._conv(L + 1 + (L - 2 - l), u[l], True)
        y = self._conv(2 * L, bb[0], False)
        return y, dict(x0=x0, e=e, h=h, u=u, b=bb)

    def backward(self, c, gy):
        """(dW, db) of <y, gy> (fp32, shapes of W and b)."""
        L = self.L
        e, h, u, bb = c['e'], c['h'], c['u'], c['b']
        dW = torch.empty_like(self.W)
        db = torch.empty_like(self.b)
        self._wgrad(2 * L, bb[0], gy, None, dW, db)
        g_b = self._dgrad(2 * L, gy, None)
        g_u = [None] * L
        for l in range(L - 1):
            layer = L + 1 + (L - 2 - l)
            self._wgrad(layer, u[l], g_b, bb[l], dW, db)
            g_u[l] = self._dgrad(layer, g_b, bb[l])
            g_b = self._down(g_u[l], 1.0)
        self._wgrad(L, e[L - 1], g_b, bb[L - 1], dW, db)
        g_e = self._dgrad(L, g_b, bb[L - 1])
        for l in range(L - 2, -1, -1):
            g_h = self._up(g_e, h[l].shape[1], h[l].shape[2], 0.25, add=g_u[l], out=g_u[l])
            self._wgrad(1 + l, e[l], g_h, h[l], dW, db)
            g_e = self._dgrad(1 + l, g_h, h[l])
        self._wgrad(0, c['x0'], g_e, e[0], dW, db)
        return dW, db

    def pack_input(self, r, feat, r_idx=None, f_idx=None, rs=1.0, B=None):
        """Network input (B, H, W, 32) bf16 from complex residuals r (rows
        r_idx of an (P, H, W) pool, or r itself) and features (rows f_idx)."""
        if r.dim() == 2:
            r = r.unsqueeze(0)
        H, W = r.shape[1], r.shape[2]
        if B is None:
            B = r.shape[0] if r_idx is None else r_idx.numel()
        code = _lib.CV_C64 if r.dtype == torch.complex64 else _lib.CV_C128
        x0 = torch.empty(B, H, W, CH, dtype=torch.bfloat16, device=self.device)
        check(lib().cv_cnn_pack_input(code, ptr(r), ptr(r_idx), None, float(rs), ptr(feat), ptr(f_idx),
                                      self.nf if feat is not None else 0, ptr(x0), B, H, W, stream_ptr()),
              'pack_input')
        return x0

    def desc(self, feat):
        """cv_unet descriptor for the native application (keep `feat` alive)."""
        return _lib.UNetDesc(self.wpack.data_ptr(), self.b.data_ptr(), 0 if feat is None else feat.data_ptr(),
                             self.nf if feat is not None else 0, self.L)

    def apply(self, r, feat, rs, os_):
        """os * net(rs * r) for one complex (H, W) field, through cv_unet_apply."""
        H, W = r.shape
        r = r.contiguous()
        code = _lib.CV_C64 if r.dtype == torch.complex64 else _lib.CV_C128
        z = torch.empty_like(r)
        need = lib().cv_unet_workspace_bytes(self.L, 1, H, W)
        ws = torch.empty(need, dtype=torch.uint8, device=self.device)
        d = self.desc(feat)
        check(lib().cv_unet_apply(ctypes.byref(d), code, ptr(r), float(rs), float(os_), ptr(z), H, W, ptr(ws), need,
                                  stream_ptr()), 'unet_apply')
        return z


# --------------------------------------------------------- one instance
class Problem:
    """One 2D Ez FDFD instance on the GPU in both precisions (scaled system
    c A, c = mu0 dL^2, see FdfdOperator)."""

    def __init__(self, omega, dL, eps, npml, device=None):
        eps = np.asarray(eps, dtype=np.float64)
        self.shape = eps.shape
        self.spec = dict(omega=omega, dL=dL, shape=eps.shape, npml=list(npml))
        self.op64 = FdfdOperator('ez', [self.spec], torch.complex128, device)
        self.op32 = FdfdOperator('ez', [self.spec], torch.complex64, device)
        e = torch.from_numpy(eps.ravel())
        self.m0_64, self.m1_64 = self.op64.materials(e)
        self.m0_32, self.m1_32 = self.op32.materials(e)
        self.c = float(self.op64.sys_scale[0])
        self.feat = torch.from_numpy(features(self.spec, eps).reshape(-1, NF)).to(self.op64.device)
        self.device = self.op64.device

    def rhs(self, src):
        """Scaled right-hand side c * (i omega J) for a Jz source (numpy)."""
        b = 1j * self.spec['omega'] * np.asarray(src, dtype=np.complex128).ravel() * self.c
        return torch.from_numpy(b).to(self.device)

    def residual(self, bs, x):
        return bs - self.op64.apply(x, self.m0_64, self.m1_64)

    def bicgstab(self, bs, tol=1e-6, max_iter=400000, x0=None, check_every=32):
        """Plain BiCGSTAB (complex128) -> (x, info); no exception on failure."""
        op = self.op64
        x = torch.zeros_like(bs) if x0 is None else x0.clone()
        info = (_lib.SolveInfo * 1)()
        ws = op._workspace(False)
        rc = lib().cv_fdfd_solve(op.code, ptr(op.desc), op.desc_host.ctypes.data_as(ctypes.c_void_p), 1,
                                 ptr(op.tiles), op.n_tiles, ptr(op.prof), ptr(self.m0_64), ptr(self.m1_64), ptr(bs),
                                 ptr(x), 0, float(tol), int(max_iter), int(check_every), ptr(ws), ws.numel(), info,
                                 stream_ptr())
        if rc not in (_lib.CV_OK, _lib.CV_ERR_NOT_CONVERGED):
            check(rc, 'bicgstab')
        return x, dict(iterations=info[0].iterations, rel_residual=info[0].rel_residual, status=info[0].status)

    def cocg(self, bs, tol=1e-6, max_iter=400000, x0=None, check_every=32):
        """Plain Jacobi-preconditioned COCG on the complex-symmetric form
        (complex128) -> (x, info); no exception on failure."""
        op = self.op64
        x = torch.zeros_like(bs) if x0 is None else x0.clone()
        info = (_lib.SolveInfo * 1)()
        ws = op._workspace(False)
        rc = lib().cv_fdfd_cocg(op.code, ptr(op.desc), op.desc_host.ctypes.data_as(ctypes.c_void_p), 1,
                                ptr(op.tiles), op.n_tiles, ptr(op.prof), ptr(self.m0_64), ptr(self.m1_64), ptr(bs),
                                ptr(x), 0, float(tol), int(max_iter), int(check_every), ptr(ws), ws.numel(), info,
                                stream_ptr())
        if rc not in (_lib.CV_OK, _lib.CV_ERR_NOT_CONVERGED):
            check(rc, 'cocg')
        return x, dict(iterations=info[0].iterations, rel_residual=info[0].rel_residual, status=info[0].status)

    def fgmres(self, bs, tol=1e-6, restart=30, max_iter=100000, inner_rtol=1e-3, precond='jacobi', net=None,
               gain=1.0, x0=None):
        """Flexible GMRES (complex64 Arnoldi, complex128 iterate and FP64 true
        residual) with precond 'none', 'jacobi' or 'unet' (needs `net`)."""
        kind = {'none': _lib.CV_PRECOND_NONE, 'jacobi': _lib.CV_PRECOND_JACOBI, 'unet': _lib.CV_PRECOND_UNET}[precond]
        nx, ny = self.shape
        levels = net.L if kind == _lib.CV_PRECOND_UNET else 0
        need = lib().cv_fdfd_fgmres_workspace_bytes(nx, ny, int(restart), kind, levels)
        if need <= 0:
            raise ValueError('invalid FGMRES configuration')
        ws = torch.empty(need, dtype=torch.uint8, device=self.device)
        x = torch.zeros_like(bs) if x0 is None else x0.clone()
        info = (_lib.SolveInfo * 1)()
        d = net.desc(self.feat) if kind == _lib.CV_PRECOND_UNET else None
        rc = lib().cv_fdfd_fgmres(ptr(self.op64.desc), self.op64.desc_host.ctypes.data_as(ctypes.c_void_p),
                                  ptr(self.op64.prof), ptr(self.m0_64), ptr(self.m1_64), ptr(self.op32.prof),
                                  ptr(self.m0_32), ptr(self.m1_32), ptr(bs), ptr(x), float(tol), int(restart),
                                  int(max_iter), float(inner_rtol), kind,
                                  ctypes.byref(d) if d is not None else None, float(gain), ptr(ws), need, info,
                                  stream_ptr())
        out = dict(iterations=info[0].iterations, cycles=info[0].restarts, rel_residual=info[0].rel_residual,
                   status=info[0].status, fell_back=bool(info[0].pad))
        if rc not in (_lib.CV_OK, _lib.CV_ERR_NOT_CONVERGED):
            check(rc, 'fgmres')
        return x, out


def problem_from(w, device=None):
    """Problem and Jz source from a workloads.fdfd_learned_blobs dict."""
    return Problem(w['omega'], w['dL'], w['eps'], w['npml'], device), w['src']


# ---------------------------------------------------------- training data
class PairSet:
    """Residual / correction pairs on the GPU: r (P, H, W) complex64 with unit
    rms, e (P, H, W) complex64 = correction / (rms(r) * gain), inv_norm (P,)
    = 1 / |e|^2, struct (P,) int32 row of feat (S, H*W, nf) and instance of
    `op` (a complex64 operator over the S structures, materials m0, m1), so
    that A_struct (gain e) = r for every pair."""

    def __init__(self, r, e, inv_norm, struct, feat, gain, op=None, m0=None, m1=None):
        self.r, self.e, self.inv_norm, self.struct, self.feat, self.gain = r, e, inv_norm, struct, feat, gain
        self.op, self.m0, self.m1 = op, m0, m1

    @property
    def n(self):
        return self.r.shape[0]


def _random_source(rng, N, npml):
    src = np.zeros((N, N), np.complex128)
    lo, hi = npml + 4, N - npml - 4
    for _ in range(int(rng.integers(1, 5))):
        i, j = rng.integers(lo, hi, 2)
        src[i, j] += rng.standard_normal() + 1j * rng.standard_normal()
    if rng.random() < 0.5:
        c = N // 2
        src[30, c - 8:c + 8] += rng.standard_normal() + 1j * rng.standard_normal()
    return src


def generate_pairs(n_struct, N=256, snaps=24, chunk=64, seed0=0, tol=1e-8, max_iter=200000, gain=None,
                   device=None, log=None):
    """Runs plain BiCGSTAB (restarted every `chunk` iterations) on n_struct
    seeded structures (workloads.fdfd_learned_blobs, seeds seed0 ..) with
    random sources, and keeps up to `snaps` iterates x_k per run, x_0 = 0
    included, as pairs (r_k = b - A x_k, e_k = x* - x_k)."""
    from . import workloads
    rs, es, structs, feats, ratios = [], [], [], [], []
    specs, epss = [], []
    stats = dict(iterations=[], failed=0)
    for s in range(n_struct):
        w = workloads.fdfd_learned_blobs(N, seed=seed0 + s)
        pb, _ = problem_from(w, device)
        rng = np.random.default_rng(seed0 + s + 7919)
        bs = pb.rhs(_random_source(rng, N, w['npml'][0]))
        x = torch.zeros_like(bs)
        traj = [x.clone()]      # x_0 = 0: the pair (b, x*) is the first one FGMRES meets
        total = 0
        ok = False
        while total < max_iter:
            x, info = pb.bicgstab(bs, tol=tol, max_iter=chunk, x0=x)
            total += info['iterations']
            if info['status'] == _lib.CV_OK:
                ok = True
                break
            if info['status'] != _lib.CV_ERR_NOT_CONVERGED:
                break
            traj.append(x.clone())
        if not ok or len(traj) == 0:
            stats['failed'] += 1
            continue
        stats['iterations'].append(total)
        pick = np.unique(np.linspace(0, len(traj) - 1, min(snaps, len(traj))).round().astype(int))
        for k in pick:
            xk = traj[k]
            r = pb.residual(bs, xk)
            e = x - xk
            srms = torch.sqrt(torch.mean(torch.abs(r) ** 2)).item()
            rs.append((r / srms).to(torch.complex64).reshape(N, N))
            es.append((e / srms).to(torch.complex64).reshape(N, N))
            ratios.append(torch.linalg.vector_norm(e).item() / torch.linalg.vector_norm(r).item())
            structs.append(len(feats))
        feats.append(pb.feat)
        specs.append(dict(pb.spec))
        epss.append(np.asarray(w['eps'], dtype=np.float64).ravel())
        del traj
        if log and (s + 1) % max(1, n_struct // 10) == 0:
            log('pairs: %d structures, %d pairs, mean BiCGSTAB iterations %.0f'
                % (s + 1, len(rs), np.mean(stats['iterations'])))
    if gain is None:
        # median |e| / |r|: dividing by it gives targets of unit rms on average
        gain = float(np.median(ratios)) if ratios else 1.0
    r = torch.stack(rs)
    e = torch.stack(es) / gain
    inv_norm = (1.0 / (torch.abs(e) ** 2).sum(dim=(1, 2))).to(torch.float32)
    dev = r.device
    op = FdfdOperator('ez', specs, torch.complex64, dev)
    m0, m1 = op.materials(torch.from_numpy(np.concatenate(epss)))
    return PairSet(r.contiguous(), e.contiguous(), inv_norm.contiguous(),
                   torch.tensor(structs, dtype=torch.int32, device=dev), torch.stack(feats).contiguous(), gain,
                   op, m0, m1), stats


def train(net, pairs, steps, batch=16, lr=2e-3, seed=0, log=None, log_every=100, w_res=1.0, w_err=0.1):
    """Adam on  w_res * |A z - r|^2 / |r|^2 + w_err * |z - e|^2 / |e|^2  (batch
    means, z = gain * network output): the residual the correction leaves,
    which is what a preconditioner is judged by, plus the error against the
    solver's correction.  Gradients come from the network's own backward
    kernels and the batched operator kernels (A and A^H per sample)."""
    net.W.requires_grad_(True)
    net.b.requires_grad_(True)
    opt = torch.optim.Adam([net.W, net.b], lr=lr)
    sched = torch.optim.lr_scheduler.CosineAnnealingLR(opt, T_max=max(1, steps), eta_min=lr * 0.05)
    gen = torch.Generator(device=net.device).manual_seed(seed)
    H, W = pairs.r.shape[1], pairs.r.shape[2]
    n = H * W
    loss = torch.zeros(2, dtype=torch.float64, device=net.device)
    L = lib()
    op = pairs.op
    use_res = w_res > 0 and op is not None
    z = torch.empty(batch, n, dtype=torch.complex64, device=net.device)
    wv = torch.empty_like(z)
    gz = torch.empty_like(z)
    hist = []
    for it in range(steps):
        idx = torch.randint(0, pairs.n, (batch,), generator=gen, device=net.device, dtype=torch.int32)
        f_idx = pairs.struct[idx.long()].contiguous()
        x0 = net.pack_input(pairs.r, pairs.feat, r_idx=idx, f_idx=f_idx, B=batch)
        y, cache = net.forward(x0)
        gy = torch.empty_like(y)
        loss.zero_()
        if use_res:
            desc_h = op.desc_host.ctypes.data_as(ctypes.c_void_p)
            check(L.cv_cnn_unpack(_lib.CV_C64, ptr(y), None, float(pairs.gain), ptr(z), batch, H, W, stream_ptr()),
                  'unpack')
            check(L.cv_fdfd_batch_apply(ptr(op.desc), desc_h, op.n_inst, ptr(op.prof), ptr(pairs.m0), ptr(pairs.m1),
                                        ptr(f_idx), ptr(z), ptr(pairs.r), ptr(idx), ptr(wv), ptr(loss[1:]),
                                        float(w_res / (n * batch)), 0, batch, stream_ptr()), 'batch apply')
            check(L.cv_fdfd_batch_apply(ptr(op.desc), desc_h, op.n_inst, ptr(op.prof), ptr(pairs.m0), ptr(pairs.m1),
                                        ptr(f_idx), ptr(wv), None, None, ptr(gz), None, 0.0, 1, batch, stream_ptr()),
                  'batch apply H')
        check(L.cv_cnn_loss(ptr(y), ptr(pairs.e), ptr(idx), ptr(pairs.inv_norm), ptr(gy), ptr(loss),
                            float(w_err / batch), ptr(gz) if use_res else None,
                            float(pairs.gain * 2.0 * w_res / (n * batch)), batch, H, W, stream_ptr()), 'loss')
        dW, db = net.backward(cache, gy)
        with torch.no_grad():
            net.W.grad = dW
            net.b.grad = db
            opt.step()
            sched.step()
            net.pack()
        if (it + 1) % log_every == 0 or it == 0:
            le, lr_ = loss.tolist()
            hist.append((it + 1, le + lr_, le, lr_))
            if log:
                log('train step %d loss %.4f (error part %.4f, residual part %.4f)' % (it + 1, le + lr_, le, lr_))
    net.W.requires_grad_(False)
    net.b.requires_grad_(False)
    return hist


def evaluate(net, pairs, count=64, seed=123):
    """Mean |A z - r| / |r| and |z - e| / |e| of the network correction z on
    `count` pairs (the residual ratio below 1 is what makes it a
    preconditioner)."""
    H, W = pairs.r.shape[1], pairs.r.shape[2]
    n = H * W
    gen = torch.Generator(device=net.device).manual_seed(seed)
    idx = torch.randint(0, pairs.n, (count,), generator=gen, device=net.device, dtype=torch.int32)
    f_idx = pairs.struct[idx.long()].contiguous()
    res, err = [], []
    for c0 in range(0, count, 16):
        i = idx[c0:c0 + 16].contiguous()
        f = f_idx[c0:c0 + 16].contiguous()
        B = i.numel()
        y, _ = net.forward(net.pack_input(pairs.r, pairs.feat, r_idx=i, f_idx=f, B=B))
        z = torch.empty(B, n, dtype=torch.complex64, device=net.device)
        w = torch.empty_like(z)
        check(lib().cv_cnn_unpack(_lib.CV_C64, ptr(y), None, float(pairs.gain), ptr(z), B, H, W, stream_ptr()),
              'unpack')
        op = pairs.op
        check(lib().cv_fdfd_batch_apply(ptr(op.desc), op.desc_host.ctypes.data_as(ctypes.c_void_p), op.n_inst,
                                        ptr(op.prof), ptr(pairs.m0), ptr(pairs.m1), ptr(f), ptr(z), ptr(pairs.r),
                                        ptr(i), ptr(w), None, 0.0, 0, B, stream_ptr()), 'batch apply')
        r = pairs.r[i.long()].reshape(B, n)
        e = pairs.e[i.long()].reshape(B, n) * pairs.gain
        res += (torch.linalg.vector_norm(w, dim=1) / torch.linalg.vector_norm(r, dim=1)).tolist()
        err += (torch.linalg.vector_norm(z - e, dim=1) / torch.linalg.vector_norm(e, dim=1)).tolist()
    return float(np.mean(res)), float(np.mean(err))
