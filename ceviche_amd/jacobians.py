"""jacobian(fun, argnum=0, mode='reverse'|'forward'|'numerical', step_size=1e-6)
with the reference's shapes and complex convention (ceviche/jacobians.py:16-90).

``fun`` is written with torch operations (the GPU primitives of this package
carry torch autograd).  Returned jacobians are numpy arrays of shape
(n_out, m_in):
  * reverse: one vector-Jacobian product per output element; for a complex
    input the entry is dL/dx - i dL/dy (autograd's convention; torch's
    gradient is conjugated to match, jacobians.py:29-35).
  * forward: one Jacobian-vector product per input basis vector with torch
    forward-mode AD (the primitives' ``jvp`` rules); a complex input combines
    the real and imaginary directions as real - 1j * imag (jacobians.py:38-51).
  * numerical: one-sided differences of step ``step_size`` (jacobians.py:54-73).
"""
import numpy as np
import torch
import torch.autograd.forward_ad as fwAD

from . import _lib


def jacobian(fun, argnum=0, mode='reverse', step_size=1e-6):
    if mode == 'reverse':
        return lambda *args: jacobian_reverse(fun, argnum, *args)
    elif mode == 'forward':
        return lambda *args: jacobian_forward(fun, argnum, *args)
    elif mode == 'numerical':
        return lambda *args: jacobian_numerical(fun, argnum, *args, step_size=step_size)
    raise ValueError("'mode' kwarg must be either 'reverse' or 'forward' or 'numerical', given {}".format(mode))


def _as_tensor(x):
    if isinstance(x, torch.Tensor):
        t = x.detach()
    else:
        t = torch.as_tensor(np.asarray(x))
    if not (t.is_floating_point() or t.is_complex()):
        t = t.to(torch.float64)
    if t.dtype == torch.float32:
        t = t.to(torch.float64)
    if t.dtype == torch.complex64:
        t = t.to(torch.complex128)
    return t


def _real_out(out):
    out = out if isinstance(out, torch.Tensor) else torch.as_tensor(out)
    if out.is_complex():
        raise ValueError('jacobian needs a real-valued function output (take abs / real first)')
    return out.reshape(-1)


def jacobian_reverse(fun, argnum, *args):
    args = list(args)
    x = _as_tensor(args[argnum]).clone().requires_grad_(True)
    args[argnum] = x
    out = _real_out(fun(*args))
    m, n = x.numel(), out.numel()
    rows = []
    for k in range(n):
        g, = torch.autograd.grad(out[k], x, retain_graph=k < n - 1, allow_unused=True)
        if g is None:
            g = torch.zeros_like(x)
        g = g.detach().reshape(-1)
        if g.is_complex():
            g = torch.conj(g).resolve_conj()
        rows.append(g.cpu().numpy())
    return np.stack(rows).reshape(n, m)


def jacobian_forward(fun, argnum, *args):
    args = list(args)
    x = _as_tensor(args[argnum])
    m = x.numel()
    cols = []
    basis = [(k, 1.0) for k in range(m)]
    if x.is_complex():
        basis = [(k, ph) for k in range(m) for ph in (1.0, 1j)]
    vals = {}
    for k, ph in basis:
        t = torch.zeros(m, dtype=x.dtype)
        t[k] = ph
        with fwAD.dual_level():
            d = fwAD.make_dual(x, t.reshape(x.shape))
            a2 = list(args)
            a2[argnum] = d
            out = fun(*a2)
            tan = fwAD.unpack_dual(out).tangent
            if tan is None:
                tan = torch.zeros_like(out)
            vals[(k, ph)] = tan.detach().reshape(-1).cpu().numpy()
    for k in range(m):
        if x.is_complex():
            cols.append(vals[(k, 1.0)] - 1j * vals[(k, 1j)])
        else:
            cols.append(vals[(k, 1.0)])
    return np.stack(cols, axis=1)


def jacobian_numerical(fun, argnum, *args, step_size=1e-7):
    args = list(args)
    x0 = np.asarray(args[argnum].detach().cpu().numpy() if isinstance(args[argnum], torch.Tensor)
                    else args[argnum])
    flat = np.atleast_1d(x0).astype(x0.dtype if np.iscomplexobj(x0) else float).ravel()

    def f(v):
        a2 = list(args)
        a2[argnum] = v.reshape(np.shape(x0)) if np.ndim(x0) else v[0]
        o = fun(*a2)
        o = o.detach().cpu().numpy() if isinstance(o, torch.Tensor) else np.asarray(o)
        return np.atleast_1d(o).ravel()

    out0 = f(flat)
    jac = np.zeros((out0.size, flat.size))
    for i in range(flat.size):
        xi = flat.copy()
        xi[i] += step_size
        jac[:, i] = np.real((f(xi) - out0) / step_size)
    return jac
