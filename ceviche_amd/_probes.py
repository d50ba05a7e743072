"""Weighted field probes: sum(field * weights) per probe and step, the
reduction of ceviche's measure_fields (ceviche/utils.py:316-332), evaluated
by the CSR weighted-sum kernel of csrc/probe.hip."""
import numpy as np
import torch

from . import _lib
from ._lib import check, lib, ptr, stream_ptr

_CODE = {torch.float32: 0, torch.float64: 1}   # CV_F32, CV_F64


def _weights_on_grid(pos, shape):
    w = pos.detach().cpu().numpy() if isinstance(pos, torch.Tensor) else np.asarray(pos)
    if np.iscomplexobj(w):
        raise ValueError('probe weights must be real')
    w = w.reshape(tuple(w.shape) + (1,) * (3 - w.ndim)) if w.ndim < 3 else w
    if w.ndim > 3:
        raise ValueError('probe weights of shape %s do not fit the grid %s' % (w.shape, shape))
    try:
        return np.broadcast_to(w.astype(np.float64), shape).reshape(-1)
    except ValueError:
        raise ValueError('probe weights of shape %s do not broadcast to the grid %s' % (w.shape, shape))


def _csr(rows, n_cols, device):
    rowptr = np.zeros(len(rows) + 1, dtype=np.int64)
    cols, vals = [], []
    for r, (c, v) in enumerate(rows):
        rowptr[r + 1] = rowptr[r] + len(c)
        cols.append(np.asarray(c, dtype=np.int64))
        vals.append(np.asarray(v, dtype=np.float64))
    col = np.concatenate(cols) if cols else np.zeros(0, np.int64)
    val = np.concatenate(vals) if vals else np.zeros(0)
    if col.size and (col.min() < 0 or col.max() >= n_cols):
        raise ValueError('probe column out of range')
    to = lambda a: torch.from_numpy(a).to(device)
    return to(rowptr), to(col), to(val)


class ProbeWeights:
    """Probes [(field name, position or weight array), ...] on a grid of
    ``shape`` as a CSR matrix over the union of their (component, cell)
    supports.  ``points`` lists that union as point probes, the columns of
    the recorded matrix the CSR rows combine."""

    def __init__(self, probes, shape, device):
        shape = tuple(int(v) for v in shape)
        self.n_probes = len(probes)
        index = {}
        points = []
        rows = []
        for name, pos in probes:
            if isinstance(pos, (tuple, list)):
                p = tuple(int(v) for v in pos) + (0,) * (3 - len(pos))
                cells, vals = [np.ravel_multi_index(p, shape)], [1.0]
            else:
                w = _weights_on_grid(pos, shape)
                cells = np.flatnonzero(w)
                vals = w[cells]
            cols = []
            for c in cells:
                key = (name, int(c))
                if key not in index:
                    index[key] = len(points)
                    points.append((name, np.unravel_index(int(c), shape)))
                cols.append(index[key])
            rows.append((cols, vals))
        self.points = [(n, tuple(int(v) for v in p)) for n, p in points]
        self.n_points = len(points)
        self.fwd = _csr(rows, self.n_points, device)
        trows = [([], []) for _ in range(self.n_points)]
        for r, (cols, vals) in enumerate(rows):
            for c, v in zip(cols, vals):
                trows[c][0].append(r)
                trows[c][1].append(v)
        self.bwd = _csr(trows, self.n_probes, device)


def weighted_sums(field_rows, csr, n_out, out=None, accumulate=False):
    """out[t, p] = sum_k w[p, k] field_rows[t, col[k]] for a (T, F) tensor."""
    x = field_rows.contiguous()
    T, F = x.shape
    if out is None:
        out = torch.empty((T, n_out), dtype=x.dtype, device=x.device)
    if F == 0 or T == 0 or csr[1].numel() == 0:      # every probe has an empty support
        if not accumulate:
            out.zero_()
        return out
    rowptr, col, w = csr
    check(lib().cv_probe_weights(_CODE[x.dtype], ptr(x), F, T, ptr(rowptr), ptr(col), ptr(w), n_out, ptr(out),
                                 out.stride(0), int(bool(accumulate)), stream_ptr()), 'probe_weights')
    return out


class ProbeSums(torch.autograd.Function):
    """(T, n_probes) weighted sums of a (T, n_points) record; linear, so the
    backward applies the transposed weights and the jvp the weights."""

    @staticmethod
    def forward(ctx, rec, weights):
        ctx.weights = weights
        return weighted_sums(rec, weights.fwd, weights.n_probes)

    @staticmethod
    def jvp(ctx, drec, _):
        return weighted_sums(drec, ctx.weights.fwd, ctx.weights.n_probes)

    @staticmethod
    def backward(ctx, g):
        return weighted_sums(g, ctx.weights.bwd, ctx.weights.n_points), None
