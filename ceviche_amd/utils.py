"""Probe and spectrum post-processing with the reference's API
(ceviche/utils.py:316-332, 350-399):

    measured = measure_fields(F, source, steps, probes, component='Ez')
    X = my_fft(x)
    freqs, signal_f = get_spectrum(series, dt)
    f_max = get_max_power_freq(series, dt)
    freqs, power = get_spectral_power(series, dt)

measure_fields steps the GPU FDTD one step at a time (the source is an
arbitrary callable of the step index, as in the reference) and accumulates
sum(field * probe) on the device with the CSR weighted-sum kernel, with no
host synchronisation until the end.  For a point source with an amplitude
series, fdtd.run with weighted probes does the same measurement fused into the
multi-step kernels and differentiably.

The transforms run on the device through torch.fft; numpy in gives numpy
out.  As in the reference, my_fft transforms along the LAST axis (numpy's
default), and get_spectrum applies it to the (steps, n) windowed series, so
for a single series (n = 1) the transform is the identity along the probe
axis and the result is the Hamming-windowed series itself; this is
reproduced as written.
"""
import numpy as np
import torch

from . import _lib
from ._probes import ProbeWeights, weighted_sums


def measure_fields(F, source, steps, probes, component='Ez'):
    """(steps, n_probes) float64 time series of sum(F.fields[component] * probe)
    after each step of F driven by Jz = source(t_index) (utils.py:316-332).
    Resets F's fields first.  Progress is printed every steps // 20 steps as
    the reference does (for steps < 20, where the reference divides by zero,
    nothing is printed)."""
    F.initialize_fields()
    if not isinstance(probes, list):
        probes = [probes]
    g = F._grid
    pw = ProbeWeights([(component, p) for p in probes], g.shape, g.device)
    cols = torch.tensor([np.ravel_multi_index(pos, g.shape) for _, pos in pw.points], dtype=torch.int64,
                        device=g.device)
    # columns of the CSR are indices into `cols`; map them to grid cells once
    rowptr, col, w = pw.fwd
    csr = (rowptr, cols[col] if col.numel() else col, w)
    out = torch.zeros((steps, len(probes)), dtype=g.dtype, device=g.device)
    with torch.no_grad():
        for t_index in range(steps):
            if steps >= 20 and t_index % (steps // 20) == 0:
                print('{:.2f} % done'.format(float(t_index) / steps * 100.0))
            F._advance(Jz=source(t_index))
            field = F._device_field(component).reshape(1, -1)
            weighted_sums(field, csr, len(probes), out=out[t_index:t_index + 1])
    measured = out.to(torch.float64)
    return measured if F._as_torch else measured.cpu().numpy()


def _as_device(x):
    if isinstance(x, torch.Tensor):
        return x.to(_lib.device()), True
    return torch.from_numpy(np.ascontiguousarray(x)).to(_lib.device()), False


def my_fft(x):
    """FFT along the last axis (numpy.fft.fft's default), utils.py:350-357.
    Differentiable in both modes through torch (the transform is linear,
    utils.py:360-370)."""
    t, was_torch = _as_device(x)
    if not torch.is_complex(t):
        t = t.to(torch.float64)
    y = torch.fft.fft(t, dim=-1)
    return y if was_torch else y.cpu().numpy()


def _hamming(M):
    """numpy.hamming: 0.54 - 0.46 cos(2 pi n / (M - 1)), [1.0] for M == 1."""
    if M == 1:
        return np.ones(1)
    n = np.arange(M)
    return 0.54 - 0.46 * np.cos(2.0 * np.pi * n / (M - 1))


def _fftfreq(n, d):
    """numpy.fft.fftfreq: [0, 1, ..., ceil(n/2)-1, -floor(n/2), ..., -1] / (d n)."""
    k = np.concatenate([np.arange(0, (n - 1) // 2 + 1), np.arange(-(n // 2), 0)])
    return k / (d * n)


def get_spectrum(series, dt):
    """(freqs, signal_f) of the Hamming-windowed series (utils.py:373-387)."""
    steps = len(series)
    t, was_torch = _as_device(series)
    t = t.reshape((steps, -1))
    window = torch.from_numpy(_hamming(steps).reshape((steps, 1))).to(t.device)
    signal_f = my_fft(window * t)
    freqs = _fftfreq(steps, dt)
    if was_torch:
        return torch.from_numpy(freqs).to(t.device), signal_f
    return freqs, signal_f.cpu().numpy() if isinstance(signal_f, torch.Tensor) else signal_f


def _argmax_lexicographic(z):
    """numpy.argmax of a complex array: the first maximum in (real, imag)
    lexicographic order, over the flattened array."""
    flat = z.reshape(-1)
    if not torch.is_complex(flat):
        return int(torch.argmax(flat))
    re, im = flat.real, flat.imag
    on_top = re == re.max()
    im_masked = torch.where(on_top, im, torch.full_like(im, -float('inf')))
    best = im_masked.max()
    hit = torch.nonzero(on_top & (im_masked == best))
    return int(hit[0, 0])


def get_max_power_freq(series, dt):
    """freqs[argmax(signal_f)] (utils.py:390-393), numpy's complex ordering."""
    freqs, signal_f = get_spectrum(series, dt)
    sf = signal_f if isinstance(signal_f, torch.Tensor) else torch.from_numpy(signal_f)
    return freqs[_argmax_lexicographic(sf)]


def get_spectral_power(series, dt):
    """(freqs, |signal_f|^2) (utils.py:396-399)."""
    freqs, signal_f = get_spectrum(series, dt)
    if isinstance(signal_f, torch.Tensor):
        return freqs, torch.square(torch.abs(signal_f))
    return freqs, np.square(np.abs(signal_f))
