"""CPU restatement of ceviche's probe and spectrum helpers (test
infrastructure only: imported by tests/ and never by the product).

measure_fields (ceviche/utils.py:316-332): reset, then per step
F.forward(Jz=source(t)) and sum(fields[component] * probe) per probe, on the
numpy FDTD oracle.  get_spectrum / get_max_power_freq / get_spectral_power
(:373-399) and my_fft (:350-357): numpy.fft.fft along the last axis of the
Hamming-windowed (steps, n) series, fftfreq(steps, dt) frequencies, argmax in
numpy's complex (lexicographic) order.

Pinned by tests/test_oracle_spectrum.py: measure_fields with one-cell probes
equals the oracle's point-probe run, probes are linear in their weights;
the window equals the closed form 0.54 - 0.46 cos(2 pi n / (M - 1)) and
scipy's Hamming window; my_fft equals a direct DFT sum; a pure tone's
spectral power (transformed along time) peaks at the tone's frequency bin;
the (steps, 1) reshaping makes get_spectrum's transform the identity
along the probe axis.
"""
import numpy as np

from .fdtd_ref import FDTD


def measure_fields(F, source, steps, probes, component='Ez'):
    F.reset()
    if not isinstance(probes, list):
        probes = [probes]
    measured = np.zeros((steps, len(probes)))
    for t in range(steps):
        fields = F.forward(Jz=source(t))
        for p, probe in enumerate(probes):
            measured[t, p] = np.sum(fields[component] * probe)
    return measured


def my_fft(x):
    return np.fft.fft(x)


def get_spectrum(series, dt):
    steps = len(series)
    series = np.asarray(series).reshape((steps, -1))
    window = np.hamming(steps).reshape((steps, 1))
    return np.fft.fftfreq(steps, d=dt), my_fft(window * series)


def get_max_power_freq(series, dt):
    freqs, signal_f = get_spectrum(series, dt)
    return freqs[np.argmax(signal_f)]


def get_spectral_power(series, dt):
    freqs, signal_f = get_spectrum(series, dt)
    return freqs, np.square(np.abs(signal_f))


__all__ = ['FDTD', 'measure_fields', 'my_fft', 'get_spectrum', 'get_max_power_freq', 'get_spectral_power']
