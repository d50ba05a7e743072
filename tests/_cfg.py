"""Shared seeded configurations of the parity tests (the benchmark workloads)."""
import numpy as np

from oracle.fdfd_ref import C_0


def cfg1():
    """Plain 2D fdfd_ez, 120x120, dL = 50 nm, npml = 20, lambda = 1.55 um; waveguide
    along x (10 cells, eps 12.25) and a 40x40 seeded design box; Jz line source on
    x-index 30; objective |Ez[90, 60]|^2."""
    N = 120
    omega = 2 * np.pi * C_0 / 1.55e-6
    eps = np.ones((N, N))
    eps[:, 55:65] = 12.25
    eps[40:80, 40:80] = np.random.default_rng(0).uniform(1, 12.25, (40, 40))
    src = np.zeros((N, N))
    src[30, :] = 1.0
    return dict(omega=omega, dL=50e-9, eps=eps, npml=[20, 20], src=src, probe=(90, 60))
