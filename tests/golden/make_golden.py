"""Writes the golden fixtures of the parity tests from the CPU oracle.

    python tests/golden/make_golden.py

waveguide_box_120.npz   Ez (complex128) of the 120x120 Ez problem by
                        scipy spsolve on the reference operator, and the eps
                        gradient of |Ez[probe]|^2 by the closed-form adjoint.
tm_blobs_400_trace.npy  (2000, 4) float64 Ez probe trace of the 400x400 TM FDTD
                        problem from the numpy restatement of fdtd.py.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from ceviche_amd.workloads import fdfd_waveguide_box, fdtd_tm_blobs  # noqa: E402
from oracle import fdfd_ref, fdtd_ref  # noqa: E402


def waveguide_box():
    c = fdfd_waveguide_box()
    hx, hy, ez = fdfd_ref.solve_ez(c['omega'], c['dL'], c['eps'], c['npml'], c['src'])
    cot = np.zeros_like(ez)
    cot[c['probe']] = 2 * ez[c['probe']]
    g, _ = fdfd_ref.grad_eps_ez(c['omega'], c['dL'], c['eps'], c['npml'], c['src'], cot)
    np.savez_compressed(os.path.join(HERE, 'waveguide_box_120.npz'), Ez=ez, grad_eps=g)


def tm_blobs():
    c = fdtd_tm_blobs()
    tr = fdtd_ref.run_probes(c['eps'], c['dL'], c['npml'], c['src'], lambda t: c['amp'][t], c['steps'],
                             [p for _, p in c['probes']], component='Ez')
    np.save(os.path.join(HERE, 'tm_blobs_400_trace.npy'), tr)


if __name__ == '__main__':
    waveguide_box()
    tm_blobs()
    print('golden fixtures written to', HERE)
