"""Writes the golden fixtures of the parity tests from the CPU oracle.

    python tests/golden/make_golden.py

waveguide_box_120.npz   Ez (complex128) of the 120x120 Ez problem by
                        scipy spsolve on the reference operator, and the eps
                        gradient of |Ez[probe]|^2 by the closed-form adjoint.
tm_blobs_400_trace.npy  (2000, 4) float64 Ez probe trace of the 400x400 TM FDTD
                        problem from the numpy restatement of fdtd.py.
tm_blobs_400_adjoint.npz
                        probe energy L = sum over steps and probes of Ez^2 of
                        the same run and its directional derivatives dL/deps . d
                        along ADJOINT_DIRECTIONS (standard normal over the
                        grid from np.random.default_rng(seed)), by complex step
                        (h = 1e-30) of the complex128 restatement.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from ceviche_amd.workloads import fdfd_waveguide_box, fdtd_tm_blobs  # noqa: E402
from oracle import fdfd_ref, fdtd_ref  # noqa: E402

ADJOINT_DIRECTIONS = (400, 401)


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


def tm_blobs_adjoint():
    c = fdtd_tm_blobs()
    probes = [p for _, p in c['probes']]
    tr = fdtd_ref.run_probes(c['eps'], c['dL'], c['npml'], c['src'], lambda t: c['amp'][t], c['steps'], probes,
                             component='Ez')
    L = float(np.sum(tr ** 2))
    h = 1e-30
    dd = []
    for seed in ADJOINT_DIRECTIONS:
        d = np.random.default_rng(seed).standard_normal(c['eps'].shape)
        tc = fdtd_ref.run_probes(c['eps'] + 1j * h * d, c['dL'], c['npml'], c['src'], lambda t: c['amp'][t] + 0j,
                                 c['steps'], probes, component='Ez', dtype=np.complex128)
        dd.append(float(np.imag(np.sum(tc * tc)) / h))
    np.savez(os.path.join(HERE, 'tm_blobs_400_adjoint.npz'), L=L, seeds=np.array(ADJOINT_DIRECTIONS),
             directional_derivatives=np.array(dd))


if __name__ == '__main__':
    which = sys.argv[1:] or ['waveguide_box', 'tm_blobs', 'tm_blobs_adjoint']
    for name in which:
        globals()[name]()
    print('golden fixtures written to', HERE)
