"""Seeded synthetic problem definitions used by the benchmark and the tests.

Each builder returns plain numpy inputs (host setup only) for a named problem:

fdfd_waveguide_box   2D Ez FDFD, 120x120, dL 50 nm, npml 20, lambda 1.55 um; a
                     straight waveguide along x (eps 12.25, 10 cells) and a
                     40x40 design box of seeded uniform eps in [1, 12.25];
                     Jz line source at x-index 30; objective |Ez[90, 60]|^2.
fdtd_tm_blobs        2D TM FDTD, NxN (400 by default), dL 20 nm, npml 20;
                     eps = Gaussian-filtered noise (sigma 4 cells) thresholded
                     to {1, 12.25}; Gaussian-pulse Jz at the centre (centre
                     step 200, width 50 steps); Ez probes at (N/4, N/2),
                     (3N/4, N/2), (N/2, N/4), (N/2, 3N/4).
fdfd_blob_batch      batch of 2D Ez FDFD instances, sizes drawn from
                     {200, 300, 400, 600}^2, dL 40 nm, npml 20, wavelengths
                     evenly spaced over 1.50-1.60 um; eps = waveguide along x
                     plus a central binary-blob design (sigma 3 cells);
                     Jz line source across the input waveguide.
fdfd_mf_isolator     multi-frequency Ez FDFD, the modulated-waveguide isolator
                     of ceviche's examples/fdfd_mf_intro.ipynb: 1100x80 cells,
                     dL 50 nm, npml 10, omega 243e12 rad/s, omega_mod 132e12
                     rad/s, 3 sidebands (Nsb 1), one modulation harmonic;
                     eps 12.25 waveguides (22 cells) joined by a 40-cell-wide
                     centre section; two 380-cell modulated strips on the
                     upper half of the guide, phases 0 and pi/2, depth 1.
fdtd_3d_blobs        3D FDTD, N^3 (256 by default), dL 25 nm, npml 10;
                     background eps 4.41 with a centred (N/4)^3 design region
                     of seeded binary blobs {1, 12.25}; z-dipole Gaussian
                     pulse; flux-plane probes.
"""
import numpy as np

C_0 = 1 / np.sqrt(8.85418782e-12 * 1.25663706e-6)


def fdfd_waveguide_box(N=120):
    omega = 2 * np.pi * C_0 / 1.55e-6
    eps = np.ones((N, N))
    c = N // 2
    eps[:, c - 5:c + 5] = 12.25
    q = N // 3
    eps[q:2 * q, q:2 * q] = np.random.default_rng(0).uniform(1, 12.25, (2 * q - q, 2 * q - q))
    src = np.zeros((N, N))
    src[N // 4, :] = 1.0
    return dict(omega=omega, dL=50e-9, eps=eps, npml=[20, 20], src=src, probe=(3 * N // 4, N // 2),
                box=(slice(q, 2 * q), slice(q, 2 * q)))


def gaussian_pulse(steps, t0, width):
    t = np.arange(steps)
    return np.exp(-(t - t0) ** 2 / (2.0 * width ** 2))


def fdtd_tm_blobs(N=400, steps=2000, seed=1):
    from scipy.ndimage import gaussian_filter
    noise = gaussian_filter(np.random.default_rng(seed).standard_normal((N, N)), 4)
    eps = np.where(noise > 0, 12.25, 1.0)[:, :, None]
    c = N // 2
    probes = [('Ez', (N // 4, c, 0)), ('Ez', (3 * N // 4, c, 0)), ('Ez', (c, N // 4, 0)), ('Ez', (c, 3 * N // 4, 0))]
    return dict(eps=eps, dL=20e-9, npml=[20, 20, 0], src=(c, c, 0), amp=gaussian_pulse(steps, 200, 50),
                probes=probes, steps=steps)


def fdtd_3d_blobs(N=256, steps=1000, seed=2):
    from scipy.ndimage import gaussian_filter
    eps = np.full((N, N, N), 2.1 ** 2)
    d = N // 4
    lo = N // 2 - d // 2
    blob = gaussian_filter(np.random.default_rng(seed).standard_normal((d, d, d)), 3) > 0
    eps[lo:lo + d, lo:lo + d, lo:lo + d] = np.where(blob, 12.25, 1.0)
    c = N // 2
    src = (c, c, max(1, int(round(N * 60 / 256))))
    zp = int(round(N * 200 / 256))
    probes = []
    for comp in ('Ex', 'Ey', 'Hx', 'Hy'):
        for i in range(lo, lo + d):
            for j in range(lo, lo + d):
                probes.append((comp, (i, j, zp)))
    return dict(eps=eps, dL=25e-9, npml=[10, 10, 10], src=src, amp=gaussian_pulse(steps, 200, 50),
                probes=probes, steps=steps, plane=(slice(lo, lo + d), slice(lo, lo + d), zp))


def fdfd_blob_batch(n_inst=64, n_lambda=8, seed=3, sizes=(200, 300, 400, 600)):
    """List of dicts (omega, dL, eps, npml, src) for a multi-frequency batch."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    lams = np.linspace(1.50e-6, 1.60e-6, n_lambda)
    out = []
    for k in range(n_inst):
        N = int(rng.choice(sizes))
        eps = np.ones((N, N))
        c = N // 2
        eps[:, c - 5:c + 5] = 12.25
        d = N // 3
        blob = gaussian_filter(rng.standard_normal((d, d)), 3) > 0
        eps[c - d // 2:c - d // 2 + d, c - d // 2:c - d // 2 + d] = np.where(blob, 12.25, 1.0)
        src = np.zeros((N, N))
        src[25, c - 8:c + 8] = 1.0
        for lam in lams:
            out.append(dict(omega=2 * np.pi * C_0 / lam, dL=40e-9, eps=eps, npml=[20, 20], src=src))
    return out


def fdfd_learned_blobs(N=256, seed=0, lam=None):
    """2D Ez FDFD instance for the learned preconditioner: NxN, dL 40 nm,
    npml 20; background eps 1, a waveguide along x (eps 12.25, 10 cells) and
    a central (N/2)^2 region of seeded binary blobs {1, 12.25} (Gaussian-
    filtered noise, sigma 3 cells, thresholded at 0); wavelength drawn
    uniformly from [1.45, 1.65] um unless given; Jz line source across the
    waveguide at x-index 30.  Seeds 0 .. 9999 are training structures,
    seeds >= 10000 are held out for evaluation."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    if lam is None:
        lam = rng.uniform(1.45e-6, 1.65e-6)
    eps = np.ones((N, N))
    c = N // 2
    eps[:, c - 5:c + 5] = 12.25
    d = N // 2
    blob = gaussian_filter(rng.standard_normal((d, d)), 3) > 0
    eps[c - d // 2:c - d // 2 + d, c - d // 2:c - d // 2 + d] = np.where(blob, 12.25, 1.0)
    src = np.zeros((N, N))
    src[30, c - 8:c + 8] = 1.0
    return dict(omega=2 * np.pi * C_0 / lam, dL=40e-9, eps=eps, npml=[20, 20], src=src, lam=lam)


def fdfd_mf_isolator():
    """Geometry of the modulated-waveguide isolator example (sizes in cells
    from its lengths: L 55 um, W 4 um, waveguide 1.1 um, centre 2 um x
    4.75 um, modulation 19 um, at dL 50 nm).  The source is the fundamental
    mode on the input cross section (x = npml + 1, y over 3 waveguide widths),
    on the centre sideband; build it with modes.insert_mode."""
    nx, ny, npml = 1100, 80, 10
    eps = np.ones((nx, ny))
    l2, r1 = nx // 2 - 47, nx // 2 + 47
    top, bot = ny // 2 + 11, ny // 2 - 11
    eps[:l2, bot:top] = 12.25
    eps[r1:, bot:top] = 12.25
    eps[l2:r1, ny // 2 - 20:ny // 2 + 20] = 12.25
    delta = np.zeros((1, nx, ny))
    phi = np.zeros((1, nx, ny))
    delta[:, l2 - 380:l2, ny // 2:top] = 1.0
    delta[:, r1:r1 + 380, ny // 2:top] = 1.0
    phi[:, r1:r1 + 380, ny // 2:top] = np.pi / 2
    return dict(omega=243e12, omega_mod=132e12, Nsb=1, dL=50e-9, npml=[npml, npml], eps=eps, delta=delta,
                phi=phi, src_x=npml + 1, src_y=np.arange(ny // 2 - 33, ny // 2 + 33))
