"""Utilities with the reference's API (ceviche/utils.py).

Sparse-matrix helpers (utils.py:13-73): make_sparse, get_entries_indices,
transpose_indices, block_4, make_IO_matrices; random test data
(utils.py:78-103): make_rand, make_rand_complex, make_rand_indeces,
make_rand_entries_indices, make_rand_sparse, make_rand_sparse_density;
numerical derivatives (utils.py:108-150, 239-276): der_num, grad_num,
jac_num, vjp_maker_num; value and shape helpers (utils.py:200-236):
float_2_array, reshape_to_ND, get_value, get_shape.  These are host
utilities like the reference's (scipy / numpy formats); torch inputs to the
sparse helpers stay on their device.  Not carried over: the 3D FDFD grid
helpers (grid_center_to_xyz, grid_xyz_to_center, vec_zz_to_xy, for the
reference's 3D FDFD stub) and the matplotlib animation aniplot.

Probe and spectrum post-processing (utils.py:316-332, 350-399):

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
import scipy.sparse as sp
import torch

from . import _lib
from ._probes import ProbeWeights, weighted_sums


# ------------------------------------------------ sparse matrix utilities
def make_sparse(entries, indices, shape):
    """CSR matrix of the COO triplets with duplicates summed, complex128
    (utils.py:13-25).  numpy inputs give a scipy.sparse CSR matrix like the
    reference; torch inputs give a torch sparse CSR tensor on their device,
    coalesced (duplicates summed) by torch."""
    if isinstance(entries, torch.Tensor) or isinstance(indices, torch.Tensor):
        e = torch.as_tensor(entries).to(torch.complex128)
        idx = torch.as_tensor(indices, device=e.device).to(torch.int64)
        coo = torch.sparse_coo_tensor(idx, e, size=tuple(int(v) for v in shape)).coalesce()
        return coo.to_sparse_csr()
    coo = sp.coo_matrix((entries, indices), shape=shape, dtype=np.complex128)
    return coo.tocsr()


def transpose_indices(indices):
    """Index set of the transposed matrix: rows and columns exchanged (utils.py:37-39)."""
    if isinstance(indices, torch.Tensor):
        return torch.flip(indices, dims=(0,))
    return np.flip(indices, axis=0)


def get_entries_indices(csr_matrix):
    """(entries, indices) of a CSR matrix in storage order, indices = [rows;
    cols] (utils.py:27-35); a torch sparse CSR tensor gives torch tensors."""
    if isinstance(csr_matrix, torch.Tensor):
        crow = csr_matrix.crow_indices()
        cols = csr_matrix.col_indices()
        counts = crow[1:] - crow[:-1]
        rows = torch.repeat_interleave(torch.arange(counts.numel(), device=crow.device), counts)
        return csr_matrix.values(), torch.stack((rows, cols))
    coo = csr_matrix.tocoo()
    return csr_matrix.data, np.vstack((coo.row, coo.col))


def block_4(A, B, C, D):
    """[[A, B], [C, D]] from four sparse blocks (utils.py:41-48)."""
    return sp.hstack([sp.vstack([A, C]), sp.vstack([B, D])])


def make_IO_matrices(indices, N):
    """One-hot incidence matrices (I, O) of an index set (utils.py:50-73):
    column k of I (N x M) selects entry k's row, O = J^T (M x N) its column,
    so that A = I diag(a) O and a = diag(I^T A O^T)."""
    idx = np.asarray(indices.detach().cpu().numpy() if isinstance(indices, torch.Tensor) else indices)
    M = idx.shape[1]
    ones = np.ones(M)
    I = make_sparse(ones, np.vstack((idx[0], np.arange(M))), shape=(N, M))
    J = make_sparse(ones, np.vstack((idx[1], np.arange(M))), shape=(N, M))
    return I, J.T


# --------------------------------------------------- data generation
def make_rand(N):
    """N uniform values in [-0.5, 0.5) from numpy's global generator (utils.py:78-80)."""
    return np.random.random(N) - 0.5


def make_rand_complex(N):
    """Complex values with real and imaginary parts from make_rand (utils.py:82-84)."""
    return make_rand(N) + 1j * make_rand(N)


def make_rand_indeces(N, M):
    """(2, M) random indices into an N x N matrix (utils.py:86-88; the reference's spelling)."""
    return np.random.randint(low=0, high=N, size=(2, M))


def make_rand_entries_indices(N, M):
    """M random complex entries and their random indices (utils.py:90-94)."""
    entries = make_rand_complex(M)
    indices = make_rand_indeces(N, M)
    return entries, indices


def make_rand_sparse(N, M):
    """Random N x N CSR matrix from M random triplets (utils.py:96-99)."""
    entries, indices = make_rand_entries_indices(N, M)
    return make_sparse(entries, indices, shape=(N, N))


def make_rand_sparse_density(N, density=1):
    """Random complex N x N sparse matrix of a given density (utils.py:101-103)."""
    return sp.random(N, N, density=density) + 1j * sp.random(N, N, density=density)


# ------------------------------------------------- numerical derivatives
def der_num(fn, arg, index, delta):
    """Central difference (fn(arg + delta/2 e_i) - fn(arg - delta/2 e_i)) / delta
    along entry ``index`` of a numpy array (utils.py:108-115); a complex delta
    steps the imaginary part (discarded, with numpy's ComplexWarning, when arg
    is real, as in the reference)."""
    arg_i_for = arg.copy()
    arg_i_back = arg.copy()
    arg_i_for[index] += delta / 2
    arg_i_back[index] -= delta / 2
    return (fn(arg_i_for) - fn(arg_i_back)) / delta


def grad_num(fn, arg, delta=1e-6):
    """Complex numerical gradient dfn/dRe - i dfn/dIm of every entry, complex128
    (utils.py:117-125): autograd's cotangent convention."""
    N = arg.size
    grad = np.zeros((N,), dtype=np.complex128)
    for i in range(N):
        grad[i] = der_num(fn, arg, i, delta)
        grad[i] += der_num(fn, arg, i, 1j * delta)
    return grad


def jac_num(fn, arg, step_size=1e-7):
    """Forward-difference Jacobian of fn at arg, laid out (input entries,
    output entries) as in utils.py:127-150: row i is (fn(arg + h e_i) -
    fn(arg)) / h over the flattened output (real float64).  Scalars count
    as one-entry arrays."""
    x0 = float_2_array(arg)
    flat = np.array(x0, dtype=np.float64).ravel()
    f0 = np.ravel(get_value(float_2_array(fn(arg))))
    jac = np.zeros((flat.size, f0.size))
    for i in range(flat.size):
        xi = flat.copy()
        xi[i] += step_size
        fi = np.ravel(get_value(fn(xi.reshape(np.shape(x0)))))
        jac[i, :] = np.real(get_value((fi - f0) / step_size))
    return jac


def float_2_array(x):
    """x unchanged when it is an array (numpy or torch), else a one-entry
    numpy array holding it (utils.py:200-204)."""
    if isinstance(x, (np.ndarray, torch.Tensor)):
        return x
    return np.array([x])


def reshape_to_ND(arr, N):
    """arr with trailing axes of length 1 added up to N dimensions; ValueError
    when it already has more (utils.py:206-215)."""
    nd = len(arr.shape)
    if nd > N:
        raise ValueError("array is larger than {} dimensional, given shape {}".format(N, tuple(arr.shape)))
    return arr.reshape(tuple(arr.shape) + (1,) * (N - nd))


def get_value(x):
    """The plain value of x: a torch tensor detached from its autograd graph
    (the counterpart of unboxing an autograd ArrayBox, utils.py:220-224),
    anything else unchanged."""
    if isinstance(x, torch.Tensor):
        return x.detach()
    return x


def get_shape(x):
    """Shape of x also for non-arrays (utils.py:229-236): (1,) for a Python
    int or float, (len(x),) for a tuple or list, else tuple(x.shape)."""
    if isinstance(x, (float, int)):
        return (1,)
    if isinstance(x, (tuple, list)):
        return (len(x),)
    return tuple(np.shape(x)) if not isinstance(x, torch.Tensor) else tuple(x.shape)


def vjp_maker_num(fn, arg_inds, steps):
    """Numerical VJP makers for fn with respect to the arguments at positions
    arg_inds, forward differences with the matching steps (utils.py:239-276):
    one maker per argument, called as maker(fn_out, *args) and returning
    vjp(v) = [sum(v * (fn(args with entry p stepped) - fn_out) / step) for
    every entry p], real float64."""
    def one(arg_ind, step):
        def vjp_maker(fn_out, *args):
            shape = np.shape(args[arg_ind])
            size = int(np.prod(shape)) if shape else 1

            def vjp(v):
                out = np.zeros(size)
                for ip in range(size):
                    moved = list(args)
                    flat = np.array(args[arg_ind], dtype=np.float64).ravel()
                    flat[ip] += step
                    moved[arg_ind] = flat.reshape(shape)
                    out[ip] = np.real(np.sum(v * (fn(*moved) - fn_out) / step))
                return out

            return vjp

        return vjp_maker

    return tuple(one(arg_inds[ia], steps[ia]) for ia in range(len(arg_inds)))


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
