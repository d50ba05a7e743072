"""ctypes binding of the C-ABI in include/ceviche_amd.h.

The shared library is loaded after ``import torch`` so that it binds to the
HIP runtime torch already loaded (both are SONAME libamdhip64.so.7).  There is
no CPU fallback: if the library is missing or no GPU is present, the product
entry points raise.
"""
import ctypes
import os

import numpy as np
import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libceviche_amd.so')

CV_OK = 0
CV_ERR_INVALID = 1
CV_ERR_HIP = 2
CV_ERR_NOT_CONVERGED = 3
CV_ERR_NONFINITE = 4
CV_ERR_BREAKDOWN = 5
CV_ERR_SINGULAR = 6

CV_C64 = 0
CV_C128 = 1
CV_POL_EZ = 0
CV_POL_HZ = 1
CV_FLAG_X_IDENTITY = 1
CV_FLAG_Y_IDENTITY = 2
CV_FLAG_CPLX_MAT = 4
CV_TILE_CELLS = 256
CV_NOALLOC_SCRATCH_BYTES = 32 << 20
CV_NOALLOC_MAX_INST = 64
CV_KRYLOV_BICGSTAB = 0
CV_KRYLOV_COCG = 1
CV_FACTOR_BTD = 1
CV_FACTOR_BCR = 2

_STATUS_NAMES = {
    CV_ERR_INVALID: 'invalid argument',
    CV_ERR_HIP: 'HIP runtime error',
    CV_ERR_NOT_CONVERGED: 'solver did not converge',
    CV_ERR_NONFINITE: 'non-finite value in solver',
    CV_ERR_BREAKDOWN: 'Krylov breakdown',
    CV_ERR_SINGULAR: 'singular matrix',
}

DESC_DTYPE = np.dtype([
    ('nx', '<i4'), ('ny', '<i4'), ('pol', '<i4'), ('flags', '<i4'),
    ('cell_off', '<i8'), ('prof_off', '<i8'),
    ('tile_first', '<i4'), ('tile_count', '<i4'),
    ('scale', '<f8'), ('dshift', '<f8'),
    ('phx_re', '<f8'), ('phx_im', '<f8'), ('phy_re', '<f8'), ('phy_im', '<f8'),
    ('omega', '<f8'),
    ('n_cpl', '<i4'), ('blk', '<i4'), ('nblk', '<i4'), ('pad', '<i4'),
    ('cpl_off', '<i8'), ('w2_off', '<i8'),
])
assert DESC_DTYPE.itemsize == 128
TILE_DTYPE = np.dtype([('inst', '<i4'), ('first', '<i4')])


class SolveInfo(ctypes.Structure):
    _fields_ = [('iterations', ctypes.c_int32), ('status', ctypes.c_int32),
                ('restarts', ctypes.c_int32), ('pad', ctypes.c_int32),
                ('rel_residual', ctypes.c_double)]


class FdtdDesc(ctypes.Structure):
    _fields_ = [('nx', ctypes.c_int32), ('ny', ctypes.c_int32), ('nz', ctypes.c_int32),
                ('pad', ctypes.c_int32), ('dt', ctypes.c_double), ('dL', ctypes.c_double),
                ('c0', ctypes.c_double), ('eps0', ctypes.c_double)]


CV_PRECOND_NONE = 0
CV_PRECOND_JACOBI = 1
CV_PRECOND_UNET = 2
CV_CNN_PACKED_ELEMS = 9216


class UNetDesc(ctypes.Structure):
    _fields_ = [('wpack', ctypes.c_void_p), ('bias', ctypes.c_void_p), ('feat', ctypes.c_void_p),
                ('nf', ctypes.c_int32), ('levels', ctypes.c_int32)]


class CevicheError(RuntimeError):
    """Raised for a non-zero status from the HIP library."""

    def __init__(self, status, msg):
        super().__init__('%s (status %d): %s' % (_STATUS_NAMES.get(status, 'error'), status, msg))
        self.status = status


class NotConvergedError(CevicheError):
    pass


_lib = None

P = ctypes.c_void_p
I = ctypes.c_int
I64 = ctypes.c_int64
D = ctypes.c_double
F = ctypes.c_float

_SIGS = {
    'cv_version': (ctypes.c_char_p, []),
    'cv_last_error': (ctypes.c_char_p, []),
    'cv_fdfd_apply': (I, [I, P, P, I, P, P, P, P, P, I, P]),
    'cv_fdfd_hz_materials': (I, [I, P, P, I, P, P, P, P]),
    'cv_fdfd_hz_materials_grad': (I, [I, P, P, I, P, P, P, P, P, P]),
    'cv_fdfd_deriv': (I, [I, P, P, I, P, P, P, P, I, I, D, D, I, P]),
    'cv_fdfd_deriv_matgrad': (I, [I, P, P, I, P, P, P, P, I, D, D, P]),
    'cv_fdfd_material_grad': (I, [I, P, P, I, P, P, P, P, P, D, P]),
    'cv_fdfd_material_apply': (I, [I, P, P, I, P, P, P, P, P, I, P]),
    'cv_fdfd_hz_materials_jvp': (I, [I, P, P, I, P, P, P, P, P, P]),
    'cv_fdfd_coupling_apply': (I, [I, P, P, I, P, P, P, I, I, P]),
    'cv_probe_weights': (I, [I, P, I64, I, P, P, P, I, P, I64, I, P]),
    'cv_fdfd_btd_factor_bytes': (I64, [I, I]),
    'cv_fdfd_btd_factor': (I, [P, P, P, P, P, P, I64, P]),
    'cv_fdfd_btd_apply': (I, [P, P, P, P, P, P, P, P, I, P]),
    'cv_fdfd_bcr_factor_bytes': (I64, [I, I]),
    'cv_fdfd_bcr_factor': (I, [P, P, P, P, P, P, I64, P]),
    'cv_fdfd_bcr_factor_batch': (I, [I, P, P, P, P, P, P, I64, P]),
    'cv_fdfd_bcr_apply': (I, [P, P, P, P, P, I, P]),
    'cv_fdfd_fgmres_blockdiag': (I, [P, P, P, P, P, P, P, P, P, P, P, P, I, D, I, I, P, I64, P, P]),
    'cv_fdfd_fgmres_direct': (I, [P, P, P, P, P, P, P, P, I, P, P, P, I, D, I, I, P, I64, P, P]),
    'cv_fdfd_coupling_grad': (I, [I, P, P, I, P, P, P, P, D, P]),
    'cv_fdfd_solve_workspace_bytes': (I64, [I, I64, I, I]),
    'cv_fdfd_solve': (I, [I, P, P, I, P, I, P, P, P, P, P, I, D, I, I, P, I64, P, P]),
    'cv_noalloc_init': (I, []),
    'cv_fdfd_solve_noalloc': (I, [I, P, P, I, P, I, P, P, P, P, P, I, D, I, I, P, P]),
    'cv_fdfd_solve_mixed_workspace_bytes': (I64, [I64, I, I]),
    'cv_fdfd_cocg': (I, [I, P, P, I, P, I, P, P, P, P, P, I, D, I, I, P, I64, P, P]),
    'cv_fdfd_solve_mixed': (I, [P, P, I, P, I, P, P, P, P, P, P, P, P, I, I, D, D, I, I, I, P, I64, P, P]),
    'cv_fdtd_step': (I, [I, P, P, P, P, P, P, P, D, I, I64, P, I, P, P, I, P, P]),
    'cv_fdtd_efields': (I, [I, P, P, P, P, P, P]),
    'cv_fdtd_step_tangent': (I, [I, P, P, P, P, P, P, P, P, P, P, P, P, D, I, I64, P, P, I, P, P, I, P, P, P]),
    'cv_fdtd_efields_tangent': (I, [I, P, P, P, P, P, P, P, P]),
    'cv_fdtd_efields_grad': (I, [I, P, P, P, P, P, P, P, P]),
    'cv_fdtd_adjoint_step': (I, [I, P, P, P, P, P, P, P, P, P, I, P, P, I, I64, P, P]),
    'cv_fdtd_eps_grad': (I, [I, P, P, P, P, P, P]),
    'cv_fdtd_run_tm': (I, [I, P, P, P, P, P, I64, P, I, I, P, P, I, P, I, I, P]),
    'cv_fdtd_tm_configs': (I, [P, P, P, I]),
    'cv_fdtd_run_fused': (I, [I, P, P, P, P, P, I, I64, P, I, I, P, P, I, P, I, P]),
    'cv_segsum': (I, [I64, P, P, P, P, P, I, I, D, D, P, I, P, P]),
    'cv_csr_solve_workspace_bytes': (I64, [I64]),
    'cv_csr_solve': (I, [I64, I, P, P, P, P, P, D, I, I, P, I64, P, P]),
    'cv_nd_plan_create': (P, [I, I]),
    'cv_nd_plan_destroy': (None, [P]),
    'cv_nd_plan_bytes': (I64, [P]),
    'cv_nd_factor_bytes': (I64, [P]),
    'cv_nd_plan_flops': (D, [P]),
    'cv_nd_plan_stats': (I, [P, P, P, P, P, I]),
    'cv_nd_plan_upload': (I, [P, P, I64, P]),
    'cv_nd_factor': (I, [P, P, P, P, P, P, P, P, I64, P]),
    'cv_nd_apply': (I, [P, P, P, P, P, I, P]),
    'cv_fdfd_fgmres_nd': (I, [P, P, P, P, P, P, P, P, P, P, P, P, P, I, D, I, I, P, I64, P, P]),
    'cv_dense_lu': (I, [I, P, P, P, P]),
    'cv_dense_lu_solve': (I, [I, P, P, P, I, I, P]),
    'cv_cnn_pack_weights': (I, [P, P, I, I, P]),
    'cv_conv3x3_fwd_bf16': (I, [P, P, P, P, I, I, I, I, P]),
    'cv_conv3x3_bwd_data_bf16': (I, [P, P, P, P, I, I, I, I, P]),
    'cv_conv3x3_bwd_weight_workspace_bytes': (I64, [I, I, I]),
    'cv_conv3x3_bwd_weight_bf16': (I, [P, P, P, P, P, I, I, I, I, P, I64, P]),
    'cv_cnn_down2': (I, [P, P, I, I, I, F, P]),
    'cv_cnn_up2': (I, [P, P, P, I, I, I, F, P]),
    'cv_cnn_pack_input': (I, [I, P, P, P, F, P, P, I, P, I, I, I, P]),
    'cv_cnn_unpack': (I, [I, P, P, F, P, I, I, I, P]),
    'cv_cnn_loss': (I, [P, P, P, P, P, P, F, P, F, I, I, I, P]),
    'cv_fdfd_batch_apply': (I, [P, P, I, P, P, P, P, P, P, P, P, P, D, I, I, P]),
    'cv_unet_workspace_bytes': (I64, [I, I, I, I]),
    'cv_unet_apply': (I, [P, I, P, F, F, P, I, I, P, I64, P]),
    'cv_fdfd_fgmres_workspace_bytes': (I64, [I, I, I, I, I]),
    'cv_fdfd_fgmres': (I, [P, P, P, P, P, P, P, P, P, P, D, I, I, D, I, P, F, P, I64, P, P]),
}


def lib():
    """The loaded library (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError('ceviche_amd HIP library not built: %s is missing '
                              '(run __graft_entry__.build() or python -m ceviche_amd._build)' % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return list(_SIGS)


def status_name(rc):
    return _STATUS_NAMES.get(rc, 'status %d' % rc)


def check(rc, what=''):
    if rc != CV_OK:
        msg = lib().cv_last_error().decode(errors='replace')
        if rc == CV_ERR_NOT_CONVERGED:
            raise NotConvergedError(rc, (what + ': ' + msg) if what else msg)
        raise CevicheError(rc, (what + ': ' + msg) if what else msg)


def device():
    """The GPU every product path runs on; raises when none is present."""
    if not torch.cuda.is_available():
        raise RuntimeError('ceviche_amd needs a ROCm GPU (torch.cuda.is_available() is False); '
                           'there is no CPU fallback')
    return torch.device('cuda', torch.cuda.current_device())


def stream_ptr():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t):
    """Device pointer of a tensor (None -> NULL)."""
    if t is None:
        return ctypes.c_void_p(0)
    return ctypes.c_void_p(t.data_ptr())
