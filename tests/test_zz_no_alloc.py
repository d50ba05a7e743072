"""The allocation-free BiCGSTAB entry point (cv_fdfd_solve_noalloc) against
cv_fdfd_solve, both called through the C-ABI on the same operator batches.
The two run the same kernels in the same order (the one directly, the other
through a captured graph), so the fields and the iteration counts agree bit
for bit; two calls in a row, the second on a smaller batch, check that the
shared scratch array carries nothing over."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import fdfd_ref as R

pytestmark = pytest.mark.gpu


def _batch(shapes, seed, dtype):
    from ceviche_amd._fdfd_op import FdfdOperator
    omega = 2 * np.pi * R.C_0 / 1.55e-6
    specs = [dict(omega=omega * (1 + 0.02 * k), dL=50e-9, shape=s, npml=(10, 10)) for k, s in enumerate(shapes)]
    op = FdfdOperator('ez', specs, dtype=dtype)
    rng = np.random.default_rng(seed)
    eps = np.concatenate([rng.uniform(1, 12.25, s).ravel() for s in shapes])
    src = np.zeros(op.total_cells, dtype=np.complex128)
    off = 0
    for (nx, ny) in shapes:
        src[off + (nx // 4) * ny: off + (nx // 4) * ny + ny] = 1.0
        off += nx * ny
    m0, m1 = op.materials(torch.from_numpy(eps))
    b = torch.from_numpy(src).to(op.device).to(dtype) * op.scale_vector(dtype)
    return op, m0, m1, b.contiguous()


def _noalloc(op, m0, m1, b, transpose, tol, max_iter=200000, check_every=32):
    from ceviche_amd._lib import lib, ptr, stream_ptr, SolveInfo, check
    x = torch.zeros_like(b)
    info = (SolveInfo * op.n_inst)()
    rc = lib().cv_fdfd_solve_noalloc(op.code, ptr(op.desc), op.desc_host.ctypes.data_as(ctypes.c_void_p),
                                     op.n_inst, ptr(op.tiles), op.n_tiles, ptr(op.prof), ptr(m0), ptr(m1), ptr(b),
                                     ptr(x), int(transpose), float(tol), int(max_iter), int(check_every), info,
                                     stream_ptr())
    check(rc, 'fdfd solve noalloc')
    return x, [(i.iterations, i.status, i.rel_residual) for i in info]


@pytest.mark.parametrize('dtype,tol', [(torch.complex128, 1e-10), (torch.complex64, 1e-5)])
def test_noalloc_solve_matches_workspace_solve(dtype, tol):
    from ceviche_amd._fdfd_op import SolverOptions
    from ceviche_amd._lib import lib
    assert lib().cv_noalloc_init() == 0
    # first call: three instances (with a transposed solve), then one smaller instance
    for shapes, transpose in ([(96, 80), (64, 64), (50, 70)], False), ([(40, 36)], True):
        print('noalloc', dtype, shapes, 'transpose', transpose, flush=True)
        op, m0, m1, b = _batch(shapes, seed=len(shapes), dtype=dtype)
        x_ref = op.solve_scaled(b, m0, m1, transpose=transpose,
                                opts=SolverOptions(tol=tol, method='bicgstab', check_every=32))
        ref_info = [(i['iterations'], i['status'], i['rel_residual']) for i in op.last_info]
        x, info = _noalloc(op, m0, m1, b, transpose, tol)
        torch.cuda.synchronize()
        assert info == ref_info, (info, ref_info)
        assert all(st == 0 and r <= tol for _, st, r in info)
        assert torch.equal(x, x_ref)


def test_noalloc_rejects_oversized_batch():
    from ceviche_amd._lib import CevicheError, CV_ERR_INVALID
    op, m0, m1, b = _batch([(600, 600)], seed=0, dtype=torch.complex128)
    with pytest.raises(CevicheError) as e:
        _noalloc(op, m0, m1, b, False, 1e-6)
    assert e.value.status == CV_ERR_INVALID
    assert 'allocation-free limit' in str(e.value)
