/*
 * ceviche_amd C-ABI: MI355X (gfx950) kernels for ceviche's simulation hot path.
 *
 * Every entry point takes caller-owned DEVICE pointers (plain void* + sizes),
 * an explicit hipStream_t (passed as void*), and returns an int status
 * (CV_OK on success).  Data buffers are never allocated by the library: all
 * workspace is passed in by the caller, sized by the *_workspace_bytes
 * helpers.  The iterative solvers do ask the HIP runtime for resources of
 * their own on each call: cv_fdfd_solve captures its loop body on a private
 * stream and instantiates a HIP graph (both released before it returns), the
 * FGMRES entry points keep host-side vectors, and the nested-dissection
 * factorisation calls rocSOLVER / rocBLAS through one rocBLAS handle per
 * (device, stream), created on first use and kept for the process (at most
 * 256 pairs; CV_ERR_INVALID beyond), whose device workspace the libraries
 * manage.
 * cv_fdfd_solve_noalloc is the variant that allocates nothing.  No torch
 * types cross this boundary.
 *
 * Layouts
 *   FDFD fields: complex, C order over (nx, ny) (y contiguous), packed over a
 *   batch of instances at desc[k].cell_off.  Complex values are interleaved
 *   (re, im) pairs of float (CV_C64) or double (CV_C128).
 *   FDTD fields: real float/double, C order over (nx, ny, nz).
 *
 * The reference interface each entry point replaces is cited (file:line in
 * fancompute/ceviche 0.1.3).
 */
#ifndef CEVICHE_AMD_H
#define CEVICHE_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status codes */
#define CV_OK                 0
#define CV_ERR_INVALID        1   /* bad argument / shape                      */
#define CV_ERR_HIP            2   /* a HIP runtime call failed                 */
#define CV_ERR_NOT_CONVERGED  3   /* Krylov hit max_iter above tolerance       */
#define CV_ERR_NONFINITE      4   /* NaN/Inf appeared in a reduction           */
#define CV_ERR_BREAKDOWN      5   /* Krylov breakdown that restarts could not fix */
#define CV_ERR_SINGULAR       6   /* zero pivot in the dense LU                */

/* ------------------------------------------------------------------ dtypes */
#define CV_C64   0   /* complex64 vectors, float32 materials  */
#define CV_C128  1   /* complex128 vectors, float64 materials */
#define CV_F32   0   /* FDTD float32 */
#define CV_F64   1   /* FDTD float64 */

/* ---------------------------------------------------------- FDFD operators */
#define CV_POL_EZ 0
#define CV_POL_HZ 1
#define CV_FLAG_X_IDENTITY 1   /* Nx == 1: Dx* is the identity (derivatives.py:76-77) */
#define CV_FLAG_Y_IDENTITY 2   /* Ny == 1: Dy* is the identity (derivatives.py:78-79) */
#define CV_FLAG_CPLX_MAT   4   /* Hz with complex eps: m0, m1 (and the eps, material
                                  tangent and gradient arrays of the Hz material
                                  entry points) hold interleaved complex values
                                  (fdfd.py:200-215 on complex eps); not accepted by
                                  the direct factorisations */

/* One FDFD instance of a batch.  Plain data, 128 bytes, device-resident array.
 *   pol == CV_POL_EZ:  A = scale*(Dxf Dxb + Dyf Dyb) + diag(dshift*(m0 + i m1))
 *                      scale = -1/mu0, dshift = -eps0 w^2, m0 = Re eps,
 *                      m1 = Im eps or NULL for real eps           (fdfd.py:172-185)
 *   pol == CV_POL_HZ:  A = scale*(Dxb diag(m1) Dxf + Dyb diag(m0) Dyf) + dshift*I
 *                      scale = 1/eps0, dshift = mu0 w^2,
 *                      m0 = 1/(eps_xx+1e-5), m1 = 1/(eps_yy+1e-5)      (fdfd.py:211-233)
 * Profiles at prof_off: kxf[nx], kxb[nx], kyf[ny], kyb[ny] (complex, operator
 * precision) holding 1/(s*dL), or 1/s on an identity axis (derivatives.py:34-59).
 * Materials at cell_off in m0 (and m1 for Hz), real, operator precision.
 *
 * Frequency-coupled systems (the multi-frequency Ez operator, fdfd.py:284-324):
 * the nblk blocks of one linear system are consecutive instances (blk = 0 ..
 * nblk-1) with equal grids and consecutive cells; each has its own dshift
 * (-eps0 W_blk^2) and shares the Krylov state and reductions of its block 0
 * (tile_first / tile_count then span all blocks' tiles).  Row block m adds
 *   sum_{h < n_cpl} w2[m] * ( gp_h .* x_{m+h+1} + gm_h .* x_{m-h-1} )
 * (terms whose block falls outside 0 .. nblk-1 dropped), with the complex
 * tables at prof[cpl_off]: gp_0[n], gm_0[n], gp_1[n], gm_1[n], ... (n = nx*ny)
 * and w2 at prof[w2_off + m] (real part used).  An ordinary instance has
 * n_cpl = 0, blk = 0, nblk = 1. */
typedef struct cv_fdfd_desc {
    int32_t nx, ny;
    int32_t pol, flags;
    int64_t cell_off;
    int64_t prof_off;
    int32_t tile_first, tile_count;  /* tile range of this instance's linear system */
    double  scale, dshift;
    double  phx_re, phx_im, phy_re, phy_im;   /* Bloch phasors exp(i*theta)       */
    double  omega;
    int32_t n_cpl, blk, nblk, pad;
    int64_t cpl_off, w2_off;
} cv_fdfd_desc;

/* Tile table entry: one workgroup processes CV_TILE_CELLS consecutive cells of
 * one instance starting at cell `first` (instance-local index). */
typedef struct cv_tile { int32_t inst, first; } cv_tile;
#define CV_TILE_CELLS 256

/* y = A x (transpose == 0) or y = A^T x (transpose != 0) for every instance.
 * Replaces the COO assembly + CSR SpMV of fdfd.py:172-185 / :211-233 and
 * primitives.py:58-72 on the FDFD operator. */
int cv_fdfd_apply(int dtype, const cv_fdfd_desc* desc, const cv_tile* tiles, int n_tiles,
                  const void* prof, const void* m0, const void* m1,
                  const void* x, void* y, int transpose, void* stream);

/* Hz materials: m0 = 1/((eps + roll(eps,1,axis=1))/2 + 1e-5),
 *               m1 = 1/((eps + roll(eps,1,axis=0))/2 + 1e-5).   fdfd.py:200-215 */
int cv_fdfd_hz_materials(int dtype, const cv_fdfd_desc* desc, const cv_tile* tiles, int n_tiles,
                         const void* eps, void* m0, void* m1, void* stream);

/* One PML-scaled Yee difference operator D in {Dxf, Dxb, Dyf, Dyb}
 * (which = 0, 1, 2, 3; derivatives.py:34-124), optionally with a real
 * per-cell material `mat` (may be NULL):
 *   mode 0:  out = coef * mat .* (D u)
 *   mode 1:  out = coef * D^T (mat .* u)
 *   mode 2:  out = conj(coef) * D^H (mat .* u)
 * accumulate != 0 adds into out.  Field recovery (fdfd.py:135-162) is two
 * mode-0 calls; its reverse mode uses mode 2.  Replaces the sp_mult of the
 * D matrices (fdfd.py:96-99 -> primitives.py:58-72). */
int cv_fdfd_deriv(int dtype, const cv_fdfd_desc* desc, const cv_tile* tiles, int n_tiles,
                  const void* prof, const void* mat, const void* u, void* out, int which,
                  int mode, double coef_re, double coef_im, int accumulate, void* stream);

/* Matrix-free gradient of the operator entries that depend on the materials:
 *   g0[k] += scale * Re( lam^T (dA/dm0[k]) x ),   g1[k] likewise for m1,
 * i.e. the gather adj[i]*x[j] of primitives.py:121-128 restricted to the
 * entries that depend on eps (the diagonal for Ez, fdfd.py:179; the
 * Dxb diag(m1) Dxf / Dyb diag(m0) Dyf products for Hz, fdfd.py:219-223).
 * For Ez, g1 (the Im eps part) may be NULL.  Real outputs in operator
 * precision, accumulated.  The caller may scale the whole system (scale and
 * dshift multiplied by one constant); the kernels use the descriptor as is. */
int cv_fdfd_material_grad(int dtype, const cv_fdfd_desc* desc, const cv_tile* tiles, int n_tiles,
                          const void* prof, const void* lam, const void* x,
                          void* g0, void* g1, double scale, void* stream);

/* Reverse mode of cv_fdfd_hz_materials: g_eps += d(m0, m1)/d eps ^T (g_m0, g_m1)
 * (g_m0 or g_m1 may be NULL).  fdfd.py:200-215. */
int cv_fdfd_hz_materials_grad(int dtype, const cv_fdfd_desc* desc, const cv_tile* tiles, int n_tiles,
                              const void* m0, const void* m1, const void* g_m0, const void* g_m1,
                              void* g_eps, void* stream);

/* Reverse mode of cv_fdfd_deriv (mode 0) with respect to its material:
 * g_mat[k] += Re( conj(coef (D u)[k]) g[k] ).  Used for the eps dependence
 * of the Hz field recovery (fdfd.py:150-157). */
int cv_fdfd_deriv_matgrad(int dtype, const cv_fdfd_desc* desc, const cv_tile* tiles, int n_tiles,
                          const void* prof, const void* u, const void* g, void* g_mat, int which,
                          double coef_re, double coef_im, void* stream);

/* Forward-mode (tangent) pieces of the operator (the defjvp rules of
 * primitives.py:137-148 specialised to the FDFD operator):
 *   y (+)= (dA/dm0 . dm0 + dA/dm1 . dm1) x
 *   Ez: dshift (dm0 + i dm1) .* x;  Hz: scale (Dxb (dm1 .* Dxf x) + Dyb (dm0 .* Dyf x)). */
int cv_fdfd_material_apply(int dtype, const cv_fdfd_desc* desc, const cv_tile* tiles, int n_tiles,
                           const void* prof, const void* dm0, const void* dm1, const void* x, void* y,
                           int accumulate, void* stream);

/* Tangent of cv_fdfd_hz_materials: dm0 = -m0^2 (de + roll(de,1,axis=1))/2,
 * dm1 = -m1^2 (de + roll(de,1,axis=0))/2. */
int cv_fdfd_hz_materials_jvp(int dtype, const cv_fdfd_desc* desc, const cv_tile* tiles, int n_tiles,
                             const void* m0, const void* m1, const void* d_eps, void* dm0, void* dm1,
                             void* stream);

/* Coupling part of frequency-coupled systems alone, with the tables read from
 * `tab` (same offsets as prof; e.g. a tangent table):
 *   y (+)= C x  (transpose == 0)  or  C^T x  (transpose != 0)
 * where C holds only the inter-block terms described at cv_fdfd_desc.  Used for
 * the forward-mode rule of the multi-frequency solve (the coupling slice of
 * the defjvp of primitives.py:137-148; entries from fdfd.py:303-324). */
int cv_fdfd_coupling_apply(int dtype, const cv_fdfd_desc* desc, const cv_tile* tiles, int n_tiles,
                           const void* tab, const void* x, void* y, int transpose, int accumulate,
                           void* stream);

/* Gradient of lam^T A x with respect to the coupling tables (complex, no
 * conjugation): for every system (its block-0 instance) and cell,
 *   g[cpl_off + 2h n + c]   += scale * sum_m w2[m] lam_m[c] x_{m+h+1}[c]
 *   g[cpl_off + (2h+1) n + c] += scale * sum_m w2[m] lam_m[c] x_{m-h-1}[c]
 * g has prof's layout.  The reverse-mode gather adj[i]*x[j] of
 * primitives.py:121-128 restricted to the coupling entries of fdfd.py:303-324. */
int cv_fdfd_coupling_grad(int dtype, const cv_fdfd_desc* desc, const cv_tile* tiles, int n_tiles,
                          const void* prof, const void* lam, const void* x, void* g, double scale,
                          void* stream);

/* ------------------------------------------------------------- Krylov solve */
#define CV_KRYLOV_BICGSTAB 0
#define CV_KRYLOV_COCG     1

/* Per-instance solver telemetry written back to a host array. */
typedef struct cv_solve_info {
    int32_t iterations;     /* Krylov iterations performed                       */
    int32_t status;         /* CV_OK / CV_ERR_NOT_CONVERGED / ...                */
    int32_t restarts;       /* breakdown restarts + refinement sweeps            */
    int32_t pad;
    double  rel_residual;   /* final ||b - A x|| / ||b||, recomputed in FP64      */
} cv_solve_info;

/* Device workspace in bytes for cv_fdfd_solve with `total_cells` packed cells,
 * `n_inst` instances and `n_tiles` tiles. */
int64_t cv_fdfd_solve_workspace_bytes(int dtype, int64_t total_cells, int n_inst, int n_tiles);

/* Solve A x = b (or A^T x = b) for every instance of the batch with BiCGSTAB
 * (fused stencil + axpy + FP64 dot kernels).  x holds the initial guess on
 * entry.  With dtype CV_C64 and refine != 0, the outer residual and x are kept
 * in complex128 (x64, b64 given as complex128 buffers) and each sweep solves a
 * complex64 correction (mixed-precision iterative refinement).
 * Replaces solvers.py:37-75 (spsolve / scipy Krylov) behind primitives.py:103-119.
 * A frequency-coupled system (nblk > 1) is one solve over all its blocks; the
 * info entry of every block reports that solve.
 * info: host array of n_inst entries. */
int cv_fdfd_solve(int dtype, const cv_fdfd_desc* desc, const cv_fdfd_desc* desc_host, int n_inst,
                  const cv_tile* tiles, int n_tiles,
                  const void* prof, const void* m0, const void* m1,
                  const void* b, void* x, int transpose,
                  double tol, int max_iter, int check_every,
                  void* workspace, int64_t workspace_bytes,
                  cv_solve_info* info, void* stream);

/* The same batched solve by Jacobi-preconditioned COCG (conjugate orthogonal
 * CG) on the complex-symmetric form W A of the operator, W = 1/(kxf kyf)
 * per cell for Ez and 1/(kxb kyb) for Hz (the forward / backward SC-PML
 * factors, derivatives.py:34-59): one stencil pass and two fused kernels per
 * iteration, FP64 reductions, convergence on the true FP64 residual of the
 * original system.  A^T solves run as W A u = b, x = W u.  Needs zero Bloch
 * phases and uncoupled instances (CV_ERR_INVALID otherwise).  Replaces the
 * same reference solve as cv_fdfd_solve (solvers.py:37-75), the Krylov
 * family being scipy's (solvers.py:20-30). */
int cv_fdfd_cocg(int dtype, const cv_fdfd_desc* desc, const cv_fdfd_desc* desc_host, int n_inst,
                 const cv_tile* tiles, int n_tiles,
                 const void* prof, const void* m0, const void* m1,
                 const void* b, void* x, int transpose,
                 double tol, int max_iter, int check_every,
                 void* workspace, int64_t workspace_bytes,
                 cv_solve_info* info, void* stream);

/* Allocation-free variant of cv_fdfd_solve.  No counterpart in the reference:
 * it is for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid allocation).  It calls no hipMalloc /
 * hipFree, creates no stream and instantiates no graph: the BiCGSTAB vectors,
 * solver states and reduction partials live in one file-scope device array of
 * CV_NOALLOC_SCRATCH_BYTES (one copy per GPU, held until the process ends),
 * and the loop kernels are launched directly on `stream`.  A batch whose
 * solver scratch (cv_fdfd_solve_workspace_bytes) exceeds that size, or with
 * more than CV_NOALLOC_MAX_INST instances, is rejected with CV_ERR_INVALID
 * (complex128: up to about 290 000 packed cells, e.g. one 512 x 512 grid).
 * Every call on a GPU shares that GPU's array: calls that overlap on one GPU
 * (two streams, two host threads) race and give wrong results; serialise
 * them.  The counters the kernels need at zero are cleared at the start of
 * every call; nothing else carries over between calls.
 * cv_noalloc_init looks the array's device address up on the current device
 * (done by the first call on each device otherwise). */
#define CV_NOALLOC_SCRATCH_BYTES (32ll << 20)
#define CV_NOALLOC_MAX_INST 64
int cv_noalloc_init(void);
int cv_fdfd_solve_noalloc(int dtype, const cv_fdfd_desc* desc, const cv_fdfd_desc* desc_host, int n_inst,
                          const cv_tile* tiles, int n_tiles,
                          const void* prof, const void* m0, const void* m1,
                          const void* b, void* x, int transpose,
                          double tol, int max_iter, int check_every,
                          cv_solve_info* info, void* stream);

/* Mixed precision: complex128 x/b/operator (prof64, m0_64, m1_64) for the
 * residual, complex64 copies (prof32, m0_32, m1_32) for the inner Krylov
 * solve, method CV_KRYLOV_BICGSTAB or CV_KRYLOV_COCG. */
int64_t cv_fdfd_solve_mixed_workspace_bytes(int64_t total_cells, int n_inst, int n_tiles);
int cv_fdfd_solve_mixed(const cv_fdfd_desc* desc, const cv_fdfd_desc* desc_host, int n_inst,
                        const cv_tile* tiles, int n_tiles,
                        const void* prof64, const void* m0_64, const void* m1_64,
                        const void* prof32, const void* m0_32, const void* m1_32,
                        const void* b64, void* x64, int transpose, int method,
                        double tol, double inner_tol, int max_iter, int max_sweeps, int check_every,
                        void* workspace, int64_t workspace_bytes,
                        cv_solve_info* info, void* stream);

/* -------------------------------------------------------------------- FDTD */
/* Grid and constants of one FDTD problem (fdtd.py:12-28, 213-222). */
typedef struct cv_fdtd_desc {
    int32_t nx, ny, nz;
    int32_t pad;
    double  dt, dL, c0, eps0;
} cv_fdtd_desc;

/* sig: 6 concatenated real profiles sigH_x[nx], sigH_y[ny], sigH_z[nz],
 *      sigD_x[nx], sigD_y[ny], sigD_z[nz]  (fdtd.py:224-263: the sigma
 *      tensors vary along their own axis only).
 * eps: relative permittivity [nx*ny*nz].
 * state: 18 arrays of nx*ny*nz (C order): H(3) D(3) ICE(3) IH(3) ICH(3) ID(3);
 *      E = D / eps_aa is formed on demand (cv_fdtd_efields).
 * All real arrays are float (CV_F32) or double (CV_F64). */

/* n_steps Yee steps (fdtd.py:74-144), each: H update, D update, D += jscale*J
 * (dense Jx/Jy/Jz, any may be NULL) + amp[t] at (src_comp, src_cell) when
 * src_comp >= 0 (amp: host array of n_steps), then records n_probes values
 * into rec[t*n_probes + p] (probe_comp: 0..2 E, 3..5 D, 6..8 H).
 * Replaces the per-step numpy passes of fdtd.py:77-137. */
int cv_fdtd_step(int dtype, const cv_fdtd_desc* d, const void* sig, const void* eps, void* state,
                 const void* Jx, const void* Jy, const void* Jz, double jscale, int src_comp,
                 int64_t src_cell, const double* amp, int n_steps, const int32_t* probe_comp,
                 const int64_t* probe_cell, int n_probes, void* rec, void* stream);

/* E[3][n] = D / eps_aa (fdtd.py:134-137, utils.py:153-176). */
/* Forward mode (tangent-linear) of cv_fdtd_step: advances the primal `state`
 * (dense sources Jx/Jy/Jz and point source amp as cv_fdtd_step) and its
 * tangent `tstate` (same layout) along the permittivity tangent d_eps, the
 * dense source tangents dJx/dJy/dJz (any may be NULL) and the amplitude
 * tangent d_amp (host array of n_steps, may be NULL); E_a = D_a / eps_aa gives
 * dE_a = dD_a / eps_aa - D_a d(eps_aa) / eps_aa^2, the rest of the update is
 * linear.  Records the primal probes into rec (may be NULL) and their
 * tangents into trec.  The forward-mode rule the reference gets from
 * autograd through fdtd.py:74-144 (jacobians.py:38-51). */
int cv_fdtd_step_tangent(int dtype, const cv_fdtd_desc* d, const void* sig, const void* eps, void* state,
                         void* tstate, const void* d_eps, const void* Jx, const void* Jy, const void* Jz,
                         const void* dJx, const void* dJy, const void* dJz, double jscale, int src_comp,
                         int64_t src_cell, const double* amp, const double* d_amp, int n_steps,
                         const int32_t* probe_comp, const int64_t* probe_cell, int n_probes, void* rec, void* trec,
                         void* stream);
/* dE = tangent of E = D / eps_aa along (tstate's D, d_eps), all cells (3 x n). */
int cv_fdtd_efields_tangent(int dtype, const cv_fdtd_desc* d, const void* sig, const void* eps, const void* state,
                            const void* tstate, const void* d_eps, void* dE, void* stream);

int cv_fdtd_efields(int dtype, const cv_fdtd_desc* d, const void* sig, const void* eps,
                    const void* state, void* E, void* stream);

/* Weighted probe sums over T field snapshots (ceviche/utils.py:316-332: the
 * sum(fields[component] * probe) of measure_fields, per probe and step):
 *   out[t*out_stride + p] (+)= sum_{k = rowptr[p]}^{rowptr[p+1]-1} w[k] * field[t*field_stride + col[k]]
 * for p < n_rows, t < T; the probe weights are a CSR matrix over the field's
 * cells (col < field_stride is the caller's contract).  field and out are
 * float (CV_F32) or double (CV_F64); accumulation is FP64.  Used once per step
 * by measure_fields and once per run for weighted probes recorded in-kernel. */
int cv_probe_weights(int dtype, const void* field, int64_t field_stride, int T, const int64_t* rowptr,
                     const int64_t* col, const double* w, int n_rows, void* out, int64_t out_stride,
                     int accumulate, void* stream);

/* One reverse-time step of the exact discrete adjoint (no reference
 * counterpart: ceviche differentiates FDTD through autograd's tape).
 * lam: 18-array adjoint state after step t (updated in place to before step t);
 * tmp: 3-array scratch; D_prev: D after step t-1; D_t: D after step t (for E
 * probes); g_mE (3 arrays, may be NULL) accumulates dL/d(1/eps_aa);
 * g_probe[n_probes]: cotangent of the step-t probe record, injected first.
 * g_src (device, one real, may be NULL): receives lam(D_src_comp, src_cell)
 * after the injection, the cotangent of a point-source amplitude added to D
 * at this step (fdtd.py:122-124). */
int cv_fdtd_adjoint_step(int dtype, const cv_fdtd_desc* d, const void* sig, const void* eps,
                         void* lam, void* tmp, const void* D_prev, void* g_mE,
                         const int32_t* probe_comp, const int64_t* probe_cell, int n_probes,
                         const void* g_probe, const void* D_t, int src_comp, int64_t src_cell,
                         void* g_src, void* stream);

/* Reverse of cv_fdtd_efields for a dense cotangent gE[3][n]:
 * gD[3][n] += gE / eps_aa;  g_mE[3][n] += D * gE (g_mE may be NULL). */
int cv_fdtd_efields_grad(int dtype, const cv_fdtd_desc* d, const void* sig, const void* eps,
                         const void* state, const void* gE, void* gD, void* g_mE, void* stream);

/* Fused, temporally blocked 2D TM runner (nz == 1, TE fields zero, Jz point
 * source): n_steps steps starting at global step t0, D_z[src_cell] += amp_dev[t]
 * (device float64 array indexed by global step), probes (comp 2 Ez, 6 Hx,
 * 7 Hy; at most 256) recorded into rec[(t - rec_t0)*n_probes + p].  config
 * selects the tile (rows per thread, steps per launch, waves per workgroup)
 * from cv_fdtd_tm_configs.  Launches alternate between `state` and the
 * caller's second buffer `state_alt` (same layout; its contents are
 * overwritten), and the result always ends in `state`.  Same arithmetic as
 * cv_fdtd_step. */
int cv_fdtd_run_tm(int dtype, const cv_fdtd_desc* d, const void* sig, const void* eps, void* state,
                   void* state_alt, int64_t src_cell, const double* amp_dev, int t0, int n_steps,
                   const int32_t* probe_comp, const int64_t* probe_cell, int n_probes, void* rec,
                   int rec_t0, int config, void* stream);

/* Fused single-pass Yee steps for any grid (3D or 2D, point source J[comp]
 * at src_cell with amplitude amp_dev[t] (device float64, global step t), or
 * comp = -1 for none): each step reads state_a and writes state_b (18-array
 * layout), then the roles swap; after an odd n_steps the result is in state_b.
 * Probes as cv_fdtd_step, rec row t - rec_t0.  Same arithmetic as
 * cv_fdtd_step; one kernel per step instead of two.  3D grids with nz a
 * multiple of 64 up to 256 use the slab-marching kernel (whole z lines per
 * workgroup, bands of rows marched along x), other 3D grids the x-marching
 * tile kernel.  Replaces the per-step loop of ceviche/fdtd.py:74-144 (the
 * forward() calls of a run). */
int cv_fdtd_run_fused(int dtype, const cv_fdtd_desc* d, const void* sig, const void* eps, void* state_a,
                      void* state_b, int comp, int64_t src_cell, const double* amp_dev, int t0,
                      int n_steps, const int32_t* probe_comp, const int64_t* probe_cell, int n_probes,
                      void* rec, int rec_t0, void* stream);

/* Launch-shape override of the slab-marching kernel of cv_fdtd_run_fused,
 * process-wide, for tuning sweeps: band_rows 1, 2 or 4 (0 = automatic: 2
 * when ny is even, else 1), slices per chunk (0 = 8), target workgroup
 * count that sets the chunk instead (0 = unused), map_bands 1 gives each
 * XCD a range of bands of every chunk, 0 a range of whole chunks.  No
 * counterpart in the reference (a tuning knob of this kernel). */
int cv_fdtd_slab_tuning(int band_rows, int chunk, int target_workgroups, int map_bands);

/* Tile configurations of cv_fdtd_run_tm: fills rpt[i], ts[i], waves[i]
 * (rows per thread, steps per launch, waves per workgroup; the region is
 * waves*rpt rows by 64 columns) for i < max_n and returns their number. */
int cv_fdtd_tm_configs(int* rpt, int* ts, int* waves, int max_n);

/* g_eps += chain rule of mE_a = 2/(eps + roll(eps, 1, a)) applied to g_mE. */
int cv_fdtd_eps_grad(int dtype, const cv_fdtd_desc* d, const void* sig, const void* eps,
                     const void* g_mE, void* g_eps, void* stream);

/* ---------------------------------------------------- generic sparse (COO) */
/* Segmented product sum over complex128 vectors:
 *   out[o(s)] (+)= alpha * sum_{q = ptr[s]}^{ptr[s+1]-1} opu(u[iu[q]]) * opv(v[iv[q]])
 * o(s) = oidx[s] (or s when oidx is NULL), iu / iv NULL mean the identity,
 * opu / opv conjugate when conj_u / conj_v.  Deterministic (storage order).
 * With host-built groupings of the COO entries this is the CSR SpMV of
 * sp_mult (primitives.py:58-72), its A^T / A^H products and the gather
 * v[i]*x[j] (:74-86), the SpGEMM numerics of spsp_mult (:153-176) and its
 * VJP/JVP rules (:178-257), and the duplicate summation of make_sparse
 * (utils.py:13-25). */
int cv_segsum(int64_t n_seg, const int64_t* ptr, const int64_t* iu, const int64_t* iv,
              const void* u, const void* v, int conj_u, int conj_v, double alpha_re, double alpha_im,
              const int64_t* oidx, int accumulate, void* out, void* stream);

/* Solves A x = b for a general sparse complex128 matrix in CSR form
 * (indptr[n+1], indices, data: the arrays of a scipy.sparse.csr_matrix, int32
 * or int64 as index_bits says; duplicates in a row are summed) by
 * Jacobi-preconditioned BiCGSTAB with FP64 reductions, to
 * ||b - A x|| <= tol ||b|| checked on a freshly computed residual.  x holds
 * the initial guess.  Replaces spsolve (solvers.py:50-62) and the Krylov
 * branch (solvers.py:64-75) behind solve_linear, the empty hook
 * _solve_cuda(A, b) (solvers.py:77-79), and the factorisation of sp_solve
 * (primitives.py:103-148) for matrices above the dense-LU size.  info: one
 * host entry.  CV_ERR_NOT_CONVERGED when max_iter is reached. */
int64_t cv_csr_solve_workspace_bytes(int64_t n);
int cv_csr_solve(int64_t n, int index_bits, const void* indptr, const void* indices, const void* data,
                 const void* b, void* x, double tol, int max_iter, int check_every, void* workspace,
                 int64_t workspace_bytes, cv_solve_info* info, void* stream);

/* In-place LU with partial pivoting of a column-major n x n complex128
 * matrix (P A = L U, unit L), pivot rows in piv[n] (device int32); *info_dev
 * (device int32) is set to k+1 for the first exactly zero pivot, else 0.
 * Direct solve for small general systems: spl.spsolve at solvers.py:62. */
int cv_dense_lu(int n, void* A, int32_t* piv, int32_t* info_dev, void* stream);

/* Solves A X = B (trans 0), A^T X = B (1) or A^H X = B (2) in place with the
 * cv_dense_lu factors; B is n x nrhs column-major.  The transposed solves
 * serve the reverse mode of sp_solve (primitives.py:121-135) without a
 * second factorisation. */
int cv_dense_lu_solve(int n, const void* LU, const int32_t* piv, void* B, int nrhs, int trans, void* stream);

/* ------------------------------------------- convolutional preconditioner
 * No counterpart in the reference: ceviche factorises the sparse system
 * (solvers.py:37-62); these kernels implement the learned preconditioner used
 * by cv_fdfd_fgmres.  Activations are NHWC bf16 with 32 channels; weights are
 * fp32 masters W[layer][tap][ci][co] (tap = 3 dy + dx) packed to bf16 matrix
 * core operands by cv_cnn_pack_weights (flip = 1 packs the mirrored,
 * channel-swapped layer whose forward pass is the data gradient). */
#define CV_UNET_MAX_LEVELS 10
#define CV_CNN_CHANNELS 32
#define CV_CNN_PACKED_ELEMS 9216   /* bf16 elements of one packed layer */

typedef struct {
    const void* wpack;   /* (2 levels + 1) packed layers, forward orientation */
    const float* bias;   /* (2 levels + 1) x 32 */
    const float* feat;   /* (H*W) x nf per-cell features (channels 2 .. 2+nf-1) */
    int nf;
    int levels;
} cv_unet;

int cv_cnn_pack_weights(const float* w, void* wpack, int flip, int n_layers, void* stream);
/* out = act(conv3x3(in) + bias), zero padding; relu != 0 applies max(., 0) */
int cv_conv3x3_fwd_bf16(const void* in, const void* wpack, const float* bias, void* out, int B, int H, int W,
                        int relu, void* stream);
/* gx (+)= conv3x3^T(gy (.) [ymask > 0]) using the flipped packing */
int cv_conv3x3_bwd_data_bf16(const void* gy, const void* ymask, const void* wpack_flip, void* gx, int B, int H,
                             int W, int accumulate, void* stream);
int64_t cv_conv3x3_bwd_weight_workspace_bytes(int B, int H, int W);
/* dw[tap][ci][co] (+)= sum_p x[p + tap][ci] g[p][co], db (+)= sum_p g[p], g = gy (.) [ymask > 0] */
int cv_conv3x3_bwd_weight_bf16(const void* x, const void* gy, const void* ymask, float* dw, float* db, int B,
                               int H, int W, int accumulate, void* workspace, int64_t workspace_bytes,
                               void* stream);
/* coarse = scale * sum of each 2x2 block (ceil sizes); fine = scale * up(coarse) + add */
int cv_cnn_down2(const void* in, void* out, int B, int H, int W, float scale, void* stream);
int cv_cnn_up2(const void* in, const void* add, void* out, int B, int H, int W, float scale, void* stream);
int cv_cnn_pack_input(int dtype, const void* r, const int* r_idx, const float* rscale, float rs, const float* feat,
                      const int* f_idx, int nf, void* out, int B, int H, int W, void* stream);
int cv_cnn_unpack(int dtype, const void* net, const float* oscale, float os, void* z, int B, int H, int W,
                  void* stream);
/* loss += gscale * inv_norm * |y[0:2] - e|^2; g = d loss / d y (bf16, 32
 * channels), plus extra_scale * (Re, Im) of the complex64 `extra` (may be NULL)
 * added to channels 0, 1 */
int cv_cnn_loss(const void* net, const void* e, const int* e_idx, const float* inv_norm, void* g, double* loss,
                float gscale, const void* extra, float extra_scale, int B, int H, int W, void* stream);
/* Training-batch operator (complex64): sample b (at x + b n) uses instance
 * inst[b] of a multi-instance operator whose instances all have n cells.
 * herm = 0: y_b = A x_b - r[r_idx[b]] (r may be NULL), loss += lscale |y_b|^2;
 * herm = 1: y_b = A^H x_b.  (loss may be NULL.) */
int cv_fdfd_batch_apply(const cv_fdfd_desc* desc, const cv_fdfd_desc* desc_host, int n_inst, const void* prof,
                        const void* m0, const void* m1, const int* inst, const void* x, const void* r,
                        const int* r_idx, void* y, double* loss, double lscale, int herm, int B, void* stream);
int64_t cv_unet_workspace_bytes(int levels, int B, int H, int W);
/* z = os * net(rs * r) for one H x W complex field (dtype CV_C64 / CV_C128) */
int cv_unet_apply(const cv_unet* net, int dtype, const void* r, float rs, float os, void* z, int H, int W,
                  void* workspace, int64_t workspace_bytes, void* stream);

/* ----------------------------------------------------- flexible GMRES
 * Replaces the sparse solve of sp_solve for one FDFD instance
 * (ceviche/primitives.py:103-119 -> solvers.py:37-62), with a preconditioner
 * that may change between applications.  x (complex128) is the initial guess
 * on entry; the FP64 true residual decides convergence (tol relative to |b|).
 * `iterations` counts Arnoldi steps, `restarts` restart cycles; with the
 * network preconditioner, `pad` = 1 reports that three network cycles in a
 * row reduced the true residual by less than 1 % and the solve finished with
 * Jacobi. */
#define CV_PRECOND_NONE 0
#define CV_PRECOND_JACOBI 1
#define CV_PRECOND_UNET 2
int64_t cv_fdfd_fgmres_workspace_bytes(int nx, int ny, int restart, int precond, int levels);
int cv_fdfd_fgmres(const cv_fdfd_desc* desc, const cv_fdfd_desc* desc_host, const void* prof64, const void* m0_64,
                   const void* m1_64, const void* prof32, const void* m0_32, const void* m1_32, const void* b,
                   void* x, double tol, int restart, int max_iter, double inner_rtol, int precond,
                   const cv_unet* net, float net_gain, void* workspace, int64_t workspace_bytes,
                   cv_solve_info* info, void* stream);

/* Block-tridiagonal direct factorisation of one 2D FDFD instance with
 * ny <= 128 (the reference's SuperLU solve, ceviche/solvers.py:50-62, done as
 * block Gaussian elimination along x): A without its periodic x-wrap entries
 * is block tridiagonal with diagonal off-diagonal blocks; the dense inverses
 * of the Schur complements S_i = D_i - diag(lo_i) S_{i-1}^-1 diag(up_{i-1})
 * are stored in complex64 (nx * ny^2 entries, cv_fdfd_btd_factor_bytes).
 * Partial pivoting inside each S_i, none across blocks.  Returns
 * CV_ERR_SINGULAR for a zero pivot. */
int64_t cv_fdfd_btd_factor_bytes(int nx, int ny);
int cv_fdfd_btd_factor(const cv_fdfd_desc* desc, const cv_fdfd_desc* desc_host, const void* prof32,
                       const void* m0_32, const void* m1_32, void* factor, int64_t factor_bytes, void* stream);
/* z = M^-1 v (transpose == 0) or M^-T v with the stored factorisation (v, z
 * complex64, not aliased): one forward and one backward sweep. */
int cv_fdfd_btd_apply(const cv_fdfd_desc* desc, const cv_fdfd_desc* desc_host, const void* prof32,
                      const void* m0_32, const void* m1_32, const void* factor, const void* v, void* z,
                      int transpose, void* stream);
/* The same block-tridiagonal matrix factorised by block cyclic reduction:
 * ceil(log2 nx) levels, each inverting its odd blocks in parallel (one
 * workgroup per block) and forming the next level with batched complex64
 * products; a solve is one reduction sweep and one back-substitution sweep of
 * batched matrix-vector products per level (A or A^T).  The factor buffer
 * holds the plan, the blocks and the solve vectors (one solve at a time). */
int64_t cv_fdfd_bcr_factor_bytes(int nx, int ny);
int cv_fdfd_bcr_factor(const cv_fdfd_desc* desc, const cv_fdfd_desc* desc_host, const void* prof32,
                       const void* m0_32, const void* m1_32, void* factor, int64_t factor_bytes, void* stream);
int cv_fdfd_bcr_apply(const cv_fdfd_desc* desc, const cv_fdfd_desc* desc_host, const void* factor, const void* v,
                      void* z, int transpose, void* stream);
/* nb instances of one shape factorised together (instance q: desc[q],
 * desc_host[q], prof32[q], m0_32[q], m1_32[q] -> factors[q], each of
 * factor_bytes): each level's inversions and products of all instances run
 * in the same launches.  Used for the sideband blocks of a frequency-coupled
 * system, which the reference factorises as one matrix
 * (ceviche/fdfd.py:320-329 -> solvers.py:50-62). */
int cv_fdfd_bcr_factor_batch(int nb, const cv_fdfd_desc* const* desc, const cv_fdfd_desc* desc_host,
                             const void* const* prof32, const void* const* m0_32, const void* const* m1_32,
                             void* const* factors, int64_t factor_bytes, void* stream);

/* Solve A x = b (transpose == 0) or A^T x = b to the FP64 relative residual
 * tol by flexible GMRES preconditioned with a factorisation of kind
 * CV_FACTOR_BTD (cv_fdfd_btd_factor) or CV_FACTOR_BCR (cv_fdfd_bcr_factor)
 * (complex64 Arnoldi vectors, complex128 x and residual).  The dropped x-wrap
 * and the complex64 rounding are corrected by the outer iteration.
 * Workspace: cv_fdfd_fgmres_workspace_bytes(nx, ny, restart, CV_PRECOND_NONE, 0). */
/* A frequency-coupled system (desc / desc_host: its nblk consecutive blocks,
 * prof with the coupling tables) solved by the same FP64-refined FGMRES,
 * preconditioned block-diagonally: block m with the block-cyclic-reduction
 * factor factors[m] of that block alone (block_desc_host[m]: the block as a
 * single instance, ny <= 128); the inter-block coupling of fdfd.py:303-324 is
 * left to the outer iteration.  info: nblk entries (all the system's). */
int cv_fdfd_fgmres_blockdiag(const cv_fdfd_desc* desc, const cv_fdfd_desc* desc_host, const void* prof64,
                             const void* m0_64, const void* m1_64, const void* prof32, const void* m0_32,
                             const void* m1_32, const cv_fdfd_desc* block_desc_host, const void* const* factors,
                             const void* b, void* x, int transpose, double tol, int restart, int max_iter,
                             void* workspace, int64_t workspace_bytes, cv_solve_info* info, void* stream);

#define CV_FACTOR_BTD 1
#define CV_FACTOR_BCR 2
int cv_fdfd_fgmres_direct(const cv_fdfd_desc* desc, const cv_fdfd_desc* desc_host, const void* prof64,
                          const void* m0_64, const void* m1_64, const void* prof32, const void* m0_32,
                          const void* m1_32, int factor_kind, const void* factor, const void* b, void* x,
                          int transpose, double tol, int restart, int max_iter, void* workspace,
                          int64_t workspace_bytes, cv_solve_info* info, void* stream);

/* ------------------------------------------------------------------ misc */
/* ------------------------------------- nested-dissection direct factor
 * Direct factorisation of one FDFD instance of any shape (complex64, the
 * periodic wrap dropped), the preconditioner of FP64-refined FGMRES for grids
 * wider than the block-tridiagonal path takes: recursive bisection into
 * leaf boxes of <= 32 cells, each tree node eliminating the cells its
 * children share from a dense frontal matrix, nodes of equal height in
 * strided batches: fronts of <= 64 eliminated cells by one Gauss-Jordan
 * sweep with partial pivoting per node (whole front up to 112 rows, else the
 * pivot panel and two Schur-complement GEMMs), larger ones up to 2048 cells
 * by a blocked Gauss-Jordan on [F_II | I F_IB] with 16-column LU panels
 * (partial pivoting) and the same complement products.  Replaces
 * the sparse direct factorisation of solvers.py:50-62 (spsolve / pardiso)
 * behind primitives.py:103-119 for those grids.
 *   cv_nd_plan_create: symbolic analysis (host) of an nx x ny grid, NULL on
 *     error (cv_last_error); cv_nd_plan_destroy frees it.
 *   cv_nd_plan_upload: the plan's index arrays into a device buffer of
 *     cv_nd_plan_bytes (once per plan; blocks until the copies are done).
 *   cv_nd_factor: numeric factorisation of the complex64 operator (prof32,
 *     m0_32, m1_32 as for cv_fdfd_bcr_factor) into a device buffer of
 *     cv_nd_factor_bytes; CV_ERR_SINGULAR for a zero pivot.
 *   cv_nd_apply: z = M^-1 v (transpose 0) or M^-T v (1), complex64, v and z
 *     may alias.
 *   cv_fdfd_fgmres_nd: cv_fdfd_fgmres_direct with this factor.
 *   cv_nd_plan_stats: per level (height) the node count and the padded
 *     interior / boundary sizes. */
typedef struct cv_nd_plan cv_nd_plan;
cv_nd_plan* cv_nd_plan_create(int nx, int ny);
void cv_nd_plan_destroy(cv_nd_plan* plan);
int64_t cv_nd_plan_bytes(const cv_nd_plan* plan);
int64_t cv_nd_factor_bytes(const cv_nd_plan* plan);
double cv_nd_plan_flops(const cv_nd_plan* plan);   /* real flops of one cv_nd_factor (unpadded) */
int cv_nd_plan_stats(const cv_nd_plan* plan, int* levels, int* nb, int* kmax, int* bmax, int max_levels);
int cv_nd_plan_upload(const cv_nd_plan* plan, void* plan_dev, int64_t plan_bytes, void* stream);
int cv_nd_factor(const cv_nd_plan* plan, const void* plan_dev, const cv_fdfd_desc* desc,
                 const cv_fdfd_desc* desc_host, const void* prof32, const void* m0_32, const void* m1_32,
                 void* factor, int64_t factor_bytes, void* stream);
int cv_nd_apply(const cv_nd_plan* plan, const void* plan_dev, const void* factor, const void* v, void* z,
                int transpose, void* stream);
int cv_fdfd_fgmres_nd(const cv_fdfd_desc* desc, const cv_fdfd_desc* desc_host, const void* prof64,
                      const void* m0_64, const void* m1_64, const void* prof32, const void* m0_32,
                      const void* m1_32, const cv_nd_plan* plan, const void* plan_dev, const void* factor,
                      const void* b, void* x, int transpose, double tol, int restart, int max_iter,
                      void* workspace, int64_t workspace_bytes, cv_solve_info* info, void* stream);

const char* cv_version(void);
const char* cv_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* CEVICHE_AMD_H */
