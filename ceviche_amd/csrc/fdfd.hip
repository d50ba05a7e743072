// FDFD operator kernels: stencil apply (A, A^T), Hz material averaging,
// single derivative operators (field recovery and its adjoints) and the
// matrix-free material gradient.  See fdfd_stencil.h for the stencil.
#include "fdfd_stencil.h"

namespace cv {

constexpr int kCellsPerThread = CV_TILE_CELLS / kBlock;

// ---------------------------------------------------------------- y = A x
template <typename R, bool TRANS>
__global__ void __launch_bounds__(kBlock)
k_apply(const cv_fdfd_desc* __restrict__ desc, const cv_tile* __restrict__ tiles,
        const cx<R>* __restrict__ prof, const R* __restrict__ m0, const R* __restrict__ m1,
        const cx<R>* __restrict__ x, cx<R>* __restrict__ y) {
    const cv_tile t = tiles[blockIdx.x];
    const cv_fdfd_desc d = load_desc(desc, t.inst);
    const int n = d.nx * d.ny;
    auto fx = [&](int64_t k) { return x[k]; };
#pragma unroll
    for (int q = 0; q < kCellsPerThread; ++q) {
        const int loc = t.first + threadIdx.x + q * kBlock;
        if (loc >= n) break;
        const int i = loc / d.ny, j = loc - i * d.ny;
        const int64_t k = d.cell_off + loc;
        y[k] = apply_any<R, TRANS>(d, prof, m0, m1, i, j, fx, x[k]);
    }
}

// ------------------------------------------------------- Hz materials
template <typename R>
__global__ void __launch_bounds__(kBlock)
k_hz_materials(const cv_fdfd_desc* __restrict__ desc, const cv_tile* __restrict__ tiles,
               const R* __restrict__ eps, R* __restrict__ m0, R* __restrict__ m1) {
    const cv_tile t = tiles[blockIdx.x];
    const cv_fdfd_desc d = load_desc(desc, t.inst);
    const int n = d.nx * d.ny;
#pragma unroll
    for (int q = 0; q < kCellsPerThread; ++q) {
        const int loc = t.first + threadIdx.x + q * kBlock;
        if (loc >= n) break;
        const int i = loc / d.ny, j = loc - i * d.ny;
        const int im = (i == 0) ? d.nx - 1 : i - 1;
        const int jm = (j == 0) ? d.ny - 1 : j - 1;
        const int64_t k = d.cell_off + loc;
        const int64_t kjm = d.cell_off + (int64_t)i * d.ny + jm, kim = d.cell_off + (int64_t)im * d.ny + j;
        if (d.flags & CV_FLAG_CPLX_MAT) {   // complex eps (fdfd.py:200-215 on complex arrays)
            const cx<R>* ce = reinterpret_cast<const cx<R>*>(eps);
            const cx<R> exx = R(0.5) * (ce[k] + ce[kjm]) + mk<R>(R(1e-5), R(0));
            const cx<R> eyy = R(0.5) * (ce[k] + ce[kim]) + mk<R>(R(1e-5), R(0));
            reinterpret_cast<cx<R>*>(m0)[k] = crecip(exx);
            reinterpret_cast<cx<R>*>(m1)[k] = crecip(eyy);
            continue;
        }
        const R e = eps[k];
        // fdfd.py:203-204: eps_xx rolls along axis 1 (y), eps_yy along axis 0 (x)
        const R exx = R(0.5) * (e + eps[kjm]);
        const R eyy = R(0.5) * (e + eps[kim]);
        m0[k] = R(1) / (exx + R(1e-5));
        m1[k] = R(1) / (eyy + R(1e-5));
    }
}

template <typename R>
__global__ void __launch_bounds__(kBlock)
k_deriv(const cv_fdfd_desc* __restrict__ desc, const cv_tile* __restrict__ tiles,
        const cx<R>* __restrict__ prof, const R* __restrict__ mat, const cx<R>* __restrict__ u,
        cx<R>* __restrict__ out, int which, int mode, double cre, double cim, int accumulate) {
    const cv_tile t = tiles[blockIdx.x];
    const cv_fdfd_desc d = load_desc(desc, t.inst);
    const int n = d.nx * d.ny;
    cx<R> coef = mk<R>(R(cre), R(cim));
    if (mode == 2) coef = conj(coef);
#pragma unroll
    for (int q = 0; q < kCellsPerThread; ++q) {
        const int loc = t.first + threadIdx.x + q * kBlock;
        if (loc >= n) break;
        const int i = loc / d.ny, j = loc - i * d.ny;
        const int64_t k = d.cell_off + loc;
        cx<R> v = coef * deriv_at<R>(d, prof, u, mat, which, mode, i, j);
        if (accumulate) v = v + out[k];
        out[k] = v;
    }
}

// -------------------------------------------------- material gradients
// out_m0[k] += scale * Re( lam^T (dA/dm0[k]) x ),  out_m1 likewise.
//   Ez: dA/dm0[k] = dshift e_k e_k^T                   -> dshift lam_k x_k
//   Hz: dA/dm1[k] = s Dxb e_k e_k^T Dxf                -> s (Dxb^T lam)_k (Dxf x)_k
//       dA/dm0[k] = s Dyb e_k e_k^T Dyf                -> s (Dyb^T lam)_k (Dyf x)_k
template <typename R>
__global__ void __launch_bounds__(kBlock)
k_material_grad(const cv_fdfd_desc* __restrict__ desc, const cv_tile* __restrict__ tiles,
                const cx<R>* __restrict__ prof, const cx<R>* __restrict__ lam,
                const cx<R>* __restrict__ x, R* __restrict__ g0, R* __restrict__ g1, double scale) {
    const cv_tile t = tiles[blockIdx.x];
    const cv_fdfd_desc d = load_desc(desc, t.inst);
    const int n = d.nx * d.ny;
#pragma unroll
    for (int q = 0; q < kCellsPerThread; ++q) {
        const int loc = t.first + threadIdx.x + q * kBlock;
        if (loc >= n) break;
        const int i = loc / d.ny, j = loc - i * d.ny;
        const int64_t k = d.cell_off + loc;
        if (d.pol == CV_POL_EZ) {
            // diag = dshift (m0 + i m1): d/dm0 -> dshift, d/dm1 -> i dshift
            const cx<R> p = lam[k] * x[k];
            g0[k] += R(scale * d.dshift) * p.re;
            if (g1) g1[k] -= R(scale * d.dshift) * p.im;
        } else {
            const cx<R> lx = deriv_at<R>(d, prof, lam, (const R*)nullptr, 1, 1, i, j);  // Dxb^T lam
            const cx<R> dx = deriv_at<R>(d, prof, x, (const R*)nullptr, 0, 0, i, j);    // Dxf x
            const cx<R> ly = deriv_at<R>(d, prof, lam, (const R*)nullptr, 3, 1, i, j);  // Dyb^T lam
            const cx<R> dy = deriv_at<R>(d, prof, x, (const R*)nullptr, 2, 0, i, j);    // Dyf x
            const R s = R(scale * d.scale);
            if (d.flags & CV_FLAG_CPLX_MAT) {
                // complex materials (holomorphic in A): the torch-convention
                // cotangent is conj(s lam^T dA/dm x)
                cx<R>* c0 = reinterpret_cast<cx<R>*>(g0);
                cx<R>* c1 = reinterpret_cast<cx<R>*>(g1);
                c1[k] = c1[k] + conj(s * (lx * dx));
                c0[k] = c0[k] + conj(s * (ly * dy));
            } else {
                g1[k] += s * (lx * dx).re;
                g0[k] += s * (ly * dy).re;
            }
        }
    }
}


// ----------------------------------------- reverse of the Hz materials
// m0 = 1/(exx + 1e-5), exx[i,j] = (eps[i,j] + eps[i,j-1])/2 (fdfd.py:203, :214)
// m1 = 1/(eyy + 1e-5), eyy[i,j] = (eps[i,j] + eps[i-1,j])/2 (fdfd.py:204, :215)
// g_eps[i,j] += 1/2 (gxx[i,j] + gxx[i,j+1]) + 1/2 (gyy[i,j] + gyy[i+1,j]),
// gxx = -m0^2 g_m0, gyy = -m1^2 g_m1.
template <typename R>
__global__ void __launch_bounds__(kBlock)
k_hz_materials_grad(const cv_fdfd_desc* __restrict__ desc, const cv_tile* __restrict__ tiles,
                    const R* __restrict__ m0, const R* __restrict__ m1, const R* __restrict__ g0,
                    const R* __restrict__ g1, R* __restrict__ geps) {
    const cv_tile t = tiles[blockIdx.x];
    const cv_fdfd_desc d = load_desc(desc, t.inst);
    const int n = d.nx * d.ny;
#pragma unroll
    for (int q = 0; q < kCellsPerThread; ++q) {
        const int loc = t.first + threadIdx.x + q * kBlock;
        if (loc >= n) break;
        const int i = loc / d.ny, j = loc - i * d.ny;
        const int ip = (i == d.nx - 1) ? 0 : i + 1;
        const int jp = (j == d.ny - 1) ? 0 : j + 1;
        const int64_t k = d.cell_off + loc;
        const int64_t kj = d.cell_off + (int64_t)i * d.ny + jp;
        const int64_t ki = d.cell_off + (int64_t)ip * d.ny + j;
        if (d.flags & CV_FLAG_CPLX_MAT) {
            // m = 1/(e + 1e-5) is holomorphic: cotangent conj(dm/de) g = -conj(m)^2 g
            const cx<R>* c0 = reinterpret_cast<const cx<R>*>(m0);
            const cx<R>* c1 = reinterpret_cast<const cx<R>*>(m1);
            const cx<R>* h0 = reinterpret_cast<const cx<R>*>(g0);
            const cx<R>* h1 = reinterpret_cast<const cx<R>*>(g1);
            auto cxx = [&](int64_t c) { return h0 ? -(conj(c0[c]) * conj(c0[c])) * h0[c] : czero<R>(); };
            auto cyy = [&](int64_t c) { return h1 ? -(conj(c1[c]) * conj(c1[c])) * h1[c] : czero<R>(); };
            cx<R>* ge = reinterpret_cast<cx<R>*>(geps);
            ge[k] = ge[k] + R(0.5) * (cxx(k) + cxx(kj)) + R(0.5) * (cyy(k) + cyy(ki));
            continue;
        }
        auto gxx = [&](int64_t c) { return g0 ? -m0[c] * m0[c] * g0[c] : R(0); };
        auto gyy = [&](int64_t c) { return g1 ? -m1[c] * m1[c] * g1[c] : R(0); };
        geps[k] += R(0.5) * (gxx(k) + gxx(kj)) + R(0.5) * (gyy(k) + gyy(ki));
    }
}

// g_mat[k] += Re( conj(coef (D u)[k]) g[k] ): reverse mode of out = coef mat .* (D u)
// with respect to the real material.
template <typename R>
__global__ void __launch_bounds__(kBlock)
k_deriv_matgrad(const cv_fdfd_desc* __restrict__ desc, const cv_tile* __restrict__ tiles,
                const cx<R>* __restrict__ prof, const cx<R>* __restrict__ u, const cx<R>* __restrict__ g,
                R* __restrict__ gmat, int which, double cre, double cim) {
    const cv_tile t = tiles[blockIdx.x];
    const cv_fdfd_desc d = load_desc(desc, t.inst);
    const int n = d.nx * d.ny;
    const cx<R> coef = mk<R>(R(cre), R(cim));
#pragma unroll
    for (int q = 0; q < kCellsPerThread; ++q) {
        const int loc = t.first + threadIdx.x + q * kBlock;
        if (loc >= n) break;
        const int i = loc / d.ny, j = loc - i * d.ny;
        const int64_t k = d.cell_off + loc;
        const cx<R> w = coef * deriv_at<R>(d, prof, u, (const R*)nullptr, which, 0, i, j);
        const cx<R> gk = g[k];
        if (d.flags & CV_FLAG_CPLX_MAT) {   // complex material: conj(w) g
            cx<R>* gm = reinterpret_cast<cx<R>*>(gmat);
            gm[k] = gm[k] + conj(w) * gk;
        } else {
            gmat[k] += w.re * gk.re + w.im * gk.im;
        }
    }
}


// ------------------------------------------------- forward-mode pieces
// y (+)= (dA/dm0 . dm0 + dA/dm1 . dm1) x, the tangent of the operator:
//   Ez: dshift (dm0 + i dm1) x
//   Hz: scale (Dxb (dm1 .* Dxf x) + Dyb (dm0 .* Dyf x))
template <typename R>
__global__ void __launch_bounds__(kBlock)
k_material_apply(const cv_fdfd_desc* __restrict__ desc, const cv_tile* __restrict__ tiles,
                 const cx<R>* __restrict__ prof, const R* __restrict__ dm0, const R* __restrict__ dm1,
                 const cx<R>* __restrict__ x, cx<R>* __restrict__ y, int accumulate) {
    const cv_tile t = tiles[blockIdx.x];
    const cv_fdfd_desc d = load_desc(desc, t.inst);
    const int n = d.nx * d.ny;
#pragma unroll
    for (int q = 0; q < kCellsPerThread; ++q) {
        const int loc = t.first + threadIdx.x + q * kBlock;
        if (loc >= n) break;
        const int i = loc / d.ny, j = loc - i * d.ny;
        const int64_t k = d.cell_off + loc;
        cx<R> v;
        if (d.pol == CV_POL_EZ) {
            const cx<R> dm = mk<R>(dm0[k], dm1 ? dm1[k] : R(0));
            v = R(d.dshift) * (dm * x[k]);
        } else {
            // w_x = dm1 .* Dxf x at (i, j) and (i-1, j); Dxb w_x at (i, j); same along y with dm0
            const int im = (i == 0) ? d.nx - 1 : i - 1;
            const int jm = (j == 0) ? d.ny - 1 : j - 1;
            const int64_t kim = d.cell_off + (int64_t)im * d.ny + j, kjm = d.cell_off + (int64_t)i * d.ny + jm;
            const cx<R>* kxb = prof + d.prof_off + d.nx;
            const cx<R>* kyb = prof + d.prof_off + 2 * d.nx + d.ny;
            const cx<R> wx = mat_times<R>(d, dm1, k, deriv_at<R>(d, prof, x, (const R*)nullptr, 0, 0, i, j), false);
            cx<R> vx;
            if (d.flags & CV_FLAG_X_IDENTITY) {
                vx = kxb[i] * wx;
            } else {
                cx<R> wxm = mat_times<R>(d, dm1, kim, deriv_at<R>(d, prof, x, (const R*)nullptr, 0, 0, im, j), false);
                if (i == 0) wxm = wxm * conj(mk<R>(R(d.phx_re), R(d.phx_im)));
                vx = kxb[i] * (wx - wxm);
            }
            const cx<R> wy = mat_times<R>(d, dm0, k, deriv_at<R>(d, prof, x, (const R*)nullptr, 2, 0, i, j), false);
            cx<R> vy;
            if (d.flags & CV_FLAG_Y_IDENTITY) {
                vy = kyb[j] * wy;
            } else {
                cx<R> wym = mat_times<R>(d, dm0, kjm, deriv_at<R>(d, prof, x, (const R*)nullptr, 2, 0, i, jm), false);
                if (j == 0) wym = wym * conj(mk<R>(R(d.phy_re), R(d.phy_im)));
                vy = kyb[j] * (wy - wym);
            }
            v = R(d.scale) * (vx + vy);
        }
        y[k] = accumulate ? y[k] + v : v;
    }
}

// dm0 = -m0^2 (de + de[j-1]) / 2, dm1 = -m1^2 (de + de[i-1]) / 2  (tangent of the Hz materials)
template <typename R>
__global__ void __launch_bounds__(kBlock)
k_hz_materials_jvp(const cv_fdfd_desc* __restrict__ desc, const cv_tile* __restrict__ tiles,
                   const R* __restrict__ m0, const R* __restrict__ m1, const R* __restrict__ de,
                   R* __restrict__ dm0, R* __restrict__ dm1) {
    const cv_tile t = tiles[blockIdx.x];
    const cv_fdfd_desc d = load_desc(desc, t.inst);
    const int n = d.nx * d.ny;
#pragma unroll
    for (int q = 0; q < kCellsPerThread; ++q) {
        const int loc = t.first + threadIdx.x + q * kBlock;
        if (loc >= n) break;
        const int i = loc / d.ny, j = loc - i * d.ny;
        const int im = (i == 0) ? d.nx - 1 : i - 1;
        const int jm = (j == 0) ? d.ny - 1 : j - 1;
        const int64_t k = d.cell_off + loc;
        const int64_t kjm = d.cell_off + (int64_t)i * d.ny + jm, kim = d.cell_off + (int64_t)im * d.ny + j;
        if (d.flags & CV_FLAG_CPLX_MAT) {   // holomorphic tangent dm = -m^2 de_avg
            const cx<R>* c0 = reinterpret_cast<const cx<R>*>(m0);
            const cx<R>* c1 = reinterpret_cast<const cx<R>*>(m1);
            const cx<R>* ce = reinterpret_cast<const cx<R>*>(de);
            const cx<R> dxx = R(0.5) * (ce[k] + ce[kjm]);
            const cx<R> dyy = R(0.5) * (ce[k] + ce[kim]);
            reinterpret_cast<cx<R>*>(dm0)[k] = -(c0[k] * c0[k]) * dxx;
            reinterpret_cast<cx<R>*>(dm1)[k] = -(c1[k] * c1[k]) * dyy;
            continue;
        }
        const R dxx = R(0.5) * (de[k] + de[kjm]);
        const R dyy = R(0.5) * (de[k] + de[kim]);
        dm0[k] = -m0[k] * m0[k] * dxx;
        dm1[k] = -m1[k] * m1[k] * dyy;
    }
}

// ------------------------------------------------ frequency coupling
// y (+)= C x (or C^T x): only the inter-block terms of coupled systems, with
// the tables read from `tab` (prof's layout).  Instances without coupling
// write zero (or leave y unchanged when accumulating).
template <typename R, bool TRANS>
__global__ void __launch_bounds__(kBlock)
k_coupling_apply(const cv_fdfd_desc* __restrict__ desc, const cv_tile* __restrict__ tiles,
                 const cx<R>* __restrict__ tab, const cx<R>* __restrict__ x, cx<R>* __restrict__ y, int accumulate) {
    const cv_tile t = tiles[blockIdx.x];
    const cv_fdfd_desc d = load_desc(desc, t.inst);
    const int n = d.nx * d.ny;
    auto fx = [&](int64_t k) { return x[k]; };
#pragma unroll
    for (int q = 0; q < kCellsPerThread; ++q) {
        const int loc = t.first + threadIdx.x + q * kBlock;
        if (loc >= n) break;
        const int i = loc / d.ny, j = loc - i * d.ny;
        const int64_t k = d.cell_off + loc;
        const cx<R> v = (d.n_cpl > 0) ? coupling_row<R, TRANS>(d, tab, i, j, fx) : czero<R>();
        y[k] = accumulate ? y[k] + v : v;
    }
}

// g[table(h, +/-)][c] += scale * sum_m w2[m] lam_m[c] x_{m +/- (h+1)}[c], one
// thread per cell of block 0 of every coupled system (other tiles return).
template <typename R>
__global__ void __launch_bounds__(kBlock)
k_coupling_grad(const cv_fdfd_desc* __restrict__ desc, const cv_tile* __restrict__ tiles,
                const cx<R>* __restrict__ prof, const cx<R>* __restrict__ lam, const cx<R>* __restrict__ x,
                cx<R>* __restrict__ g, R scale) {
    const cv_tile t = tiles[blockIdx.x];
    const cv_fdfd_desc d = load_desc(desc, t.inst);
    if (d.blk != 0 || d.n_cpl == 0) return;
    const int n = d.nx * d.ny;
    const cx<R>* w2 = prof + d.w2_off;
#pragma unroll
    for (int q = 0; q < kCellsPerThread; ++q) {
        const int loc = t.first + threadIdx.x + q * kBlock;
        if (loc >= n) break;
        const int64_t k0 = d.cell_off + loc;
        for (int h = 0; h < d.n_cpl; ++h) {
            const int dd = h + 1;
            cx<R> up = czero<R>(), dn = czero<R>();
            for (int m = 0; m < d.nblk; ++m) {
                const cx<R> lm = w2[m].re * lam[k0 + (int64_t)m * n];
                if (m + dd < d.nblk) up = up + lm * x[k0 + (int64_t)(m + dd) * n];
                if (m - dd >= 0) dn = dn + lm * x[k0 + (int64_t)(m - dd) * n];
            }
            cx<R>* gp = g + d.cpl_off + (int64_t)(2 * h) * n + loc;
            cx<R>* gm = g + d.cpl_off + (int64_t)(2 * h + 1) * n + loc;
            *gp = *gp + scale * up;
            *gm = *gm + scale * dn;
        }
    }
}

}  // namespace cv

using namespace cv;

extern "C" int cv_fdfd_apply(int dtype, const cv_fdfd_desc* desc, const cv_tile* tiles, int n_tiles,
                             const void* prof, const void* m0, const void* m1, const void* x, void* y,
                             int transpose, void* stream) {
    CV_CHECK(n_tiles >= 0, "n_tiles < 0");
    CV_CHECK(desc && tiles && prof && m0 && x && y, "null pointer argument");
    if (n_tiles == 0) return CV_OK;
    hipStream_t s = as_stream(stream);
    dim3 g(n_tiles), b(kBlock);
    if (dtype == CV_C128) {
        if (transpose)
            k_apply<double, true><<<g, b, 0, s>>>(desc, tiles, (const cx<double>*)prof, (const double*)m0,
                                                  (const double*)m1, (const cx<double>*)x, (cx<double>*)y);
        else
            k_apply<double, false><<<g, b, 0, s>>>(desc, tiles, (const cx<double>*)prof, (const double*)m0,
                                                   (const double*)m1, (const cx<double>*)x, (cx<double>*)y);
    } else if (dtype == CV_C64) {
        if (transpose)
            k_apply<float, true><<<g, b, 0, s>>>(desc, tiles, (const cx<float>*)prof, (const float*)m0,
                                                 (const float*)m1, (const cx<float>*)x, (cx<float>*)y);
        else
            k_apply<float, false><<<g, b, 0, s>>>(desc, tiles, (const cx<float>*)prof, (const float*)m0,
                                                  (const float*)m1, (const cx<float>*)x, (cx<float>*)y);
    } else {
        CV_CHECK(false, "unknown dtype %d", dtype);
    }
    CV_LAUNCH_CHECK();
    return CV_OK;
}

extern "C" int cv_fdfd_hz_materials(int dtype, const cv_fdfd_desc* desc, const cv_tile* tiles, int n_tiles,
                                    const void* eps, void* m0, void* m1, void* stream) {
    CV_CHECK(desc && tiles && eps && m0 && m1, "null pointer argument");
    if (n_tiles == 0) return CV_OK;
    hipStream_t s = as_stream(stream);
    if (dtype == CV_C128)
        k_hz_materials<double><<<n_tiles, kBlock, 0, s>>>(desc, tiles, (const double*)eps, (double*)m0, (double*)m1);
    else if (dtype == CV_C64)
        k_hz_materials<float><<<n_tiles, kBlock, 0, s>>>(desc, tiles, (const float*)eps, (float*)m0, (float*)m1);
    else
        CV_CHECK(false, "unknown dtype %d", dtype);
    CV_LAUNCH_CHECK();
    return CV_OK;
}

extern "C" int cv_fdfd_deriv(int dtype, const cv_fdfd_desc* desc, const cv_tile* tiles, int n_tiles,
                             const void* prof, const void* mat, const void* u, void* out, int which,
                             int mode, double coef_re, double coef_im, int accumulate, void* stream) {
    CV_CHECK(desc && tiles && prof && u && out, "null pointer argument");
    CV_CHECK(which >= 0 && which < 4, "which must be 0..3 (xf, xb, yf, yb)");
    CV_CHECK(mode >= 0 && mode < 3, "mode must be 0 (N), 1 (T) or 2 (H)");
    if (n_tiles == 0) return CV_OK;
    hipStream_t s = as_stream(stream);
    if (dtype == CV_C128)
        k_deriv<double><<<n_tiles, kBlock, 0, s>>>(desc, tiles, (const cx<double>*)prof, (const double*)mat,
                                                   (const cx<double>*)u, (cx<double>*)out, which, mode,
                                                   coef_re, coef_im, accumulate);
    else if (dtype == CV_C64)
        k_deriv<float><<<n_tiles, kBlock, 0, s>>>(desc, tiles, (const cx<float>*)prof, (const float*)mat,
                                                  (const cx<float>*)u, (cx<float>*)out, which, mode,
                                                  coef_re, coef_im, accumulate);
    else
        CV_CHECK(false, "unknown dtype %d", dtype);
    CV_LAUNCH_CHECK();
    return CV_OK;
}

extern "C" int cv_fdfd_material_grad(int dtype, const cv_fdfd_desc* desc, const cv_tile* tiles,
                                     int n_tiles, const void* prof, const void* lam, const void* x,
                                     void* g0, void* g1, double scale, void* stream) {
    CV_CHECK(desc && tiles && prof && lam && x && g0, "null pointer argument");
    if (n_tiles == 0) return CV_OK;
    hipStream_t s = as_stream(stream);
    if (dtype == CV_C128)
        k_material_grad<double><<<n_tiles, kBlock, 0, s>>>(desc, tiles, (const cx<double>*)prof,
                                                           (const cx<double>*)lam, (const cx<double>*)x,
                                                           (double*)g0, (double*)g1, scale);
    else if (dtype == CV_C64)
        k_material_grad<float><<<n_tiles, kBlock, 0, s>>>(desc, tiles, (const cx<float>*)prof,
                                                          (const cx<float>*)lam, (const cx<float>*)x,
                                                          (float*)g0, (float*)g1, scale);
    else
        CV_CHECK(false, "unknown dtype %d", dtype);
    CV_LAUNCH_CHECK();
    return CV_OK;
}

extern "C" int cv_fdfd_hz_materials_grad(int dtype, const cv_fdfd_desc* desc, const cv_tile* tiles, int n_tiles,
                                         const void* m0, const void* m1, const void* g0, const void* g1,
                                         void* geps, void* stream) {
    CV_CHECK(desc && tiles && m0 && m1 && geps, "null pointer argument");
    if (n_tiles == 0) return CV_OK;
    hipStream_t s = as_stream(stream);
    if (dtype == CV_C128)
        k_hz_materials_grad<double><<<n_tiles, kBlock, 0, s>>>(desc, tiles, (const double*)m0, (const double*)m1,
                                                               (const double*)g0, (const double*)g1,
                                                               (double*)geps);
    else if (dtype == CV_C64)
        k_hz_materials_grad<float><<<n_tiles, kBlock, 0, s>>>(desc, tiles, (const float*)m0, (const float*)m1,
                                                              (const float*)g0, (const float*)g1, (float*)geps);
    else
        CV_CHECK(false, "unknown dtype %d", dtype);
    CV_LAUNCH_CHECK();
    return CV_OK;
}

extern "C" int cv_fdfd_deriv_matgrad(int dtype, const cv_fdfd_desc* desc, const cv_tile* tiles, int n_tiles,
                                     const void* prof, const void* u, const void* g, void* gmat, int which,
                                     double coef_re, double coef_im, void* stream) {
    CV_CHECK(desc && tiles && prof && u && g && gmat, "null pointer argument");
    CV_CHECK(which >= 0 && which < 4, "which must be 0..3 (xf, xb, yf, yb)");
    if (n_tiles == 0) return CV_OK;
    hipStream_t s = as_stream(stream);
    if (dtype == CV_C128)
        k_deriv_matgrad<double><<<n_tiles, kBlock, 0, s>>>(desc, tiles, (const cx<double>*)prof,
                                                           (const cx<double>*)u, (const cx<double>*)g,
                                                           (double*)gmat, which, coef_re, coef_im);
    else if (dtype == CV_C64)
        k_deriv_matgrad<float><<<n_tiles, kBlock, 0, s>>>(desc, tiles, (const cx<float>*)prof,
                                                          (const cx<float>*)u, (const cx<float>*)g, (float*)gmat,
                                                          which, coef_re, coef_im);
    else
        CV_CHECK(false, "unknown dtype %d", dtype);
    CV_LAUNCH_CHECK();
    return CV_OK;
}

extern "C" int cv_fdfd_material_apply(int dtype, const cv_fdfd_desc* desc, const cv_tile* tiles, int n_tiles,
                                      const void* prof, const void* dm0, const void* dm1, const void* x, void* y,
                                      int accumulate, void* stream) {
    CV_CHECK(desc && tiles && prof && dm0 && x && y, "null pointer argument");
    if (n_tiles == 0) return CV_OK;
    hipStream_t s = as_stream(stream);
    if (dtype == CV_C128)
        k_material_apply<double><<<n_tiles, kBlock, 0, s>>>(desc, tiles, (const cx<double>*)prof, (const double*)dm0,
                                                            (const double*)dm1, (const cx<double>*)x,
                                                            (cx<double>*)y, accumulate);
    else if (dtype == CV_C64)
        k_material_apply<float><<<n_tiles, kBlock, 0, s>>>(desc, tiles, (const cx<float>*)prof, (const float*)dm0,
                                                           (const float*)dm1, (const cx<float>*)x, (cx<float>*)y,
                                                           accumulate);
    else
        CV_CHECK(false, "unknown dtype %d", dtype);
    CV_LAUNCH_CHECK();
    return CV_OK;
}

extern "C" int cv_fdfd_hz_materials_jvp(int dtype, const cv_fdfd_desc* desc, const cv_tile* tiles, int n_tiles,
                                        const void* m0, const void* m1, const void* d_eps, void* dm0, void* dm1,
                                        void* stream) {
    CV_CHECK(desc && tiles && m0 && m1 && d_eps && dm0 && dm1, "null pointer argument");
    if (n_tiles == 0) return CV_OK;
    hipStream_t s = as_stream(stream);
    if (dtype == CV_C128)
        k_hz_materials_jvp<double><<<n_tiles, kBlock, 0, s>>>(desc, tiles, (const double*)m0, (const double*)m1,
                                                              (const double*)d_eps, (double*)dm0, (double*)dm1);
    else if (dtype == CV_C64)
        k_hz_materials_jvp<float><<<n_tiles, kBlock, 0, s>>>(desc, tiles, (const float*)m0, (const float*)m1,
                                                             (const float*)d_eps, (float*)dm0, (float*)dm1);
    else
        CV_CHECK(false, "unknown dtype %d", dtype);
    CV_LAUNCH_CHECK();
    return CV_OK;
}

extern "C" int cv_fdfd_coupling_apply(int dtype, const cv_fdfd_desc* desc, const cv_tile* tiles, int n_tiles,
                                      const void* tab, const void* x, void* y, int transpose, int accumulate,
                                      void* stream) {
    CV_CHECK(desc && tiles && tab && x && y, "null pointer argument");
    CV_CHECK(n_tiles >= 0, "n_tiles < 0");
    if (n_tiles == 0) return CV_OK;
    hipStream_t s = as_stream(stream);
    if (dtype == CV_C128) {
        if (transpose)
            k_coupling_apply<double, true><<<n_tiles, kBlock, 0, s>>>(desc, tiles, (const cx<double>*)tab,
                                                                     (const cx<double>*)x, (cx<double>*)y, accumulate);
        else
            k_coupling_apply<double, false><<<n_tiles, kBlock, 0, s>>>(desc, tiles, (const cx<double>*)tab,
                                                                      (const cx<double>*)x, (cx<double>*)y,
                                                                      accumulate);
    } else if (dtype == CV_C64) {
        if (transpose)
            k_coupling_apply<float, true><<<n_tiles, kBlock, 0, s>>>(desc, tiles, (const cx<float>*)tab,
                                                                    (const cx<float>*)x, (cx<float>*)y, accumulate);
        else
            k_coupling_apply<float, false><<<n_tiles, kBlock, 0, s>>>(desc, tiles, (const cx<float>*)tab,
                                                                     (const cx<float>*)x, (cx<float>*)y, accumulate);
    } else {
        CV_CHECK(false, "unknown dtype %d", dtype);
    }
    CV_LAUNCH_CHECK();
    return CV_OK;
}

extern "C" int cv_fdfd_coupling_grad(int dtype, const cv_fdfd_desc* desc, const cv_tile* tiles, int n_tiles,
                                     const void* prof, const void* lam, const void* x, void* g, double scale,
                                     void* stream) {
    CV_CHECK(desc && tiles && prof && lam && x && g, "null pointer argument");
    CV_CHECK(n_tiles >= 0, "n_tiles < 0");
    if (n_tiles == 0) return CV_OK;
    hipStream_t s = as_stream(stream);
    if (dtype == CV_C128)
        k_coupling_grad<double><<<n_tiles, kBlock, 0, s>>>(desc, tiles, (const cx<double>*)prof,
                                                           (const cx<double>*)lam, (const cx<double>*)x,
                                                           (cx<double>*)g, scale);
    else if (dtype == CV_C64)
        k_coupling_grad<float><<<n_tiles, kBlock, 0, s>>>(desc, tiles, (const cx<float>*)prof, (const cx<float>*)lam,
                                                          (const cx<float>*)x, (cx<float>*)g, (float)scale);
    else
        CV_CHECK(false, "unknown dtype %d", dtype);
    CV_LAUNCH_CHECK();
    return CV_OK;
}
