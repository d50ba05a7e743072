// Matrix-free SC-PML Helmholtz stencils for the 2D FDFD operators.
//
// The reference assembles A as a COO/CSR matrix on every solve
// (ceviche/fdfd.py:172-185 for Ez, :211-233 for Hz) from the PML-scaled Yee
// difference matrices of ceviche/derivatives.py:34-124.  Here each row of A
// is a 5-point stencil whose coefficients are formed on the fly from 1D
// profiles k*f[i] = 1/(s_f[i] dL), k*b[i] = 1/(s_b[i] dL) (length nx / ny)
// and the per-cell material array(s), so one application moves only the
// input vector, the material and the output vector through HBM.
//
// Row (i, j) of the x-part, with ip = i+1 and im = i-1 (periodic), Bloch
// phasor ph applied on the wrapped entries exactly as derivatives.py:94-108:
//   Dxf:  (Dxf u)[i] = kxf[i] * (phf_i u[ip] - u[i]),  phf_i = ph if i == nx-1
//   Dxb:  (Dxb u)[i] = kxb[i] * (u[i] - phb_i u[im]),  phb_i = conj(ph) if i == 0
// Identity axes (nx == 1, derivatives.py:76-79): Dxf = diag(kxf), Dxb = diag(kxb)
// with the profiles holding 1/s (no 1/dL).
#pragma once

#include "common.h"

namespace cv {

template <typename R>
struct Tri {
    cx<R> m, c, p;
};

// Ez x-part (Dxf Dxb) at row i.  F = kxf, B = kxb, ph = Bloch phasor.
template <typename R>
__host__ __device__ __forceinline__ Tri<R> axis_ez(const cx<R>* __restrict__ F, const cx<R>* __restrict__ B,
                                          int n, int i, cx<R> ph, bool identity) {
    Tri<R> t;
    if (identity) {
        t.m = czero<R>();
        t.p = czero<R>();
        t.c = F[i] * B[i];
        return t;
    }
    const int ip = (i == n - 1) ? 0 : i + 1;
    const cx<R> Fi = F[i], Bi = B[i], Bp = B[ip];
    const bool wrap_hi = (i == n - 1);
    const bool wrap_lo = (i == 0);
    // coefficient on u[ip]: F[i] * phf_i * B[ip]
    cx<R> fp = wrap_hi ? Fi * ph : Fi;
    t.p = fp * Bp;
    // coefficient on u[i]: -F[i] phf_i B[ip] phb_ip - F[i] B[i]   (phb_ip = conj(ph) iff ip == 0)
    cx<R> q = t.p;
    if (wrap_hi) q = q * conj(ph);
    t.c = -(q + Fi * Bi);
    // coefficient on u[im]: F[i] B[i] phb_i
    cx<R> fb = Fi * Bi;
    t.m = wrap_lo ? fb * conj(ph) : fb;
    return t;
}

// Hz x-part (Dxb diag(u) Dxf) at row i; u_i = u[i], u_m = u[im] (material on the
// axis' own rows; U = R for real materials, cx<R> for complex ones).
// F = kxf, B = kxb.
template <typename R, typename U>
__host__ __device__ __forceinline__ Tri<R> axis_hz(const cx<R>* __restrict__ F, const cx<R>* __restrict__ B,
                                          int n, int i, cx<R> ph, U u_i, U u_m, bool identity) {
    Tri<R> t;
    if (identity) {
        t.m = czero<R>();
        t.p = czero<R>();
        t.c = u_i * (B[i] * F[i]);
        return t;
    }
    const int im = (i == 0) ? n - 1 : i - 1;
    const cx<R> Bi = B[i], Fi = F[i], Fm = F[im];
    const bool wrap_hi = (i == n - 1);   // phf_i
    const bool wrap_lo = (i == 0);       // phb_i and phf_im
    // (Dxb u Dxf x)[i] = B[i] ( u_i (Dxf x)[i] - phb_i u_m (Dxf x)[im] )
    //   (Dxf x)[i]  = F[i]  (phf_i  x[ip] - x[i])
    //   (Dxf x)[im] = F[im] (phf_im x[i]  - x[im])
    cx<R> a = u_i * (Bi * Fi);                 // B u_i F
    t.p = wrap_hi ? a * ph : a;
    cx<R> b = u_m * (Bi * Fm);                 // B u_m F[im]
    if (wrap_lo) b = b * conj(ph);             // phb_i
    t.m = b;
    cx<R> bc = wrap_lo ? b * ph : b;           // phb_i * phf_im
    t.c = -(a + bc);
    return t;
}

// Full-row coefficients of A at (i, j).
template <typename R>
struct Row5 {
    cx<R> c, xm, xp, ym, yp;
};

template <typename R>
__host__ __device__ __forceinline__ Row5<R> row_coefs(const cv_fdfd_desc& d, const cx<R>* __restrict__ prof,
                                             const R* __restrict__ m0, const R* __restrict__ m1,
                                             int i, int j) {
    const int nx = d.nx, ny = d.ny;
    const cx<R>* kxf = prof + d.prof_off;
    const cx<R>* kxb = kxf + nx;
    const cx<R>* kyf = kxb + nx;
    const cx<R>* kyb = kyf + ny;
    const cx<R> phx = mk<R>(R(d.phx_re), R(d.phx_im));
    const cx<R> phy = mk<R>(R(d.phy_re), R(d.phy_im));
    const bool idx = d.flags & CV_FLAG_X_IDENTITY;
    const bool idy = d.flags & CV_FLAG_Y_IDENTITY;
    const R s = R(d.scale), sh = R(d.dshift);
    const int64_t k = d.cell_off + (int64_t)i * ny + j;
    Tri<R> tx, ty;
    R diag, diag_im = R(0);
    if (d.pol == CV_POL_EZ) {
        tx = axis_ez<R>(kxf, kxb, nx, i, phx, idx);
        ty = axis_ez<R>(kyf, kyb, ny, j, phy, idy);
        diag = sh * m0[k];
        if (m1) diag_im = sh * m1[k];  // complex eps: m1 = Im(eps)
    } else {
        const int im = (i == 0) ? nx - 1 : i - 1;
        const int jm = (j == 0) ? ny - 1 : j - 1;
        const int64_t kxm = d.cell_off + (int64_t)im * ny + j, kym = d.cell_off + (int64_t)i * ny + jm;
        if (d.flags & CV_FLAG_CPLX_MAT) {   // complex eps: m0, m1 hold interleaved complex values
            const cx<R>* c0 = reinterpret_cast<const cx<R>*>(m0);
            const cx<R>* c1 = reinterpret_cast<const cx<R>*>(m1);
            tx = axis_hz<R, cx<R>>(kxf, kxb, nx, i, phx, c1[k], c1[kxm], idx);
            ty = axis_hz<R, cx<R>>(kyf, kyb, ny, j, phy, c0[k], c0[kym], idy);
        } else {
            tx = axis_hz<R, R>(kxf, kxb, nx, i, phx, m1[k], m1[kxm], idx);
            ty = axis_hz<R, R>(kyf, kyb, ny, j, phy, m0[k], m0[kym], idy);
        }
        diag = sh;
    }
    Row5<R> r;
    r.c = s * (tx.c + ty.c);
    r.c.re += diag;
    r.c.im += diag_im;
    r.xm = s * tx.m;
    r.xp = s * tx.p;
    r.ym = s * ty.m;
    r.yp = s * ty.p;
    return r;
}

// Inter-block terms of a frequency-coupled system at row (i, j) of this
// block (A, or A^T when TRANS); see cv_fdfd_desc.  Row block m of A couples to
// block m + (h+1) with w2[m] gp_h and to block m - (h+1) with w2[m] gm_h, so
// row m of A^T takes w2[m-(h+1)] gp_h from block m - (h+1) and
// w2[m+(h+1)] gm_h from block m + (h+1).
template <typename R, bool TRANS, typename Fetch>
__host__ __device__ __forceinline__ cx<R> coupling_row(const cv_fdfd_desc& d, const cx<R>* __restrict__ tab,
                                              int i, int j, Fetch f) {
    const int64_t n = (int64_t)d.nx * d.ny;
    const int64_t loc = (int64_t)i * d.ny + j;
    const int64_t k = d.cell_off + loc;
    const cx<R>* g = tab + d.cpl_off;
    const cx<R>* w2 = tab + d.w2_off;
    const R w_self = w2[d.blk].re;
    cx<R> acc = czero<R>();
    for (int h = 0; h < d.n_cpl; ++h) {
        const int dd = h + 1;
        const int up = d.blk + dd, dn = d.blk - dd;
        if (up < d.nblk) {
            const cx<R> c = TRANS ? g[(int64_t)(2 * h + 1) * n + loc] : g[(int64_t)(2 * h) * n + loc];
            acc = acc + (TRANS ? w2[up].re : w_self) * (c * f(k + dd * n));
        }
        if (dn >= 0) {
            const cx<R> c = TRANS ? g[(int64_t)(2 * h) * n + loc] : g[(int64_t)(2 * h + 1) * n + loc];
            acc = acc + (TRANS ? w2[dn].re : w_self) * (c * f(k - dd * n));
        }
    }
    return acc;
}

// (A u)[i,j] with u fetched through `f(k)` (global cell index k).
template <typename R, typename Fetch>
__host__ __device__ __forceinline__ cx<R> apply_row(const cv_fdfd_desc& d, const cx<R>* __restrict__ prof,
                                           const R* __restrict__ m0, const R* __restrict__ m1,
                                           int i, int j, Fetch f, cx<R> u_c) {
    const Row5<R> r = row_coefs<R>(d, prof, m0, m1, i, j);
    const int nx = d.nx, ny = d.ny;
    const int64_t base = d.cell_off;
    cx<R> acc = r.c * u_c;
    if (!(d.flags & CV_FLAG_X_IDENTITY)) {
        const int ip = (i == nx - 1) ? 0 : i + 1;
        const int im = (i == 0) ? nx - 1 : i - 1;
        acc = acc + r.xp * f(base + (int64_t)ip * ny + j);
        acc = acc + r.xm * f(base + (int64_t)im * ny + j);
    }
    if (!(d.flags & CV_FLAG_Y_IDENTITY)) {
        const int jp = (j == ny - 1) ? 0 : j + 1;
        const int jm = (j == 0) ? ny - 1 : j - 1;
        acc = acc + r.yp * f(base + (int64_t)i * ny + jp);
        acc = acc + r.ym * f(base + (int64_t)i * ny + jm);
    }
    if (d.n_cpl > 0) acc = acc + coupling_row<R, false>(d, prof, i, j, f);
    return acc;
}

// (A^T u)[i,j]: column (i,j) of A, i.e. the neighbours' coefficients that
// point back at (i,j).
template <typename R, typename Fetch>
__host__ __device__ __forceinline__ cx<R> apply_row_T(const cv_fdfd_desc& d, const cx<R>* __restrict__ prof,
                                             const R* __restrict__ m0, const R* __restrict__ m1,
                                             int i, int j, Fetch f, cx<R> u_c) {
    const int nx = d.nx, ny = d.ny;
    const int64_t base = d.cell_off;
    const Row5<R> r = row_coefs<R>(d, prof, m0, m1, i, j);
    cx<R> acc = r.c * u_c;
    if (!(d.flags & CV_FLAG_X_IDENTITY)) {
        const int ip = (i == nx - 1) ? 0 : i + 1;
        const int im = (i == 0) ? nx - 1 : i - 1;
        const Row5<R> rm = row_coefs<R>(d, prof, m0, m1, im, j);
        const Row5<R> rp = row_coefs<R>(d, prof, m0, m1, ip, j);
        acc = acc + rm.xp * f(base + (int64_t)im * ny + j);
        acc = acc + rp.xm * f(base + (int64_t)ip * ny + j);
    }
    if (!(d.flags & CV_FLAG_Y_IDENTITY)) {
        const int jp = (j == ny - 1) ? 0 : j + 1;
        const int jm = (j == 0) ? ny - 1 : j - 1;
        const Row5<R> rm = row_coefs<R>(d, prof, m0, m1, i, jm);
        const Row5<R> rp = row_coefs<R>(d, prof, m0, m1, i, jp);
        acc = acc + rm.yp * f(base + (int64_t)i * ny + jm);
        acc = acc + rp.ym * f(base + (int64_t)i * ny + jp);
    }
    if (d.n_cpl > 0) acc = acc + coupling_row<R, true>(d, prof, i, j, f);
    return acc;
}

template <typename R, bool TRANS, typename Fetch>
__host__ __device__ __forceinline__ cx<R> apply_any(const cv_fdfd_desc& d, const cx<R>* __restrict__ prof,
                                           const R* __restrict__ m0, const R* __restrict__ m1,
                                           int i, int j, Fetch f, cx<R> u_c) {
    if (TRANS) return apply_row_T<R>(d, prof, m0, m1, i, j, f, u_c);
    return apply_row<R>(d, prof, m0, m1, i, j, f, u_c);
}

// ------------------------------------------------- single derivative ops
// which: 0 xf, 1 xb, 2 yf, 3 yb.  mode: 0 out = coef * mat .* (D u)
//                                      1 out = coef * D^T (mat .* u)
//                                      2 out = conj(coef) * D^H (conj(mat) .* u)
// mat is real, or interleaved complex when the instance has CV_FLAG_CPLX_MAT.
template <typename R>
__host__ __device__ __forceinline__ cx<R> mat_times(const cv_fdfd_desc& d, const R* __restrict__ mat, int64_t k,
                                           cx<R> v, bool conj_mat) {
    if (d.flags & CV_FLAG_CPLX_MAT) {
        const cx<R> m = reinterpret_cast<const cx<R>*>(mat)[k];
        return (conj_mat ? conj(m) : m) * v;
    }
    return mat[k] * v;
}

template <typename R>
__host__ __device__ __forceinline__ cx<R> deriv_at(const cv_fdfd_desc& d, const cx<R>* __restrict__ prof,
                                          const cx<R>* __restrict__ u, const R* __restrict__ mat,
                                          int which, int mode, int i, int j) {
    const int nx = d.nx, ny = d.ny;
    const bool xaxis = which < 2;
    const bool fwd = (which & 1) == 0;
    const int n = xaxis ? nx : ny;
    const int a = xaxis ? i : j;
    const cx<R>* kxf = prof + d.prof_off;
    const cx<R>* K = kxf + (xaxis ? 0 : 2 * nx) + (fwd ? 0 : n);
    const bool ident = xaxis ? (d.flags & CV_FLAG_X_IDENTITY) : (d.flags & CV_FLAG_Y_IDENTITY);
    cx<R> ph = xaxis ? mk<R>(R(d.phx_re), R(d.phx_im)) : mk<R>(R(d.phy_re), R(d.phy_im));
    const bool herm = (mode == 2);
    const int64_t base = d.cell_off;
    auto idx = [&](int aa) -> int64_t {
        return xaxis ? base + (int64_t)aa * ny + j : base + (int64_t)i * ny + aa;
    };
    auto val = [&](int aa) -> cx<R> {
        cx<R> v = u[idx(aa)];
        if (mode != 0 && mat) v = mat_times<R>(d, mat, idx(aa), v, mode == 2);
        return v;
    };
    auto kk = [&](int aa) -> cx<R> { return herm ? conj(K[aa]) : K[aa]; };
    if (herm) ph = conj(ph);
    const int ap = (a == n - 1) ? 0 : a + 1;
    const int am = (a == 0) ? n - 1 : a - 1;
    cx<R> out;
    if (ident) {
        out = kk(a) * val(a);
    } else if (mode == 0) {
        if (fwd) {  // K[a] (phf_a u[ap] - u[a])
            cx<R> up = val(ap);
            if (a == n - 1) up = up * ph;
            out = K[a] * (up - val(a));
        } else {    // K[a] (u[a] - phb_a u[am])
            cx<R> um = val(am);
            if (a == 0) um = um * conj(ph);
            out = K[a] * (val(a) - um);
        }
    } else {
        // transpose (or conjugate transpose with conj'd K and ph)
        if (fwd) {  // (Df^T w)[a] = -K[a] w[a] + K[am] phf_am w[am],  phf_am = ph iff a == 0
            cx<R> t = kk(am) * val(am);
            if (a == 0) t = t * ph;
            out = t - kk(a) * val(a);
        } else {    // (Db^T w)[a] = K[a] w[a] - K[ap] phb_ap w[ap],  phb_ap = conj(ph) iff a == n-1
            cx<R> t = kk(ap) * val(ap);
            if (a == n - 1) t = t * conj(ph);
            out = kk(a) * val(a) - t;
        }
    }
    if (mode == 0 && mat) out = mat_times<R>(d, mat, idx(a), out, false);
    return out;
}

// Loads the instance descriptor of this workgroup's tile into registers.
__host__ __device__ __forceinline__ cv_fdfd_desc load_desc(const cv_fdfd_desc* __restrict__ desc, int inst) {
    return desc[inst];
}

// Index of the instance that holds the Krylov state of `inst`'s linear system
// (block 0 of a frequency-coupled system, the instance itself otherwise).
__host__ __device__ __forceinline__ int system_of(const cv_fdfd_desc& d, int inst) { return inst - d.blk; }

static_assert(sizeof(cv_fdfd_desc) == 128, "cv_fdfd_desc layout");

}  // namespace cv
