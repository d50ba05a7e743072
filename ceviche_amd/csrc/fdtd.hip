// FDTD Yee update with UPML integral terms (ceviche/fdtd.py:74-144) and its
// exact discrete adjoint, for float32 / float64 state.
//
// State (18 arrays of nx*ny*nz, C order, z fastest):
//   0..2 H   3..5 D   6..8 ICE   9..11 IH   12..14 ICH   15..17 ID
// E is not stored: E_a = D_a / eps_aa with eps_aa = (eps + roll(eps, 1, a)) / 2
// (utils.py:153-176), formed on the fly from eps.  The PML integrals are only
// touched where their coefficient is non-zero (ICE_a where sigmaH_a != 0, IH_a
// where sigmaH_b sigmaH_c != 0, likewise for D): elsewhere the reference
// multiplies them by an exact zero.
//
// Coefficients (fdtd.py:265-311) are functions of the per-axis sigma profiles
// sigH_a[n_a], sigD_a[n_a] (fdtd.py:224-263) and are formed in-kernel with the
// reference's formula; interior cells (all sigma zero) use host-computed
// constants.
#include "common.h"

#include <vector>
#include <algorithm>

namespace cv {

template <typename R>
struct Fdtd {
    int nx, ny, nz;
    int64_t n;
    R inv_dL;
    double dt, c0, eps0;
    const R* sig;      // sigHx[nx] sigHy[ny] sigHz[nz] sigDx[nx] sigDy[ny] sigDz[nz]
    const R* eps;      // relative permittivity, n
    R* st;             // 18 * n
    // interior coefficients (sigma = 0): m1 and m2 for H and D
    R h1, h2, d1, d2;
};

template <typename R>
__host__ __device__ __forceinline__ const R* sigH(const Fdtd<R>& f, int a) {
    return f.sig + (a == 0 ? 0 : (a == 1 ? f.nx : f.nx + f.ny));
}
template <typename R>
__host__ __device__ __forceinline__ const R* sigD(const Fdtd<R>& f, int a) {
    return f.sig + (f.nx + f.ny + f.nz) + (a == 0 ? 0 : (a == 1 ? f.nx : f.nx + f.ny));
}

struct Coef4 {
    double m1, m2, m3, m4;
};

// fdtd.py:272-311 for one component; sa is the component's own-axis sigma,
// sb, sc the other two.  sign = -1 for H (mH2, mH3 negative), +1 for D.
template <typename R>
__host__ __device__ __forceinline__ void coef4(R sa, R sb, R sc, R dt, R c0, R e0, R sign, R& m1, R& m2, R& m3, R& m4) {
    const R m0 = (R(1) / dt + (sb + sc) / R(2) / e0 + sb * sc * dt / R(4) / (e0 * e0));
    m1 = (R(1) / m0 * (R(1) / dt - (sb + sc) / R(2) / e0 - sb * sc * dt / R(4) / (e0 * e0)));
    m2 = sign / m0 * c0;
    m3 = sign / m0 * c0 * dt * sa / e0;
    m4 = R(-1) / m0 * dt * sb * sc / (e0 * e0);
}

__host__ __device__ __forceinline__ int wrapm(int i, int n) { return i == 0 ? n - 1 : i - 1; }
__host__ __device__ __forceinline__ int wrapp(int i, int n) { return i == n - 1 ? 0 : i + 1; }

template <typename R>
__host__ __device__ __forceinline__ int64_t cell(const Fdtd<R>& f, int i, int j, int k) {
    return ((int64_t)i * f.ny + j) * f.nz + k;
}

// (i, j, k) of flat cell c; 32-bit unsigned division when the grid allows it
// (a 64-bit division by a runtime value is a long instruction sequence)
template <typename R>
__host__ __device__ __forceinline__ void decompose(const Fdtd<R>& f, int64_t c, int& i, int& j, int& k) {
    if (f.n <= (int64_t)0xffffffffLL) {
        const uint32_t u = (uint32_t)c, nz = (uint32_t)f.nz, ny = (uint32_t)f.ny;
        const uint32_t ij = u / nz;
        k = (int)(u - ij * nz);
        const uint32_t ii = ij / ny;
        j = (int)(ij - ii * ny);
        i = (int)ii;
    } else {
        const int64_t ij = c / f.nz;
        k = (int)(c - ij * f.nz);
        i = (int)(ij / f.ny);
        j = (int)(ij - (int64_t)i * f.ny);
    }
}

// E_a at (i, j, k) = D_a / ((eps + eps[-e_a]) / 2); D points at the D block (Dx, Dy, Dz)
template <typename R>
__host__ __device__ __forceinline__ R efield(const Fdtd<R>& f, const R* __restrict__ D, int a, int i, int j, int k) {
    const int64_t c = cell(f, i, j, k);
    D += a * f.n;
    int64_t cm;
    if (a == 0) cm = cell(f, wrapm(i, f.nx), j, k);
    else if (a == 1) cm = cell(f, i, wrapm(j, f.ny), k);
    else cm = cell(f, i, j, wrapm(k, f.nz));
    const R eaa = (f.eps[c] + f.eps[cm]) / R(2);
    return (R(1) / eaa) * D[c];
}

// Tangent of E_a = D_a / eps_aa along (dD, d_eps): dD_a / eps_aa -
// D_a d(eps_aa) / eps_aa^2, d(eps_aa) = (d_eps + d_eps[-e_a]) / 2.
template <typename R>
__host__ __device__ __forceinline__ R efield_tangent(const Fdtd<R>& f, const R* __restrict__ D,
                                                     const R* __restrict__ dD, const R* __restrict__ deps, int a,
                                                     int i, int j, int k) {
    const int64_t c = cell(f, i, j, k);
    int64_t cm;
    if (a == 0) cm = cell(f, wrapm(i, f.nx), j, k);
    else if (a == 1) cm = cell(f, i, wrapm(j, f.ny), k);
    else cm = cell(f, i, j, wrapm(k, f.nz));
    const R inv = R(1) / ((f.eps[c] + f.eps[cm]) / R(2));
    const R deaa = (deps[c] + deps[cm]) / R(2);
    return inv * dD[a * f.n + c] - (deaa * inv * inv) * D[a * f.n + c];
}

// ----------------------------------------------------------------- H phase
// H update of cell c of the state f.st with E read through `E(a, i, j, k)`
// (the primal E of f.st, or the tangent E of a forward-mode step).
template <typename R, typename EGet>
__host__ __device__ __forceinline__ void h_cell_with(const Fdtd<R>& f, int64_t c, EGet E) {
    int i, j, k;
    decompose(f, c, i, j, k);
    const int64_t n = f.n;
    R* H = f.st;
    R* ICE = f.st + 6 * n;
    R* IH = f.st + 9 * n;
    const int ip = wrapp(i, f.nx), jp = wrapp(j, f.ny), kp = wrapp(k, f.nz);
    const R ex = E(0, i, j, k), ey = E(1, i, j, k), ez = E(2, i, j, k);
    R ce[3];
    // curl_E, ceviche/derivatives.py:16-22 (forward differences, periodic)
    ce[0] = ((E(2, i, jp, k) - ez) - (E(1, i, j, kp) - ey)) * f.inv_dL;
    ce[1] = ((E(0, i, j, kp) - ex) - (E(2, ip, j, k) - ez)) * f.inv_dL;
    ce[2] = ((E(1, ip, j, k) - ey) - (E(0, i, jp, k) - ex)) * f.inv_dL;
    const int pos[3] = {i, j, k};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int ax_b = (a == 0) ? 1 : 0;
        const int ax_c = (a == 2) ? 1 : 2;
        const R sa = sigH(f, a)[pos[a]], sb = sigH(f, ax_b)[pos[ax_b]], sc = sigH(f, ax_c)[pos[ax_c]];
        const int64_t ca = a * n + c;
        const R h = H[ca];
        R hn;
        if (sa == R(0) && sb == R(0) && sc == R(0)) {
            hn = f.h1 * h + f.h2 * ce[a];
        } else {
            R m1, m2, m3, m4;
            coef4<R>(sa, sb, sc, R(f.dt), R(f.c0), R(f.eps0), R(-1), m1, m2, m3, m4);
            hn = m1 * h + m2 * ce[a];
            if (sa != R(0)) {
                const R ice = ICE[ca] + ce[a];
                ICE[ca] = ice;
                hn += m3 * ice;
            }
            if (sb != R(0) && sc != R(0)) {
                const R ih = IH[ca] + h;
                IH[ca] = ih;
                hn += m4 * ih;
            }
        }
        H[ca] = hn;
    }
}

template <typename R>
__host__ __device__ __forceinline__ void h_cell(const Fdtd<R>& f, int64_t c) {
    const R* D = f.st + 3 * f.n;
    h_cell_with(f, c, [&](int a, int i, int j, int k) { return efield(f, D, a, i, j, k); });
}

template <typename R>
__global__ void __launch_bounds__(256) k_h(Fdtd<R> f) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= f.n) return;
    h_cell(f, c);
}

// tangent H update: ft.st is the tangent state, D the primal D of the same step
template <typename R>
__global__ void __launch_bounds__(256) k_h_tangent(Fdtd<R> ft, const R* __restrict__ D, const R* __restrict__ deps) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ft.n) return;
    const R* dD = ft.st + 3 * ft.n;
    h_cell_with(ft, c, [&](int a, int i, int j, int k) { return efield_tangent(ft, D, dD, deps, a, i, j, k); });
}

// ----------------------------------------------------------------- D phase
struct PointSrc {
    int comp;         // 0..2 = Jx, Jy, Jz; -1 none
    int64_t cell;
    double amp;
};

template <typename R>
__host__ __device__ __forceinline__ void d_cell(const Fdtd<R>& f, const R* __restrict__ Jx, const R* __restrict__ Jy,
                                           const R* __restrict__ Jz, R jscale, PointSrc ps, int64_t c) {
    int i, j, k;
    decompose(f, c, i, j, k);
    const int64_t n = f.n;
    const R* H = f.st;
    R* D = f.st + 3 * n;
    R* ICH = f.st + 12 * n;
    R* ID = f.st + 15 * n;
    const int64_t cim = cell(f, wrapm(i, f.nx), j, k), cjm = cell(f, i, wrapm(j, f.ny), k),
                  ckm = cell(f, i, j, wrapm(k, f.nz));
    const R* Hx = H;
    const R* Hy = H + n;
    const R* Hz = H + 2 * n;
    R ch[3];
    // curl_H, ceviche/derivatives.py:24-30 (backward differences, periodic)
    ch[0] = ((Hz[c] - Hz[cjm]) - (Hy[c] - Hy[ckm])) * f.inv_dL;
    ch[1] = ((Hx[c] - Hx[ckm]) - (Hz[c] - Hz[cim])) * f.inv_dL;
    ch[2] = ((Hy[c] - Hy[cim]) - (Hx[c] - Hx[cjm])) * f.inv_dL;
    const int pos[3] = {i, j, k};
    const R* J[3] = {Jx, Jy, Jz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int ax_b = (a == 0) ? 1 : 0;
        const int ax_c = (a == 2) ? 1 : 2;
        const R sa = sigD(f, a)[pos[a]], sb = sigD(f, ax_b)[pos[ax_b]], sc = sigD(f, ax_c)[pos[ax_c]];
        const int64_t ca = a * n + c;
        const R d = D[ca];
        R dn;
        if (sa == R(0) && sb == R(0) && sc == R(0)) {
            dn = f.d1 * d + f.d2 * ch[a];
        } else {
            R m1, m2, m3, m4;
            coef4<R>(sa, sb, sc, R(f.dt), R(f.c0), R(f.eps0), R(1), m1, m2, m3, m4);
            dn = m1 * d + m2 * ch[a];
            if (sa != R(0)) {
                const R ich = ICH[ca] + ch[a];
                ICH[ca] = ich;
                dn += m3 * ich;
            }
            if (sb != R(0) && sc != R(0)) {
                const R idv = ID[ca] + d;
                ID[ca] = idv;
                dn += m4 * idv;
            }
        }
        if (J[a]) dn += jscale * J[a][c];
        if (ps.comp == a && ps.cell == c) dn += R(ps.amp);
        D[ca] = dn;
    }
}

template <typename R>
__global__ void __launch_bounds__(256) k_d(Fdtd<R> f, const R* __restrict__ Jx, const R* __restrict__ Jy,
                                           const R* __restrict__ Jz, R jscale, PointSrc ps) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= f.n) return;
    d_cell(f, Jx, Jy, Jz, jscale, ps, c);
}

// ------------------------------------------------------------ probe record
// comp: 0..2 E, 3..5 D, 6..8 H
template <typename R>
__global__ void k_record(Fdtd<R> f, const int32_t* __restrict__ pcomp, const int64_t* __restrict__ pcell,
                         int np, R* __restrict__ out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const int comp = pcomp[p];
    const int64_t c = pcell[p];
    R v;
    if (comp < 3) {
        int i, j, k;
        decompose(f, c, i, j, k);
        v = efield(f, f.st + 3 * f.n, comp, i, j, k);
    } else if (comp < 6) {
        v = f.st[(int64_t)comp * f.n + c];
    } else {
        v = f.st[(int64_t)(comp - 6) * f.n + c];
    }
    out[p] = v;
}

// tangent probes: E comps from efield_tangent (primal D of f.st, tangent of
// ft.st), D / H comps straight from the tangent state
template <typename R>
__global__ void k_record_tangent(Fdtd<R> f, Fdtd<R> ft, const R* __restrict__ deps,
                                 const int32_t* __restrict__ pcomp, const int64_t* __restrict__ pcell, int np,
                                 R* __restrict__ out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const int comp = pcomp[p];
    const int64_t c = pcell[p];
    R v;
    if (comp < 3) {
        int i, j, k;
        decompose(f, c, i, j, k);
        v = efield_tangent(f, f.st + 3 * f.n, ft.st + 3 * f.n, deps, comp, i, j, k);
    } else if (comp < 6) {
        v = ft.st[(int64_t)comp * f.n + c];
    } else {
        v = ft.st[(int64_t)(comp - 6) * f.n + c];
    }
    out[p] = v;
}

// E = D / eps_aa for all cells (the E fields the reference returns)
template <typename R>
__global__ void __launch_bounds__(256) k_efields(Fdtd<R> f, R* __restrict__ E) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= f.n) return;
    int i, j, k;
    decompose(f, c, i, j, k);
#pragma unroll
    for (int a = 0; a < 3; ++a) E[a * f.n + c] = efield(f, f.st + 3 * f.n, a, i, j, k);
}

// ============================================================== adjoint
// Adjoint state lam (18 arrays, same layout) is the cotangent of the state
// after step t; T holds lam_H_tot between the two phases (3 arrays).
//   phase A:  muCH = (n2+n3) lamD + lamICH           (cells c, c+e)
//             T    = lamH + curlE(muCH)
//             lamICE = m3 T + lamICE;  lamIH = m4 T + lamIH;  lamH = m1 T + lamIH
//   phase B:  muCE = m2 T + lamICE                    (cells c, c-e)
//             q    = curlH(muCE)
//             lamID = n4 lamD + lamID; lamICH = n3 lamD + lamICH
//             lamD = n1 lamD + lamID + mE q;   g_mE += D_{t-1} q
// (m*: H coefficients, n*: D coefficients; curl_E^T = curl_H for periodic
// differences, so the adjoint is a time-reversed Yee step.)
template <typename R>
__host__ __device__ __forceinline__ void dcoef(const Fdtd<R>& f, int kind, int a, const int pos[3], R& m1, R& m2, R& m3,
                                      R& m4, R& sa, R& sb, R& sc) {
    const int ax_b = (a == 0) ? 1 : 0;
    const int ax_c = (a == 2) ? 1 : 2;
    if (kind == 0) {
        sa = sigH(f, a)[pos[a]];
        sb = sigH(f, ax_b)[pos[ax_b]];
        sc = sigH(f, ax_c)[pos[ax_c]];
    } else {
        sa = sigD(f, a)[pos[a]];
        sb = sigD(f, ax_b)[pos[ax_b]];
        sc = sigD(f, ax_c)[pos[ax_c]];
    }
    if (sa == R(0) && sb == R(0) && sc == R(0)) {
        m1 = kind == 0 ? f.h1 : f.d1;
        m2 = kind == 0 ? f.h2 : f.d2;
        m3 = R(0);
        m4 = R(0);
    } else {
        coef4<R>(sa, sb, sc, R(f.dt), R(f.c0), R(f.eps0), kind == 0 ? R(-1) : R(1), m1, m2, m3, m4);
    }
}

template <typename R>
__host__ __device__ __forceinline__ R mu_ch(const Fdtd<R>& f, const R* __restrict__ lam, int a, int i, int j, int k) {
    const int pos[3] = {i, j, k};
    R m1, m2, m3, m4, sa, sb, sc;
    dcoef<R>(f, 1, a, pos, m1, m2, m3, m4, sa, sb, sc);
    const int64_t c = cell(f, i, j, k);
    const R lD = lam[(3 + a) * f.n + c];
    R v = (m2 + m3) * lD;
    // lamICH is only ever fed where sigmaD_a != 0 (it is zero elsewhere)
    if (sa != R(0)) v += lam[(12 + a) * f.n + c];
    return v;
}

template <typename R>
__host__ __device__ __forceinline__ void adj_a_cell(const Fdtd<R>& f, R* __restrict__ lam, R* __restrict__ T, int64_t c) {
    int i, j, k;
    decompose(f, c, i, j, k);
    const int ip = wrapp(i, f.nx), jp = wrapp(j, f.ny), kp = wrapp(k, f.nz);
    const int64_t n = f.n;
    // curl_E applied to muCH (forward differences)
    const R mx = mu_ch(f, lam, 0, i, j, k), my = mu_ch(f, lam, 1, i, j, k), mz = mu_ch(f, lam, 2, i, j, k);
    R ce[3];
    ce[0] = ((mu_ch(f, lam, 2, i, jp, k) - mz) - (mu_ch(f, lam, 1, i, j, kp) - my)) * f.inv_dL;
    ce[1] = ((mu_ch(f, lam, 0, i, j, kp) - mx) - (mu_ch(f, lam, 2, ip, j, k) - mz)) * f.inv_dL;
    ce[2] = ((mu_ch(f, lam, 1, ip, j, k) - my) - (mu_ch(f, lam, 0, i, jp, k) - mx)) * f.inv_dL;
    const int pos[3] = {i, j, k};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        R m1, m2, m3, m4, sa, sb, sc;
        dcoef<R>(f, 0, a, pos, m1, m2, m3, m4, sa, sb, sc);
        const int64_t ca = a * n + c;
        const R t = lam[ca] + ce[a];
        T[ca] = t;
        R lih = lam[(9 + a) * n + c];
        if (sb != R(0) && sc != R(0)) {
            lih = m4 * t + lih;
            lam[(9 + a) * n + c] = lih;
        }
        if (sa != R(0)) lam[(6 + a) * n + c] = m3 * t + lam[(6 + a) * n + c];
        lam[ca] = m1 * t + ((sb != R(0) && sc != R(0)) ? lih : R(0));
    }
}

template <typename R>
__global__ void __launch_bounds__(256) k_adj_a(Fdtd<R> f, R* __restrict__ lam, R* __restrict__ T) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= f.n) return;
    adj_a_cell(f, lam, T, c);
}

template <typename R>
__host__ __device__ __forceinline__ R mu_ce(const Fdtd<R>& f, const R* __restrict__ lam, const R* __restrict__ T, int a,
                                   int i, int j, int k) {
    const int pos[3] = {i, j, k};
    R m1, m2, m3, m4, sa, sb, sc;
    dcoef<R>(f, 0, a, pos, m1, m2, m3, m4, sa, sb, sc);
    const int64_t c = cell(f, i, j, k);
    R v = m2 * T[a * f.n + c];
    if (sa != R(0)) v += lam[(6 + a) * f.n + c];  // updated lamICE = m3 T + lamICE'
    return v;
}

template <typename R>
__host__ __device__ __forceinline__ void adj_b_cell(const Fdtd<R>& f, R* __restrict__ lam, const R* __restrict__ T,
                                               const R* __restrict__ Dprev, R* __restrict__ gmE, int64_t c) {
    int i, j, k;
    decompose(f, c, i, j, k);
    const int im = wrapm(i, f.nx), jm = wrapm(j, f.ny), km = wrapm(k, f.nz);
    const int64_t n = f.n;
    const R ux = mu_ce(f, lam, T, 0, i, j, k), uy = mu_ce(f, lam, T, 1, i, j, k), uz = mu_ce(f, lam, T, 2, i, j, k);
    R q[3];
    // curl_H applied to muCE (backward differences)
    q[0] = ((uz - mu_ce(f, lam, T, 2, i, jm, k)) - (uy - mu_ce(f, lam, T, 1, i, j, km))) * f.inv_dL;
    q[1] = ((ux - mu_ce(f, lam, T, 0, i, j, km)) - (uz - mu_ce(f, lam, T, 2, im, j, k))) * f.inv_dL;
    q[2] = ((uy - mu_ce(f, lam, T, 1, im, j, k)) - (ux - mu_ce(f, lam, T, 0, i, jm, k))) * f.inv_dL;
    const int pos[3] = {i, j, k};
    const int64_t cm[3] = {cell(f, im, j, k), cell(f, i, jm, k), cell(f, i, j, km)};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        R m1, m2, m3, m4, sa, sb, sc;
        dcoef<R>(f, 1, a, pos, m1, m2, m3, m4, sa, sb, sc);
        const int64_t ca = (3 + a) * n + c;
        const R lD = lam[ca];
        R lid = R(0);
        if (sb != R(0) && sc != R(0)) {
            lid = m4 * lD + lam[(15 + a) * n + c];
            lam[(15 + a) * n + c] = lid;
        }
        if (sa != R(0)) lam[(12 + a) * n + c] = m3 * lD + lam[(12 + a) * n + c];
        const R eaa = (f.eps[c] + f.eps[cm[a]]) / R(2);
        const R mE = R(1) / eaa;
        lam[ca] = m1 * lD + lid + mE * q[a];
        if (gmE) gmE[a * n + c] += Dprev[a * n + c] * q[a];
    }
}

template <typename R>
__global__ void __launch_bounds__(256) k_adj_b(Fdtd<R> f, R* __restrict__ lam, const R* __restrict__ T,
                                               const R* __restrict__ Dprev, R* __restrict__ gmE) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= f.n) return;
    adj_b_cell(f, lam, T, Dprev, gmE, c);
}

// Probe cotangent injection: lam(comp, cell) += g[p]; for E probes also
// lamD += mE g and g_mE += D_t g.  One thread per probe; atomics make
// repeated (comp, cell) pairs accumulate.
template <typename R>
__global__ void k_inject(Fdtd<R> f, R* __restrict__ lam, const int32_t* __restrict__ pcomp,
                         const int64_t* __restrict__ pcell, int np, const R* __restrict__ g,
                         const R* __restrict__ Dt, R* __restrict__ gmE) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const int comp = pcomp[p];
    const int64_t c = pcell[p];
    const R gv = g[p];
    if (comp < 3) {
        int i, j, k;
        decompose(f, c, i, j, k);
        int64_t cm;
        if (comp == 0) cm = cell(f, wrapm(i, f.nx), j, k);
        else if (comp == 1) cm = cell(f, i, wrapm(j, f.ny), k);
        else cm = cell(f, i, j, wrapm(k, f.nz));
        const R mE = R(1) / ((f.eps[c] + f.eps[cm]) / R(2));
        atomicAdd(&lam[(3 + comp) * f.n + c], mE * gv);
        if (gmE) atomicAdd(&gmE[comp * f.n + c], Dt[comp * f.n + c] * gv);
    } else if (comp < 6) {
        atomicAdd(&lam[(int64_t)comp * f.n + c], gv);
    } else {
        atomicAdd(&lam[(int64_t)(comp - 6) * f.n + c], gv);
    }
}

// g_eps from g_mE: mE_a = 1/eps_aa, eps_aa = (eps + roll(eps, 1, a))/2
//   g_eps[c] += sum_a 1/2 (g_aa[c] + g_aa[c + e_a]),  g_aa = -mE_a^2 g_mE_a
template <typename R>
__global__ void __launch_bounds__(256) k_eps_grad(Fdtd<R> f, const R* __restrict__ gmE, R* __restrict__ geps) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= f.n) return;
    int i, j, k;
    decompose(f, c, i, j, k);
    auto gaa = [&](int a, int ii, int jj, int kk) -> R {
        const int64_t cc = cell(f, ii, jj, kk);
        int64_t cm;
        if (a == 0) cm = cell(f, wrapm(ii, f.nx), jj, kk);
        else if (a == 1) cm = cell(f, ii, wrapm(jj, f.ny), kk);
        else cm = cell(f, ii, jj, wrapm(kk, f.nz));
        const R mE = R(1) / ((f.eps[cc] + f.eps[cm]) / R(2));
        return -mE * mE * gmE[a * f.n + cc];
    };
    R g = R(0.5) * (gaa(0, i, j, k) + gaa(0, wrapp(i, f.nx), j, k));
    g += R(0.5) * (gaa(1, i, j, k) + gaa(1, i, wrapp(j, f.ny), k));
    g += R(0.5) * (gaa(2, i, j, k) + gaa(2, i, j, wrapp(k, f.nz)));
    geps[c] += g;
}

// Reverse of E = D / eps_aa for a dense cotangent gE[3][n]:
//   gD[a] += mE_a gE[a],  g_mE[a] += D[a] gE[a]
template <typename R>
__global__ void __launch_bounds__(256) k_efields_grad(Fdtd<R> f, const R* __restrict__ gE, R* __restrict__ gD,
                                                      R* __restrict__ gmE) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= f.n) return;
    int i, j, k;
    decompose(f, c, i, j, k);
    const int64_t cm[3] = {cell(f, wrapm(i, f.nx), j, k), cell(f, i, wrapm(j, f.ny), k), cell(f, i, j, wrapm(k, f.nz))};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const R g = gE[a * f.n + c];
        const R mE = R(1) / ((f.eps[c] + f.eps[cm[a]]) / R(2));
        gD[a * f.n + c] += mE * g;
        if (gmE) gmE[a * f.n + c] += f.st[(3 + a) * f.n + c] * g;
    }
}

// ================================================ 2D TM temporal blocking
// Fused, temporally blocked Yee update for nz == 1 with only Hx, Hy, Dz live
// (TM drive: Jz point source, TE fields zero).  One workgroup of WV waves owns
// a region of RR = WV*RPT rows (x) by 64 columns (y, contiguous): lane =
// column, each wave a band of RPT rows.  Cell values (Hx, Hy, Dz, 1/eps and the PML integrals
// ICEx, ICEy, IDz) live in registers; neighbour values are exchanged through
// LDS.  TS steps are advanced per launch; after TS steps the cells at least TS
// away from the region edge are exact (the dependency cone of one step is one
// cell on each side), so the owned tile is the region minus a TS-wide halo.
// Region loads use periodic indices, which also makes small domains exact.
template <typename R>
struct TmArgs {
    int nx, ny;
    int tiles_x, tiles_y;      // tile grid
    int t0, nsteps;            // first global step of this launch, steps (<= TS)
    int64_t src;               // flat source cell or -1
    int np;                    // probes
    R inv_dL;
    R h1, h2, d1, d2;          // interior coefficients
    double dt, c0, eps0;
    const R* sig;
    const R* eps;
    const R* st_in;            // state at the start of the launch (18 arrays)
    R* st_out;                 // state after the launch (a different buffer)
    const double* amp;         // amplitude per global step
    const int32_t* pcomp;      // probe components: 2 Ez, 6 Hx, 7 Hy
    const int64_t* pcell;      // probe flat cells
    R* rec;                    // (steps, np), row t - rec_t0
    int rec_t0;
};

__host__ __device__ __forceinline__ int pmod(int a, int n) {
    int r = a % n;
    return r < 0 ? r + n : r;
}

constexpr int kTmMaxProbes = 256;

template <typename R, int WV, int RPT, int TS>
__global__ void __launch_bounds__(64 * WV) k_tm_block(TmArgs<R> a) {
    constexpr int RR = WV * RPT;
    constexpr int NT = 64 * WV;
    constexpr int RC = 64;
    constexpr int OX = RR - 2 * TS, OY = RC - 2 * TS;  // owned tile
    __shared__ R sEz[RR][RC];
    __shared__ R sHx[RR][RC];
    __shared__ R sHy[RR][RC];
    __shared__ int s_np;
    __shared__ int s_pidx[kTmMaxProbes];
    __shared__ short s_pr[kTmMaxProbes], s_pc[kTmMaxProbes], s_pk[kTmMaxProbes];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    // XCD-aware tile order: consecutive workgroups go to different XCDs, so
    // give each XCD a contiguous run of tiles (better L2 reuse of halos).
    const int nb = gridDim.x;
    int b = blockIdx.x;
    if ((nb & 7) == 0) b = (b & 7) * (nb >> 3) + (b >> 3);
    const int ti = b / a.tiles_y, tj = b - ti * a.tiles_y;
    const int i0 = ti * OX, j0 = tj * OY;
    const int nx = a.nx, ny = a.ny;
    const int64_t n = (int64_t)nx * ny;
    const int gj = pmod(j0 - TS + lane, ny);
    const R* sigHx = a.sig;
    const R* sigHy = a.sig + nx;
    const R* sigDx = a.sig + nx + ny + 1;
    const R* sigDy = a.sig + 2 * nx + ny + 1;
    const R shy = sigHy[gj], sdy = sigDy[gj];
    // Tiles read their halo from st_in and write only their owned cells to
    // st_out: a tile that finishes early never changes what a later tile of
    // the same launch reads.
    const R* Hx = a.st_in;
    const R* Hy = a.st_in + n;
    const R* Dz = a.st_in + 5 * n;
    const R* ICEx = a.st_in + 6 * n;
    const R* ICEy = a.st_in + 7 * n;
    const R* IDz = a.st_in + 17 * n;
    R* oHx = a.st_out;
    R* oHy = a.st_out + n;
    R* oDz = a.st_out + 5 * n;
    R* oICEx = a.st_out + 6 * n;
    R* oICEy = a.st_out + 7 * n;
    R* oIDz = a.st_out + 17 * n;
    const R dt = R(a.dt), c0 = R(a.c0), e0 = R(a.eps0);

    if (threadIdx.x == 0) s_np = 0;
    __syncthreads();
    for (int p = threadIdx.x; p < a.np; p += NT) {
        const int64_t c = a.pcell[p];
        const int pi = (int)(c / ny), pj = (int)(c - (int64_t)pi * ny);
        const int ri = pi - i0, rj = pj - j0;
        if (ri >= 0 && ri < OX && rj >= 0 && rj < OY) {
            const int slot = atomicAdd(&s_np, 1);
            if (slot < kTmMaxProbes) {
                s_pidx[slot] = p;
                s_pr[slot] = (short)(ri + TS);
                s_pc[slot] = (short)(rj + TS);
                s_pk[slot] = (short)a.pcomp[p];
            }
        }
    }

    R hx[RPT], hy[RPT], dz[RPT], me[RPT], icx[RPT], icy[RPT], idz[RPT], shx[RPT], sdx[RPT];
    // update coefficients per cell, formed once per launch (the PML cells'
    // divisions stay out of the time loop): Hx = xa hx + xb cex + xc icx,
    // Hy likewise, Dz = da dz + db chz + dc idz; xc / yc / dc are 0 where the
    // auxiliary term is absent, so one code path serves every cell.
    R xa[RPT], xb[RPT], xc[RPT], ya[RPT], yb[RPT], yc[RPT], da[RPT], db[RPT], dc[RPT];
    int gi[RPT];
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        const int r = w * RPT + q;
        gi[q] = pmod(i0 - TS + r, nx);
        const int64_t c = (int64_t)gi[q] * ny + gj;
        hx[q] = Hx[c];
        hy[q] = Hy[c];
        dz[q] = Dz[c];
        const R e = a.eps[c];
        me[q] = R(1) / ((e + e) / R(2));
        shx[q] = sigHx[gi[q]];
        sdx[q] = sigDx[gi[q]];
        icx[q] = shx[q] != R(0) ? ICEx[c] : R(0);
        icy[q] = shy != R(0) ? ICEy[c] : R(0);
        idz[q] = (sdx[q] != R(0) && sdy != R(0)) ? IDz[c] : R(0);
        sEz[r][lane] = me[q] * dz[q];
        if (shx[q] == R(0) && shy == R(0)) {
            xa[q] = a.h1; xb[q] = a.h2; xc[q] = R(0);
            ya[q] = a.h1; yb[q] = a.h2; yc[q] = R(0);
        } else {
            R m1, m2, m3, m4;
            coef4<R>(shx[q], shy, R(0), dt, c0, e0, R(-1), m1, m2, m3, m4);   // Hx: own sigma sHx
            xa[q] = m1; xb[q] = m2; xc[q] = shx[q] != R(0) ? m3 : R(0);
            coef4<R>(shy, shx[q], R(0), dt, c0, e0, R(-1), m1, m2, m3, m4);   // Hy: own sigma sHy
            ya[q] = m1; yb[q] = m2; yc[q] = shy != R(0) ? m3 : R(0);
        }
        if (sdx[q] == R(0) && sdy == R(0)) {
            da[q] = a.d1; db[q] = a.d2; dc[q] = R(0);
        } else {
            R m1, m2, m3, m4;
            coef4<R>(R(0), sdx[q], sdy, dt, c0, e0, R(1), m1, m2, m3, m4);
            da[q] = m1; db[q] = m2; dc[q] = (sdx[q] != R(0) && sdy != R(0)) ? m4 : R(0);
        }
    }
    const int src_i = a.src >= 0 ? (int)(a.src / ny) : -1;
    const int src_j = a.src >= 0 ? (int)(a.src - (int64_t)src_i * ny) : -1;
    __syncthreads();
    const int np_tile = min(s_np, kTmMaxProbes);

    for (int s = 0; s < a.nsteps; ++s) {
        const int t = a.t0 + s;
        // ---- H update: CEx = (Ez[j+1] - Ez)/dL, CEy = -(Ez[i+1] - Ez)/dL
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int r = w * RPT + q;
            const R ez = me[q] * dz[q];
            const R ez_ip = sEz[min(r + 1, RR - 1)][lane];
            const R ez_jp = sEz[r][min(lane + 1, RC - 1)];
            const R cex = ((ez_jp - ez) - (R(0) - R(0))) * a.inv_dL;
            const R cey = ((R(0) - R(0)) - (ez_ip - ez)) * a.inv_dL;
            icx[q] += cex;
            icy[q] += cey;
            R v = xa[q] * hx[q] + xb[q] * cex;
            v += xc[q] * icx[q];
            hx[q] = v;
            v = ya[q] * hy[q] + yb[q] * cey;
            v += yc[q] * icy[q];
            hy[q] = v;
            sHx[r][lane] = hx[q];
            sHy[r][lane] = hy[q];
        }
        __syncthreads();
        // ---- D update: CHz = (Hy - Hy[i-1])/dL - (Hx - Hx[j-1])/dL
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int r = w * RPT + q;
            const R hy_im = sHy[max(r - 1, 0)][lane];
            const R hx_jm = sHx[r][max(lane - 1, 0)];
            const R chz = ((hy[q] - hy_im) - (hx[q] - hx_jm)) * a.inv_dL;
            idz[q] += dz[q];
            R v = da[q] * dz[q] + db[q] * chz;
            v += dc[q] * idz[q];
            if (gi[q] == src_i && gj == src_j) v += R(a.amp[t]);
            dz[q] = v;
            sEz[r][lane] = me[q] * v;
        }
        __syncthreads();
        if (np_tile > 0) {   // uniform over the workgroup
            for (int p = threadIdx.x; p < np_tile; p += NT) {
                const int k = s_pk[p];
                const R v =
                    (k == 2) ? sEz[s_pr[p]][s_pc[p]] : (k == 6 ? sHx[s_pr[p]][s_pc[p]] : sHy[s_pr[p]][s_pc[p]]);
                a.rec[(int64_t)(t - a.rec_t0) * a.np + s_pidx[p]] = v;
            }
            __syncthreads();   // the next H phase overwrites sHx / sHy
        }
    }
    // ---- store the owned tile
    const int oj = lane - TS;
    if (oj >= 0 && oj < OY && j0 + oj < ny) {
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int r = w * RPT + q;
            const int oi = r - TS;
            if (oi < 0 || oi >= OX || i0 + oi >= nx) continue;
            const int64_t c = (int64_t)(i0 + oi) * ny + (j0 + oj);
            oHx[c] = hx[q];
            oHy[c] = hy[q];
            oDz[c] = dz[q];
            if (shx[q] != R(0)) oICEx[c] = icx[q];
            if (shy != R(0)) oICEy[c] = icy[q];
            if (sdx[q] != R(0) && sdy != R(0)) oIDz[c] = idz[q];
        }
    }
}

// The six arrays the TM kernel reads and writes (Hx, Hy, Dz, ICEx, ICEy, IDz).
static const int kTmArrays[6] = {0, 1, 5, 6, 7, 17};

template <typename R>
static int copy_tm_arrays(R* dst, const R* src, int64_t n, hipStream_t s) {
    for (int k = 0; k < 6; ++k)
        CV_HIP(hipMemcpyAsync(dst + kTmArrays[k] * n, src + kTmArrays[k] * n, sizeof(R) * n,
                              hipMemcpyDeviceToDevice, s));
    return CV_OK;
}

template <typename R, int WV, int RPT, int TS>
static int launch_tm(const Fdtd<R>& f, R* alt, const double* amp_dev, int64_t src, int t0, int n_steps,
                     const int32_t* pcomp, const int64_t* pcell, int np, R* rec, int rec_t0, hipStream_t s) {
    static_assert(WV * RPT > 2 * TS && 64 > 2 * TS, "the owned tile must not be empty");
    TmArgs<R> a;
    a.nx = f.nx;
    a.ny = f.ny;
    constexpr int OX = WV * RPT - 2 * TS, OY = 64 - 2 * TS;
    a.tiles_x = (f.nx + OX - 1) / OX;
    a.tiles_y = (f.ny + OY - 1) / OY;
    a.src = src;
    a.np = np;
    a.inv_dL = f.inv_dL;
    a.h1 = f.h1;
    a.h2 = f.h2;
    a.d1 = f.d1;
    a.d2 = f.d2;
    a.dt = f.dt;
    a.c0 = f.c0;
    a.eps0 = f.eps0;
    a.sig = f.sig;
    a.eps = f.eps;
    a.st_in = f.st;
    a.st_out = alt;
    a.amp = amp_dev;
    a.pcomp = pcomp;
    a.pcell = pcell;
    a.rec = rec;
    a.rec_t0 = rec_t0;
    const int nblocks = a.tiles_x * a.tiles_y;
    if (n_steps <= 0) return CV_OK;
    // alt starts equal to the state, so cells a launch does not write (PML
    // integrals where sigma is 0) agree in both buffers
    int rc = copy_tm_arrays<R>(alt, f.st, f.n, s);
    if (rc != CV_OK) return rc;
    R* cur = f.st;
    R* nxt = alt;
    for (int t = 0; t < n_steps; t += TS) {
        a.t0 = t0 + t;
        a.nsteps = (n_steps - t < TS) ? (n_steps - t) : TS;
        a.st_in = cur;
        a.st_out = nxt;
        k_tm_block<R, WV, RPT, TS><<<nblocks, 64 * WV, 0, s>>>(a);
        R* tmp = cur;
        cur = nxt;
        nxt = tmp;
    }
    if (cur != f.st) return copy_tm_arrays<R>(f.st, cur, f.n, s);
    return CV_OK;
}

template <typename R>
static int dispatch_tm(int cfg, const Fdtd<R>& f, R* alt, const double* amp_dev, int64_t src, int t0, int n_steps,
                       const int32_t* pcomp, const int64_t* pcell, int np, R* rec, int rec_t0, hipStream_t s) {
    switch (cfg) {
        case 0: return launch_tm<R, 4, 4, 2>(f, alt, amp_dev, src, t0, n_steps, pcomp, pcell, np, rec, rec_t0, s);
        case 1: return launch_tm<R, 4, 4, 4>(f, alt, amp_dev, src, t0, n_steps, pcomp, pcell, np, rec, rec_t0, s);
        case 2: return launch_tm<R, 4, 8, 4>(f, alt, amp_dev, src, t0, n_steps, pcomp, pcell, np, rec, rec_t0, s);
        case 3: return launch_tm<R, 4, 8, 8>(f, alt, amp_dev, src, t0, n_steps, pcomp, pcell, np, rec, rec_t0, s);
        case 4: return launch_tm<R, 4, 16, 8>(f, alt, amp_dev, src, t0, n_steps, pcomp, pcell, np, rec, rec_t0, s);
        case 5: return launch_tm<R, 4, 16, 4>(f, alt, amp_dev, src, t0, n_steps, pcomp, pcell, np, rec, rec_t0, s);
        case 6: return launch_tm<R, 8, 4, 8>(f, alt, amp_dev, src, t0, n_steps, pcomp, pcell, np, rec, rec_t0, s);
        case 7: return launch_tm<R, 16, 4, 16>(f, alt, amp_dev, src, t0, n_steps, pcomp, pcell, np, rec, rec_t0, s);
        case 8: return launch_tm<R, 16, 2, 8>(f, alt, amp_dev, src, t0, n_steps, pcomp, pcell, np, rec, rec_t0, s);
        case 9: return launch_tm<R, 16, 4, 8>(f, alt, amp_dev, src, t0, n_steps, pcomp, pcell, np, rec, rec_t0, s);
        case 10: return launch_tm<R, 8, 4, 4>(f, alt, amp_dev, src, t0, n_steps, pcomp, pcell, np, rec, rec_t0, s);
        case 11: return launch_tm<R, 16, 2, 4>(f, alt, amp_dev, src, t0, n_steps, pcomp, pcell, np, rec, rec_t0, s);
        default: break;
    }
    CV_CHECK(false, "unknown TM tile configuration %d", cfg);
}

// ================================================ fused single-pass 3D step
// One kernel per step: the thread of cell c recomputes the new H at c and at
// its three backward neighbours (c - e_x, c - e_y, c - e_z) from the OLD state
// A, then forms the new D at c, writing H, D and the PML integrals of c to the
// NEW state B (ping-pong, so no cell reads a value another thread has already
// advanced).  DRAM traffic is one read of H, D, eps and one write of H, D per
// cell-step; the neighbour values come from cache.
template <typename R>
__host__ __device__ __forceinline__ R h_comp(const Fdtd<R>& f, const R* __restrict__ A, int a, int i, int j, int k,
                                             R& ice_out, R& ih_out) {
    const int64_t n = f.n;
    const R* D = A + 3 * n;
    const int ip = wrapp(i, f.nx), jp = wrapp(j, f.ny), kp = wrapp(k, f.nz);
    R ce;
    if (a == 0)
        ce = ((efield(f, D, 2, i, jp, k) - efield(f, D, 2, i, j, k)) -
              (efield(f, D, 1, i, j, kp) - efield(f, D, 1, i, j, k))) * f.inv_dL;
    else if (a == 1)
        ce = ((efield(f, D, 0, i, j, kp) - efield(f, D, 0, i, j, k)) -
              (efield(f, D, 2, ip, j, k) - efield(f, D, 2, i, j, k))) * f.inv_dL;
    else
        ce = ((efield(f, D, 1, ip, j, k) - efield(f, D, 1, i, j, k)) -
              (efield(f, D, 0, i, jp, k) - efield(f, D, 0, i, j, k))) * f.inv_dL;
    const int pos[3] = {i, j, k};
    const int ax_b = (a == 0) ? 1 : 0;
    const int ax_c = (a == 2) ? 1 : 2;
    const R sa = sigH(f, a)[pos[a]], sb = sigH(f, ax_b)[pos[ax_b]], sc = sigH(f, ax_c)[pos[ax_c]];
    const int64_t ca = a * n + cell(f, i, j, k);
    const R h = A[ca];
    if (sa == R(0) && sb == R(0) && sc == R(0)) return f.h1 * h + f.h2 * ce;
    R m1, m2, m3, m4;
    coef4<R>(sa, sb, sc, R(f.dt), R(f.c0), R(f.eps0), R(-1), m1, m2, m3, m4);
    R hn = m1 * h + m2 * ce;
    if (sa != R(0)) {
        ice_out = A[6 * n + ca] + ce;
        hn += m3 * ice_out;
    }
    if (sb != R(0) && sc != R(0)) {
        ih_out = A[9 * n + ca] + h;
        hn += m4 * ih_out;
    }
    return hn;
}

template <typename R>
__host__ __device__ __forceinline__ void fused_cell(const Fdtd<R>& f, const R* __restrict__ A, R* __restrict__ B,
                                                    PointSrc ps, int64_t c) {
    int i, j, k;
    decompose(f, c, i, j, k);
    const int64_t n = f.n;
    const int im = wrapm(i, f.nx), jm = wrapm(j, f.ny), km = wrapm(k, f.nz);
    R h[3], ice[3], ih[3];
    for (int a = 0; a < 3; ++a) {
        ice[a] = R(0);
        ih[a] = R(0);
        h[a] = h_comp(f, A, a, i, j, k, ice[a], ih[a]);
    }
    R dummy0, dummy1;
    // backward neighbours' new H components used by curl_H at c
    const R hz_jm = h_comp(f, A, 2, i, jm, k, dummy0, dummy1);
    const R hy_km = h_comp(f, A, 1, i, j, km, dummy0, dummy1);
    const R hx_km = h_comp(f, A, 0, i, j, km, dummy0, dummy1);
    const R hz_im = h_comp(f, A, 2, im, j, k, dummy0, dummy1);
    const R hy_im = h_comp(f, A, 1, im, j, k, dummy0, dummy1);
    const R hx_jm = h_comp(f, A, 0, i, jm, k, dummy0, dummy1);
    R ch[3];
    ch[0] = ((h[2] - hz_jm) - (h[1] - hy_km)) * f.inv_dL;
    ch[1] = ((h[0] - hx_km) - (h[2] - hz_im)) * f.inv_dL;
    ch[2] = ((h[1] - hy_im) - (h[0] - hx_jm)) * f.inv_dL;
    const int pos[3] = {i, j, k};
    for (int a = 0; a < 3; ++a) {
        const int64_t ca = a * n + c;
        B[ca] = h[a];
        const int ax_b = (a == 0) ? 1 : 0;
        const int ax_c = (a == 2) ? 1 : 2;
        {
            const R sa = sigH(f, a)[pos[a]], sb = sigH(f, ax_b)[pos[ax_b]], sc = sigH(f, ax_c)[pos[ax_c]];
            if (sa != R(0)) B[6 * n + ca] = ice[a];
            if (sb != R(0) && sc != R(0)) B[9 * n + ca] = ih[a];
        }
        const R sa = sigD(f, a)[pos[a]], sb = sigD(f, ax_b)[pos[ax_b]], sc = sigD(f, ax_c)[pos[ax_c]];
        const R d = A[3 * n + ca];
        R dn;
        if (sa == R(0) && sb == R(0) && sc == R(0)) {
            dn = f.d1 * d + f.d2 * ch[a];
        } else {
            R m1, m2, m3, m4;
            coef4<R>(sa, sb, sc, R(f.dt), R(f.c0), R(f.eps0), R(1), m1, m2, m3, m4);
            dn = m1 * d + m2 * ch[a];
            if (sa != R(0)) {
                const R ich = A[12 * n + ca] + ch[a];
                B[12 * n + ca] = ich;
                dn += m3 * ich;
            }
            if (sb != R(0) && sc != R(0)) {
                const R idv = A[15 * n + ca] + d;
                B[15 * n + ca] = idv;
                dn += m4 * idv;
            }
        }
        if (ps.comp == a && ps.cell == c) dn += R(ps.amp);
        B[3 * n + ca] = dn;
    }
}

template <typename R>
__global__ void __launch_bounds__(256) k_fused(Fdtd<R> f, const R* __restrict__ A, R* __restrict__ B, PointSrc ps,
                                               const double* __restrict__ amp_dev, int t) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= f.n) return;
    if (ps.comp >= 0) ps.amp = amp_dev[t];
    fused_cell(f, A, B, ps, c);
}

// =================================== x-marching single-pass 3D step
// One workgroup (16 / kMRpt waves) owns a region of 16 rows (y) x 64 lanes
// (z) of (y, z) columns and marches along x over a chunk of slices; each
// thread keeps kMRpt cells of the region (consecutive rows, its lane) in
// registers:
// per slice i it forms E(i+1) from the loads of D(i+1) and eps, the new H(i)
// from E(i) (own cell, and the y+1 / z+1 neighbours through LDS) and E(i+1),
// then the new D(i) from H(i) (own, and y-1 / z-1 through LDS) and H(i-1),
// which it carried from the previous slice.  H is valid on region rows 0..14
// and lanes 0..62, D on rows 1..14 and lanes 1..62: the owned 14 x 62 cells.
// Reads come from the old state A only and writes go to B (ping-pong), so
// workgroup order does not matter; a chunk starts by recomputing H(i0-1).
// DRAM traffic per cell-step: H, D, eps read and H, D written (52 B in
// float32) plus the 1 / 14 and 1 / 62 halo fractions.  Same arithmetic as
// h_cell / d_cell.
constexpr int kMRows = 16, kMLanes = 64;
constexpr int kMOwnY = kMRows - 2, kMOwnZ = kMLanes - 2;

template <typename R, int kMRpt>
__global__ void __launch_bounds__(64 * kMRows / kMRpt) k_march(Fdtd<R> f, const R* __restrict__ A, R* __restrict__ B,
                                                             PointSrc ps, const double* __restrict__ amp_dev, int t,
                                                             int tiles_y, int tiles_z, int chunk) {
    __shared__ R sE[3][kMRows][kMLanes];
    __shared__ R sH[3][kMRows][kMLanes];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int b = blockIdx.x;
    const int tz = b % tiles_z;
    b /= tiles_z;
    const int ty = b % tiles_y;
    const int chk = b / tiles_y;
    const int i0 = chk * chunk;
    const int i1 = min(f.nx, i0 + chunk);
    if (i0 >= i1) return;   // uniform over the workgroup
    if (ps.comp >= 0) ps.amp = amp_dev[t];
    const int64_t n = f.n;
    const int ny = f.ny, nz = f.nz;
    const R* Dx = A + 3 * n;
    const R* Dy = A + 4 * n;
    const R* Dz = A + 5 * n;
    // region coordinates of this thread's cells
    const int zz = tz * kMOwnZ - 1 + lane;            // unwrapped z
    const int k = pmod(zz, nz);
    const int km = wrapm(k, nz);
    int j[kMRpt], jm[kMRpt];
    bool own[kMRpt];
    const bool own_z = lane >= 1 && lane <= kMLanes - 2 && zz < nz;
#pragma unroll
    for (int q = 0; q < kMRpt; ++q) {
        const int r = w * kMRpt + q;
        const int yy = ty * kMOwnY - 1 + r;
        j[q] = pmod(yy, ny);
        jm[q] = wrapm(j[q], ny);
        own[q] = own_z && r >= 1 && r <= kMRows - 2 && yy < ny;
    }
    const R sHy_k = sigH(f, 2)[k], sDz_k = sigD(f, 2)[k];
    R sHy_j[kMRpt], sDy_j[kMRpt];
#pragma unroll
    for (int q = 0; q < kMRpt; ++q) {
        sHy_j[q] = sigH(f, 1)[j[q]];
        sDy_j[q] = sigD(f, 1)[j[q]];
    }
    auto cidx = [&](int xi, int q) -> int64_t { return ((int64_t)xi * ny + j[q]) * nz + k; };

    // raw loads of slice xi (D and the four eps values E needs), issued
    // early so that they are in flight during the previous slice's work
    struct Raw {
        R d[kMRpt][3];
        R e0[kMRpt], ex[kMRpt], ey[kMRpt], ez[kMRpt];
    };
    auto load_raw = [&](int xi, Raw& rw) {
        const int xm = wrapm(xi, f.nx);
#pragma unroll
        for (int q = 0; q < kMRpt; ++q) {
            const int64_t c = cidx(xi, q);
            rw.e0[q] = f.eps[c];
            rw.ex[q] = f.eps[((int64_t)xm * ny + j[q]) * nz + k];
            rw.ey[q] = f.eps[((int64_t)xi * ny + jm[q]) * nz + k];
            rw.ez[q] = f.eps[((int64_t)xi * ny + j[q]) * nz + km];
            rw.d[q][0] = Dx[c];
            rw.d[q][1] = Dy[c];
            rw.d[q][2] = Dz[c];
        }
    };
    auto to_E = [&](const Raw& rw, R (&e)[kMRpt][3]) {
#pragma unroll
        for (int q = 0; q < kMRpt; ++q) {
            const R ea = (rw.e0[q] + rw.ex[q]) / R(2);
            const R eb = (rw.e0[q] + rw.ey[q]) / R(2);
            const R eg = (rw.e0[q] + rw.ez[q]) / R(2);
            e[q][0] = (R(1) / ea) * rw.d[q][0];
            e[q][1] = (R(1) / eb) * rw.d[q][1];
            e[q][2] = (R(1) / eg) * rw.d[q][2];
        }
    };
    auto load_H = [&](int xi, R (&h)[kMRpt][3]) {
#pragma unroll
        for (int q = 0; q < kMRpt; ++q) {
            const int64_t c = cidx(xi, q);
#pragma unroll
            for (int a = 0; a < 3; ++a) h[q][a] = A[a * n + c];
        }
    };
    // new H of slice xi at this thread's cells; eC = E(xi) (also in sE),
    // eN = E(xi+1), hO = old H(xi)
    auto h_new = [&](int xi, const R (&eC)[kMRpt][3], const R (&eN)[kMRpt][3], const R (&hO)[kMRpt][3],
                     R (&h)[kMRpt][3], bool store) {
        const R sHx = sigH(f, 0)[xi];
#pragma unroll
        for (int q = 0; q < kMRpt; ++q) {
            const int r = w * kMRpt + q;
            const int rp = min(r + 1, kMRows - 1), lp = min(lane + 1, kMLanes - 1);
            R ce[3];
            ce[0] = ((sE[2][rp][lane] - eC[q][2]) - (sE[1][r][lp] - eC[q][1])) * f.inv_dL;
            ce[1] = ((sE[0][r][lp] - eC[q][0]) - (eN[q][2] - eC[q][2])) * f.inv_dL;
            ce[2] = ((eN[q][1] - eC[q][1]) - (sE[0][rp][lane] - eC[q][0])) * f.inv_dL;
            const int64_t c = cidx(xi, q);
            const R sg[3] = {sHx, sHy_j[q], sHy_k};
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const int ax_b = (a == 0) ? 1 : 0;
                const int ax_c = (a == 2) ? 1 : 2;
                const R sa = sg[a], sb = sg[ax_b], sc = sg[ax_c];
                const int64_t ca = a * n + c;
                const R hv = hO[q][a];
                R hn;
                if (sa == R(0) && sb == R(0) && sc == R(0)) {
                    hn = f.h1 * hv + f.h2 * ce[a];
                } else {
                    R m1, m2, m3, m4;
                    coef4<R>(sa, sb, sc, R(f.dt), R(f.c0), R(f.eps0), R(-1), m1, m2, m3, m4);
                    hn = m1 * hv + m2 * ce[a];
                    if (sa != R(0)) {
                        const R ice = A[6 * n + ca] + ce[a];
                        if (store && own[q]) B[6 * n + ca] = ice;
                        hn += m3 * ice;
                    }
                    if (sb != R(0) && sc != R(0)) {
                        const R ih = A[9 * n + ca] + hv;
                        if (store && own[q]) B[9 * n + ca] = ih;
                        hn += m4 * ih;
                    }
                }
                h[q][a] = hn;
                if (store && own[q]) B[ca] = hn;
            }
        }
    };

    Raw rwA, rwB;
    R eC[kMRpt][3], eN[kMRpt][3], dC[kMRpt][3], hP[kMRpt][3], hC[kMRpt][3], hO[kMRpt][3], hON[kMRpt][3];
    // H(i0 - 1) from E(i0 - 1), E(i0) and old H(i0 - 1)
    const int xprev = wrapm(i0, f.nx);
    load_raw(xprev, rwA);
    load_raw(i0, rwB);
    load_H(xprev, hO);
    to_E(rwA, eC);
    to_E(rwB, eN);
#pragma unroll
    for (int q = 0; q < kMRpt; ++q)
#pragma unroll
        for (int a = 0; a < 3; ++a) sE[a][w * kMRpt + q][lane] = eC[q][a];
    load_raw(wrapp(i0, f.nx), rwA);     // slice i0 + 1 in flight
    __syncthreads();
    h_new(xprev, eC, eN, hO, hP, false);
    load_H(i0, hO);                     // old H(i0) in flight
#pragma unroll
    for (int q = 0; q < kMRpt; ++q)
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            eC[q][a] = eN[q][a];
            dC[q][a] = rwB.d[q][a];
        }
    for (int xi = i0; xi < i1; ++xi) {
        __syncthreads();   // every read of sE / sH of the previous slice is done
#pragma unroll
        for (int q = 0; q < kMRpt; ++q)
#pragma unroll
            for (int a = 0; a < 3; ++a) sE[a][w * kMRpt + q][lane] = eC[q][a];
        to_E(rwA, eN);                      // E(xi + 1), loaded one slice ago
        R dN[kMRpt][3];
#pragma unroll
        for (int q = 0; q < kMRpt; ++q)
#pragma unroll
            for (int a = 0; a < 3; ++a) dN[q][a] = rwA.d[q][a];
        const int xn2 = wrapp(wrapp(xi, f.nx), f.nx);
        if (xi + 1 < i1) {
            load_raw(xn2, rwA);             // slice xi + 2 in flight
            load_H(wrapp(xi, f.nx), hON);   // old H(xi + 1) in flight
        }
        __syncthreads();
        h_new(xi, eC, eN, hO, hC, true);
#pragma unroll
        for (int q = 0; q < kMRpt; ++q)
#pragma unroll
            for (int a = 0; a < 3; ++a) sH[a][w * kMRpt + q][lane] = hC[q][a];
        __syncthreads();
        const R sDx = sigD(f, 0)[xi];
#pragma unroll
        for (int q = 0; q < kMRpt; ++q) {
            const int r = w * kMRpt + q;
            const int rm = max(r - 1, 0), lm = max(lane - 1, 0);
            R ch[3];
            ch[0] = ((hC[q][2] - sH[2][rm][lane]) - (hC[q][1] - sH[1][r][lm])) * f.inv_dL;
            ch[1] = ((hC[q][0] - sH[0][r][lm]) - (hC[q][2] - hP[q][2])) * f.inv_dL;
            ch[2] = ((hC[q][1] - hP[q][1]) - (hC[q][0] - sH[0][rm][lane])) * f.inv_dL;
            const int64_t c = cidx(xi, q);
            const R sg[3] = {sDx, sDy_j[q], sDz_k};
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const int ax_b = (a == 0) ? 1 : 0;
                const int ax_c = (a == 2) ? 1 : 2;
                const R sa = sg[a], sb = sg[ax_b], sc = sg[ax_c];
                const int64_t ca = a * n + c;
                const R dv = dC[q][a];
                R dn;
                if (sa == R(0) && sb == R(0) && sc == R(0)) {
                    dn = f.d1 * dv + f.d2 * ch[a];
                } else {
                    R m1, m2, m3, m4;
                    coef4<R>(sa, sb, sc, R(f.dt), R(f.c0), R(f.eps0), R(1), m1, m2, m3, m4);
                    dn = m1 * dv + m2 * ch[a];
                    if (sa != R(0)) {
                        const R ich = A[12 * n + ca] + ch[a];
                        if (own[q]) B[12 * n + ca] = ich;
                        dn += m3 * ich;
                    }
                    if (sb != R(0) && sc != R(0)) {
                        const R idv = A[15 * n + ca] + dv;
                        if (own[q]) B[15 * n + ca] = idv;
                        dn += m4 * idv;
                    }
                }
                if (ps.comp == a && ps.cell == c) dn += R(ps.amp);
                if (own[q]) B[3 * n + ca] = dn;
            }
        }
#pragma unroll
        for (int q = 0; q < kMRpt; ++q)
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                hP[q][a] = hC[q][a];
                eC[q][a] = eN[q][a];
                dC[q][a] = dN[q][a];
                hO[q][a] = hON[q][a];
            }
    }
}

// ==================================== slab-marching single-pass 3D step
// A workgroup of nz threads (one z line each, nz a multiple of 64 up to 256)
// owns a band of TJ whole (y, z) rows and marches along x over a chunk of
// slices, so z needs no halo and y one row on each side.  Per slice xi it
// holds E(xi) and E(xi+1) of rows -1..TJ (E = D / eps_aa formed once from D
// and eps of the band), forms the new H(xi) of rows -1..TJ-1 (row -1 only
// Hx, Hz, the y-1 values the D update needs), then the new D(xi) of the band
// from H(xi) and the H(xi-1) carried from the previous slice.  The z+1
// neighbours of E and the z-1 neighbours of H go through LDS.  All reads come
// from the old state A and all writes go to B (ping-pong), so neighbouring
// bands may recompute each other's halo rows in any order; a chunk starts by
// recomputing H(i0-1).  HBM traffic per cell-step: H, D, eps read, H, D
// written (52 B in float32, plus the PML integrals inside the PML); the halo
// rows are re-read by the neighbouring band on the same XCD (consecutive
// bands of a chunk are mapped to one XCD).  Same arithmetic as h_cell /
// d_cell, term by term.
constexpr int kSlabMaxNz = 256;

// Load / store at a 32-bit byte offset through a raw buffer resource over
// one slice of one array (base and size uniform, in scalar registers; one
// vector register per address).  An offset outside the slice reads 0 and
// drops the store.  Resource word 3 = 0x00020000, the gfx9 raw-buffer format.
template <typename R>
__device__ __forceinline__ R ld_at(const R* base, uint32_t byte_off, uint32_t bytes) {
    const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, (int)bytes, 0x00020000);
    if constexpr (sizeof(R) == 4)
        return __builtin_bit_cast(R, __builtin_amdgcn_raw_buffer_load_b32(r, (int)byte_off, 0, 0));
    else
        return __builtin_bit_cast(R, __builtin_amdgcn_raw_buffer_load_b64(r, (int)byte_off, 0, 0));
}
template <typename R>
__device__ __forceinline__ void st_at(R* base, uint32_t byte_off, uint32_t bytes, R v) {
    const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, (int)bytes, 0x00020000);
    if constexpr (sizeof(R) == 4)
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, (int)byte_off, 0, 0);
    else
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0)), v),
                                              r, (int)byte_off, 0, 0);
}

template <typename R, int TJ>
__global__ void __launch_bounds__(kSlabMaxNz) k_slab(Fdtd<R> f, const R* __restrict__ A, R* __restrict__ B,
                                                     const R* __restrict__ sig, const R* __restrict__ epsv,
                                                     PointSrc ps, const double* __restrict__ amp_dev, int t,
                                                     int tiles_y, int chunk, int map_bands) {
    __shared__ R sE[2][TJ + 1][kSlabMaxNz];   // Ex of rows 0..TJ-1 at [0][r], Ey of rows -1..TJ-1 at [1][r+1]
    __shared__ R sH[2][TJ][kSlabMaxNz];       // new Hx, Hy of rows 0..TJ-1
    // Workgroup b runs on XCD b % 8; consecutive bands of a chunk go to one
    // XCD in consecutive dispatch order, so halo rows are re-read from the
    // same L2.  map_bands: when the bands split evenly over the 8 XCDs, XCD x
    // takes bands [x T/8, (x+1) T/8) of every chunk (T = bands), which
    // spreads the x-PML chunks over all XCDs; otherwise XCD x takes a range
    // of whole chunks.
    const int b = blockIdx.x;
    int ty, ch;
    if (!map_bands) {   // XCD x takes a contiguous range of whole chunks
        const int nb = gridDim.x;
        const int l = ((nb & 7) == 0) ? (b & 7) * (nb >> 3) + (b >> 3) : b;
        ty = l % tiles_y;
        ch = l / tiles_y;
    } else if ((tiles_y & 7) == 0) {
        const int per = tiles_y >> 3, j = b >> 3;
        ty = (b & 7) * per + j % per;
        ch = j / per;
    } else {
        ty = b % tiles_y;
        ch = b / tiles_y;
    }
    const int i0 = ch * chunk;
    const int i1 = min(f.nx, i0 + chunk);
    if (i0 >= i1) return;   // uniform over the workgroup
    const int64_t n = f.n;
    const int nx = f.nx, ny = f.ny, nz = f.nz;
    const int64_t plane = (int64_t)ny * nz;
    const uint32_t sbytes = (uint32_t)(plane * (int64_t)sizeof(R));   // bytes of one slice
    // the point source as (slice, offset in the slice); slice -1 = none
    int src_x = -1;
    uint32_t src_off = 0;
    R src_amp = R(0);
    if (ps.comp >= 0) {
        src_x = (int)(ps.cell / plane);
        src_off = (uint32_t)(ps.cell - (int64_t)src_x * plane) * (uint32_t)sizeof(R);
        src_amp = R(amp_dev[t]);
    }
    // sigma with workgroup-uniform indices through scalar loads (constant
    // address space), which a vector-memory wait does not wait for
    using csig_t = const __attribute__((address_space(4))) R*;
    const csig_t csig = (csig_t)sig;
    const int k = threadIdx.x;
    const int kp = (k + 1 == nz) ? 0 : k + 1;
    const int km = (k == 0) ? nz - 1 : k - 1;
    // byte offsets in a slice of rows -2..TJ of the band (wrapped) at index r + 2
    uint32_t off[TJ + 3], offm[TJ + 3];
#pragma unroll
    for (int q = 0; q < TJ + 3; ++q) {
        const uint32_t row = (uint32_t)pmod(ty * TJ + q - 2, ny) * (uint32_t)nz;
        off[q] = (row + (uint32_t)k) * (uint32_t)sizeof(R);
        offm[q] = (row + (uint32_t)km) * (uint32_t)sizeof(R);
    }
    R sHyj[TJ + 1], sDyj[TJ];   // sigma_y of rows -1..TJ-1 (H) and 0..TJ-1 (D)
#pragma unroll
    for (int q = 0; q < TJ + 1; ++q) sHyj[q] = csig[nx + pmod(ty * TJ + q - 1, ny)];
#pragma unroll
    for (int q = 0; q < TJ; ++q) sDyj[q] = csig[nx + ny + nz + nx + ty * TJ + q];
    const R sHzk = sig[nx + ny + k], sDzk = sig[2 * (nx + ny) + nz + k];

    // D of rows -1..TJ (Dy of row TJ is not needed), eps of rows -2..TJ, eps at
    // z-1 of rows -1..TJ, all of slice xi
    struct Raw {
        R d[3][TJ + 2];
        R e[TJ + 3];
        R ek[TJ + 2];
    };
    auto load_raw = [&](int xi, Raw& rw) {
        const R* ep = epsv + (int64_t)xi * plane;
        const R* dp = A + 3 * n + (int64_t)xi * plane;
#pragma unroll
        for (int q = 0; q < TJ + 3; ++q) rw.e[q] = ld_at(ep, off[q], sbytes);
#pragma unroll
        for (int q = 0; q < TJ + 2; ++q) {
            rw.ek[q] = ld_at(ep, offm[q + 1], sbytes);
            rw.d[0][q] = ld_at(dp, off[q + 1], sbytes);
            if (q < TJ + 1) rw.d[1][q] = ld_at(dp + n, off[q + 1], sbytes);
            rw.d[2][q] = ld_at(dp + 2 * n, off[q + 1], sbytes);
        }
    };
    // eps of the previous slice, rows -1..TJ, for eps_xx
    auto to_E = [&](const Raw& rw, const R (&ep)[TJ + 2], R (&e)[3][TJ + 2]) {
#pragma unroll
        for (int q = 0; q < TJ + 2; ++q) {
            const R e0 = rw.e[q + 1];
            const R exx = (e0 + ep[q]) / R(2);
            const R eyy = (e0 + rw.e[q]) / R(2);
            const R ezz = (e0 + rw.ek[q]) / R(2);
            e[0][q] = (R(1) / exx) * rw.d[0][q];
            e[1][q] = (q < TJ + 1) ? (R(1) / eyy) * rw.d[1][q] : R(0);
            e[2][q] = (R(1) / ezz) * rw.d[2][q];
        }
    };
    // old H of rows -1..TJ-1 (Hy of row -1 is not needed)
    auto load_H = [&](int xi, R (&h)[3][TJ + 1]) {
        const R* hp = A + (int64_t)xi * plane;
#pragma unroll
        for (int q = 0; q < TJ + 1; ++q) {
            h[0][q] = ld_at(hp, off[q + 1], sbytes);
            h[1][q] = (q > 0) ? ld_at(hp + n, off[q + 1], sbytes) : R(0);
            h[2][q] = ld_at(hp + 2 * n, off[q + 1], sbytes);
        }
    };
    auto put_E = [&](const R (&e)[3][TJ + 2]) {
#pragma unroll
        for (int q = 0; q < TJ + 1; ++q) {
            if (q < TJ) sE[0][q][k] = e[0][q + 1];
            sE[1][q][k] = e[1][q];
        }
    };
    // PML integrals the general H update of slice xi reads (ICE where the
    // component's own sigma is non-zero, IH where both others are), all
    // issued together; an integral that is not needed gets an offset past
    // the slice (reads 0, no memory access).
    auto h_aux_load = [&](int xi, R (&ice)[3][TJ + 1], R (&ih)[3][TJ + 1]) {
        const R* ap = A + (int64_t)xi * plane;
        const R sHx = csig[xi];
#pragma unroll
        for (int q = 0; q < TJ + 1; ++q) {
            const R sg[3] = {sHx, sHyj[q], sHzk};
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const R sa = sg[a], sb = sg[(a == 0) ? 1 : 0], sc = sg[(a == 2) ? 1 : 2];
                ice[a][q] = ld_at(ap + (6 + a) * n, sa != R(0) ? off[q + 1] : sbytes, sbytes);
                ih[a][q] = ld_at(ap + (9 + a) * n, (sb != R(0) && sc != R(0)) ? off[q + 1] : sbytes, sbytes);
            }
        }
    };
    // new H of slice xi, rows -1..TJ-1 (index r + 1), from E(xi) = eC (z+1
    // values in sE), E(xi+1) = eN, the old H and the integrals of
    // h_aux_load; `store` writes the band's rows and their PML integrals to
    // B.  Components: row -1 Hx, Hz only, or (hyz_only) Hy, Hz of rows
    // 0..TJ-1.
    auto h_new = [&](int xi, const R (&eC)[3][TJ + 2], const R (&eN)[3][TJ + 2], const R (&hO)[3][TJ + 1],
                     const R (&ice0)[3][TJ + 1], const R (&ih0)[3][TJ + 1], R (&h)[3][TJ + 1], bool store,
                     bool hyz_only) {
        R* bp = B + (int64_t)xi * plane;
        const R sHx = csig[xi];
#pragma unroll
        for (int q = 0; q < TJ + 1; ++q) {
            const int r = q - 1;
            const R ex = eC[0][q], ey = eC[1][q], ez = eC[2][q];
            R ce[3];
            // curl_E, ceviche/derivatives.py:16-22 (forward differences, periodic)
            ce[0] = ((eC[2][q + 1] - ez) - (sE[1][q][kp] - ey)) * f.inv_dL;
            ce[1] = (r >= 0) ? ((sE[0][r < 0 ? 0 : r][kp] - ex) - (eN[2][q] - ez)) * f.inv_dL : R(0);
            ce[2] = ((eN[1][q] - ey) - (eC[0][q + 1] - ex)) * f.inv_dL;
            const uint32_t o = off[q + 1];
            const R sg[3] = {sHx, sHyj[q], sHzk};
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                if ((r < 0 && a == 1) || (hyz_only && (r < 0 || a == 0))) {
                    h[a][q] = R(0);
                    continue;
                }
                const int ax_b = (a == 0) ? 1 : 0;
                const int ax_c = (a == 2) ? 1 : 2;
                const R sa = sg[a], sb = sg[ax_b], sc = sg[ax_c];
                const int64_t ca = a * n;
                const R hv = hO[a][q];
                R hn;
                if (sa == R(0) && sb == R(0) && sc == R(0)) {
                    hn = f.h1 * hv + f.h2 * ce[a];
                } else {
                    R m1, m2, m3, m4;
                    coef4<R>(sa, sb, sc, R(f.dt), R(f.c0), R(f.eps0), R(-1), m1, m2, m3, m4);
                    hn = m1 * hv + m2 * ce[a];
                    if (sa != R(0)) {
                        const R ice = ice0[a][q] + ce[a];
                        if (store && r >= 0) st_at(bp + 6 * n + ca, o, sbytes, ice);
                        hn += m3 * ice;
                    }
                    if (sb != R(0) && sc != R(0)) {
                        const R ih = ih0[a][q] + hv;
                        if (store && r >= 0) st_at(bp + 9 * n + ca, o, sbytes, ih);
                        hn += m4 * ih;
                    }
                }
                h[a][q] = hn;
                if (store && r >= 0) st_at(bp + ca, o, sbytes, hn);
            }
        }
    };

    // Fast path: a slice outside the x PML in a band whose rows are outside
    // the y PML.  There the only live PML integrals are those of the z axis
    // (ICE of Hz, ICH of Dz, on z-PML lanes), their values are prefetched a
    // slice ahead with the other loads (lanes outside the z PML pass an
    // offset past the slice: no memory access), and the coefficients depend
    // on the lane only: coef4(0, 0, s_z) for the x and y components and
    // coef4(s_z, 0, 0) for z, the same calls with the same arguments as the
    // general update, which the remaining slices use.
    bool band_ypml = false;
#pragma unroll
    for (int q = 0; q < TJ + 1; ++q) band_ypml |= sHyj[q] != R(0);
#pragma unroll
    for (int q = 0; q < TJ; ++q) band_ypml |= sDyj[q] != R(0);
    R cHxy1 = f.h1, cHxy2 = f.h2, cHz1 = f.h1, cHz2 = f.h2, cHz3 = R(0);
    R cDxy1 = f.d1, cDxy2 = f.d2, cDz1 = f.d1, cDz2 = f.d2, cDz3 = R(0);
    if (sHzk != R(0)) {
        R m4;
        coef4<R>(R(0), R(0), sHzk, R(f.dt), R(f.c0), R(f.eps0), R(-1), cHxy1, cHxy2, cHz3, m4);
        coef4<R>(sHzk, R(0), R(0), R(f.dt), R(f.c0), R(f.eps0), R(-1), cHz1, cHz2, cHz3, m4);
    }
    if (sDzk != R(0)) {
        R m4;
        coef4<R>(R(0), R(0), sDzk, R(f.dt), R(f.c0), R(f.eps0), R(1), cDxy1, cDxy2, cDz3, m4);
        coef4<R>(sDzk, R(0), R(0), R(f.dt), R(f.c0), R(f.eps0), R(1), cDz1, cDz2, cDz3, m4);
    }
    // ICE of Hz (rows -1..TJ-1) and ICH of Dz (rows 0..TJ-1) of slice xi
    auto load_zaux = [&](int xi, R (&ice)[TJ + 1], R (&ich)[TJ]) {
        const R* ap = A + (int64_t)xi * plane;
#pragma unroll
        for (int q = 0; q < TJ + 1; ++q) ice[q] = ld_at(ap + 8 * n, sHzk != R(0) ? off[q + 1] : sbytes, sbytes);
#pragma unroll
        for (int q = 0; q < TJ; ++q) ich[q] = ld_at(ap + 14 * n, sDzk != R(0) ? off[q + 2] : sbytes, sbytes);
    };
    // new H of slice xi on the fast path (same rows and components as h_new)
    auto h_fast = [&](int xi, const R (&eC)[3][TJ + 2], const R (&eN)[3][TJ + 2], const R (&hO)[3][TJ + 1],
                      const R (&ice)[TJ + 1], R (&h)[3][TJ + 1]) {
        R* bp = B + (int64_t)xi * plane;
#pragma unroll
        for (int q = 0; q < TJ + 1; ++q) {
            const int r = q - 1;
            const R ex = eC[0][q], ey = eC[1][q], ez = eC[2][q];
            const uint32_t o = off[q + 1];
            const R ce0 = ((eC[2][q + 1] - ez) - (sE[1][q][kp] - ey)) * f.inv_dL;
            const R ce2 = ((eN[1][q] - ey) - (eC[0][q + 1] - ex)) * f.inv_dL;
            h[0][q] = cHxy1 * hO[0][q] + cHxy2 * ce0;
            if (r >= 0) {
                const R ce1 = ((sE[0][r < 0 ? 0 : r][kp] - ex) - (eN[2][q] - ez)) * f.inv_dL;
                h[1][q] = cHxy1 * hO[1][q] + cHxy2 * ce1;
            } else {
                h[1][q] = R(0);
            }
            R hz = cHz1 * hO[2][q] + cHz2 * ce2;
            if (sHzk != R(0)) {
                const R ice_n = ice[q] + ce2;
                if (r >= 0) st_at(bp + 8 * n, o, sbytes, ice_n);
                hz += cHz3 * ice_n;
            }
            h[2][q] = hz;
            if (r >= 0) {
                st_at(bp, o, sbytes, h[0][q]);
                st_at(bp + n, o, sbytes, h[1][q]);
                st_at(bp + 2 * n, o, sbytes, hz);
            }
        }
    };
    R zIce[TJ + 1], zIch[TJ];        // z-PML integrals of slice xi (fast path)

    R eP[TJ + 2];                    // eps of the slice before the one last loaded, rows -1..TJ
    R eC[3][TJ + 2], eN[3][TJ + 2];  // E(xi), E(xi+1)
    R dC[3][TJ], dN[3][TJ];          // old D(xi), D(xi+1) of the band
    R hC[3][TJ + 1];                 // new H(xi)
    R hPy[TJ], hPz[TJ];              // new Hy, Hz of H(xi-1), rows 0..TJ-1
    R hO[3][TJ + 1];                 // old H(xi)
    Raw rw;
    const int xm1 = wrapm(i0, nx), xm2 = wrapm(xm1, nx);
    {
        const R* ep = epsv + (int64_t)xm2 * plane;
#pragma unroll
        for (int q = 0; q < TJ + 2; ++q) eP[q] = ld_at(ep, off[q + 1], sbytes);
    }
    // H(i0-1) from E(i0-1), E(i0) and the old H(i0-1)
    load_raw(xm1, rw);
    to_E(rw, eP, eC);
#pragma unroll
    for (int q = 0; q < TJ + 2; ++q) eP[q] = rw.e[q + 1];
    load_raw(i0, rw);
    to_E(rw, eP, eN);
#pragma unroll
    for (int q = 0; q < TJ + 2; ++q) eP[q] = rw.e[q + 1];
#pragma unroll
    for (int q = 0; q < TJ; ++q)
#pragma unroll
        for (int a = 0; a < 3; ++a) dC[a][q] = rw.d[a][q + 1];
    load_H(xm1, hO);
    put_E(eC);
    {
        R ice[3][TJ + 1], ih[3][TJ + 1];
        h_aux_load(xm1, ice, ih);
        __syncthreads();
        h_new(xm1, eC, eN, hO, ice, ih, hC, false, true);
    }
#pragma unroll
    for (int q = 0; q < TJ; ++q) {
        hPy[q] = hC[1][q + 1];
        hPz[q] = hC[2][q + 1];
    }
    // E(i0), E(i0+1), old H(i0)
    load_raw(wrapp(i0, nx), rw);
    load_H(i0, hO);
#pragma unroll
    for (int q = 0; q < TJ + 2; ++q)
#pragma unroll
        for (int a = 0; a < 3; ++a) eC[a][q] = eN[a][q];
    to_E(rw, eP, eN);
#pragma unroll
    for (int q = 0; q < TJ + 2; ++q) eP[q] = rw.e[q + 1];
#pragma unroll
    for (int q = 0; q < TJ; ++q)
#pragma unroll
        for (int a = 0; a < 3; ++a) dN[a][q] = rw.d[a][q + 1];
    load_zaux(i0, zIce, zIch);
    __syncthreads();   // every read of sE holding E(i0-1) is done
    put_E(eC);
    __syncthreads();

    for (int xi = i0; xi < i1; ++xi) {
        const bool more = xi + 1 < i1;   // uniform
        const bool fast = !band_ypml && csig[xi] == R(0) && csig[nx + ny + nz + xi] == R(0);   // uniform
        R zIceN[TJ + 1], zIchN[TJ];
        if (more) load_zaux(wrapp(xi, nx), zIceN, zIchN);    // issued before the slice xi + 2 loads
        R ice[3][TJ + 1], ih[3][TJ + 1];
        if (!fast) h_aux_load(xi, ice, ih);                  // likewise
        if (more) load_raw(wrapp(wrapp(xi, nx), nx), rw);   // slice xi + 2 in flight
        if (fast) h_fast(xi, eC, eN, hO, zIce, hC);
        else h_new(xi, eC, eN, hO, ice, ih, hC, true, false);
#pragma unroll
        for (int q = 0; q < TJ; ++q) {
            sH[0][q][k] = hC[0][q + 1];
            sH[1][q][k] = hC[1][q + 1];
        }
        if (more) load_H(wrapp(xi, nx), hO);   // old H(xi + 1) in flight
        __syncthreads();
        const int64_t p = (int64_t)xi * plane;
        R* bp = B + p;
        const R* ap = A + p;
        const R sDx = csig[nx + ny + nz + xi];
        R ich0[3][TJ], id0[3][TJ];   // PML integrals of the general D update, loaded together
        if (!fast) {
#pragma unroll
            for (int q = 0; q < TJ; ++q) {
                const R sg[3] = {sDx, sDyj[q], sDzk};
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const R sa = sg[a], sb = sg[(a == 0) ? 1 : 0], sc = sg[(a == 2) ? 1 : 2];
                    ich0[a][q] = ld_at(ap + (12 + a) * n, sa != R(0) ? off[q + 2] : sbytes, sbytes);
                    id0[a][q] = ld_at(ap + (15 + a) * n, (sb != R(0) && sc != R(0)) ? off[q + 2] : sbytes, sbytes);
                }
            }
        }
#pragma unroll
        for (int q = 0; q < TJ; ++q) {
            const int qh = q + 1;   // index of row q in hC
            R ch[3];
            // curl_H, ceviche/derivatives.py:24-30 (backward differences, periodic)
            ch[0] = ((hC[2][qh] - hC[2][qh - 1]) - (hC[1][qh] - sH[1][q][km])) * f.inv_dL;
            ch[1] = ((hC[0][qh] - sH[0][q][km]) - (hC[2][qh] - hPz[q])) * f.inv_dL;
            ch[2] = ((hC[1][qh] - hPy[q]) - (hC[0][qh] - hC[0][qh - 1])) * f.inv_dL;
            const uint32_t o = off[q + 2];
            if (fast) {
                R dn[3];
                dn[0] = cDxy1 * dC[0][q] + cDxy2 * ch[0];
                dn[1] = cDxy1 * dC[1][q] + cDxy2 * ch[1];
                dn[2] = cDz1 * dC[2][q] + cDz2 * ch[2];
                if (sDzk != R(0)) {
                    const R ich = zIch[q] + ch[2];
                    st_at(bp + 14 * n, o, sbytes, ich);
                    dn[2] += cDz3 * ich;
                }
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    if (ps.comp == a && xi == src_x && o == src_off) dn[a] += src_amp;
                    st_at(bp + 3 * n + a * n, o, sbytes, dn[a]);
                }
                continue;
            }
            const R sg[3] = {sDx, sDyj[q], sDzk};
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const int ax_b = (a == 0) ? 1 : 0;
                const int ax_c = (a == 2) ? 1 : 2;
                const R sa = sg[a], sb = sg[ax_b], sc = sg[ax_c];
                const int64_t ca = a * n;
                const R dv = dC[a][q];
                R dn;
                if (sa == R(0) && sb == R(0) && sc == R(0)) {
                    dn = f.d1 * dv + f.d2 * ch[a];
                } else {
                    R m1, m2, m3, m4;
                    coef4<R>(sa, sb, sc, R(f.dt), R(f.c0), R(f.eps0), R(1), m1, m2, m3, m4);
                    dn = m1 * dv + m2 * ch[a];
                    if (sa != R(0)) {
                        const R ich = ich0[a][q] + ch[a];
                        st_at(bp + 12 * n + ca, o, sbytes, ich);
                        dn += m3 * ich;
                    }
                    if (sb != R(0) && sc != R(0)) {
                        const R idv = id0[a][q] + dv;
                        st_at(bp + 15 * n + ca, o, sbytes, idv);
                        dn += m4 * idv;
                    }
                }
                if (ps.comp == a && xi == src_x && o == src_off) dn += src_amp;
                st_at(bp + 3 * n + ca, o, sbytes, dn);
            }
        }
        if (more) {
#pragma unroll
            for (int q = 0; q < TJ; ++q) {
                hPy[q] = hC[1][q + 1];
                hPz[q] = hC[2][q + 1];
            }
#pragma unroll
            for (int q = 0; q < TJ + 2; ++q)
#pragma unroll
                for (int a = 0; a < 3; ++a) eC[a][q] = eN[a][q];
#pragma unroll
            for (int q = 0; q < TJ; ++q)
#pragma unroll
                for (int a = 0; a < 3; ++a) dC[a][q] = dN[a][q];
            to_E(rw, eP, eN);   // E(xi + 2)
#pragma unroll
            for (int q = 0; q < TJ + 2; ++q) eP[q] = rw.e[q + 1];
#pragma unroll
            for (int q = 0; q < TJ; ++q)
#pragma unroll
                for (int a = 0; a < 3; ++a) dN[a][q] = rw.d[a][q + 1];
#pragma unroll
            for (int q = 0; q < TJ + 1; ++q) zIce[q] = zIceN[q];
#pragma unroll
            for (int q = 0; q < TJ; ++q) zIch[q] = zIchN[q];
            put_E(eC);
            __syncthreads();
        }
    }
}

// slab kernel applicability and launch shape: nz a multiple of 64 up to
// kSlabMaxNz; bands of 2 rows when ny is even, else 1 (4-row bands hold
// more registers: fewer waves, slower at 256^3, and float64 spills); chunks
// of 8 slices (the fastest of 8, 16, 32, 64 measured at 256^3), or the
// tuning override (g_slab_target then sets the chunk from a target
// workgroup count)
static int g_slab_rows = 0, g_slab_chunk = 0, g_slab_target = 0, g_slab_map = 0;

static bool slab_shape(int nx, int ny, int nz, int elem, int& tj, int& chunk, int& nb) {
    if (nz % 64 != 0 || nz > kSlabMaxNz) return false;
    if ((int64_t)ny * nz * elem >= ((int64_t)1 << 32)) return false;   // byte offsets in a slice fit 32 bits
    tj = (ny % 2 == 0) ? 2 : 1;
    if (g_slab_rows > 0 && ny % g_slab_rows == 0) tj = g_slab_rows;
    const int tiles_y = ny / tj;
    int nch = (nx + 7) / 8;
    if (g_slab_target > 0) nch = std::max(1, std::min((g_slab_target + tiles_y - 1) / tiles_y, nx / 8));
    chunk = (nx + nch - 1) / nch;
    if (g_slab_chunk > 0) chunk = std::min(nx, g_slab_chunk);
    nch = (nx + chunk - 1) / chunk;
    nb = tiles_y * nch;
    return true;
}

template <typename R>
static void launch_slab(const Fdtd<R>& f, const R* A, R* B, PointSrc ps, const double* amp_dev, int t, int tj,
                        int chunk, int nb, hipStream_t s) {
    const int tiles_y = f.ny / tj;
    if (tj == 4) k_slab<R, 4><<<nb, f.nz, 0, s>>>(f, A, B, f.sig, f.eps, ps, amp_dev, t, tiles_y, chunk, g_slab_map);
    else if (tj == 2) k_slab<R, 2><<<nb, f.nz, 0, s>>>(f, A, B, f.sig, f.eps, ps, amp_dev, t, tiles_y, chunk, g_slab_map);
    else k_slab<R, 1><<<nb, f.nz, 0, s>>>(f, A, B, f.sig, f.eps, ps, amp_dev, t, tiles_y, chunk, g_slab_map);
}

}  // namespace cv

// Launch-shape override of the slab-marching 3D step (band rows 1, 2 or 4,
// slices per chunk, target workgroup count; 0 keeps the automatic choice;
// map_bands 1 gives each XCD a range of bands of every chunk, 0 a range of
// whole chunks), process-wide, for tuning sweeps.
extern "C" int cv_fdtd_slab_tuning(int band_rows, int chunk, int target_workgroups, int map_bands) {
    CV_CHECK(band_rows == 0 || band_rows == 1 || band_rows == 2 || band_rows == 4, "band rows must be 0, 1, 2 or 4");
    CV_CHECK(chunk >= 0 && target_workgroups >= 0, "chunk and target must be >= 0");
    cv::g_slab_rows = band_rows;
    cv::g_slab_chunk = chunk;
    cv::g_slab_target = target_workgroups;
    cv::g_slab_map = map_bands ? 1 : 0;
    return CV_OK;
}

namespace cv {

// ------------------------------------------------------------ host glue
template <typename R>
static int make_fdtd(const cv_fdtd_desc* d, const void* sig, const void* eps, void* state, Fdtd<R>& f) {
    CV_CHECK(d && sig && eps && state, "null pointer argument");
    CV_CHECK(d->nx > 0 && d->ny > 0 && d->nz > 0, "empty grid");
    f.nx = d->nx;
    f.ny = d->ny;
    f.nz = d->nz;
    f.n = (int64_t)d->nx * d->ny * d->nz;
    f.inv_dL = R(1.0 / d->dL);
    f.dt = d->dt;
    f.c0 = d->c0;
    f.eps0 = d->eps0;
    f.sig = (const R*)sig;
    f.eps = (const R*)eps;
    f.st = (R*)state;
    // interior coefficients with the reference formula at sigma = 0, in R
    const R dt = R(d->dt), c0 = R(d->c0), e0 = R(d->eps0);
    const R m0 = (R(1) / dt + (R(0) + R(0)) / R(2) / e0 + R(0) * R(0) * dt / R(4) / (e0 * e0));
    f.h1 = (R(1) / m0 * (R(1) / dt - (R(0) + R(0)) / R(2) / e0 - R(0) * R(0) * dt / R(4) / (e0 * e0)));
    f.h2 = R(-1) / m0 * c0;
    f.d1 = f.h1;
    f.d2 = R(1) / m0 * c0;
    return CV_OK;
}

static inline unsigned int nblk(int64_t n) { return (unsigned int)((n + 255) / 256); }

}  // namespace cv

using namespace cv;

extern "C" int cv_fdtd_step(int dtype, const cv_fdtd_desc* d, const void* sig, const void* eps, void* state,
                            const void* Jx, const void* Jy, const void* Jz, double jscale, int src_comp,
                            int64_t src_cell, const double* amp, int n_steps, const int32_t* probe_comp,
                            const int64_t* probe_cell, int n_probes, void* rec, void* stream) {
    hipStream_t s = as_stream(stream);
    CV_CHECK(n_steps >= 0, "n_steps < 0");
    CV_CHECK(src_comp >= -1 && src_comp < 3, "src_comp must be -1 (none) or 0..2");
    CV_CHECK(n_probes == 0 || (probe_comp && probe_cell && rec), "probes need comp, cell and rec buffers");
    if (dtype == CV_F32) {
        Fdtd<float> f;
        int rc = make_fdtd<float>(d, sig, eps, state, f);
        if (rc) return rc;
        CV_CHECK(src_comp < 0 || (src_cell >= 0 && src_cell < f.n), "source cell out of range");
        for (int t = 0; t < n_steps; ++t) {
            PointSrc ps{src_comp, src_cell, (src_comp >= 0 && amp) ? amp[t] : 0.0};
            k_h<float><<<nblk(f.n), 256, 0, s>>>(f);
            k_d<float><<<nblk(f.n), 256, 0, s>>>(f, (const float*)Jx, (const float*)Jy, (const float*)Jz,
                                                 float(jscale), ps);
            if (n_probes > 0)
                k_record<float><<<(n_probes + 63) / 64, 64, 0, s>>>(f, probe_comp, probe_cell, n_probes,
                                                                    (float*)rec + (int64_t)t * n_probes);
        }
    } else if (dtype == CV_F64) {
        Fdtd<double> f;
        int rc = make_fdtd<double>(d, sig, eps, state, f);
        if (rc) return rc;
        CV_CHECK(src_comp < 0 || (src_cell >= 0 && src_cell < f.n), "source cell out of range");
        for (int t = 0; t < n_steps; ++t) {
            PointSrc ps{src_comp, src_cell, (src_comp >= 0 && amp) ? amp[t] : 0.0};
            k_h<double><<<nblk(f.n), 256, 0, s>>>(f);
            k_d<double><<<nblk(f.n), 256, 0, s>>>(f, (const double*)Jx, (const double*)Jy, (const double*)Jz,
                                                  jscale, ps);
            if (n_probes > 0)
                k_record<double><<<(n_probes + 63) / 64, 64, 0, s>>>(f, probe_comp, probe_cell, n_probes,
                                                                     (double*)rec + (int64_t)t * n_probes);
        }
    } else {
        CV_CHECK(false, "unknown dtype %d", dtype);
    }
    CV_LAUNCH_CHECK();
    return CV_OK;
}

template <typename R>
static int step_tangent_impl(const cv_fdtd_desc* d, const void* sig, const void* eps, void* state, void* tstate,
                             const void* deps, const void* Jx, const void* Jy, const void* Jz, const void* dJx,
                             const void* dJy, const void* dJz, double jscale, int src_comp, int64_t src_cell,
                             const double* amp, const double* damp, int n_steps, const int32_t* probe_comp,
                             const int64_t* probe_cell, int n_probes, void* rec, void* trec, hipStream_t s) {
    Fdtd<R> f, ft;
    int rc = make_fdtd<R>(d, sig, eps, state, f);
    if (rc) return rc;
    rc = make_fdtd<R>(d, sig, eps, tstate, ft);
    if (rc) return rc;
    CV_CHECK(src_comp < 0 || (src_cell >= 0 && src_cell < f.n), "source cell out of range");
    const R* D = f.st + 3 * f.n;
    for (int t = 0; t < n_steps; ++t) {
        PointSrc ps{src_comp, src_cell, (src_comp >= 0 && amp) ? amp[t] : 0.0};
        PointSrc pt{src_comp, src_cell, (src_comp >= 0 && damp) ? damp[t] : 0.0};
        k_h<R><<<nblk(f.n), 256, 0, s>>>(f);
        k_h_tangent<R><<<nblk(f.n), 256, 0, s>>>(ft, D, (const R*)deps);   // primal D still at step t
        k_d<R><<<nblk(f.n), 256, 0, s>>>(f, (const R*)Jx, (const R*)Jy, (const R*)Jz, R(jscale), ps);
        k_d<R><<<nblk(f.n), 256, 0, s>>>(ft, (const R*)dJx, (const R*)dJy, (const R*)dJz, R(jscale), pt);
        if (n_probes > 0) {
            if (rec)
                k_record<R><<<(n_probes + 63) / 64, 64, 0, s>>>(f, probe_comp, probe_cell, n_probes,
                                                                (R*)rec + (int64_t)t * n_probes);
            k_record_tangent<R><<<(n_probes + 63) / 64, 64, 0, s>>>(f, ft, (const R*)deps, probe_comp, probe_cell,
                                                                    n_probes, (R*)trec + (int64_t)t * n_probes);
        }
    }
    CV_LAUNCH_CHECK();
    return CV_OK;
}

extern "C" int cv_fdtd_step_tangent(int dtype, const cv_fdtd_desc* d, const void* sig, const void* eps,
                                    void* state, void* tstate, const void* d_eps, const void* Jx, const void* Jy,
                                    const void* Jz, const void* dJx, const void* dJy, const void* dJz,
                                    double jscale, int src_comp, int64_t src_cell, const double* amp,
                                    const double* d_amp, int n_steps, const int32_t* probe_comp,
                                    const int64_t* probe_cell, int n_probes, void* rec, void* trec, void* stream) {
    CV_CHECK(n_steps >= 0, "n_steps < 0");
    CV_CHECK(state && tstate && d_eps && state != tstate, "null or aliased state / tangent buffers");
    CV_CHECK(src_comp >= -1 && src_comp < 3, "src_comp must be -1 (none) or 0..2");
    CV_CHECK(n_probes == 0 || (probe_comp && probe_cell && trec), "probes need comp, cell and trec buffers");
    hipStream_t s = as_stream(stream);
    if (dtype == CV_F32)
        return step_tangent_impl<float>(d, sig, eps, state, tstate, d_eps, Jx, Jy, Jz, dJx, dJy, dJz, jscale,
                                        src_comp, src_cell, amp, d_amp, n_steps, probe_comp, probe_cell, n_probes,
                                        rec, trec, s);
    if (dtype == CV_F64)
        return step_tangent_impl<double>(d, sig, eps, state, tstate, d_eps, Jx, Jy, Jz, dJx, dJy, dJz, jscale,
                                         src_comp, src_cell, amp, d_amp, n_steps, probe_comp, probe_cell, n_probes,
                                         rec, trec, s);
    CV_CHECK(false, "unknown dtype %d", dtype);
}

template <typename R>
__global__ void __launch_bounds__(256) k_efields_tangent(Fdtd<R> f, const R* __restrict__ dD,
                                                         const R* __restrict__ deps, R* __restrict__ dE) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= f.n) return;
    int i, j, k;
    decompose(f, c, i, j, k);
#pragma unroll
    for (int a = 0; a < 3; ++a) dE[a * f.n + c] = efield_tangent(f, f.st + 3 * f.n, dD, deps, a, i, j, k);
}

extern "C" int cv_fdtd_efields_tangent(int dtype, const cv_fdtd_desc* d, const void* sig, const void* eps,
                                       const void* state, const void* tstate, const void* d_eps, void* dE,
                                       void* stream) {
    CV_CHECK(state && tstate && d_eps && dE, "null pointer argument");
    hipStream_t s = as_stream(stream);
    if (dtype == CV_F32) {
        Fdtd<float> f;
        int rc = make_fdtd<float>(d, sig, eps, const_cast<void*>(state), f);
        if (rc) return rc;
        k_efields_tangent<float><<<nblk(f.n), 256, 0, s>>>(f, (const float*)tstate + 3 * f.n, (const float*)d_eps,
                                                           (float*)dE);
    } else if (dtype == CV_F64) {
        Fdtd<double> f;
        int rc = make_fdtd<double>(d, sig, eps, const_cast<void*>(state), f);
        if (rc) return rc;
        k_efields_tangent<double><<<nblk(f.n), 256, 0, s>>>(f, (const double*)tstate + 3 * f.n,
                                                            (const double*)d_eps, (double*)dE);
    } else {
        CV_CHECK(false, "unknown dtype %d", dtype);
    }
    CV_LAUNCH_CHECK();
    return CV_OK;
}

extern "C" int cv_fdtd_efields(int dtype, const cv_fdtd_desc* d, const void* sig, const void* eps,
                               const void* state, void* E, void* stream) {
    hipStream_t s = as_stream(stream);
    CV_CHECK(E, "null E");
    if (dtype == CV_F32) {
        Fdtd<float> f;
        int rc = make_fdtd<float>(d, sig, eps, (void*)state, f);
        if (rc) return rc;
        k_efields<float><<<nblk(f.n), 256, 0, s>>>(f, (float*)E);
    } else if (dtype == CV_F64) {
        Fdtd<double> f;
        int rc = make_fdtd<double>(d, sig, eps, (void*)state, f);
        if (rc) return rc;
        k_efields<double><<<nblk(f.n), 256, 0, s>>>(f, (double*)E);
    } else {
        CV_CHECK(false, "unknown dtype %d", dtype);
    }
    CV_LAUNCH_CHECK();
    return CV_OK;
}

// g_src[0] = lam(D_comp, cell): the cotangent of a point source amplitude
// added to D at this step (the lam after the step's probe injection).
template <typename R>
__global__ void k_gather_src(const R* __restrict__ lam, int64_t idx, R* __restrict__ g_src) {
    if (threadIdx.x == 0 && blockIdx.x == 0) g_src[0] = lam[idx];
}

template <typename R>
static int adjoint_step_impl(const cv_fdtd_desc* d, const void* sig, const void* eps, void* lam_, void* tmp,
                             const void* D_prev, void* g_mE, const int32_t* probe_comp, const int64_t* probe_cell,
                             int n_probes, const void* g_probe, const void* D_t, int src_comp, int64_t src_cell,
                             void* g_src, hipStream_t s) {
    Fdtd<R> f;
    int rc = make_fdtd<R>(d, sig, eps, lam_, f);
    if (rc) return rc;
    R* lam = (R*)lam_;
    if (n_probes > 0)
        k_inject<R><<<(n_probes + 255) / 256, 256, 0, s>>>(f, lam, probe_comp, probe_cell, n_probes,
                                                           (const R*)g_probe, (const R*)D_t, (R*)g_mE);
    if (g_src) {
        CV_CHECK(src_comp >= 0 && src_comp < 3 && src_cell >= 0 && src_cell < f.n,
                 "source gradient requested for an invalid source (%d, %lld)", src_comp, (long long)src_cell);
        k_gather_src<R><<<1, 64, 0, s>>>(lam, (int64_t)(3 + src_comp) * f.n + src_cell, (R*)g_src);
    }
    k_adj_a<R><<<nblk(f.n), 256, 0, s>>>(f, lam, (R*)tmp);
    k_adj_b<R><<<nblk(f.n), 256, 0, s>>>(f, lam, (const R*)tmp, (const R*)D_prev, (R*)g_mE);
    CV_LAUNCH_CHECK();
    return CV_OK;
}

extern "C" int cv_fdtd_adjoint_step(int dtype, const cv_fdtd_desc* d, const void* sig, const void* eps,
                                    void* lam, void* tmp, const void* D_prev, void* g_mE, const int32_t* probe_comp,
                                    const int64_t* probe_cell, int n_probes, const void* g_probe,
                                    const void* D_t, int src_comp, int64_t src_cell, void* g_src, void* stream) {
    hipStream_t s = as_stream(stream);
    CV_CHECK(lam && tmp, "null adjoint buffers");
    if (dtype == CV_F32)
        return adjoint_step_impl<float>(d, sig, eps, lam, tmp, D_prev, g_mE, probe_comp, probe_cell, n_probes,
                                        g_probe, D_t, src_comp, src_cell, g_src, s);
    if (dtype == CV_F64)
        return adjoint_step_impl<double>(d, sig, eps, lam, tmp, D_prev, g_mE, probe_comp, probe_cell, n_probes,
                                         g_probe, D_t, src_comp, src_cell, g_src, s);
    CV_CHECK(false, "unknown dtype %d", dtype);
}

extern "C" int cv_fdtd_eps_grad(int dtype, const cv_fdtd_desc* d, const void* sig, const void* eps,
                                const void* g_mE, void* g_eps, void* stream) {
    hipStream_t s = as_stream(stream);
    CV_CHECK(g_mE && g_eps, "null gradient buffers");
    if (dtype == CV_F32) {
        Fdtd<float> f;
        int rc = make_fdtd<float>(d, sig, eps, (void*)g_mE, f);
        if (rc) return rc;
        k_eps_grad<float><<<nblk(f.n), 256, 0, s>>>(f, (const float*)g_mE, (float*)g_eps);
    } else if (dtype == CV_F64) {
        Fdtd<double> f;
        int rc = make_fdtd<double>(d, sig, eps, (void*)g_mE, f);
        if (rc) return rc;
        k_eps_grad<double><<<nblk(f.n), 256, 0, s>>>(f, (const double*)g_mE, (double*)g_eps);
    } else {
        CV_CHECK(false, "unknown dtype %d", dtype);
    }
    CV_LAUNCH_CHECK();
    return CV_OK;
}

extern "C" int cv_fdtd_efields_grad(int dtype, const cv_fdtd_desc* d, const void* sig, const void* eps,
                                    const void* state, const void* gE, void* gD, void* g_mE, void* stream) {
    hipStream_t s = as_stream(stream);
    CV_CHECK(gE && gD, "null gradient buffers");
    if (dtype == CV_F32) {
        Fdtd<float> f;
        int rc = make_fdtd<float>(d, sig, eps, (void*)state, f);
        if (rc) return rc;
        k_efields_grad<float><<<nblk(f.n), 256, 0, s>>>(f, (const float*)gE, (float*)gD, (float*)g_mE);
    } else if (dtype == CV_F64) {
        Fdtd<double> f;
        int rc = make_fdtd<double>(d, sig, eps, (void*)state, f);
        if (rc) return rc;
        k_efields_grad<double><<<nblk(f.n), 256, 0, s>>>(f, (const double*)gE, (double*)gD, (double*)g_mE);
    } else {
        CV_CHECK(false, "unknown dtype %d", dtype);
    }
    CV_LAUNCH_CHECK();
    return CV_OK;
}

extern "C" int cv_fdtd_tm_configs(int* rpt, int* ts, int* waves, int max_n) {
    static const int R_[12] = {4, 4, 8, 8, 16, 16, 4, 4, 2, 4, 4, 2};
    static const int T_[12] = {2, 4, 4, 8, 8, 4, 8, 16, 8, 8, 4, 4};
    static const int W_[12] = {4, 4, 4, 4, 4, 4, 8, 16, 16, 16, 8, 16};
    for (int i = 0; i < 12 && i < max_n; ++i) {
        if (rpt) rpt[i] = R_[i];
        if (ts) ts[i] = T_[i];
        if (waves) waves[i] = W_[i];
    }
    return 12;
}

extern "C" int cv_fdtd_run_tm(int dtype, const cv_fdtd_desc* d, const void* sig, const void* eps, void* state,
                              void* state_alt, int64_t src_cell, const double* amp_dev, int t0, int n_steps,
                              const int32_t* probe_comp, const int64_t* probe_cell, int n_probes, void* rec,
                              int rec_t0, int config, void* stream) {
    hipStream_t s = as_stream(stream);
    CV_CHECK(state && state_alt && state != state_alt, "two distinct state buffers are needed");
    CV_CHECK(d && d->nz == 1, "the TM kernel needs nz == 1");
    CV_CHECK(n_probes <= kTmMaxProbes, "the TM kernel records at most %d probes", kTmMaxProbes);
    CV_CHECK(n_steps >= 0 && t0 >= 0, "bad step range");
    CV_CHECK(n_probes == 0 || (probe_comp && probe_cell && rec), "probes need comp, cell and rec buffers");
    CV_CHECK(src_cell < 0 || amp_dev, "a source needs its amplitude array");
    if (dtype == CV_F32) {
        Fdtd<float> f;
        int rc = make_fdtd<float>(d, sig, eps, state, f);
        if (rc) return rc;
        CV_CHECK(src_cell < f.n, "source cell out of range");
        rc = dispatch_tm<float>(config, f, (float*)state_alt, amp_dev, src_cell, t0, n_steps, probe_comp, probe_cell,
                                n_probes, (float*)rec, rec_t0, s);
        if (rc) return rc;
    } else if (dtype == CV_F64) {
        Fdtd<double> f;
        int rc = make_fdtd<double>(d, sig, eps, state, f);
        if (rc) return rc;
        CV_CHECK(src_cell < f.n, "source cell out of range");
        rc = dispatch_tm<double>(config, f, (double*)state_alt, amp_dev, src_cell, t0, n_steps, probe_comp,
                                 probe_cell, n_probes, (double*)rec, rec_t0, s);
        if (rc) return rc;
    } else {
        CV_CHECK(false, "unknown dtype %d", dtype);
    }
    CV_LAUNCH_CHECK();
    return CV_OK;
}

extern "C" int cv_fdtd_run_fused(int dtype, const cv_fdtd_desc* d, const void* sig, const void* eps, void* state_a,
                                 void* state_b, int comp, int64_t src_cell, const double* amp_dev, int t0,
                                 int n_steps, const int32_t* probe_comp, const int64_t* probe_cell, int n_probes,
                                 void* rec, int rec_t0, void* stream) {
    hipStream_t s = as_stream(stream);
    CV_CHECK(state_a && state_b && state_a != state_b, "two distinct state buffers are needed");
    CV_CHECK(comp >= -1 && comp < 3, "comp must be -1 (none) or 0..2");
    CV_CHECK(comp < 0 || amp_dev, "a source needs its amplitude array");
    CV_CHECK(n_probes == 0 || (probe_comp && probe_cell && rec), "probes need comp, cell and rec buffers");
    if (dtype == CV_F32) {
        Fdtd<float> f;
        int rc = make_fdtd<float>(d, sig, eps, state_a, f);
        if (rc) return rc;
        CV_CHECK(comp < 0 || (src_cell >= 0 && src_cell < f.n), "source cell out of range");
        float* A = (float*)state_a;
        float* B = (float*)state_b;
        for (int t = 0; t < n_steps; ++t) {
            PointSrc ps{comp, src_cell, 0.0};
            int tj, chunk_s, nb_s;
            if (f.nz > 1 && slab_shape(f.nx, f.ny, f.nz, 4, tj, chunk_s, nb_s)) {
                launch_slab<float>(f, A, B, ps, amp_dev, t0 + t, tj, chunk_s, nb_s, s);
            } else if (f.nz > 1) {
                const int ty = (f.ny + kMOwnY - 1) / kMOwnY, tz = (f.nz + kMOwnZ - 1) / kMOwnZ;
                const int nch = std::max(1, std::min(f.nx / 4 > 0 ? f.nx / 4 : 1, (2048 + ty * tz - 1) / (ty * tz)));
                const int chunk = (f.nx + nch - 1) / nch;
                const int nb = ty * tz * ((f.nx + chunk - 1) / chunk);
                k_march<float, 1><<<nb, 64 * kMRows, 0, s>>>(f, A, B, ps, amp_dev, t0 + t, ty, tz, chunk);
            } else {
                k_fused<float><<<nblk(f.n), 256, 0, s>>>(f, A, B, ps, amp_dev, t0 + t);
            }
            if (n_probes > 0) {
                Fdtd<float> g = f;
                g.st = B;
                k_record<float><<<(n_probes + 63) / 64, 64, 0, s>>>(g, probe_comp, probe_cell, n_probes,
                                                                    (float*)rec + (int64_t)(t0 + t - rec_t0) * n_probes);
            }
            float* tmp = A;
            A = B;
            B = tmp;
        }
    } else if (dtype == CV_F64) {
        Fdtd<double> f;
        int rc = make_fdtd<double>(d, sig, eps, state_a, f);
        if (rc) return rc;
        CV_CHECK(comp < 0 || (src_cell >= 0 && src_cell < f.n), "source cell out of range");
        double* A = (double*)state_a;
        double* B = (double*)state_b;
        for (int t = 0; t < n_steps; ++t) {
            PointSrc ps{comp, src_cell, 0.0};
            int tj, chunk_s, nb_s;
            if (f.nz > 1 && slab_shape(f.nx, f.ny, f.nz, 8, tj, chunk_s, nb_s)) {
                launch_slab<double>(f, A, B, ps, amp_dev, t0 + t, tj, chunk_s, nb_s, s);
            } else if (f.nz > 1) {
                const int ty = (f.ny + kMOwnY - 1) / kMOwnY, tz = (f.nz + kMOwnZ - 1) / kMOwnZ;
                const int nch = std::max(1, std::min(f.nx / 4 > 0 ? f.nx / 4 : 1, (2048 + ty * tz - 1) / (ty * tz)));
                const int chunk = (f.nx + nch - 1) / nch;
                const int nb = ty * tz * ((f.nx + chunk - 1) / chunk);
                k_march<double, 1><<<nb, 64 * kMRows, 0, s>>>(f, A, B, ps, amp_dev, t0 + t, ty, tz, chunk);
            } else {
                k_fused<double><<<nblk(f.n), 256, 0, s>>>(f, A, B, ps, amp_dev, t0 + t);
            }
            if (n_probes > 0) {
                Fdtd<double> g = f;
                g.st = B;
                k_record<double><<<(n_probes + 63) / 64, 64, 0, s>>>(g, probe_comp, probe_cell, n_probes,
                                                                     (double*)rec + (int64_t)(t0 + t - rec_t0) * n_probes);
            }
            double* tmp = A;
            A = B;
            B = tmp;
        }
    } else {
        CV_CHECK(false, "unknown dtype %d", dtype);
    }
    CV_LAUNCH_CHECK();
    return CV_OK;
}
