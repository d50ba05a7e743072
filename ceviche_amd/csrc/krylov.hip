// Device-resident batched BiCGSTAB for the FDFD operators.
//
// Replaces the sparse direct solve behind ceviche's sp_solve
// (ceviche/primitives.py:103-119 -> solvers.py:37-62).  One iteration is
// three fused kernels over all instances of the batch:
//   K1  p' = r + beta (p - omega v)  (or p' = r on restart), v' = A p',
//       partial (rh, v')                                   -> alpha
//   K2  s = r - alpha v' (on the fly), t = A s, partial (t, s), (t, t) -> omega
//   K3  x += alpha p' + omega s, r = s - omega t, partial (rh, r), (r, r)
//                                                          -> beta, convergence
// Reductions are FP64 regardless of the vector precision; the last workgroup
// of each instance turns the partial sums into the next scalar (no host round
// trip).  p and v are double-buffered because K1 reads neighbours of the old
// p / v while writing the new ones.
#include "fdfd_stencil.h"

#include <vector>
#include <algorithm>
#include <mutex>

namespace cv {

constexpr int kCPT = CV_TILE_CELLS / kBlock;

enum : int32_t { kRunning = 0, kConverged = 1, kMaxIter = 2, kNonFinite = 3, kBreakdown = 4 };

struct KState {
    cx<double> rho, alpha, omega, beta;
    double rr, bb, tol2, rhrh;
    double rr_true, scale;        // true residual norm^2 at the last check; rhs scale (mixed)
    int32_t iter, done, restart, max_iter;
    int32_t restarts, skip, pad1, pad2;   // skip: K2 / K3 pass over a restarted iteration
};

// An exact (rh, v) = 0 restarts the iteration with rh = r; after this many
// restarts in one solve it is reported as a breakdown.
constexpr int kMaxRestarts = 64;

// Scratch of the two-level last-workgroup reductions: per-tile partials,
// per-group partials and counters (indexed by the group's first tile) and one
// counter per instance.  Counters must be zero before the first kernel.
struct RedBuf {
    double* partials;      // n_tiles * 4
    double* gpartials;     // n_tiles * 4
    unsigned int* counters;  // n_inst
    unsigned int* gcount;    // n_tiles
};

__global__ void k_zero_u32(unsigned int* p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// mode 0: fresh start (r = b - A x, state reset, iteration count reset)
// mode 1: verification of instances flagged converged (true residual; restart
//         the instance if it is above tolerance), report-only for the others
template <typename R, bool TRANS>
__global__ void __launch_bounds__(kBlock)
k_init(const cv_fdfd_desc* __restrict__ desc, const cv_tile* __restrict__ tiles,
       const cx<R>* __restrict__ prof, const R* __restrict__ m0, const R* __restrict__ m1,
       const cx<R>* __restrict__ b, const cx<R>* __restrict__ x, cx<R>* __restrict__ r,
       cx<R>* __restrict__ rh, RedBuf rb,
       KState* __restrict__ state, double tol, int max_iter, int mode) {
    __shared__ double lds[kWaves * 2];
    const cv_tile t = tiles[blockIdx.x];
    const cv_fdfd_desc d = load_desc(desc, t.inst);
    const int sys = system_of(d, t.inst);
    KState* st = state + sys;
    const int done0 = (mode == 0) ? kRunning : st->done;
    if (mode == 1 && done0 == kRunning) return;
    const bool write_r = (mode == 0) || (done0 == kConverged);
    const int n = d.nx * d.ny;
    auto fx = [&](int64_t k) { return x[k]; };
    double acc[2] = {0.0, 0.0};
#pragma unroll
    for (int q = 0; q < kCPT; ++q) {
        const int loc = t.first + threadIdx.x + q * kBlock;
        if (loc >= n) break;
        const int i = loc / d.ny, j = loc - i * d.ny;
        const int64_t k = d.cell_off + loc;
        const cx<R> bk = b[k];
        const cx<R> rk = bk - apply_any<R, TRANS>(d, prof, m0, m1, i, j, fx, x[k]);
        if (write_r) {
            r[k] = rk;
            rh[k] = rk;
        }
        acc[0] += norm2(ccast<double>(rk));
        acc[1] += norm2(ccast<double>(bk));
    }
    block_sum<2>(acc, lds);
    double tot[2];
    if (!last_block_reduce<2>(acc, rb.partials, rb.gpartials, rb.counters + sys, rb.gcount, blockIdx.x, d.tile_first, d.tile_count,
                              lds, tot))
        return;
    if (threadIdx.x != 0) return;
    const double rr = tot[0], bb = tot[1];
    if (mode == 0) {
        st->bb = bb;
        st->tol2 = tol * tol * bb;
        st->max_iter = max_iter;
        st->iter = 0;
        st->restarts = 0;
    }
    st->rr_true = rr;
    if (mode == 1 && done0 != kConverged) return;  // report only
    st->rr = rr;
    st->rho = mk(rr, 0.0);
    st->rhrh = rr;
    st->restart = 1;
    st->skip = 0;
    if (!isfinite(rr)) {
        st->done = kNonFinite;
    } else if (rr <= st->tol2) {
        st->done = kConverged;
    } else {
        st->done = kRunning;
        if (mode == 1) st->restarts += 1;
        if (st->iter >= st->max_iter) st->done = kMaxIter;
    }
}

// K1: p' = r + beta (p - omega v) (p' = r and rh = r on restart); v' = A p'.
template <typename R, bool TRANS>
__global__ void __launch_bounds__(kBlock)
k_bicg_p(const cv_fdfd_desc* __restrict__ desc, const cv_tile* __restrict__ tiles,
         const cx<R>* __restrict__ prof, const R* __restrict__ m0, const R* __restrict__ m1,
         const cx<R>* __restrict__ r, cx<R>* __restrict__ rh, const cx<R>* __restrict__ p_old,
         const cx<R>* __restrict__ v_old, cx<R>* __restrict__ p_new, cx<R>* __restrict__ v_new,
         RedBuf rb, KState* __restrict__ state) {
    __shared__ double lds[kWaves * 2];
    const cv_tile t = tiles[blockIdx.x];
    const cv_fdfd_desc d = load_desc(desc, t.inst);
    const int sys = system_of(d, t.inst);
    KState* st = state + sys;
    if (st->done != kRunning) return;
    const bool restart = st->restart != 0;
    const cx<R> beta = ccast<R>(st->beta), omega = ccast<R>(st->omega);
    const int n = d.nx * d.ny;
    auto fp = [&](int64_t k) -> cx<R> {
        if (restart) return r[k];
        return r[k] + beta * (p_old[k] - omega * v_old[k]);
    };
    double acc[2] = {0.0, 0.0};
#pragma unroll
    for (int q = 0; q < kCPT; ++q) {
        const int loc = t.first + threadIdx.x + q * kBlock;
        if (loc >= n) break;
        const int i = loc / d.ny, j = loc - i * d.ny;
        const int64_t k = d.cell_off + loc;
        const cx<R> pk = fp(k);
        p_new[k] = pk;
        const cx<R> vk = apply_any<R, TRANS>(d, prof, m0, m1, i, j, fp, pk);
        v_new[k] = vk;
        cx<R> rhk;
        if (restart) {
            rhk = r[k];
            rh[k] = rhk;
        } else {
            rhk = rh[k];
        }
        // (rh, v) = sum conj(rh) v
        acc[0] += (double)rhk.re * (double)vk.re + (double)rhk.im * (double)vk.im;
        acc[1] += (double)rhk.re * (double)vk.im - (double)rhk.im * (double)vk.re;
    }
    block_sum<2>(acc, lds);
    double tot[2];
    if (!last_block_reduce<2>(acc, rb.partials, rb.gpartials, rb.counters + sys, rb.gcount, blockIdx.x, d.tile_first, d.tile_count,
                              lds, tot))
        return;
    if (threadIdx.x != 0) return;
    const cx<double> rhv = mk(tot[0], tot[1]);
    st->restart = 0;
    if (!cfinite(rhv)) {
        st->done = kNonFinite;
        return;
    }
    if (rhv.re == 0.0 && rhv.im == 0.0) {
        // (rh, v) = 0 exactly: restart with rh = p = r (a shadow residual with
        // a single non-zero, as a point right-hand side gives, meets this)
        st->skip = 1;
        st->restart = 1;
        st->rho = mk(st->rr, 0.0);
        st->rhrh = st->rr;
        st->restarts += 1;
        if (st->restarts > kMaxRestarts) st->done = kBreakdown;
        return;
    }
    st->skip = 0;
    st->alpha = cdiv(st->rho, rhv);
}

// K2: s = r - alpha v'; t = A s; (t, s), (t, t) -> omega.
template <typename R, bool TRANS>
__global__ void __launch_bounds__(kBlock)
k_bicg_s(const cv_fdfd_desc* __restrict__ desc, const cv_tile* __restrict__ tiles,
         const cx<R>* __restrict__ prof, const R* __restrict__ m0, const R* __restrict__ m1,
         const cx<R>* __restrict__ r, const cx<R>* __restrict__ v, cx<R>* __restrict__ tv,
         RedBuf rb, KState* __restrict__ state) {
    __shared__ double lds[kWaves * 3];
    const cv_tile t = tiles[blockIdx.x];
    const cv_fdfd_desc d = load_desc(desc, t.inst);
    const int sys = system_of(d, t.inst);
    KState* st = state + sys;
    if (st->done != kRunning || st->skip) return;
    const cx<R> alpha = ccast<R>(st->alpha);
    const int n = d.nx * d.ny;
    auto fs = [&](int64_t k) -> cx<R> { return r[k] - alpha * v[k]; };
    double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < kCPT; ++q) {
        const int loc = t.first + threadIdx.x + q * kBlock;
        if (loc >= n) break;
        const int i = loc / d.ny, j = loc - i * d.ny;
        const int64_t k = d.cell_off + loc;
        const cx<R> sk = fs(k);
        const cx<R> tk = apply_any<R, TRANS>(d, prof, m0, m1, i, j, fs, sk);
        tv[k] = tk;
        acc[0] += (double)tk.re * (double)sk.re + (double)tk.im * (double)sk.im;   // conj(t) s
        acc[1] += (double)tk.re * (double)sk.im - (double)tk.im * (double)sk.re;
        acc[2] += norm2(ccast<double>(tk));
    }
    block_sum<3>(acc, lds);
    double tot[3];
    if (!last_block_reduce<3>(acc, rb.partials, rb.gpartials, rb.counters + sys, rb.gcount, blockIdx.x, d.tile_first, d.tile_count,
                              lds, tot))
        return;
    if (threadIdx.x != 0) return;
    if (!(isfinite(tot[0]) && isfinite(tot[1]) && isfinite(tot[2]))) {
        st->done = kNonFinite;
        return;
    }
    st->omega = (tot[2] > 0.0) ? mk(tot[0] / tot[2], tot[1] / tot[2]) : mk(0.0, 0.0);
}

// K3: x += alpha p' + omega s; r = s - omega t; (rh, r), (r, r).
template <typename R>
__global__ void __launch_bounds__(kBlock)
k_bicg_x(const cv_fdfd_desc* __restrict__ desc, const cv_tile* __restrict__ tiles,
         cx<R>* __restrict__ x, const cx<R>* __restrict__ p, cx<R>* __restrict__ r,
         const cx<R>* __restrict__ v, const cx<R>* __restrict__ tv, const cx<R>* __restrict__ rh,
         RedBuf rb, KState* __restrict__ state) {
    __shared__ double lds[kWaves * 3];
    const cv_tile t = tiles[blockIdx.x];
    const cv_fdfd_desc d = load_desc(desc, t.inst);
    const int sys = system_of(d, t.inst);
    KState* st = state + sys;
    if (st->done != kRunning || st->skip) return;
    const cx<R> alpha = ccast<R>(st->alpha), omega = ccast<R>(st->omega);
    const int n = d.nx * d.ny;
    double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < kCPT; ++q) {
        const int loc = t.first + threadIdx.x + q * kBlock;
        if (loc >= n) break;
        const int64_t k = d.cell_off + loc;
        const cx<R> sk = r[k] - alpha * v[k];
        x[k] = x[k] + alpha * p[k] + omega * sk;
        const cx<R> rk = sk - omega * tv[k];
        r[k] = rk;
        const cx<R> hk = rh[k];
        acc[0] += (double)hk.re * (double)rk.re + (double)hk.im * (double)rk.im;
        acc[1] += (double)hk.re * (double)rk.im - (double)hk.im * (double)rk.re;
        acc[2] += norm2(ccast<double>(rk));
    }
    block_sum<3>(acc, lds);
    double tot[3];
    if (!last_block_reduce<3>(acc, rb.partials, rb.gpartials, rb.counters + sys, rb.gcount, blockIdx.x, d.tile_first, d.tile_count,
                              lds, tot))
        return;
    if (threadIdx.x != 0) return;
    const cx<double> rho_new = mk(tot[0], tot[1]);
    const double rr = tot[2];
    st->iter += 1;
    st->rr = rr;
    if (!(cfinite(rho_new) && isfinite(rr))) {
        st->done = kNonFinite;
    } else if (rr <= st->tol2) {
        st->done = kConverged;
    } else if (st->iter >= st->max_iter) {
        st->done = kMaxIter;
    } else {
        const cx<double> om = st->omega;
        const bool om_zero = (om.re == 0.0 && om.im == 0.0);
        const bool rho_zero = (rho_new.re == 0.0 && rho_new.im == 0.0);
        // (rh, r) / (|rh| |r|) decays naturally in BiCGSTAB, so only exact
        // breakdowns restart (with rh = r, p = r)
        if (om_zero || rho_zero) {
            st->restart = 1;
            st->rho = mk(rr, 0.0);
            st->rhrh = rr;
            st->restarts += 1;
        } else {
            st->beta = cdiv(rho_new, st->rho) * cdiv(st->alpha, om);
            st->rho = rho_new;
        }
    }
}

// ------------------------------------------------- mixed-precision helpers
// rhs32 = r64 / sqrt(rr) (per instance), x32 = 0.  Instances that have already
// converged get rhs32 = 0, so the inner solve marks them converged at once.
__global__ void __launch_bounds__(kBlock)
k_mixed_rhs(const cv_fdfd_desc* __restrict__ desc, const cv_tile* __restrict__ tiles,
            const cx<double>* __restrict__ r64, cx<float>* __restrict__ b32, cx<float>* __restrict__ x32,
            KState* __restrict__ outer) {
    const cv_tile t = tiles[blockIdx.x];
    const cv_fdfd_desc d = load_desc(desc, t.inst);
    KState* st = outer + system_of(d, t.inst);
    const bool active = st->done == kRunning;
    const int n = d.nx * d.ny;
    const double sc = active ? 1.0 / sqrt(st->rr_true) : 0.0;
#pragma unroll
    for (int q = 0; q < kCPT; ++q) {
        const int loc = t.first + threadIdx.x + q * kBlock;
        if (loc >= n) break;
        const int64_t k = d.cell_off + loc;
        const cx<double> rv = active ? r64[k] : czero<double>();
        b32[k] = mk<float>(float(rv.re * sc), float(rv.im * sc));
        x32[k] = czero<float>();
    }
}

// x64 += sqrt(rr) * x32 (per instance, only for instances still running)
__global__ void __launch_bounds__(kBlock)
k_mixed_update(const cv_fdfd_desc* __restrict__ desc, const cv_tile* __restrict__ tiles,
               cx<double>* __restrict__ x64, const cx<float>* __restrict__ x32,
               const KState* __restrict__ outer) {
    const cv_tile t = tiles[blockIdx.x];
    const cv_fdfd_desc d = load_desc(desc, t.inst);
    const KState* st = outer + system_of(d, t.inst);
    if (st->done != kRunning) return;
    const int n = d.nx * d.ny;
    const double sc = sqrt(st->rr_true);
#pragma unroll
    for (int q = 0; q < kCPT; ++q) {
        const int loc = t.first + threadIdx.x + q * kBlock;
        if (loc >= n) break;
        const int64_t k = d.cell_off + loc;
        const cx<float> dv = x32[k];
        const cx<double> xv = x64[k];
        x64[k] = mk(xv.re + sc * (double)dv.re, xv.im + sc * (double)dv.im);
    }
}

// outer residual r64 = b64 - A64 x64 with per-instance norms; sets outer
// state: done = converged if ||r|| <= tol ||b||, else running.
template <bool TRANS>
__global__ void __launch_bounds__(kBlock)
k_mixed_resid(const cv_fdfd_desc* __restrict__ desc, const cv_tile* __restrict__ tiles,
              const cx<double>* __restrict__ prof, const double* __restrict__ m0,
              const double* __restrict__ m1, const cx<double>* __restrict__ b,
              const cx<double>* __restrict__ x, cx<double>* __restrict__ r, RedBuf rb,
              KState* __restrict__ outer, double tol, int first) {
    __shared__ double lds[kWaves * 2];
    const cv_tile t = tiles[blockIdx.x];
    const cv_fdfd_desc d = load_desc(desc, t.inst);
    const int sys = system_of(d, t.inst);
    KState* st = outer + sys;
    if (!first && st->done != kRunning) return;
    const int n = d.nx * d.ny;
    auto fx = [&](int64_t k) { return x[k]; };
    double acc[2] = {0.0, 0.0};
#pragma unroll
    for (int q = 0; q < kCPT; ++q) {
        const int loc = t.first + threadIdx.x + q * kBlock;
        if (loc >= n) break;
        const int i = loc / d.ny, j = loc - i * d.ny;
        const int64_t k = d.cell_off + loc;
        const cx<double> bk = b[k];
        const cx<double> rk = bk - apply_any<double, TRANS>(d, prof, m0, m1, i, j, fx, x[k]);
        r[k] = rk;
        acc[0] += norm2(rk);
        acc[1] += norm2(bk);
    }
    block_sum<2>(acc, lds);
    double tot[2];
    if (!last_block_reduce<2>(acc, rb.partials, rb.gpartials, rb.counters + sys, rb.gcount, blockIdx.x, d.tile_first, d.tile_count,
                              lds, tot))
        return;
    if (threadIdx.x != 0) return;
    if (first) {
        st->bb = tot[1];
        st->tol2 = tol * tol * tot[1];
        st->iter = 0;
        st->restarts = 0;
    }
    st->rr_true = tot[0];
    st->rr = tot[0];
    if (!isfinite(tot[0]))
        st->done = kNonFinite;
    else if (tot[0] <= st->tol2)
        st->done = kConverged;
    else
        st->done = kRunning;
}

// ------------------------------------------------------------------ COCG
// Conjugate orthogonal CG on the complex-symmetric form B = W A of the
// operator (zero Bloch phases, no coupled blocks): with the forward SC-PML
// factors s (kf = 1/(s dL)), W = 1/(kxf_i kyf_j) for Ez and 1/(kxb_i kyb_j)
// for Hz makes W A symmetric (A = S^-1 D0 S^-1 D0 with D0b = -D0f^T).
// A x = b is solved as B x = W b; A^T y = b (A^T = B W^-1) as B u = b with
// y = W u.  Jacobi preconditioning on B, M = diag(B) = W_i A_ii; the
// recurrences use the unconjugated bilinear form (x, y) = sum x_i y_i.  One
// iteration is two fused kernels:
//   C1  p' = z + beta p (p' = z on restart) at every stencil neighbour,
//       q = B p' = W (A p'), partial (p', q)                  -> alpha
//   C2  x += alpha p' (W p' for A^T), z -= alpha q / M, partials (r, z)
//       with r = M z, and |rho|^2 of the true residual rho = r / W (A) or
//       r (A^T)                                     -> beta, convergence
// z = M^-1 r is the only residual vector kept.
// W is normalised to the product of the two s factors (O(1)): the profiles
// hold 1/(s dL) (1/s on an identity axis) and |scale| = dL^2 for both
// polarisations, so float vectors stay far from the denormal range.
template <typename R>
__device__ __forceinline__ cx<R> sym_weight(const cv_fdfd_desc& d, const cx<R>* __restrict__ prof, int i, int j) {
    const cx<R>* kxf = prof + d.prof_off;
    const cx<R>* kxb = kxf + d.nx;
    const cx<R>* kyf = kxb + d.nx;
    const cx<R>* kyb = kyf + d.ny;
    const R dL = R(sqrt(fabs(d.scale)));
    cx<R> a = (d.pol == CV_POL_EZ) ? kxf[i] * kyf[j] : kxb[i] * kyb[j];
    if (!(d.flags & CV_FLAG_X_IDENTITY)) a = dL * a;
    if (!(d.flags & CV_FLAG_Y_IDENTITY)) a = dL * a;
    return crecip(a);
}

// mode 0: fresh start from the initial guess x; mode 1: verification of a
// converged instance with a freshly computed residual (restart when above
// tolerance).  z = M^-1 r, r = W (b - A x) (A) or b - A^T x (A^T).
template <typename R, bool TRANS>
__global__ void __launch_bounds__(kBlock)
k_cocg_init(const cv_fdfd_desc* __restrict__ desc, const cv_tile* __restrict__ tiles,
            const cx<R>* __restrict__ prof, const R* __restrict__ m0, const R* __restrict__ m1,
            const cx<R>* __restrict__ b, const cx<R>* __restrict__ x, cx<R>* __restrict__ z, RedBuf rb,
            KState* __restrict__ state, double tol, int max_iter, int mode) {
    __shared__ double lds[kWaves * 4];
    const cv_tile t = tiles[blockIdx.x];
    const cv_fdfd_desc d = load_desc(desc, t.inst);
    KState* st = state + t.inst;
    const int done0 = (mode == 0) ? kRunning : st->done;
    if (mode == 1 && done0 == kRunning) return;
    const bool write_z = (mode == 0) || (done0 == kConverged);
    const int n = d.nx * d.ny;
    auto fx = [&](int64_t k) { return x[k]; };
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < kCPT; ++q) {
        const int loc = t.first + threadIdx.x + q * kBlock;
        if (loc >= n) break;
        const int i = loc / d.ny, j = loc - i * d.ny;
        const int64_t k = d.cell_off + loc;
        const cx<R> bk = b[k];
        const cx<R> rho = bk - apply_any<R, TRANS>(d, prof, m0, m1, i, j, fx, x[k]);
        const cx<R> w = sym_weight<R>(d, prof, i, j);
        const cx<R> mii = w * row_coefs<R>(d, prof, m0, m1, i, j).c;
        const cx<R> r = TRANS ? rho : w * rho;
        const cx<R> zk = r * crecip(mii);
        if (write_z) z[k] = zk;
        const cx<double> rz = ccast<double>(r) * ccast<double>(zk);   // products in FP64
        acc[0] += rz.re;
        acc[1] += rz.im;
        acc[2] += norm2(ccast<double>(rho));
        acc[3] += norm2(ccast<double>(bk));
    }
    block_sum<4>(acc, lds);
    double tot[4];
    if (!last_block_reduce<4>(acc, rb.partials, rb.gpartials, rb.counters + t.inst, rb.gcount, blockIdx.x,
                              d.tile_first, d.tile_count, lds, tot))
        return;
    if (threadIdx.x != 0) return;
    if (mode == 0) {
        st->bb = tot[3];
        st->tol2 = tol * tol * tot[3];
        st->max_iter = max_iter;
        st->iter = 0;
        st->restarts = 0;
    }
    st->rr_true = tot[2];
    if (mode == 1 && done0 != kConverged) return;  // report only
    st->rr = tot[2];
    st->rho = mk(tot[0], tot[1]);
    st->restart = 1;
    st->skip = 0;
    if (!(isfinite(tot[0]) && isfinite(tot[1]) && isfinite(tot[2]))) {
        st->done = kNonFinite;
    } else if (tot[2] <= st->tol2) {
        st->done = kConverged;
    } else {
        st->done = kRunning;
        if (mode == 1) st->restarts += 1;
        if (st->iter >= st->max_iter) st->done = kMaxIter;
    }
}

// C1: p' = z + beta p (z on restart); q = W (A p'); (p', q) -> alpha.
template <typename R>
__global__ void __launch_bounds__(kBlock)
k_cocg_p(const cv_fdfd_desc* __restrict__ desc, const cv_tile* __restrict__ tiles,
         const cx<R>* __restrict__ prof, const R* __restrict__ m0, const R* __restrict__ m1,
         const cx<R>* __restrict__ z, const cx<R>* __restrict__ p_old, cx<R>* __restrict__ p_new,
         cx<R>* __restrict__ qv, RedBuf rb, KState* __restrict__ state) {
    __shared__ double lds[kWaves * 2];
    const cv_tile t = tiles[blockIdx.x];
    const cv_fdfd_desc d = load_desc(desc, t.inst);
    KState* st = state + t.inst;
    if (st->done != kRunning) return;
    const bool restart = st->restart != 0;
    const cx<R> beta = ccast<R>(st->beta);
    const int n = d.nx * d.ny;
    auto fp = [&](int64_t k) -> cx<R> {
        if (restart) return z[k];
        return z[k] + beta * p_old[k];
    };
    double acc[2] = {0.0, 0.0};
#pragma unroll
    for (int q = 0; q < kCPT; ++q) {
        const int loc = t.first + threadIdx.x + q * kBlock;
        if (loc >= n) break;
        const int i = loc / d.ny, j = loc - i * d.ny;
        const int64_t k = d.cell_off + loc;
        const cx<R> pk = fp(k);
        p_new[k] = pk;
        const cx<R> qk = sym_weight<R>(d, prof, i, j) * apply_row<R>(d, prof, m0, m1, i, j, fp, pk);
        qv[k] = qk;
        const cx<double> pq = ccast<double>(pk) * ccast<double>(qk);
        acc[0] += pq.re;
        acc[1] += pq.im;
    }
    block_sum<2>(acc, lds);
    double tot[2];
    if (!last_block_reduce<2>(acc, rb.partials, rb.gpartials, rb.counters + t.inst, rb.gcount, blockIdx.x,
                              d.tile_first, d.tile_count, lds, tot))
        return;
    if (threadIdx.x != 0) return;
    const cx<double> sigma = mk(tot[0], tot[1]);
    st->restart = 0;
    if (!cfinite(sigma)) {
        st->done = kNonFinite;
        return;
    }
    if (sigma.re == 0.0 && sigma.im == 0.0) {   // (p, B p) = 0: restart with p = z
        st->skip = 1;
        st->restart = 1;
        st->restarts += 1;
        if (st->restarts > kMaxRestarts) st->done = kBreakdown;
        return;
    }
    st->skip = 0;
    st->alpha = cdiv(st->rho, sigma);
}

// C2: x += alpha p' (W p' for A^T); z -= alpha q / M; (r, z), |rho|^2 -> beta.
template <typename R, bool TRANS>
__global__ void __launch_bounds__(kBlock)
k_cocg_x(const cv_fdfd_desc* __restrict__ desc, const cv_tile* __restrict__ tiles,
         const cx<R>* __restrict__ prof, const R* __restrict__ m0, const R* __restrict__ m1,
         cx<R>* __restrict__ x, const cx<R>* __restrict__ p, cx<R>* __restrict__ z,
         const cx<R>* __restrict__ qv, RedBuf rb, KState* __restrict__ state) {
    __shared__ double lds[kWaves * 3];
    const cv_tile t = tiles[blockIdx.x];
    const cv_fdfd_desc d = load_desc(desc, t.inst);
    KState* st = state + t.inst;
    if (st->done != kRunning || st->skip) return;
    const cx<R> alpha = ccast<R>(st->alpha);
    const int n = d.nx * d.ny;
    double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < kCPT; ++q) {
        const int loc = t.first + threadIdx.x + q * kBlock;
        if (loc >= n) break;
        const int i = loc / d.ny, j = loc - i * d.ny;
        const int64_t k = d.cell_off + loc;
        const cx<R> w = sym_weight<R>(d, prof, i, j);
        const cx<R> mii = w * row_coefs<R>(d, prof, m0, m1, i, j).c;
        const cx<R> pk = p[k];
        x[k] = x[k] + alpha * (TRANS ? w * pk : pk);
        const cx<R> zk = z[k] - alpha * (qv[k] * crecip(mii));
        z[k] = zk;
        const cx<double> r = ccast<double>(mii) * ccast<double>(zk);
        const cx<double> rz = r * ccast<double>(zk);
        const cx<double> rho = TRANS ? r : cdiv(r, ccast<double>(w));
        acc[0] += rz.re;
        acc[1] += rz.im;
        acc[2] += norm2(rho);
    }
    block_sum<3>(acc, lds);
    double tot[3];
    if (!last_block_reduce<3>(acc, rb.partials, rb.gpartials, rb.counters + t.inst, rb.gcount, blockIdx.x,
                              d.tile_first, d.tile_count, lds, tot))
        return;
    if (threadIdx.x != 0) return;
    const cx<double> rz_new = mk(tot[0], tot[1]);
    const double rr = tot[2];
    st->iter += 1;
    st->rr = rr;
    if (!(cfinite(rz_new) && isfinite(rr))) {
        st->done = kNonFinite;
    } else if (rr <= st->tol2) {
        st->done = kConverged;
    } else if (st->iter >= st->max_iter) {
        st->done = kMaxIter;
    } else if (rz_new.re == 0.0 && rz_new.im == 0.0) {
        st->restart = 1;   // (r, z) = 0 without convergence: restart with p = z
        st->restarts += 1;
        if (st->restarts > kMaxRestarts) st->done = kBreakdown;
        st->rho = rz_new;
    } else {
        st->beta = cdiv(rz_new, st->rho);
        st->rho = rz_new;
    }
}

// --------------------------------------------------------------- host side
struct Work {
    void* r;
    void* rh;
    void* p[2];
    void* v[2];
    void* t;
    KState* state;
    RedBuf rb;
};

static int64_t align256(int64_t x) { return (x + 255) & ~int64_t(255); }

// A captured, instantiated loop body.  Capture runs on a private stream of
// default priority (the caller's stream may be the legacy null stream, which
// cannot be captured); the executable graph is launched on the caller's stream.
struct LoopGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t cap = nullptr;
    template <typename Body>
    hipError_t capture(Body&& body) {
        hipError_t e = hipStreamCreateWithFlags(&cap, hipStreamNonBlocking);
        if (e != hipSuccess) return e;
        e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return e;
        body(cap);
        const hipError_t le = hipGetLastError();
        e = hipStreamEndCapture(cap, &graph);
        if (le != hipSuccess) return le;
        if (e != hipSuccess) return e;
        return hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    }
    ~LoopGraph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (cap) (void)hipStreamDestroy(cap);
    }
};

static int64_t layout(int64_t cs, int64_t cells, int n_inst, int n_tiles, char* base, Work* w) {
    int64_t off = 0;
    auto take = [&](int64_t bytes) -> void* {
        void* p = base ? (void*)(base + off) : nullptr;
        off += align256(bytes);
        return p;
    };
    Work tmp;
    Work* o = w ? w : &tmp;
    o->r = take(cs * cells);
    o->rh = take(cs * cells);
    o->p[0] = take(cs * cells);
    o->p[1] = take(cs * cells);
    o->v[0] = take(cs * cells);
    o->v[1] = take(cs * cells);
    o->t = take(cs * cells);
    o->state = (KState*)take(sizeof(KState) * (int64_t)n_inst);
    o->rb.counters = (unsigned int*)take(sizeof(unsigned int) * ((int64_t)n_inst + n_tiles));
    o->rb.gcount = o->rb.counters ? o->rb.counters + n_inst : nullptr;
    o->rb.partials = (double*)take(sizeof(double) * 4 * (int64_t)n_tiles);
    o->rb.gpartials = (double*)take(sizeof(double) * 4 * (int64_t)n_tiles);
    return off;
}

template <typename R, bool TRANS>
static int solve_impl(const cv_fdfd_desc* desc, const cv_fdfd_desc* desc_host, int n_inst, const cv_tile* tiles,
                      int n_tiles, const void* prof_, const void* m0_, const void* m1_, const void* b_, void* x_,
                      double tol, int max_iter, int check_every, Work& w, KState* hs,
                      hipStream_t s, bool use_graph = true) {
    const cx<R>* prof = (const cx<R>*)prof_;
    const R* m0 = (const R*)m0_;
    const R* m1 = (const R*)m1_;
    const cx<R>* b = (const cx<R>*)b_;
    cx<R>* x = (cx<R>*)x_;
    cx<R>* r = (cx<R>*)w.r;
    cx<R>* rh = (cx<R>*)w.rh;
    cx<R>* p[2] = {(cx<R>*)w.p[0], (cx<R>*)w.p[1]};
    cx<R>* v[2] = {(cx<R>*)w.v[0], (cx<R>*)w.v[1]};
    cx<R>* tv = (cx<R>*)w.t;
    k_zero_u32<<<(n_inst + n_tiles + 255) / 256, 256, 0, s>>>(w.rb.counters, n_inst + n_tiles);
    // p and v start at zero (they are only read when restart == 0, which never
    // happens before the first K1), so no clear is needed.
    k_init<R, TRANS><<<n_tiles, kBlock, 0, s>>>(desc, tiles, prof, m0, m1, b, x, r, rh, w.rb,
                                                w.state, tol, max_iter, 0);
    CV_LAUNCH_CHECK();
    const int max_checks = 8;
    int checks = 0;
    // The check_every iterations between host checks (an even count, so the
    // p / v buffers are back in place at the end) are captured once into a
    // graph and replayed: every scalar lives in device memory, so the launch
    // arguments never change, and one graph launch replaces 3 * check_every
    // kernel launches.
    auto body = [&](hipStream_t q) {
        int cur = 0;
        for (int c = 0; c < check_every; ++c) {
            const int nxt = cur ^ 1;
            k_bicg_p<R, TRANS><<<n_tiles, kBlock, 0, q>>>(desc, tiles, prof, m0, m1, r, rh, p[cur], v[cur], p[nxt],
                                                         v[nxt], w.rb, w.state);
            k_bicg_s<R, TRANS><<<n_tiles, kBlock, 0, q>>>(desc, tiles, prof, m0, m1, r, v[nxt], tv, w.rb, w.state);
            k_bicg_x<R><<<n_tiles, kBlock, 0, q>>>(desc, tiles, x, p[nxt], r, v[nxt], tv, rh, w.rb, w.state);
            cur = nxt;
        }
    };
    // Without the graph (use_graph false: the allocation-free entry point,
    // since instantiating a graph allocates) the same launches are issued
    // directly on the caller's stream.
    LoopGraph graph;
    if (use_graph) CV_HIP(graph.capture(body));
    while (true) {
        if (use_graph) {
            CV_HIP(hipGraphLaunch(graph.exec, s));
        } else {
            body(s);
            CV_LAUNCH_CHECK();
        }
        CV_HIP(hipMemcpyAsync(hs, w.state, sizeof(KState) * n_inst, hipMemcpyDeviceToHost, s));
        CV_HIP(hipStreamSynchronize(s));
        bool all_done = true;
        for (int k = 0; k < n_inst; ++k)
            if (desc_host[k].blk == 0) all_done &= (hs[k].done != kRunning);
        if (!all_done) continue;
        // verify converged instances with the true residual (restarts them if needed)
        k_init<R, TRANS><<<n_tiles, kBlock, 0, s>>>(desc, tiles, prof, m0, m1, b, x, r, rh, w.rb,
                                                    w.state, tol, max_iter, 1);
        CV_LAUNCH_CHECK();
        CV_HIP(hipMemcpyAsync(hs, w.state, sizeof(KState) * n_inst, hipMemcpyDeviceToHost, s));
        CV_HIP(hipStreamSynchronize(s));
        bool again = false;
        for (int k = 0; k < n_inst; ++k)
            if (desc_host[k].blk == 0) again |= (hs[k].done == kRunning);
        if (!again || ++checks >= max_checks) break;
    }
    return CV_OK;
}

// COCG over the batch with the same host protocol as solve_impl (the
// iterations between host checks captured once into a graph, or launched
// directly when use_graph is false).  Work: z in w.r, p in w.p[0..1], q in w.v[0].
template <typename R, bool TRANS>
static int cocg_impl(const cv_fdfd_desc* desc, const cv_fdfd_desc* desc_host, int n_inst, const cv_tile* tiles,
                     int n_tiles, const void* prof_, const void* m0_, const void* m1_, const void* b_, void* x_,
                     double tol, int max_iter, int check_every, Work& w, KState* hs, hipStream_t s,
                     bool use_graph = true) {
    const cx<R>* prof = (const cx<R>*)prof_;
    const R* m0 = (const R*)m0_;
    const R* m1 = (const R*)m1_;
    const cx<R>* b = (const cx<R>*)b_;
    cx<R>* x = (cx<R>*)x_;
    cx<R>* z = (cx<R>*)w.r;
    cx<R>* p[2] = {(cx<R>*)w.p[0], (cx<R>*)w.p[1]};
    cx<R>* qv = (cx<R>*)w.v[0];
    k_zero_u32<<<(n_inst + n_tiles + 255) / 256, 256, 0, s>>>(w.rb.counters, n_inst + n_tiles);
    k_cocg_init<R, TRANS><<<n_tiles, kBlock, 0, s>>>(desc, tiles, prof, m0, m1, b, x, z, w.rb, w.state, tol,
                                                     max_iter, 0);
    CV_LAUNCH_CHECK();
    auto body = [&](hipStream_t q) {
        int cur = 0;
        for (int c = 0; c < check_every; ++c) {
            const int nxt = cur ^ 1;
            k_cocg_p<R><<<n_tiles, kBlock, 0, q>>>(desc, tiles, prof, m0, m1, z, p[cur], p[nxt], qv, w.rb, w.state);
            k_cocg_x<R, TRANS><<<n_tiles, kBlock, 0, q>>>(desc, tiles, prof, m0, m1, x, p[nxt], z, qv, w.rb,
                                                          w.state);
            cur = nxt;
        }
    };
    LoopGraph graph;
    if (use_graph) CV_HIP(graph.capture(body));
    int checks = 0;
    while (true) {
        if (use_graph) {
            CV_HIP(hipGraphLaunch(graph.exec, s));
        } else {
            body(s);
            CV_LAUNCH_CHECK();
        }
        CV_HIP(hipMemcpyAsync(hs, w.state, sizeof(KState) * n_inst, hipMemcpyDeviceToHost, s));
        CV_HIP(hipStreamSynchronize(s));
        bool all_done = true;
        for (int k = 0; k < n_inst; ++k) all_done &= (hs[k].done != kRunning);
        if (!all_done) continue;
        k_cocg_init<R, TRANS><<<n_tiles, kBlock, 0, s>>>(desc, tiles, prof, m0, m1, b, x, z, w.rb, w.state, tol,
                                                         max_iter, 1);
        CV_LAUNCH_CHECK();
        CV_HIP(hipMemcpyAsync(hs, w.state, sizeof(KState) * n_inst, hipMemcpyDeviceToHost, s));
        CV_HIP(hipStreamSynchronize(s));
        bool again = false;
        for (int k = 0; k < n_inst; ++k) again |= (hs[k].done == kRunning);
        if (!again || ++checks >= 8) break;
    }
    return CV_OK;
}

// COCG applies to instances with zero Bloch phases and no coupled blocks.
static int check_cocg(const cv_fdfd_desc* dh, int n_inst) {
    for (int k = 0; k < n_inst; ++k) {
        const cv_fdfd_desc& d = dh[k];
        CV_CHECK(d.nblk == 1 && d.n_cpl == 0, "instance %d: COCG does not take frequency-coupled systems", k);
        CV_CHECK(d.phx_re == 1.0 && d.phx_im == 0.0 && d.phy_re == 1.0 && d.phy_im == 0.0,
                 "instance %d: COCG needs zero Bloch phases (the operator is complex symmetric only then)", k);
    }
    return CV_OK;
}

static int status_of(const KState& k) {
    switch (k.done) {
        case kConverged: return CV_OK;
        case kMaxIter: return CV_ERR_NOT_CONVERGED;
        case kNonFinite: return CV_ERR_NONFINITE;
        case kBreakdown: return CV_ERR_BREAKDOWN;
        default: return CV_ERR_NOT_CONVERGED;
    }
}

// Checks the batch layout the kernels rely on; returns the number of packed
// cells in *cells.  Blocks of a frequency-coupled system must be consecutive
// instances with equal grids, consecutive cells and the system's tile range,
// so that every inter-block access stays inside the system's cells.
static int check_batch(const cv_fdfd_desc* dh, int n_inst, int n_tiles, int64_t* cells) {
    int64_t total = 0;
    for (int k = 0; k < n_inst; ++k) {
        const cv_fdfd_desc& d = dh[k];
        CV_CHECK(d.nx > 0 && d.ny > 0, "instance %d has an empty grid", k);
        CV_CHECK(d.tile_first >= 0 && d.tile_count > 0 && d.tile_first + d.tile_count <= n_tiles,
                 "instance %d: tile range out of the table", k);
        total = std::max<int64_t>(total, d.cell_off + (int64_t)d.nx * d.ny);
        CV_CHECK(d.nblk >= 1 && d.blk >= 0 && d.blk < d.nblk && d.n_cpl >= 0, "instance %d: bad block fields", k);
        if (d.nblk == 1) {
            CV_CHECK(d.n_cpl == 0, "instance %d: coupling tables on a single-block system", k);
            continue;
        }
        CV_CHECK(d.pol == CV_POL_EZ, "instance %d: coupled blocks must be Ez", k);
        const int root = k - d.blk;
        CV_CHECK(root >= 0 && root + d.nblk <= n_inst, "instance %d: coupled blocks out of the batch", k);
        const cv_fdfd_desc& r = dh[root];
        const int64_t n = (int64_t)r.nx * r.ny;
        CV_CHECK(d.nx == r.nx && d.ny == r.ny && d.nblk == r.nblk && d.n_cpl == r.n_cpl &&
                     d.cell_off == r.cell_off + d.blk * n && d.tile_first == r.tile_first &&
                     d.tile_count == r.tile_count && d.cpl_off == r.cpl_off && d.w2_off == r.w2_off,
                 "instance %d: block %d does not match block 0 of its system", k, d.blk);
    }
    *cells = total;
    return CV_OK;
}

static void fill_info(const cv_fdfd_desc* dh, const KState* hs, cv_solve_info* info, int n_inst,
                      int* worst) {
    *worst = CV_OK;
    for (int k = 0; k < n_inst; ++k) {
        const KState& h = hs[k - dh[k].blk];
        info[k].iterations = h.iter;
        info[k].status = status_of(h);
        info[k].restarts = h.restarts;
        info[k].pad = 0;
        info[k].rel_residual = h.bb > 0 ? sqrt(h.rr_true / h.bb) : 0.0;
        if (info[k].status != CV_OK && *worst == CV_OK) *worst = info[k].status;
    }
}

}  // namespace cv

using namespace cv;

extern "C" int64_t cv_fdfd_solve_workspace_bytes(int dtype, int64_t total_cells, int n_inst, int n_tiles) {
    const int64_t cs = (dtype == CV_C128) ? 16 : 8;
    return layout(cs, total_cells, n_inst, n_tiles, nullptr, nullptr);
}

// Dispatch of one batched Krylov solve on (method, precision, orientation).
static int run_krylov(int method, int dtype, int transpose, const cv_fdfd_desc* desc, const cv_fdfd_desc* desc_host,
                      int n_inst, const cv_tile* tiles, int n_tiles, const void* prof, const void* m0, const void* m1,
                      const void* b, void* x, double tol, int max_iter, int check_every, Work& w, KState* hs,
                      hipStream_t s, bool use_graph) {
#define CV_KRYLOV_ARGS desc, desc_host, n_inst, tiles, n_tiles, prof, m0, m1, b, x, tol, max_iter, check_every, w, hs, \
                       s, use_graph
    if (method == CV_KRYLOV_COCG) {
        if (dtype == CV_C128)
            return transpose ? cocg_impl<double, true>(CV_KRYLOV_ARGS) : cocg_impl<double, false>(CV_KRYLOV_ARGS);
        return transpose ? cocg_impl<float, true>(CV_KRYLOV_ARGS) : cocg_impl<float, false>(CV_KRYLOV_ARGS);
    }
    if (dtype == CV_C128)
        return transpose ? solve_impl<double, true>(CV_KRYLOV_ARGS) : solve_impl<double, false>(CV_KRYLOV_ARGS);
    return transpose ? solve_impl<float, true>(CV_KRYLOV_ARGS) : solve_impl<float, false>(CV_KRYLOV_ARGS);
#undef CV_KRYLOV_ARGS
}

// Shared by cv_fdfd_solve and cv_fdfd_solve_noalloc: checks the arguments,
// lays the workspace out at `workspace`, runs the batched BiCGSTAB and fills
// `info` from the host copy `hs` of the n_inst solver states.
static int solve_entry(int dtype, const cv_fdfd_desc* desc, const cv_fdfd_desc* desc_host, int n_inst,
                       const cv_tile* tiles, int n_tiles, const void* prof, const void* m0, const void* m1,
                       const void* b, void* x, int transpose, double tol, int max_iter, int check_every,
                       void* workspace, int64_t workspace_bytes, cv_solve_info* info, KState* hs, bool use_graph,
                       hipStream_t s, int method = CV_KRYLOV_BICGSTAB) {
    int64_t cells = 0;
    {
        const int rc = check_batch(desc_host, n_inst, n_tiles, &cells);
        if (rc != CV_OK) return rc;
    }
    CV_CHECK(method == CV_KRYLOV_BICGSTAB || method == CV_KRYLOV_COCG, "unknown Krylov method %d", method);
    if (method == CV_KRYLOV_COCG) {
        const int rc = check_cocg(desc_host, n_inst);
        if (rc != CV_OK) return rc;
    }
    for (int k = 0; k < n_inst; ++k)
        if (desc_host[k].pol == CV_POL_HZ) CV_CHECK(m1 != nullptr, "Hz operator needs m1");
    const int64_t cs = (dtype == CV_C128) ? 16 : 8;
    Work w;
    const int64_t need = layout(cs, cells, n_inst, n_tiles, (char*)workspace, &w);
    CV_CHECK(workspace_bytes >= need, "workspace too small: %lld < %lld", (long long)workspace_bytes,
             (long long)need);
    if (check_every <= 0) check_every = 16;
    check_every += check_every & 1;
    const int rc = run_krylov(method, dtype, transpose, desc, desc_host, n_inst, tiles, n_tiles, prof, m0, m1, b, x,
                              tol, max_iter, check_every, w, hs, s, use_graph);
    if (rc != CV_OK) return rc;
    int worst;
    fill_info(desc_host, hs, info, n_inst, &worst);
    return worst;
}

extern "C" int cv_fdfd_solve(int dtype, const cv_fdfd_desc* desc, const cv_fdfd_desc* desc_host, int n_inst,
                             const cv_tile* tiles, int n_tiles, const void* prof, const void* m0, const void* m1,
                             const void* b, void* x, int transpose, double tol, int max_iter, int check_every,
                             void* workspace, int64_t workspace_bytes, cv_solve_info* info, void* stream) {
    CV_CHECK(desc && desc_host && tiles && prof && m0 && b && x && workspace && info, "null pointer argument");
    CV_CHECK(n_inst > 0 && n_tiles > 0, "empty batch");
    CV_CHECK(dtype == CV_C64 || dtype == CV_C128, "unknown dtype %d", dtype);
    CV_CHECK(tol > 0 && max_iter > 0, "tol and max_iter must be positive");
    std::vector<KState> hs(n_inst);
    return solve_entry(dtype, desc, desc_host, n_inst, tiles, n_tiles, prof, m0, m1, b, x, transpose, tol, max_iter,
                       check_every, workspace, workspace_bytes, info, hs.data(), true, as_stream(stream));
}

extern "C" int cv_fdfd_cocg(int dtype, const cv_fdfd_desc* desc, const cv_fdfd_desc* desc_host, int n_inst,
                            const cv_tile* tiles, int n_tiles, const void* prof, const void* m0, const void* m1,
                            const void* b, void* x, int transpose, double tol, int max_iter, int check_every,
                            void* workspace, int64_t workspace_bytes, cv_solve_info* info, void* stream) {
    CV_CHECK(desc && desc_host && tiles && prof && m0 && b && x && workspace && info, "null pointer argument");
    CV_CHECK(n_inst > 0 && n_tiles > 0, "empty batch");
    CV_CHECK(dtype == CV_C64 || dtype == CV_C128, "unknown dtype %d", dtype);
    CV_CHECK(tol > 0 && max_iter > 0, "tol and max_iter must be positive");
    std::vector<KState> hs(n_inst);
    return solve_entry(dtype, desc, desc_host, n_inst, tiles, n_tiles, prof, m0, m1, b, x, transpose, tol, max_iter,
                       check_every, workspace, workspace_bytes, info, hs.data(), true, as_stream(stream),
                       CV_KRYLOV_COCG);
}

// ------------------------------------------------ allocation-free solve
// The workspace of cv_fdfd_solve_noalloc lives in this file-scope device
// array (one copy per GPU, held until the process ends).  Host code never
// passes `scratch` itself to a kernel or a copy: on the host the name stands
// for a host-side shadow, not the device copy.  The device address of each
// GPU's copy is looked up once per device with hipGetSymbolAddress and kept
// in g_scratch_ptr.
static __device__ __align__(256) char scratch[CV_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// Device address of the current device's scratch array (looked up on the
// first call on that device); CV_OK or CV_ERR_HIP / CV_ERR_INVALID.
static int scratch_ptr(void** out) {
    int dev = -1;
    CV_HIP(hipGetDevice(&dev));
    CV_CHECK(dev >= 0 && dev < kMaxDevices, "device index %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        CV_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        CV_CHECK(p != nullptr, "no device address for the scratch array on device %d", dev);
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return CV_OK;
}

extern "C" int cv_noalloc_init(void) {
    void* p = nullptr;
    return scratch_ptr(&p);
}

extern "C" int cv_fdfd_solve_noalloc(int dtype, const cv_fdfd_desc* desc, const cv_fdfd_desc* desc_host, int n_inst,
                                     const cv_tile* tiles, int n_tiles, const void* prof, const void* m0,
                                     const void* m1, const void* b, void* x, int transpose, double tol, int max_iter,
                                     int check_every, cv_solve_info* info, void* stream) {
    CV_CHECK(desc && desc_host && tiles && prof && m0 && b && x && info, "null pointer argument");
    CV_CHECK(n_inst > 0 && n_tiles > 0, "empty batch");
    CV_CHECK(n_inst <= CV_NOALLOC_MAX_INST, "%d instances exceed the allocation-free limit of %d", n_inst,
             CV_NOALLOC_MAX_INST);
    CV_CHECK(dtype == CV_C64 || dtype == CV_C128, "unknown dtype %d", dtype);
    CV_CHECK(tol > 0 && max_iter > 0, "tol and max_iter must be positive");
    int64_t cells = 0;
    {
        const int rc = check_batch(desc_host, n_inst, n_tiles, &cells);
        if (rc != CV_OK) return rc;
    }
    const int64_t need = layout(dtype == CV_C128 ? 16 : 8, cells, n_inst, n_tiles, nullptr, nullptr);
    CV_CHECK(need <= (int64_t)CV_NOALLOC_SCRATCH_BYTES,
             "problem needs %lld bytes of solver scratch, above the allocation-free limit of %lld bytes "
             "(%lld packed cells, %d tiles); use cv_fdfd_solve with a workspace",
             (long long)need, (long long)CV_NOALLOC_SCRATCH_BYTES, (long long)cells, n_tiles);
    void* ws = nullptr;
    {
        const int rc = scratch_ptr(&ws);
        if (rc != CV_OK) return rc;
    }
    KState hs[CV_NOALLOC_MAX_INST];
    return solve_entry(dtype, desc, desc_host, n_inst, tiles, n_tiles, prof, m0, m1, b, x, transpose, tol, max_iter,
                       check_every, ws, need, info, hs, false, as_stream(stream));
}

// ------------------------------------------------------------ mixed solve
extern "C" int64_t cv_fdfd_solve_mixed_workspace_bytes(int64_t total_cells, int n_inst, int n_tiles) {
    int64_t inner = layout(8, total_cells, n_inst, n_tiles, nullptr, nullptr);
    // r64, b32, x32, outer state (the outer residual kernel shares the inner
    // solver's reduction scratch: the two never run at the same time)
    return inner + align256(16 * total_cells) + 2 * align256(8 * total_cells) +
           align256(sizeof(KState) * (int64_t)n_inst);
}

extern "C" int cv_fdfd_solve_mixed(const cv_fdfd_desc* desc, const cv_fdfd_desc* desc_host, int n_inst,
                                   const cv_tile* tiles, int n_tiles, const void* prof64, const void* m0_64,
                                   const void* m1_64, const void* prof32, const void* m0_32, const void* m1_32,
                                   const void* b64, void* x64, int transpose, int method, double tol, double inner_tol,
                                   int max_iter, int max_sweeps, int check_every, void* workspace,
                                   int64_t workspace_bytes, cv_solve_info* info, void* stream) {
    CV_CHECK(desc && desc_host && tiles && prof64 && m0_64 && prof32 && m0_32 && b64 && x64 && workspace && info,
             "null pointer argument");
    CV_CHECK(n_inst > 0 && n_tiles > 0, "empty batch");
    int64_t cells = 0;
    {
        const int rc = check_batch(desc_host, n_inst, n_tiles, &cells);
        if (rc != CV_OK) return rc;
    }
    const int64_t need = cv_fdfd_solve_mixed_workspace_bytes(cells, n_inst, n_tiles);
    CV_CHECK(workspace_bytes >= need, "workspace too small: %lld < %lld", (long long)workspace_bytes,
             (long long)need);
    CV_CHECK(method == CV_KRYLOV_BICGSTAB || method == CV_KRYLOV_COCG, "unknown Krylov method %d", method);
    if (method == CV_KRYLOV_COCG) {
        const int rc = check_cocg(desc_host, n_inst);
        if (rc != CV_OK) return rc;
    }
    if (check_every <= 0) check_every = 16;
    check_every += check_every & 1;
    if (max_sweeps <= 0) max_sweeps = 10;
    char* base = (char*)workspace;
    Work w;
    int64_t off = layout(8, cells, n_inst, n_tiles, base, &w);
    cx<double>* r64 = (cx<double>*)(base + off);
    off += align256(16 * cells);
    cx<float>* b32 = (cx<float>*)(base + off);
    off += align256(8 * cells);
    cx<float>* x32 = (cx<float>*)(base + off);
    off += align256(8 * cells);
    KState* outer = (KState*)(base + off);
    hipStream_t s = as_stream(stream);
    k_zero_u32<<<(n_inst + n_tiles + 255) / 256, 256, 0, s>>>(w.rb.counters, n_inst + n_tiles);
    std::vector<KState> ho(n_inst), hi(n_inst);
    std::vector<int> iters(n_inst, 0), sweeps(n_inst, 0);
    auto resid = [&](int first) -> int {
        if (transpose)
            k_mixed_resid<true><<<n_tiles, kBlock, 0, s>>>(desc, tiles, (const cx<double>*)prof64,
                                                           (const double*)m0_64, (const double*)m1_64,
                                                           (const cx<double>*)b64, (const cx<double>*)x64, r64,
                                                           w.rb, outer, tol, first);
        else
            k_mixed_resid<false><<<n_tiles, kBlock, 0, s>>>(desc, tiles, (const cx<double>*)prof64,
                                                            (const double*)m0_64, (const double*)m1_64,
                                                            (const cx<double>*)b64, (const cx<double>*)x64, r64,
                                                            w.rb, outer, tol, first);
        CV_LAUNCH_CHECK();
        CV_HIP(hipMemcpyAsync(ho.data(), outer, sizeof(KState) * n_inst, hipMemcpyDeviceToHost, s));
        CV_HIP(hipStreamSynchronize(s));
        return CV_OK;
    };
    int rc = resid(1);
    if (rc != CV_OK) return rc;
    int inner_status = CV_OK;
    for (int sweep = 0; sweep < max_sweeps; ++sweep) {
        bool any = false;
        for (int k = 0; k < n_inst; ++k)
            if (desc_host[k].blk == 0) any |= (ho[k].done == kRunning);
        if (!any) break;
        k_mixed_rhs<<<n_tiles, kBlock, 0, s>>>(desc, tiles, r64, b32, x32, outer);
        CV_LAUNCH_CHECK();
        // Inner solve over all tiles: instances already converged get a zero rhs
        // (b32 left from their last sweep is irrelevant: their x64 is not updated).
        rc = run_krylov(method, CV_C64, transpose, desc, desc_host, n_inst, tiles, n_tiles, prof32, m0_32, m1_32,
                        b32, x32, inner_tol, max_iter, check_every, w, hi.data(), s, true);
        if (rc != CV_OK) return rc;
        for (int k = 0; k < n_inst; ++k) {
            const int sys = k - desc_host[k].blk;
            if (ho[sys].done == kRunning) {
                iters[k] += hi[sys].iter;
                sweeps[k] += 1;
                if (hi[sys].done == kNonFinite || hi[sys].done == kBreakdown) inner_status = status_of(hi[sys]);
            }
        }
        k_mixed_update<<<n_tiles, kBlock, 0, s>>>(desc, tiles, (cx<double>*)x64, x32, outer);
        CV_LAUNCH_CHECK();
        rc = resid(0);
        if (rc != CV_OK) return rc;
    }
    int worst = CV_OK;
    for (int k = 0; k < n_inst; ++k) {
        const KState& h = ho[k - desc_host[k].blk];
        info[k].iterations = iters[k];
        info[k].restarts = sweeps[k];
        info[k].pad = 0;
        info[k].rel_residual = h.bb > 0 ? sqrt(h.rr_true / h.bb) : 0.0;
        info[k].status = (h.done == kConverged) ? CV_OK
                         : (h.done == kNonFinite) ? CV_ERR_NONFINITE
                                                  : CV_ERR_NOT_CONVERGED;
        if (info[k].status != CV_OK && worst == CV_OK) worst = info[k].status;
    }
    if (worst == CV_OK && inner_status != CV_OK) worst = CV_OK;  // outer residual is the judge
    return worst;
}
