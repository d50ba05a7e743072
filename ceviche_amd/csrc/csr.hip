// Device-resident Krylov solve of a general sparse complex128 system in CSR
// form: the iterative replacement of the sparse direct solve behind
// ceviche's sp_solve (ceviche/primitives.py:103-119 -> solvers.py:37-62) and
// of the solver hook _solve_cuda(A, b) (solvers.py:77-79) for matrices that
// are not FDFD stencils (or whose grid the caller does not describe).
//
// Jacobi-preconditioned BiCGSTAB on A M^-1 (M = diag(A), right
// preconditioning, so r is the true residual of A x = b), complex128 vectors,
// FP64 reductions.  One iteration is three kernels over the rows (one row per
// thread, the CSR row gathered in storage order):
//   K1  p' = r + beta (p - omega v) (p' = r on restart) at every column the
//       row touches, v' = A M^-1 p', partial (rh, v')           -> alpha
//   K2  s = r - alpha v' at every column, t = A M^-1 s, partial (t, s), (t, t)
//                                                               -> omega
//   K3  x += M^-1 (alpha p' + omega s), r = s - omega t, partial (rh, r),
//       (r, r)                                     -> beta, convergence
// The last workgroup of each kernel turns the partial sums into the next
// scalar (the two-level protocol of common.h), so the loop needs no host
// round trip; the host reads the state every check_every iterations and
// verifies convergence with a freshly computed residual.
#include "common.h"

#include <vector>

namespace cv {
namespace {

enum : int32_t { kRun = 0, kConv = 1, kMaxIt = 2, kNonFin = 3, kBreak = 4 };

struct CState {
    cx<double> rho, alpha, omega, beta;
    double rr, bb, tol2, rr_true;
    int32_t iter, done, restart, max_iter;
    int32_t restarts, skip, pad1, pad2;   // skip: K2 / K3 pass over a restarted iteration
};

// An exact (rh, v) = 0 restarts the iteration with rh = r; after this many
// restarts in one solve it is reported as a breakdown.
constexpr int kMaxRestarts = 64;

struct CRed {
    double* partials;        // n_blk * 4
    double* gpartials;       // n_blk * 4
    unsigned int* counter;   // 1
    unsigned int* gcount;    // n_blk
};

template <typename I>
struct CsrA {
    const I* ptr;
    const I* col;
    const cx<double>* val;
    const cx<double>* dinv;
    int64_t n;
    int n_blk;
};

// sum_k A[i, col_k] dinv[col_k] f(col_k)
template <typename I, typename F>
__device__ __forceinline__ cx<double> row_apply(const CsrA<I>& A, int64_t i, F&& f) {
    cx<double> acc = czero<double>();
    const int64_t lo = (int64_t)A.ptr[i], hi = (int64_t)A.ptr[i + 1];
    for (int64_t k = lo; k < hi; ++k) {
        const int64_t j = (int64_t)A.col[k];
        acc = acc + A.val[k] * (A.dinv[j] * f(j));
    }
    return acc;
}

__global__ void k_csr_zero(unsigned int* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// dinv[i] = 1 / A_ii (duplicates summed); 1 where the diagonal is zero.
template <typename I>
__global__ void __launch_bounds__(kBlock) k_csr_dinv(CsrA<I> A, cx<double>* __restrict__ dinv) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= A.n) return;
    cx<double> d = czero<double>();
    for (int64_t k = (int64_t)A.ptr[i]; k < (int64_t)A.ptr[i + 1]; ++k)
        if ((int64_t)A.col[k] == i) d = d + A.val[k];
    dinv[i] = (d.re == 0.0 && d.im == 0.0) ? cone<double>() : cdiv(cone<double>(), d);
}

// mode 0: fresh start (r = rh = b - A x, state reset); mode 1: verification of
// a converged state with the true residual (restart when above tolerance).
template <typename I>
__global__ void __launch_bounds__(kBlock)
k_csr_init(CsrA<I> A, const cx<double>* __restrict__ b, const cx<double>* __restrict__ x, cx<double>* __restrict__ r,
           cx<double>* __restrict__ rh, CRed rb, CState* __restrict__ st, double tol, int max_iter, int mode) {
    __shared__ double lds[kWaves * 2];
    const int done0 = (mode == 0) ? kRun : st->done;
    if (mode == 1 && done0 != kConv) return;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    double acc[2] = {0.0, 0.0};
    if (i < A.n) {
        cx<double> ax = czero<double>();
        for (int64_t k = (int64_t)A.ptr[i]; k < (int64_t)A.ptr[i + 1]; ++k) ax = ax + A.val[k] * x[A.col[k]];
        const cx<double> ri = b[i] - ax;
        r[i] = ri;
        rh[i] = ri;
        acc[0] = norm2(ri);
        acc[1] = norm2(b[i]);
    }
    block_sum<2>(acc, lds);
    double tot[2];
    if (!last_block_reduce<2>(acc, rb.partials, rb.gpartials, rb.counter, rb.gcount, blockIdx.x, 0, A.n_blk, lds,
                              tot))
        return;
    if (threadIdx.x != 0) return;
    const double rr = tot[0];
    if (mode == 0) {
        st->bb = tot[1];
        st->tol2 = tol * tol * tot[1];
        st->max_iter = max_iter;
        st->iter = 0;
        st->restarts = 0;
    }
    st->skip = 0;
    st->rr_true = rr;
    st->rr = rr;
    st->rho = mk(rr, 0.0);
    st->restart = 1;
    if (!isfinite(rr)) {
        st->done = kNonFin;
    } else if (rr <= st->tol2) {
        st->done = kConv;
    } else {
        st->done = (st->iter >= st->max_iter) ? kMaxIt : kRun;
        if (mode == 1) st->restarts += 1;
    }
}

template <typename I>
__global__ void __launch_bounds__(kBlock)
k_csr_p(CsrA<I> A, const cx<double>* __restrict__ r, cx<double>* __restrict__ rh,
        const cx<double>* __restrict__ p_old, const cx<double>* __restrict__ v_old, cx<double>* __restrict__ p_new,
        cx<double>* __restrict__ v_new, CRed rb, CState* __restrict__ st) {
    __shared__ double lds[kWaves * 2];
    if (st->done != kRun) return;
    const bool restart = st->restart != 0;
    const cx<double> beta = st->beta, omega = st->omega;
    auto fp = [&](int64_t k) -> cx<double> {
        if (restart) return r[k];
        return r[k] + beta * (p_old[k] - omega * v_old[k]);
    };
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    double acc[2] = {0.0, 0.0};
    if (i < A.n) {
        p_new[i] = fp(i);
        const cx<double> vi = row_apply(A, i, fp);
        v_new[i] = vi;
        cx<double> hi;
        if (restart) {
            hi = r[i];
            rh[i] = hi;
        } else {
            hi = rh[i];
        }
        acc[0] = hi.re * vi.re + hi.im * vi.im;   // conj(rh) v
        acc[1] = hi.re * vi.im - hi.im * vi.re;
    }
    block_sum<2>(acc, lds);
    double tot[2];
    if (!last_block_reduce<2>(acc, rb.partials, rb.gpartials, rb.counter, rb.gcount, blockIdx.x, 0, A.n_blk, lds,
                              tot))
        return;
    if (threadIdx.x != 0) return;
    const cx<double> rhv = mk(tot[0], tot[1]);
    st->restart = 0;
    if (!cfinite(rhv)) {
        st->done = kNonFin;
        return;
    }
    if (rhv.re == 0.0 && rhv.im == 0.0) {
        // (rh, v) = 0 exactly: restart with rh = p = r (a shadow residual with
        // a single non-zero, as a point right-hand side gives, meets this)
        st->skip = 1;
        st->restart = 1;
        st->rho = mk(st->rr, 0.0);
        st->restarts += 1;
        if (st->restarts > kMaxRestarts) st->done = kBreak;
        return;
    }
    st->skip = 0;
    st->alpha = cdiv(st->rho, rhv);
}

template <typename I>
__global__ void __launch_bounds__(kBlock)
k_csr_s(CsrA<I> A, const cx<double>* __restrict__ r, const cx<double>* __restrict__ v, cx<double>* __restrict__ tv,
        CRed rb, CState* __restrict__ st) {
    __shared__ double lds[kWaves * 3];
    if (st->done != kRun || st->skip) return;
    const cx<double> alpha = st->alpha;
    auto fs = [&](int64_t k) -> cx<double> { return r[k] - alpha * v[k]; };
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    double acc[3] = {0.0, 0.0, 0.0};
    if (i < A.n) {
        const cx<double> si = fs(i);
        const cx<double> ti = row_apply(A, i, fs);
        tv[i] = ti;
        acc[0] = ti.re * si.re + ti.im * si.im;   // conj(t) s
        acc[1] = ti.re * si.im - ti.im * si.re;
        acc[2] = norm2(ti);
    }
    block_sum<3>(acc, lds);
    double tot[3];
    if (!last_block_reduce<3>(acc, rb.partials, rb.gpartials, rb.counter, rb.gcount, blockIdx.x, 0, A.n_blk, lds,
                              tot))
        return;
    if (threadIdx.x != 0) return;
    if (!(isfinite(tot[0]) && isfinite(tot[1]) && isfinite(tot[2]))) {
        st->done = kNonFin;
        return;
    }
    st->omega = (tot[2] > 0.0) ? mk(tot[0] / tot[2], tot[1] / tot[2]) : mk(0.0, 0.0);
}

template <typename I>
__global__ void __launch_bounds__(kBlock)
k_csr_x(CsrA<I> A, cx<double>* __restrict__ x, const cx<double>* __restrict__ p, cx<double>* __restrict__ r,
        const cx<double>* __restrict__ v, const cx<double>* __restrict__ tv, const cx<double>* __restrict__ rh,
        CRed rb, CState* __restrict__ st) {
    __shared__ double lds[kWaves * 3];
    if (st->done != kRun || st->skip) return;
    const cx<double> alpha = st->alpha, omega = st->omega;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    double acc[3] = {0.0, 0.0, 0.0};
    if (i < A.n) {
        const cx<double> si = r[i] - alpha * v[i];
        x[i] = x[i] + A.dinv[i] * (alpha * p[i] + omega * si);
        const cx<double> ri = si - omega * tv[i];
        r[i] = ri;
        const cx<double> hi = rh[i];
        acc[0] = hi.re * ri.re + hi.im * ri.im;
        acc[1] = hi.re * ri.im - hi.im * ri.re;
        acc[2] = norm2(ri);
    }
    block_sum<3>(acc, lds);
    double tot[3];
    if (!last_block_reduce<3>(acc, rb.partials, rb.gpartials, rb.counter, rb.gcount, blockIdx.x, 0, A.n_blk, lds,
                              tot))
        return;
    if (threadIdx.x != 0) return;
    const cx<double> rho_new = mk(tot[0], tot[1]);
    const double rr = tot[2];
    st->iter += 1;
    st->rr = rr;
    if (!(cfinite(rho_new) && isfinite(rr))) {
        st->done = kNonFin;
    } else if (rr <= st->tol2) {
        st->done = kConv;
    } else if (st->iter >= st->max_iter) {
        st->done = kMaxIt;
    } else {
        const cx<double> om = st->omega;
        if ((om.re == 0.0 && om.im == 0.0) || (rho_new.re == 0.0 && rho_new.im == 0.0)) {
            st->restart = 1;   // exact breakdown: restart with rh = r, p = r
            st->rho = mk(rr, 0.0);
            st->restarts += 1;
        } else {
            st->beta = cdiv(rho_new, st->rho) * cdiv(st->alpha, om);
            st->rho = rho_new;
        }
    }
}

int64_t al(int64_t x) { return (x + 255) & ~int64_t(255); }

struct CWork {
    cx<double>* r;
    cx<double>* rh;
    cx<double>* p[2];
    cx<double>* v[2];
    cx<double>* t;
    cx<double>* dinv;
    CState* st;
    CRed rb;
};

int64_t csr_layout(int64_t n, char* base, CWork* w) {
    const int64_t n_blk = (n + kBlock - 1) / kBlock;
    int64_t off = 0;
    auto take = [&](int64_t bytes) -> void* {
        void* p = base ? (void*)(base + off) : nullptr;
        off += al(bytes);
        return p;
    };
    CWork tmp;
    CWork* o = w ? w : &tmp;
    o->r = (cx<double>*)take(16 * n);
    o->rh = (cx<double>*)take(16 * n);
    o->p[0] = (cx<double>*)take(16 * n);
    o->p[1] = (cx<double>*)take(16 * n);
    o->v[0] = (cx<double>*)take(16 * n);
    o->v[1] = (cx<double>*)take(16 * n);
    o->t = (cx<double>*)take(16 * n);
    o->dinv = (cx<double>*)take(16 * n);
    o->st = (CState*)take(sizeof(CState));
    o->rb.counter = (unsigned int*)take(sizeof(unsigned int) * (1 + n_blk));
    o->rb.gcount = o->rb.counter ? o->rb.counter + 1 : nullptr;
    o->rb.partials = (double*)take(sizeof(double) * 4 * n_blk);
    o->rb.gpartials = (double*)take(sizeof(double) * 4 * n_blk);
    return off;
}

template <typename I>
int csr_solve_impl(int64_t n, const void* indptr, const void* indices, const void* data, const void* b_, void* x_,
                   double tol, int max_iter, int check_every, void* ws, cv_solve_info* info, hipStream_t s) {
    CWork w;
    csr_layout(n, (char*)ws, &w);
    const int n_blk = (int)((n + kBlock - 1) / kBlock);
    CsrA<I> A{(const I*)indptr, (const I*)indices, (const cx<double>*)data, w.dinv, n, n_blk};
    const cx<double>* b = (const cx<double>*)b_;
    cx<double>* x = (cx<double>*)x_;
    k_csr_zero<<<(n_blk + 1 + 255) / 256, 256, 0, s>>>(w.rb.counter, n_blk + 1);
    k_csr_dinv<I><<<n_blk, kBlock, 0, s>>>(A, w.dinv);
    k_csr_init<I><<<n_blk, kBlock, 0, s>>>(A, b, x, w.r, w.rh, w.rb, w.st, tol, max_iter, 0);
    CV_LAUNCH_CHECK();
    CState h{};
    int checks = 0;
    while (true) {
        int cur = 0;
        for (int c = 0; c < check_every; ++c) {
            const int nxt = cur ^ 1;
            k_csr_p<I><<<n_blk, kBlock, 0, s>>>(A, w.r, w.rh, w.p[cur], w.v[cur], w.p[nxt], w.v[nxt], w.rb, w.st);
            k_csr_s<I><<<n_blk, kBlock, 0, s>>>(A, w.r, w.v[nxt], w.t, w.rb, w.st);
            k_csr_x<I><<<n_blk, kBlock, 0, s>>>(A, x, w.p[nxt], w.r, w.v[nxt], w.t, w.rh, w.rb, w.st);
            cur = nxt;
        }
        CV_LAUNCH_CHECK();
        CV_HIP(hipMemcpyAsync(&h, w.st, sizeof(CState), hipMemcpyDeviceToHost, s));
        CV_HIP(hipStreamSynchronize(s));
        if (h.done == kRun) continue;
        if (h.done != kConv) break;
        // verify with the true residual; a restart continues the loop
        k_csr_init<I><<<n_blk, kBlock, 0, s>>>(A, b, x, w.r, w.rh, w.rb, w.st, tol, max_iter, 1);
        CV_LAUNCH_CHECK();
        CV_HIP(hipMemcpyAsync(&h, w.st, sizeof(CState), hipMemcpyDeviceToHost, s));
        CV_HIP(hipStreamSynchronize(s));
        if (h.done != kRun || ++checks >= 8) break;
    }
    info->iterations = h.iter;
    info->restarts = h.restarts;
    info->pad = 0;
    info->rel_residual = h.bb > 0 ? sqrt(h.rr_true / h.bb) : 0.0;
    switch (h.done) {
        case kConv: info->status = CV_OK; break;
        case kNonFin: info->status = CV_ERR_NONFINITE; break;
        case kBreak: info->status = CV_ERR_BREAKDOWN; break;
        default: info->status = CV_ERR_NOT_CONVERGED; break;
    }
    if (info->status != CV_OK)
        set_error("CSR BiCGSTAB stopped after %d iterations at relative residual %.3e (tolerance %.1e)", h.iter,
                  info->rel_residual, tol);
    return info->status;
}

}  // namespace
}  // namespace cv

using namespace cv;

extern "C" int64_t cv_csr_solve_workspace_bytes(int64_t n) { return csr_layout(n, nullptr, nullptr); }

extern "C" int cv_csr_solve(int64_t n, int index_bits, const void* indptr, const void* indices, const void* data,
                            const void* b, void* x, double tol, int max_iter, int check_every, void* workspace,
                            int64_t workspace_bytes, cv_solve_info* info, void* stream) {
    CV_CHECK(indptr && indices && data && b && x && workspace && info, "null pointer argument");
    CV_CHECK(n > 0, "empty system");
    CV_CHECK(n <= (int64_t)kBlock * 0x7fffffff, "system too large");
    CV_CHECK(index_bits == 32 || index_bits == 64, "index_bits must be 32 or 64, got %d", index_bits);
    CV_CHECK(tol > 0 && max_iter > 0, "tol and max_iter must be positive");
    CV_CHECK(index_bits == 64 || n < 0x7fffffffll, "int32 indices cannot address %lld rows", (long long)n);
    const int64_t need = csr_layout(n, nullptr, nullptr);
    CV_CHECK(workspace_bytes >= need, "workspace too small: %lld < %lld", (long long)workspace_bytes,
             (long long)need);
    if (check_every <= 0) check_every = 32;
    check_every += check_every & 1;   // even: p / v double buffers back in place
    hipStream_t s = as_stream(stream);
    if (index_bits == 32)
        return csr_solve_impl<int32_t>(n, indptr, indices, data, b, x, tol, max_iter, check_every, workspace, info,
                                       s);
    return csr_solve_impl<int64_t>(n, indptr, indices, data, b, x, tol, max_iter, check_every, workspace, info, s);
}
