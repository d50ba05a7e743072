// Restarted flexible GMRES for one 2D FDFD instance, with a choice of
// preconditioner: none, Jacobi, or the convolutional network of cnn.hip.
//
// Replaces the sparse direct solve behind ceviche's sp_solve
// (ceviche/primitives.py:103-119 -> solvers.py:37-62) like the BiCGSTAB of
// krylov.hip; this solver is the one that takes a preconditioner that changes
// from application to application (a network is not a fixed linear map).
//
// Precision: the outer iterate x and the residual r = b - A x are complex128
// and the residual is recomputed from x in FP64 at every restart, so the
// convergence test never trusts a recurrence.  One restart cycle solves
// A d = r approximately in complex64 (Arnoldi vectors V, preconditioned
// vectors Z) with FP64 inner products and two passes of classical
// Gram-Schmidt; then x += Z y in FP64.  The (m+1) x m Hessenberg least
// squares problem is solved on the host with complex Givens rotations.
#include "fdfd_stencil.h"

#include <vector>
#include <complex>
#include <algorithm>

namespace cv {

constexpr int kFgCPT = 4;                    // cells per thread in the vector kernels
constexpr int kFgCells = kBlock * kFgCPT;    // cells per workgroup
constexpr int kFgMaxM = 64;                  // longest restart cycle

// Sum of NV values per workgroup into partial[blockIdx.x * NV + c] via
// wave shuffles and LDS (fixed order).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// r = b - A64 x (or A64^T x; row-major single instance), partial[wg] = {|r|^2, |b|^2}
template <bool TRANS>
__global__ void __launch_bounds__(kBlock)
k_fg_resid(const cv_fdfd_desc* __restrict__ desc, const cx<double>* __restrict__ prof, const double* __restrict__ m0,
           const double* __restrict__ m1, const cx<double>* __restrict__ b, const cx<double>* __restrict__ x,
           cx<double>* __restrict__ r, double* __restrict__ partial) {
    __shared__ double lds[kWaves * 2];
    // one instance, or the equal-sized blocks of one frequency-coupled system
    const cv_fdfd_desc d0 = load_desc(desc, 0);
    const int64_t n0 = (int64_t)d0.nx * d0.ny;
    const int64_t n = n0 * d0.nblk;
    auto fx = [&](int64_t k) { return x[k]; };
    double acc[2] = {0.0, 0.0};
#pragma unroll
    for (int q = 0; q < kFgCPT; ++q) {
        const int64_t k = (int64_t)blockIdx.x * kFgCells + q * kBlock + threadIdx.x;
        if (k >= n) break;
        const int inst = (int)(k / n0);
        const cv_fdfd_desc d = (inst == 0) ? d0 : load_desc(desc, inst);
        const int64_t loc = k - (int64_t)inst * n0;
        const int i = (int)(loc / d.ny), j = (int)(loc - (int64_t)i * d.ny);
        const cx<double> bk = b[k];
        const cx<double> rk = bk - apply_any<double, TRANS>(d, prof, m0, m1, i, j, fx, x[k]);
        r[k] = rk;
        acc[0] += norm2(rk);
        acc[1] += norm2(bk);
    }
    block_sum<2>(acc, lds);
    if (threadIdx.x == 0) {
        partial[blockIdx.x * 2] = acc[0];
        partial[blockIdx.x * 2 + 1] = acc[1];
    }
}

// out[c] = sum_g partial[g * nv + c] for c < nv (one workgroup per value;
// fixed order).  If acc_into, out[c] += that sum.
__global__ void __launch_bounds__(kBlock)
k_fg_reduce(const double* __restrict__ partial, int n_wg, int nv, double* __restrict__ out, int acc_into) {
    __shared__ double lds[kWaves];
    const int c = blockIdx.x;
    double s = 0.0;
    for (int g = threadIdx.x; g < n_wg; g += kBlock) s += partial[(int64_t)g * nv + c];
    double v[1] = {s};
    block_sum<1>(v, lds);
    if (threadIdx.x == 0) out[c] = acc_into ? out[c] + v[0] : v[0];
}

// v0 = r64 * inv (complex128 -> complex64)
__global__ void __launch_bounds__(kBlock)
k_fg_start(const cx<double>* __restrict__ r, cx<float>* __restrict__ v, double inv, int64_t n) {
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) return;
    const cx<double> a = r[k];
    v[k] = mk<float>((float)(a.re * inv), (float)(a.im * inv));
}

// w = A32 z (or A32^T z)
template <bool TRANS>
__global__ void __launch_bounds__(kBlock)
k_fg_apply(const cv_fdfd_desc* __restrict__ desc, const cx<float>* __restrict__ prof, const float* __restrict__ m0,
           const float* __restrict__ m1, const cx<float>* __restrict__ z, cx<float>* __restrict__ w) {
    const cv_fdfd_desc d0 = load_desc(desc, 0);
    const int64_t n0 = (int64_t)d0.nx * d0.ny;
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= n0 * d0.nblk) return;
    const int inst = (int)(k / n0);
    const cv_fdfd_desc d = (inst == 0) ? d0 : load_desc(desc, inst);
    const int64_t loc = k - (int64_t)inst * n0;
    const int i = (int)(loc / d.ny), j = (int)(loc - (int64_t)i * d.ny);
    auto fz = [&](int64_t q) { return z[q]; };
    w[k] = apply_any<float, TRANS>(d, prof, m0, m1, i, j, fz, z[k]);
}

// z = v / diag(A32)
__global__ void __launch_bounds__(kBlock)
k_fg_jacobi(const cv_fdfd_desc* __restrict__ desc, const cx<float>* __restrict__ prof, const float* __restrict__ m0,
            const float* __restrict__ m1, const cx<float>* __restrict__ v, cx<float>* __restrict__ z) {
    const cv_fdfd_desc d = load_desc(desc, 0);
    const int64_t n = (int64_t)d.nx * d.ny;
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) return;
    const int i = (int)(k / d.ny), j = (int)(k - (int64_t)i * d.ny);
    const Row5<float> rc = row_coefs<float>(d, prof, m0, m1, i, j);
    const cx<double> dg = ccast<double>(rc.c);
    const cx<double> q = cdiv(ccast<double>(v[k]), dg);
    z[k] = mk<float>((float)q.re, (float)q.im);
}

// partial[wg][2c .. 2c+1] = sum over the workgroup's cells of conj(V_c) w,
// c < nvec (V_c = V + c * n)
__global__ void __launch_bounds__(kBlock)
k_fg_mdot(const cx<float>* __restrict__ V, int nvec, const cx<float>* __restrict__ w, int64_t n,
          double* __restrict__ partial) {
    __shared__ double lds[kWaves][2 * kFgMaxM + 2];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    cx<double> wv[kFgCPT];
    int64_t ks[kFgCPT];
#pragma unroll
    for (int q = 0; q < kFgCPT; ++q) {
        ks[q] = (int64_t)blockIdx.x * kFgCells + q * kBlock + threadIdx.x;
        wv[q] = ks[q] < n ? ccast<double>(w[ks[q]]) : czero<double>();
    }
    for (int c = 0; c < nvec; ++c) {
        const cx<float>* vc = V + (int64_t)c * n;
        double sr = 0.0, si = 0.0;
#pragma unroll
        for (int q = 0; q < kFgCPT; ++q) {
            if (ks[q] < n) {
                const cx<double> a = ccast<double>(vc[ks[q]]);
                sr += a.re * wv[q].re + a.im * wv[q].im;
                si += a.re * wv[q].im - a.im * wv[q].re;
            }
        }
        sr = wave_sum(sr);
        si = wave_sum(si);
        if (lane == 0) {
            lds[wave][2 * c] = sr;
            lds[wave][2 * c + 1] = si;
        }
    }
    __syncthreads();
    for (int c = threadIdx.x; c < 2 * nvec; c += kBlock) {
        double s = 0.0;
#pragma unroll
        for (int wv2 = 0; wv2 < kWaves; ++wv2) s += lds[wv2][c];
        partial[(int64_t)blockIdx.x * 2 * nvec + c] = s;
    }
}

// w -= sum_c h_c V_c  (h complex double, on the device)
__global__ void __launch_bounds__(kBlock)
k_fg_maxpy(const cx<float>* __restrict__ V, int nvec, const double* __restrict__ h, cx<float>* __restrict__ w,
           int64_t n) {
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) return;
    cx<double> acc = ccast<double>(w[k]);
    for (int c = 0; c < nvec; ++c) {
        const cx<double> hc = mk(h[2 * c], h[2 * c + 1]);
        acc = acc - hc * ccast<double>(V[(int64_t)c * n + k]);
    }
    w[k] = mk<float>((float)acc.re, (float)acc.im);
}

// v = w / sqrt(nrm2[0])  (in place allowed)
__global__ void __launch_bounds__(kBlock)
k_fg_normalize(cx<float>* __restrict__ w, const double* __restrict__ nrm2, int64_t n) {
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) return;
    const double s = nrm2[0] > 0.0 ? 1.0 / sqrt(nrm2[0]) : 0.0;
    const cx<float> a = w[k];
    w[k] = mk<float>((float)(a.re * s), (float)(a.im * s));
}

// x += sum_c y_c Z_c  (complex128 accumulation)
__global__ void __launch_bounds__(kBlock)
k_fg_update(const cx<float>* __restrict__ Z, int nvec, const double* __restrict__ y, cx<double>* __restrict__ x,
            int64_t n) {
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) return;
    cx<double> acc = x[k];
    for (int c = 0; c < nvec; ++c) acc = acc + mk(y[2 * c], y[2 * c + 1]) * ccast<double>(Z[(int64_t)c * n + k]);
    x[k] = acc;
}

static int64_t al256(int64_t x) { return (x + 255) & ~int64_t(255); }

struct FgWork {
    cx<float>* V;       // (m+1) x n
    cx<float>* Z;       // m x n
    cx<double>* r;      // n
    double* partial;    // n_wg x 2(m+1)
    double* h;          // 2 x 2(m+1) (two Gram-Schmidt passes) + norm^2 + true-residual pair
    void* net_ws;
};

static int64_t fg_layout(int64_t n, int m, int64_t net_bytes, char* base, FgWork* w) {
    int64_t off = 0;
    auto take = [&](int64_t bytes) -> void* {
        void* p = base ? (void*)(base + off) : nullptr;
        off += al256(bytes);
        return p;
    };
    const int64_t n_wg = (n + kFgCells - 1) / kFgCells;
    FgWork t;
    FgWork* o = w ? w : &t;
    o->V = (cx<float>*)take(8 * n * (m + 1));
    o->Z = (cx<float>*)take(8 * n * m);
    o->r = (cx<double>*)take(16 * n);
    o->partial = (double*)take(8 * n_wg * 2 * (m + 1));
    o->h = (double*)take(8 * (4 * (m + 1) + 4));
    o->net_ws = take(net_bytes);
    return off;
}

}  // namespace cv

using namespace cv;

extern "C" int64_t cv_fdfd_fgmres_workspace_bytes(int nx, int ny, int restart, int precond, int levels) {
    if (nx <= 0 || ny <= 0 || restart < 1 || restart > kFgMaxM) return 0;
    const int64_t net = (precond == CV_PRECOND_UNET) ? cv_unet_workspace_bytes(levels, 1, nx, ny) : 0;
    return fg_layout((int64_t)nx * ny, restart, net, nullptr, nullptr);
}

namespace cv {

// With a direct factorisation as the preconditioner a solve takes a few
// restart cycles; four consecutive full cycles (all m steps taken) that reduce
// the true residual norm by less than 10 % end it as not converged (the
// caller's fallback takes over).  A cycle that stops early has met its
// estimated target, so it neither counts nor resets the count.
struct StallStop {
    int stalls = 0;
    bool operator()(double rr, double rr_prev, bool full) {
        if (full) stalls = rr > 0.81 * rr_prev ? stalls + 1 : 0;
        return stalls >= 4;
    }
};

// The restarted flexible GMRES loop shared by every preconditioner.
// precondition(v, z) writes z ~ M^-1 v; after_cycle(rr, rr_prev, full) may
// change it (full: the cycle took all m steps), and returning true ends the
// solve as not converged (unless the tolerance was just met).
template <bool TRANS, typename Precond, typename AfterCycle>
static int fgmres_core(const cv_fdfd_desc* desc, const cv_fdfd_desc& dh, const void* prof64, const void* m0_64,
                       const void* m1_64, const void* prof32, const void* m0_32, const void* m1_32, const void* b,
                       void* x, double tol, int m, int max_iter, double inner_rtol, FgWork& w, Precond&& precondition,
                       AfterCycle&& after_cycle, cv_solve_info* info, hipStream_t s) {
    const int64_t n = (int64_t)dh.nx * dh.ny * dh.nblk;
    const int n_wg = (int)((n + kFgCells - 1) / kFgCells);
    const int gv = (int)((n + kBlock - 1) / kBlock);
    const cx<double>* b64 = (const cx<double>*)b;
    cx<double>* x64 = (cx<double>*)x;
    double* h1 = w.h;                       // first Gram-Schmidt pass, then the sum
    double* h2 = w.h + 2 * (m + 1);         // second pass
    double* nrm = w.h + 4 * (m + 1);        // |w|^2 at [0]; |r|^2, |b|^2 at [2], [3]
    auto true_resid = [&](double* rr, double* bb) -> int {
        k_fg_resid<TRANS><<<n_wg, kBlock, 0, s>>>(desc, (const cx<double>*)prof64, (const double*)m0_64,
                                                  (const double*)m1_64, b64, x64, w.r, w.partial);
        k_fg_reduce<<<2, kBlock, 0, s>>>(w.partial, n_wg, 2, nrm + 2, 0);
        CV_LAUNCH_CHECK();
        double hv[2];
        CV_HIP(hipMemcpyAsync(hv, nrm + 2, sizeof(hv), hipMemcpyDeviceToHost, s));
        CV_HIP(hipStreamSynchronize(s));
        *rr = hv[0];
        *bb = hv[1];
        return CV_OK;
    };
    double rr = 0, bb = 0;
    int rc = true_resid(&rr, &bb);
    if (rc != CV_OK) return rc;
    const double target = tol * sqrt(bb);
    int iters = 0, cycles = 0, status = CV_ERR_NOT_CONVERGED;
    std::vector<std::complex<double>> H((size_t)(m + 1) * m), g(m + 1), sn(m), y(m);
    std::vector<double> cs(m), hbuf(2 * (m + 1) * 2 + 4);
    using C = std::complex<double>;
    while (true) {
        const double beta = sqrt(rr);
        if (!std::isfinite(beta)) { status = CV_ERR_NONFINITE; break; }
        if (beta <= target) { status = CV_OK; break; }
        if (iters >= max_iter) { status = CV_ERR_NOT_CONVERGED; break; }
        k_fg_start<<<gv, kBlock, 0, s>>>(w.r, w.V, 1.0 / beta, n);
        CV_LAUNCH_CHECK();
        std::fill(g.begin(), g.end(), C(0.0));
        g[0] = beta;
        const double inner_target = std::max(0.5 * target, inner_rtol * beta);
        int k = 0;
        for (int j = 0; j < m && iters < max_iter; ++j) {
            cx<float>* vj = w.V + (int64_t)j * n;
            cx<float>* zj = w.Z + (int64_t)j * n;
            cx<float>* wv = w.V + (int64_t)(j + 1) * n;
            if ((rc = precondition(vj, zj)) != CV_OK) return rc;
            k_fg_apply<TRANS><<<gv, kBlock, 0, s>>>(desc, (const cx<float>*)prof32, (const float*)m0_32,
                                             (const float*)m1_32, zj, wv);
            // two passes of classical Gram-Schmidt against V_0 .. V_j
            k_fg_mdot<<<n_wg, kBlock, 0, s>>>(w.V, j + 1, wv, n, w.partial);
            k_fg_reduce<<<2 * (j + 1), kBlock, 0, s>>>(w.partial, n_wg, 2 * (j + 1), h1, 0);
            k_fg_maxpy<<<gv, kBlock, 0, s>>>(w.V, j + 1, h1, wv, n);
            k_fg_mdot<<<n_wg, kBlock, 0, s>>>(w.V, j + 1, wv, n, w.partial);
            k_fg_reduce<<<2 * (j + 1), kBlock, 0, s>>>(w.partial, n_wg, 2 * (j + 1), h2, 0);
            k_fg_maxpy<<<gv, kBlock, 0, s>>>(w.V, j + 1, h2, wv, n);
            k_fg_mdot<<<n_wg, kBlock, 0, s>>>(wv, 1, wv, n, w.partial);
            k_fg_reduce<<<1, kBlock, 0, s>>>(w.partial, n_wg, 2, nrm, 0);
            k_fg_normalize<<<gv, kBlock, 0, s>>>(wv, nrm, n);
            CV_LAUNCH_CHECK();
            CV_HIP(hipMemcpyAsync(hbuf.data(), w.h, sizeof(double) * (4 * (m + 1) + 1), hipMemcpyDeviceToHost, s));
            CV_HIP(hipStreamSynchronize(s));
            ++iters;
            k = j + 1;
            // column j of the Hessenberg matrix
            for (int i = 0; i <= j; ++i)
                H[(size_t)i * m + j] = C(hbuf[2 * i] + hbuf[2 * (m + 1) + 2 * i],
                                         hbuf[2 * i + 1] + hbuf[2 * (m + 1) + 2 * i + 1]);
            const double hn = sqrt(std::max(hbuf[4 * (m + 1)], 0.0));
            for (int i = 0; i < j; ++i) {
                const C a = H[(size_t)i * m + j], c2 = H[(size_t)(i + 1) * m + j];
                H[(size_t)i * m + j] = cs[i] * a + sn[i] * c2;
                H[(size_t)(i + 1) * m + j] = -std::conj(sn[i]) * a + cs[i] * c2;
            }
            const C a = H[(size_t)j * m + j];
            const double r = std::sqrt(std::norm(a) + hn * hn);
            if (r == 0.0 || !std::isfinite(r)) { k = j; break; }
            if (std::abs(a) == 0.0) {
                cs[j] = 0.0;
                sn[j] = 1.0;
            } else {
                cs[j] = std::abs(a) / r;
                sn[j] = (a / std::abs(a)) * hn / r;   // conj(hn) = hn (real)
            }
            H[(size_t)j * m + j] = cs[j] * a + sn[j] * hn;
            g[j + 1] = -std::conj(sn[j]) * g[j];
            g[j] = cs[j] * g[j];
            if (std::abs(g[j + 1]) <= inner_target || hn <= 1e-30 * beta) break;
        }
        ++cycles;
        if (k == 0) { status = CV_ERR_BREAKDOWN; break; }
        // y = R^-1 g (back substitution), x += Z y
        for (int i = k - 1; i >= 0; --i) {
            C acc = g[i];
            for (int q = i + 1; q < k; ++q) acc -= H[(size_t)i * m + q] * y[q];
            y[i] = acc / H[(size_t)i * m + i];
        }
        for (int i = 0; i < k; ++i) {
            hbuf[2 * i] = y[i].real();
            hbuf[2 * i + 1] = y[i].imag();
        }
        CV_HIP(hipMemcpyAsync(h1, hbuf.data(), sizeof(double) * 2 * k, hipMemcpyHostToDevice, s));
        k_fg_update<<<gv, kBlock, 0, s>>>(w.Z, k, h1, x64, n);
        CV_LAUNCH_CHECK();
        const double rr_prev = rr;
        if ((rc = true_resid(&rr, &bb)) != CV_OK) return rc;
        if (after_cycle(rr, rr_prev, k == m) && sqrt(rr) > target) {
            status = CV_ERR_NOT_CONVERGED;
            break;
        }
    }
    info[0].iterations = iters;
    info[0].restarts = cycles;
    info[0].rel_residual = bb > 0 ? sqrt(rr / bb) : 0.0;
    info[0].status = status;
    return status;
}

}  // namespace cv

extern "C" int cv_fdfd_fgmres(const cv_fdfd_desc* desc, const cv_fdfd_desc* desc_host, const void* prof64,
                              const void* m0_64, const void* m1_64, const void* prof32, const void* m0_32,
                              const void* m1_32, const void* b, void* x, double tol, int restart, int max_iter,
                              double inner_rtol, int precond, const cv_unet* net, float net_gain, void* workspace,
                              int64_t workspace_bytes, cv_solve_info* info, void* stream) {
    CV_CHECK(desc && desc_host && prof64 && m0_64 && prof32 && m0_32 && b && x && workspace && info,
             "null pointer argument");
    const cv_fdfd_desc dh = desc_host[0];
    CV_CHECK(dh.nx > 0 && dh.ny > 0 && dh.cell_off == 0, "one instance starting at cell 0 is required");
    CV_CHECK(dh.pol == CV_POL_EZ || (m1_64 && m1_32), "Hz operator needs m1");
    CV_CHECK(dh.nblk == 1 && dh.n_cpl == 0, "frequency-coupled systems are not supported by FGMRES");
    CV_CHECK(restart >= 1 && restart <= kFgMaxM, "restart must be in [1, %d]", kFgMaxM);
    CV_CHECK(tol > 0 && max_iter > 0, "tol and max_iter must be positive");
    CV_CHECK(precond == CV_PRECOND_NONE || precond == CV_PRECOND_JACOBI || precond == CV_PRECOND_UNET,
             "unknown preconditioner %d", precond);
    CV_CHECK(precond != CV_PRECOND_UNET || net, "network preconditioner needs a network");
    const int m = restart;
    const int64_t n = (int64_t)dh.nx * dh.ny;
    const int levels = (precond == CV_PRECOND_UNET) ? net->levels : 0;
    const int64_t need = cv_fdfd_fgmres_workspace_bytes(dh.nx, dh.ny, m, precond, levels);
    CV_CHECK(need > 0 && workspace_bytes >= need, "workspace too small: %lld < %lld", (long long)workspace_bytes,
             (long long)need);
    const int64_t net_bytes = (precond == CV_PRECOND_UNET) ? cv_unet_workspace_bytes(levels, 1, dh.nx, dh.ny) : 0;
    FgWork w;
    fg_layout(n, m, net_bytes, (char*)workspace, &w);
    hipStream_t s = as_stream(stream);
    const int gv = (int)((n + kBlock - 1) / kBlock);
    const float in_scale = (float)sqrt((double)n);     // unit-norm v -> unit rms

    // The network is used while its restart cycles keep reducing the true
    // residual; three cycles in a row that reduce it by less than 1 % switch
    // the rest of the solve to Jacobi (info.pad = 1 reports the fallback).
    int use = precond, stalls = 0, fell_back = 0;
    auto precondition = [&](const cx<float>* v, cx<float>* z) -> int {
        if (use == CV_PRECOND_NONE) {
            CV_HIP(hipMemcpyAsync(z, v, 8 * n, hipMemcpyDeviceToDevice, s));
        } else if (use == CV_PRECOND_JACOBI) {
            k_fg_jacobi<<<gv, kBlock, 0, s>>>(desc, (const cx<float>*)prof32, (const float*)m0_32,
                                              (const float*)m1_32, v, z);
            CV_LAUNCH_CHECK();
        } else {
            return cv_unet_apply(net, CV_C64, v, in_scale, net_gain / in_scale, z, dh.nx, dh.ny, w.net_ws, net_bytes,
                                 stream);
        }
        return CV_OK;
    };

    auto after_cycle = [&](double rr, double rr_prev, bool) {
        if (use == CV_PRECOND_UNET) {
            stalls = (rr > 0.9801 * rr_prev) ? stalls + 1 : 0;
            if (stalls >= 3) {
                use = CV_PRECOND_JACOBI;
                fell_back = 1;
            }
        }
        return false;
    };
    const int status = fgmres_core<false>(desc, dh, prof64, m0_64, m1_64, prof32, m0_32, m1_32, b, x, tol, m, max_iter,
                                          inner_rtol, w, precondition, after_cycle, info, s);
    info[0].pad = fell_back;
    return status;
}

namespace cv {
int direct_apply(int kind, const cv_fdfd_desc* desc, const cv_fdfd_desc& dh, const void* prof32, const void* m0_32,
                 const void* m1_32, const void* factor, const void* v, void* z, int transpose, hipStream_t s);
}

extern "C" int cv_fdfd_fgmres_direct(const cv_fdfd_desc* desc, const cv_fdfd_desc* desc_host, const void* prof64,
                                     const void* m0_64, const void* m1_64, const void* prof32, const void* m0_32,
                                     const void* m1_32, int factor_kind, const void* factor, const void* b, void* x,
                                     int transpose, double tol, int restart, int max_iter, void* workspace,
                                     int64_t workspace_bytes, cv_solve_info* info, void* stream) {
    CV_CHECK(factor_kind == CV_FACTOR_BTD || factor_kind == CV_FACTOR_BCR, "unknown factor kind %d", factor_kind);
    CV_CHECK(desc && desc_host && prof64 && m0_64 && prof32 && m0_32 && factor && b && x && workspace && info,
             "null pointer argument");
    const cv_fdfd_desc dh = desc_host[0];
    CV_CHECK(dh.nx > 0 && dh.ny > 0 && dh.cell_off == 0, "one instance starting at cell 0 is required");
    CV_CHECK(dh.ny <= 128, "the block-tridiagonal preconditioner needs ny <= 128");
    CV_CHECK(dh.pol == CV_POL_EZ || (m1_64 && m1_32), "Hz operator needs m1");
    CV_CHECK(dh.nblk == 1 && dh.n_cpl == 0, "frequency-coupled systems are not supported by FGMRES");
    CV_CHECK(restart >= 1 && restart <= kFgMaxM, "restart must be in [1, %d]", kFgMaxM);
    CV_CHECK(tol > 0 && max_iter > 0, "tol and max_iter must be positive");
    const int m = restart;
    const int64_t need = cv_fdfd_fgmres_workspace_bytes(dh.nx, dh.ny, m, CV_PRECOND_NONE, 0);
    CV_CHECK(need > 0 && workspace_bytes >= need, "workspace too small: %lld < %lld", (long long)workspace_bytes,
             (long long)need);
    FgWork w;
    fg_layout((int64_t)dh.nx * dh.ny, m, 0, (char*)workspace, &w);
    hipStream_t s = as_stream(stream);
    auto precondition = [&](const cx<float>* v, cx<float>* z) -> int {
        return direct_apply(factor_kind, desc, dh, prof32, m0_32, m1_32, factor, v, z, transpose, s);
    };
    StallStop after_cycle;
    int status;
    if (transpose)
        status = fgmres_core<true>(desc, dh, prof64, m0_64, m1_64, prof32, m0_32, m1_32, b, x, tol, m, max_iter, 0.0,
                                   w, precondition, after_cycle, info, s);
    else
        status = fgmres_core<false>(desc, dh, prof64, m0_64, m1_64, prof32, m0_32, m1_32, b, x, tol, m, max_iter, 0.0,
                                    w, precondition, after_cycle, info, s);
    info[0].pad = 0;
    return status;
}

namespace cv {
int bcr_apply_batch(const cv_fdfd_desc& dh, int nb, const void* const* factors, const cx<float>* v, cx<float>* z,
                    int64_t stride, int transpose, hipStream_t s);
}

namespace cv {
int nd_apply(const cv_nd_plan* P, const void* plan_dev, const void* factor, const cx<float>* v, cx<float>* z,
             int transpose, hipStream_t s);
}

extern "C" int cv_fdfd_fgmres_nd(const cv_fdfd_desc* desc, const cv_fdfd_desc* desc_host, const void* prof64,
                                 const void* m0_64, const void* m1_64, const void* prof32, const void* m0_32,
                                 const void* m1_32, const cv_nd_plan* plan, const void* plan_dev, const void* factor,
                                 const void* b, void* x, int transpose, double tol, int restart, int max_iter,
                                 void* workspace, int64_t workspace_bytes, cv_solve_info* info, void* stream) {
    CV_CHECK(desc && desc_host && prof64 && m0_64 && prof32 && m0_32 && plan && plan_dev && factor && b && x &&
                 workspace && info,
             "null pointer argument");
    const cv_fdfd_desc dh = desc_host[0];
    CV_CHECK(dh.nx > 0 && dh.ny > 0 && dh.cell_off == 0, "one instance starting at cell 0 is required");
    CV_CHECK(dh.pol == CV_POL_EZ || (m1_64 && m1_32), "Hz operator needs m1");
    CV_CHECK(dh.nblk == 1 && dh.n_cpl == 0, "frequency-coupled systems are not supported by FGMRES");
    CV_CHECK(restart >= 1 && restart <= kFgMaxM, "restart must be in [1, %d]", kFgMaxM);
    CV_CHECK(tol > 0 && max_iter > 0, "tol and max_iter must be positive");
    const int m = restart;
    const int64_t need = cv_fdfd_fgmres_workspace_bytes(dh.nx, dh.ny, m, CV_PRECOND_NONE, 0);
    CV_CHECK(need > 0 && workspace_bytes >= need, "workspace too small: %lld < %lld", (long long)workspace_bytes,
             (long long)need);
    FgWork w;
    fg_layout((int64_t)dh.nx * dh.ny, m, 0, (char*)workspace, &w);
    hipStream_t s = as_stream(stream);
    auto precondition = [&](const cx<float>* v, cx<float>* z) -> int {
        return nd_apply(plan, plan_dev, factor, v, z, transpose, s);
    };
    StallStop after_cycle;
    int status;
    if (transpose)
        status = fgmres_core<true>(desc, dh, prof64, m0_64, m1_64, prof32, m0_32, m1_32, b, x, tol, m, max_iter, 0.0,
                                   w, precondition, after_cycle, info, s);
    else
        status = fgmres_core<false>(desc, dh, prof64, m0_64, m1_64, prof32, m0_32, m1_32, b, x, tol, m, max_iter, 0.0,
                                    w, precondition, after_cycle, info, s);
    info[0].pad = 0;
    return status;
}

extern "C" int cv_fdfd_fgmres_blockdiag(const cv_fdfd_desc* desc, const cv_fdfd_desc* desc_host, const void* prof64,
                                        const void* m0_64, const void* m1_64, const void* prof32, const void* m0_32,
                                        const void* m1_32, const cv_fdfd_desc* block_desc_host,
                                        const void* const* factors, const void* b, void* x, int transpose, double tol,
                                        int restart, int max_iter, void* workspace, int64_t workspace_bytes,
                                        cv_solve_info* info, void* stream) {
    CV_CHECK(desc && desc_host && prof64 && m0_64 && prof32 && m0_32 && block_desc_host && factors && b && x &&
                 workspace && info,
             "null pointer argument");
    const cv_fdfd_desc dh = desc_host[0];
    const int nb = dh.nblk;
    CV_CHECK(dh.nx > 0 && dh.ny > 0 && dh.cell_off == 0 && dh.blk == 0 && nb >= 1, "block 0 of a system at cell 0 is required");
    CV_CHECK(dh.ny <= 128, "the block factorisations need ny <= 128");
    const int64_t n0 = (int64_t)dh.nx * dh.ny;
    for (int m = 0; m < nb; ++m) {
        const cv_fdfd_desc& d = desc_host[m];
        CV_CHECK(d.nx == dh.nx && d.ny == dh.ny && d.blk == m && d.nblk == nb && d.cell_off == m * n0 &&
                     d.n_cpl == dh.n_cpl && d.cpl_off == dh.cpl_off && d.w2_off == dh.w2_off && d.pol == CV_POL_EZ,
                 "block %d does not match block 0 of its system", m);
        const cv_fdfd_desc& f = block_desc_host[m];
        CV_CHECK(f.nx == dh.nx && f.ny == dh.ny && f.nblk == 1 && f.cell_off == 0 && factors[m],
                 "factor %d does not describe one block of the system", m);
    }
    CV_CHECK(restart >= 1 && restart <= kFgMaxM, "restart must be in [1, %d]", kFgMaxM);
    CV_CHECK(tol > 0 && max_iter > 0, "tol and max_iter must be positive");
    const int m = restart;
    const int64_t need = cv_fdfd_fgmres_workspace_bytes(dh.nx * nb, dh.ny, m, CV_PRECOND_NONE, 0);
    CV_CHECK(need > 0 && workspace_bytes >= need, "workspace too small: %lld < %lld", (long long)workspace_bytes,
             (long long)need);
    FgWork w;
    fg_layout(n0 * nb, m, 0, (char*)workspace, &w);
    hipStream_t s = as_stream(stream);
    // block-diagonal preconditioner: each sideband block's own factorisation,
    // the inter-block coupling left to the outer iteration
    auto precondition = [&](const cx<float>* v, cx<float>* z) -> int {
        return bcr_apply_batch(block_desc_host[0], nb, factors, v, z, n0, transpose, s);
    };
    // no fallback behind this solve: it runs to max_iter
    auto after_cycle = [](double, double, bool) { return false; };
    int status;
    if (transpose)
        status = fgmres_core<true>(desc, dh, prof64, m0_64, m1_64, prof32, m0_32, m1_32, b, x, tol, m, max_iter, 0.0,
                                   w, precondition, after_cycle, info, s);
    else
        status = fgmres_core<false>(desc, dh, prof64, m0_64, m1_64, prof32, m0_32, m1_32, b, x, tol, m, max_iter, 0.0,
                                    w, precondition, after_cycle, info, s);
    for (int q = 0; q < nb; ++q) info[q] = info[0];
    info[0].pad = 0;
    return status;
}
