// Generic sparse primitives for ceviche's COO interface (entries, indices):
// a deterministic segmented product-sum kernel (SpMV, A^T / A^H products,
// outer-product gathers, SpGEMM numerics and their derivatives), and a dense
// complex128 LU with partial pivoting for small general systems.
//
// The symbolic part (grouping entries by row / column / output position) is
// done once per index set on the host, like the reference's COO->CSR
// conversion (ceviche/utils.py:13-25); every numeric operation runs here.
#include "common.h"

namespace cv {

// out[o(s)] (+)= alpha * sum_{q in [ptr[s], ptr[s+1])} opu(u[iu[q]]) * opv(v[iv[q]])
// with o(s) = oidx ? oidx[s] : s and opu / opv optional conjugation.  One
// thread per segment; the sum runs in storage order (deterministic).
__global__ void __launch_bounds__(256)
k_segsum(int64_t n_seg, const int64_t* __restrict__ ptr, const int64_t* __restrict__ iu,
         const int64_t* __restrict__ iv, const cx<double>* __restrict__ u, const cx<double>* __restrict__ v,
         int conj_u, int conj_v, cx<double> alpha, const int64_t* __restrict__ oidx, int accumulate,
         cx<double>* __restrict__ out) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_seg) return;
    cx<double> acc = czero<double>();
    for (int64_t q = ptr[s]; q < ptr[s + 1]; ++q) {
        cx<double> a = u[iu ? iu[q] : q];
        cx<double> b = v[iv ? iv[q] : q];
        if (conj_u) a = conj(a);
        if (conj_v) b = conj(b);
        acc = acc + a * b;
    }
    const int64_t o = oidx ? oidx[s] : s;
    acc = alpha * acc;
    out[o] = accumulate ? out[o] + acc : acc;
}

// ------------------------------------------------------------- dense LU
// Column-major n x n complex128.  Pivot choice: max |re| + |im| (LAPACK's
// izamax measure), lowest index on ties.
__global__ void __launch_bounds__(256)
k_lu_pivot(int n, cx<double>* __restrict__ A, int k, int32_t* __restrict__ piv, int32_t* __restrict__ info) {
    __shared__ double s_val[256];
    __shared__ int s_idx[256];
    __shared__ int s_p;
    double best = -1.0;
    int bi = n;
    for (int i = k + threadIdx.x; i < n; i += blockDim.x) {
        const cx<double> a = A[(int64_t)k * n + i];
        const double m = fabs(a.re) + fabs(a.im);
        if (m > best || (m == best && i < bi)) {
            best = m;
            bi = i;
        }
    }
    s_val[threadIdx.x] = best;
    s_idx[threadIdx.x] = bi;
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            const double ov = s_val[threadIdx.x + off];
            const int oi = s_idx[threadIdx.x + off];
            if (ov > s_val[threadIdx.x] || (ov == s_val[threadIdx.x] && oi < s_idx[threadIdx.x])) {
                s_val[threadIdx.x] = ov;
                s_idx[threadIdx.x] = oi;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        s_p = s_idx[0];
        piv[k] = s_p;
        if (s_val[0] == 0.0 && *info == 0) *info = k + 1;
    }
    __syncthreads();
    const int p = s_p;
    if (p != k) {
        for (int j = threadIdx.x; j < n; j += blockDim.x) {
            const cx<double> t = A[(int64_t)j * n + k];
            A[(int64_t)j * n + k] = A[(int64_t)j * n + p];
            A[(int64_t)j * n + p] = t;
        }
    }
    __syncthreads();
    const cx<double> d = A[(int64_t)k * n + k];
    if (d.re == 0.0 && d.im == 0.0) return;
    const cx<double> inv = cdiv(cone<double>(), d);
    for (int i = k + 1 + threadIdx.x; i < n; i += blockDim.x) A[(int64_t)k * n + i] = A[(int64_t)k * n + i] * inv;
}

// A[i, j] -= A[i, k] A[k, j] for i, j > k (i fastest: coalesced columns)
__global__ void __launch_bounds__(256) k_lu_update(int n, cx<double>* __restrict__ A, int k) {
    const int64_t m = n - k - 1;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m * m) return;
    const int i = k + 1 + (int)(e % m);
    const int j = k + 1 + (int)(e / m);
    A[(int64_t)j * n + i] = A[(int64_t)j * n + i] - A[(int64_t)k * n + i] * A[(int64_t)j * n + k];
}

// Solves with the packed factors (P A = L U) for one right-hand side per
// workgroup.  trans: 0 A x = b, 1 A^T x = b, 2 A^H x = b.
__global__ void __launch_bounds__(1024)
k_lu_solve(int n, const cx<double>* __restrict__ LU, const int32_t* __restrict__ piv, cx<double>* __restrict__ B,
           int trans) {
    cx<double>* b = B + (int64_t)blockIdx.x * n;
    const int tid = threadIdx.x, nt = blockDim.x;
    auto lu = [&](int i, int j) -> cx<double> {
        const cx<double> a = LU[(int64_t)j * n + i];
        return trans == 2 ? conj(a) : a;
    };
    if (trans == 0) {
        if (tid == 0)
            for (int k = 0; k < n; ++k) {
                const int p = piv[k];
                if (p != k) {
                    const cx<double> t = b[k];
                    b[k] = b[p];
                    b[p] = t;
                }
            }
        __syncthreads();
        for (int j = 0; j < n; ++j) {  // L y = P b (unit lower)
            const cx<double> xj = b[j];
            for (int i = j + 1 + tid; i < n; i += nt) b[i] = b[i] - lu(i, j) * xj;
            __syncthreads();
        }
        for (int j = n - 1; j >= 0; --j) {  // U x = y
            if (tid == 0) b[j] = cdiv(b[j], lu(j, j));
            __syncthreads();
            const cx<double> xj = b[j];
            for (int i = tid; i < j; i += nt) b[i] = b[i] - lu(i, j) * xj;
            __syncthreads();
        }
    } else {
        // A^T = U^T L^T P (P^T = P^-1):  U^T y = b, L^T z = y, x = P^T z
        for (int j = 0; j < n; ++j) {
            if (tid == 0) b[j] = cdiv(b[j], lu(j, j));
            __syncthreads();
            const cx<double> yj = b[j];
            for (int i = j + 1 + tid; i < n; i += nt) b[i] = b[i] - lu(j, i) * yj;
            __syncthreads();
        }
        for (int j = n - 1; j >= 0; --j) {
            const cx<double> zj = b[j];
            for (int i = tid; i < j; i += nt) b[i] = b[i] - lu(j, i) * zj;
            __syncthreads();
        }
        if (tid == 0)
            for (int k = n - 1; k >= 0; --k) {
                const int p = piv[k];
                if (p != k) {
                    const cx<double> t = b[k];
                    b[k] = b[p];
                    b[p] = t;
                }
            }
    }
}

__global__ void k_zero_i32(int32_t* p) { *p = 0; }

}  // namespace cv

using namespace cv;

extern "C" int cv_segsum(int64_t n_seg, const int64_t* ptr, const int64_t* iu, const int64_t* iv, const void* u,
                         const void* v, int conj_u, int conj_v, double alpha_re, double alpha_im,
                         const int64_t* oidx, int accumulate, void* out, void* stream) {
    CV_CHECK(n_seg >= 0, "n_seg < 0");
    CV_CHECK(ptr && u && v && out, "null pointer argument");
    if (n_seg == 0) return CV_OK;
    const int64_t nb = (n_seg + 255) / 256;
    CV_CHECK(nb < (1ll << 31), "too many segments");
    k_segsum<<<(unsigned)nb, 256, 0, as_stream(stream)>>>(n_seg, ptr, iu, iv, (const cx<double>*)u,
                                                          (const cx<double>*)v, conj_u, conj_v,
                                                          mk(alpha_re, alpha_im), oidx, accumulate,
                                                          (cx<double>*)out);
    CV_LAUNCH_CHECK();
    return CV_OK;
}

extern "C" int cv_dense_lu(int n, void* A, int32_t* piv, int32_t* info_dev, void* stream) {
    CV_CHECK(n > 0 && A && piv && info_dev, "bad LU arguments");
    hipStream_t s = as_stream(stream);
    k_zero_i32<<<1, 1, 0, s>>>(info_dev);
    for (int k = 0; k < n; ++k) {
        k_lu_pivot<<<1, 256, 0, s>>>(n, (cx<double>*)A, k, piv, info_dev);
        const int64_t m = (int64_t)(n - k - 1) * (n - k - 1);
        if (m > 0) k_lu_update<<<(unsigned)((m + 255) / 256), 256, 0, s>>>(n, (cx<double>*)A, k);
    }
    CV_LAUNCH_CHECK();
    return CV_OK;
}

extern "C" int cv_dense_lu_solve(int n, const void* LU, const int32_t* piv, void* B, int nrhs, int trans,
                                 void* stream) {
    CV_CHECK(n > 0 && LU && piv && B && nrhs > 0, "bad solve arguments");
    CV_CHECK(trans >= 0 && trans <= 2, "trans must be 0 (N), 1 (T) or 2 (H)");
    const int nt = n >= 1024 ? 1024 : (n >= 256 ? 256 : 64);
    k_lu_solve<<<nrhs, nt, 0, as_stream(stream)>>>(n, (const cx<double>*)LU, piv, (cx<double>*)B, trans);
    CV_LAUNCH_CHECK();
    return CV_OK;
}
