// Block-tridiagonal direct factorisation of the 2D FDFD operator, used as the
// preconditioner of an FP64-refined flexible GMRES (fgmres.hip).
//
// The reference solves A x = b with SuperLU (ceviche/solvers.py:50-62).  Here,
// for grids whose y extent fits one workgroup (ny <= 128), A without its
// periodic x-wrap entries is block tridiagonal in x: row block i (the ny cells
// of column i) couples to blocks i-1 and i+1 through diagonal matrices
// (lo_i = the x-minus stencil coefficients of the rows of block i, up_i = the
// x-plus ones) and its diagonal block D_i is the y-stencil (tridiagonal plus
// the periodic / Bloch corners).  Block Gaussian elimination along x,
//     S_0 = D_0,   S_i = D_i - diag(lo_i) S_{i-1}^{-1} diag(up_{i-1}),
// stores the dense inverses S_i^{-1} (complex64, nx * ny^2 entries), and one
// application of the factorisation is a forward and a backward sweep of dense
// matrix-vector products:
//     y_i = S_i^{-1} (v_i - lo_i .* y_{i-1}),   z_i = y_i - S_i^{-1} (up_i .* z_{i+1}).
// The transposed system uses the same inverses transposed (its Schur
// complements are S_i^T).  Pivoting is partial (rows) inside each S_i; there
// is none across blocks, and the outer Krylov iteration with the true FP64
// residual restores the dropped x-wrap and any loss of accuracy.
//
// Both sweeps are sequential in x, so each kernel is one workgroup of 1024
// threads; the factorisation holds S_i in registers (8 rows x 2 columns per
// thread) and inverts it by Gauss-Jordan elimination with the pivot row /
// column staged through LDS.
#include "fdfd_stencil.h"

namespace cv {

constexpr int kBtdThreads = 1024;
constexpr int kBtdMax = 128;          // largest block (ny)

struct BtdFactorArgs {
    const cv_fdfd_desc* desc;
    const cx<float>* prof;
    const float* m0;
    const float* m1;
    cx<float>* sinv;      // nx blocks of ny x ny, row-major
    cx<float>* lo;        // nx x ny: x-minus coefficients of the rows (0 in block 0)
    cx<float>* up;        // nx x ny: x-plus coefficients of the rows (0 in the last block)
    int* status;          // 0 ok, 1 + block index of a zero pivot
};

__device__ __forceinline__ float cabs2(cx<float> a) { return a.re * a.re + a.im * a.im; }

// v from the lane the DPP control selects (all lanes active)
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
__device__ __forceinline__ float lane_f(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

__device__ __forceinline__ cx<float> cdivf(cx<float> a, cx<float> b) {
    const float d = b.re * b.re + b.im * b.im;
    return mk<float>((a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d);
}

// Thread layout: wave w owns rows 8w .. 8w+7, lane l owns columns l and l+64,
// so every row test in the elimination is uniform across a wave and the
// pivot row / column move through LDS without bank conflicts.
constexpr int kFRows = 8;                     // rows per wave
constexpr int kFCols = kBtdMax / 64;          // columns per lane (2)

__global__ void __launch_bounds__(kBtdThreads) k_btd_factor(BtdFactorArgs a) {
    __shared__ cx<float> stage[kBtdMax * kBtdMax];
    __shared__ cx<float> colbuf[2][kBtdMax];
    __shared__ cx<float> rowbuf[2][kBtdMax];
    __shared__ cx<float> swapbuf[kBtdMax];
    __shared__ cx<float> upprev[kBtdMax];
    __shared__ cx<float> cdg[kBtdMax], clo[kBtdMax], cym[kBtdMax], cyp[kBtdMax];
    __shared__ int perm[kBtdMax];
    __shared__ int colmap[kBtdMax];
    __shared__ int bad;
    const cv_fdfd_desc d = load_desc(a.desc, 0);
    const int n = d.ny, nx = d.nx;
    const int t = threadIdx.x;
    // the wave index is uniform: keep it (and every row test) scalar
    const int w = __builtin_amdgcn_readfirstlane(t >> 6), l = t & 63;
    const int r0 = w * kFRows;
    cx<float> m[kFRows][kFCols];
#pragma unroll
    for (int p = 0; p < kFRows; ++p)
#pragma unroll
        for (int q = 0; q < kFCols; ++q) m[p][q] = czero<float>();
    if (t == 0) bad = 0;
    if (t < kBtdMax) {
        colbuf[0][t] = colbuf[1][t] = czero<float>();
        rowbuf[0][t] = rowbuf[1][t] = swapbuf[t] = czero<float>();
    }
    const bool xid = d.flags & CV_FLAG_X_IDENTITY;
    for (int i = 0; i < nx; ++i) {
        // S_i = D_i - diag(lo_i) S_{i-1}^{-1} diag(up_{i-1}); m holds S_{i-1}^{-1} (logical layout)
        if (t < n) {
            const Row5<float> rp = row_coefs<float>(d, a.prof, a.m0, a.m1, (i > 0) ? i - 1 : 0, t);
            upprev[t] = (i > 0 && !xid) ? rp.xp : czero<float>();
            const Row5<float> rw = row_coefs<float>(d, a.prof, a.m0, a.m1, i, t);
            cdg[t] = rw.c;
            clo[t] = (i > 0 && !xid) ? rw.xm : czero<float>();
            a.lo[(int64_t)i * n + t] = clo[t];
            a.up[(int64_t)i * n + t] = (i < nx - 1 && !xid) ? rw.xp : czero<float>();
            const bool yid = d.flags & CV_FLAG_Y_IDENTITY;
            cym[t] = yid ? czero<float>() : rw.ym;
            cyp[t] = yid ? czero<float>() : rw.yp;
        }
        __syncthreads();
#pragma unroll
        for (int p = 0; p < kFRows; ++p) {
            const int r = r0 + p;
            const cx<float> lo = (r < n) ? clo[r] : czero<float>();
#pragma unroll
            for (int q = 0; q < kFCols; ++q) {
                const int c = l + 64 * q;
                cx<float> v = czero<float>();
                if (r < n && c < n) {
                    v = czero<float>() - lo * m[p][q] * upprev[c];
                    if (c == r) v = v + cdg[r];
                    if (c == ((r == 0) ? n - 1 : r - 1)) v = v + cym[r];
                    if (c == ((r == n - 1) ? 0 : r + 1)) v = v + cyp[r];
                }
                m[p][q] = v;
            }
        }
        // Gauss-Jordan inversion in place, partial pivoting by physical row swaps.
        // Entries outside the n x n block are zero and stay zero (colbuf / rowbuf
        // padding is zero), so the update below needs no bounds tests.
        for (int k = 0; k < n; ++k) {
            const int b = k & 1;
            const int qk = k >> 6;
            if (l == (k & 63)) {
#pragma unroll
                for (int p = 0; p < kFRows; ++p) {
                    const cx<float> v0 = m[p][0], v1 = m[p][kFCols - 1];
                    colbuf[b][r0 + p] = (qk == 0) ? v0 : v1;
                }
            }
            __syncthreads();
            // Pivot: first row of largest |S[r][k]| among k <= r < n, found by
            // every wave on its own (no barrier): lane l looks at rows l and
            // l + 64, DPP combines within 16-lane rows, readlane across rows,
            // ballot picks the smallest such row.
            int pr;
            {
                float best = -1.0f;
                int arg = k;
                if (l >= k && l < n) { best = cabs2(colbuf[b][l]); arg = l; }
                if (l + 64 >= k && l + 64 < n) {
                    const float v2 = cabs2(colbuf[b][l + 64]);
                    if (v2 > best) { best = v2; arg = l + 64; }
                }
                float mx = best;
                mx = fmaxf(mx, dpp_f<0xB1>(mx));    // quad_perm [1,0,3,2]
                mx = fmaxf(mx, dpp_f<0x4E>(mx));    // quad_perm [2,3,0,1]
                mx = fmaxf(mx, dpp_f<0x141>(mx));   // row_half_mirror
                mx = fmaxf(mx, dpp_f<0x140>(mx));   // row_mirror
                const float g = fmaxf(fmaxf(lane_f(mx, 0), lane_f(mx, 16)), fmaxf(lane_f(mx, 32), lane_f(mx, 48)));
                const unsigned long long low = __ballot(best == g && arg < 64);
                const unsigned long long any = __ballot(best == g);
                if (!(g > 0.0f)) pr = k;
                else if (low) pr = __ffsll(low) - 1;
                else pr = __ffsll(any) - 1 + 64;
                pr = __builtin_amdgcn_readfirstlane(pr);
                if (t == 0) {
                    perm[k] = pr;
                    if (!(g > 0.0f) && bad == 0) bad = 1 + i;
                }
            }
            // row extraction by select chains (no indexed access to m)
            if ((pr >> 3) == w) {
                const int pp = pr - r0;
#pragma unroll
                for (int q = 0; q < kFCols; ++q) {
                    cx<float> v = m[0][q];
#pragma unroll
                    for (int p = 1; p < kFRows; ++p) v = (p == pp) ? m[p][q] : v;
                    rowbuf[b][l + 64 * q] = v;
                }
            }
            if (pr != k && (k >> 3) == w) {
                const int pk = k - r0;
#pragma unroll
                for (int q = 0; q < kFCols; ++q) {
                    cx<float> v = m[0][q];
#pragma unroll
                    for (int p = 1; p < kFRows; ++p) v = (p == pk) ? m[p][q] : v;
                    swapbuf[l + 64 * q] = v;
                }
            }
            __syncthreads();
            // rows k and pr exchange; then row k /= pivot and the other rows
            // lose their column-k component:
            //   S[k][c] = rj = (c == k ? 1 : S[k][c]) / piv
            //   S[r][c] = (c == k ? 0 : S[r][c]) - f_r rj     (f_r = S[r][k] before the step)
            const cx<float> ck = colbuf[b][pr], cp = colbuf[b][k];
            const float d2 = ck.re * ck.re + ck.im * ck.im;
            const float inv_d2 = (d2 > 0.0f) ? 1.0f / d2 : 0.0f;
            const cx<float> ipiv = mk<float>(ck.re * inv_d2, -ck.im * inv_d2);
            cx<float> rj[kFCols];
#pragma unroll
            for (int q = 0; q < kFCols; ++q) {
                const int c = l + 64 * q;
                rj[q] = ((c == k) ? mk<float>(1.0f, 0.0f) : rowbuf[b][c]) * ipiv;
            }
            const bool swapped = (pr != k);
            cx<float> sw[kFCols];
#pragma unroll
            for (int q = 0; q < kFCols; ++q) sw[q] = swapbuf[l + 64 * q];
#pragma unroll
            for (int p = 0; p < kFRows; ++p) {
                const int r = r0 + p;           // uniform across the wave
                const bool is_k = (r == k);
                const bool is_pr = swapped && (r == pr);
                const cx<float> f = is_pr ? cp : colbuf[b][r];
#pragma unroll
                for (int q = 0; q < kFCols; ++q) {
                    const int c = l + 64 * q;
                    const cx<float> cur = is_pr ? sw[q] : m[p][q];
                    const cx<float> keep = (c == k) ? czero<float>() : cur;
                    const cx<float> upd = keep - f * rj[q];
                    m[p][q] = is_k ? rj[q] : upd;
                }
            }
        }
        // undo the column interchanges: logical column j is physical column colmap[j]
        __syncthreads();
        if (t < n) {
            int v = t;
            for (int q = 0; q < n; ++q) {
                const int pq = perm[q];
                if (v == q) v = pq;
                else if (v == pq) v = q;
            }
            colmap[t] = v;
        }
#pragma unroll
        for (int p = 0; p < kFRows; ++p)
#pragma unroll
            for (int q = 0; q < kFCols; ++q) stage[(r0 + p) * kBtdMax + l + 64 * q] = m[p][q];
        __syncthreads();
        cx<float>* out = a.sinv + (int64_t)i * n * n;
#pragma unroll
        for (int p = 0; p < kFRows; ++p) {
            const int r = r0 + p;
#pragma unroll
            for (int q = 0; q < kFCols; ++q) {
                const int c = l + 64 * q;
                if (r < n && c < n) {
                    const cx<float> v = stage[r * kBtdMax + colmap[c]];
                    m[p][q] = v;
                    out[(int64_t)r * n + c] = v;
                }
            }
        }
        __syncthreads();
    }
    if (t == 0) *a.status = bad;
}

// z = M^-1 v (or M^-T v) with the stored inverses; v and z complex64, z may
// not alias v.  One workgroup; y_i is kept in z.  The 16 entries of S_i each
// thread needs are loaded one block ahead (registers), so the next block's
// loads are in flight while the current one is used.
//   A:   thread (row r = t / 8, part = t % 8) sums columns part, part + 8, ...
//        and 8 lanes combine by shuffles;
//   A^T: thread (column c = t % 128, part = t / 128) sums rows 16 part .. + 15,
//        coalesced along c, and the 8 parts combine through LDS.
template <bool TRANS>
__global__ void __launch_bounds__(kBtdThreads)
k_btd_apply(const cv_fdfd_desc* __restrict__ desc, const cx<float>* __restrict__ sinv,
            const cx<float>* __restrict__ lo, const cx<float>* __restrict__ up, const cx<float>* __restrict__ v,
            cx<float>* __restrict__ z) {
    constexpr int kParts = 8, kPer = kBtdMax / kParts;    // 16 entries per thread
    __shared__ cx<float> rhs[kBtdMax];
    __shared__ cx<float> part_sum[kParts][kBtdMax];
    const cv_fdfd_desc d = load_desc(desc, 0);
    const int n = d.ny, nx = d.nx;
    const bool xid = d.flags & CV_FLAG_X_IDENTITY;
    const int t = threadIdx.x;
    const int a_row = t / kParts, a_part = t % kParts;      // A layout
    const int t_col = t % kBtdMax, t_part = t / kBtdMax;    // A^T layout
    auto load_block = [&](int i, cx<float> (&reg)[kPer]) {
        const cx<float>* S = sinv + (int64_t)i * n * n;
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            int r, c;
            if (TRANS) { r = t_part * kPer + j; c = t_col; }
            else { r = a_row; c = a_part + kParts * j; }
            reg[j] = (r < n && c < n) ? S[(int64_t)r * n + c] : czero<float>();
        }
    };
    // out (+/-)= S rhs or S^T rhs with the block in reg
    auto matvec = [&](const cx<float> (&reg)[kPer], cx<float>* out, bool subtract) {
        __syncthreads();   // rhs ready
        cx<float> acc = czero<float>();
        if (TRANS) {
#pragma unroll
            for (int j = 0; j < kPer; ++j) acc = acc + reg[j] * rhs[t_part * kPer + j];
            part_sum[t_part][t_col] = acc;
            __syncthreads();
            if (t < n) {
                cx<float> sum = czero<float>();
#pragma unroll
                for (int g = 0; g < kParts; ++g) sum = sum + part_sum[g][t];
                out[t] = subtract ? out[t] - sum : sum;
            }
        } else {
#pragma unroll
            for (int j = 0; j < kPer; ++j) acc = acc + reg[j] * rhs[(a_part + kParts * j) & (kBtdMax - 1)];
#pragma unroll
            for (int off = 1; off < kParts; off <<= 1) {
                acc.re += __shfl_xor(acc.re, off, 64);
                acc.im += __shfl_xor(acc.im, off, 64);
            }
            if (a_part == 0 && a_row < n) out[a_row] = subtract ? out[a_row] - acc : acc;
        }
    };
    if (t < kBtdMax) rhs[t] = czero<float>();
    cx<float> cur[kPer], nxt[kPer];
    // forward sweep: y_i = S'_i^-1 (v_i - lo'_i .* y_{i-1})
    load_block(0, cur);
    for (int i = 0; i < nx; ++i) {
        if (i + 1 < nx) load_block(i + 1, nxt);
        __syncthreads();   // every thread is done with the previous rhs
        if (t < n) {
            cx<float> r = v[(int64_t)i * n + t];
            if (i > 0 && !xid) {
                // lo' = lo_i (A) or up_{i-1} (A^T), at row t
                const cx<float> c = TRANS ? up[(int64_t)(i - 1) * n + t] : lo[(int64_t)i * n + t];
                r = r - c * z[(int64_t)(i - 1) * n + t];
            }
            rhs[t] = r;
        }
        matvec(cur, z + (int64_t)i * n, false);
#pragma unroll
        for (int j = 0; j < kPer; ++j) cur[j] = nxt[j];
    }
    // backward sweep: z_i = y_i - S'_i^-1 (up'_i .* z_{i+1})
    if (nx >= 2 && !xid) {
        load_block(nx - 2, cur);
        for (int i = nx - 2; i >= 0; --i) {
            if (i >= 1) load_block(i - 1, nxt);
            __syncthreads();
            if (t < n) {
                // up' = up_i (A) or lo_{i+1} (A^T), at row t
                const cx<float> c = TRANS ? lo[(int64_t)(i + 1) * n + t] : up[(int64_t)i * n + t];
                rhs[t] = c * z[(int64_t)(i + 1) * n + t];
            }
            matvec(cur, z + (int64_t)i * n, true);
#pragma unroll
            for (int j = 0; j < kPer; ++j) cur[j] = nxt[j];
        }
    }
}

}  // namespace cv

using namespace cv;

// factor buffer: S_i^-1 blocks, lo, up (complex64), status word
static int64_t btd_inv_bytes(int nx, int ny) { return (int64_t)nx * ny * ny * 8; }
static int64_t btd_coef_bytes(int nx, int ny) { return (int64_t)nx * ny * 8; }

extern "C" int64_t cv_fdfd_btd_factor_bytes(int nx, int ny) {
    if (nx <= 0 || ny <= 0 || ny > kBtdMax) return 0;
    return btd_inv_bytes(nx, ny) + 2 * btd_coef_bytes(nx, ny) + 256;
}

extern "C" int cv_fdfd_btd_factor(const cv_fdfd_desc* desc, const cv_fdfd_desc* desc_host, const void* prof32,
                                  const void* m0_32, const void* m1_32, void* factor, int64_t factor_bytes,
                                  void* stream) {
    CV_CHECK(desc && desc_host && prof32 && m0_32 && factor, "null pointer argument");
    const cv_fdfd_desc dh = desc_host[0];
    CV_CHECK(dh.nx > 0 && dh.ny > 0 && dh.cell_off == 0, "one instance starting at cell 0 is required");
    CV_CHECK(dh.ny <= kBtdMax, "the block-tridiagonal factorisation needs ny <= %d (ny = %d)", kBtdMax, dh.ny);
    CV_CHECK(dh.nblk == 1 && dh.n_cpl == 0, "frequency-coupled systems are not supported");
    CV_CHECK(dh.pol == CV_POL_EZ || m1_32, "Hz operator needs m1");
    const int64_t need = cv_fdfd_btd_factor_bytes(dh.nx, dh.ny);
    CV_CHECK(factor_bytes >= need, "factor buffer too small: %lld < %lld", (long long)factor_bytes, (long long)need);
    BtdFactorArgs a;
    a.desc = desc;
    a.prof = (const cx<float>*)prof32;
    a.m0 = (const float*)m0_32;
    a.m1 = (const float*)m1_32;
    char* base = (char*)factor;
    a.sinv = (cx<float>*)base;
    a.lo = (cx<float>*)(base + btd_inv_bytes(dh.nx, dh.ny));
    a.up = (cx<float>*)(base + btd_inv_bytes(dh.nx, dh.ny) + btd_coef_bytes(dh.nx, dh.ny));
    a.status = (int*)(base + btd_inv_bytes(dh.nx, dh.ny) + 2 * btd_coef_bytes(dh.nx, dh.ny));
    hipStream_t s = as_stream(stream);
    k_btd_factor<<<1, kBtdThreads, 0, s>>>(a);
    CV_LAUNCH_CHECK();
    int st = 0;
    CV_HIP(hipMemcpyAsync(&st, a.status, sizeof(int), hipMemcpyDeviceToHost, s));
    CV_HIP(hipStreamSynchronize(s));
    if (st != 0) {
        set_error("zero pivot in block %d of the block-tridiagonal factorisation", st - 1);
        return CV_ERR_SINGULAR;
    }
    return CV_OK;
}

namespace cv {
int btd_apply(const cv_fdfd_desc* desc, const cv_fdfd_desc& dh, const void* prof32, const void* m0_32,
              const void* m1_32, const void* factor, const void* v, void* z, int transpose, hipStream_t s) {
    const char* base = (const char*)factor;
    const cx<float>* sinv = (const cx<float>*)base;
    const cx<float>* lo = (const cx<float>*)(base + btd_inv_bytes(dh.nx, dh.ny));
    const cx<float>* up = (const cx<float>*)(base + btd_inv_bytes(dh.nx, dh.ny) + btd_coef_bytes(dh.nx, dh.ny));
    (void)prof32;
    (void)m0_32;
    (void)m1_32;
    if (transpose)
        k_btd_apply<true><<<1, kBtdThreads, 0, s>>>(desc, sinv, lo, up, (const cx<float>*)v, (cx<float>*)z);
    else
        k_btd_apply<false><<<1, kBtdThreads, 0, s>>>(desc, sinv, lo, up, (const cx<float>*)v, (cx<float>*)z);
    CV_LAUNCH_CHECK();
    return CV_OK;
}
}  // namespace cv

extern "C" int cv_fdfd_btd_apply(const cv_fdfd_desc* desc, const cv_fdfd_desc* desc_host, const void* prof32,
                                 const void* m0_32, const void* m1_32, const void* factor, const void* v, void* z,
                                 int transpose, void* stream) {
    CV_CHECK(desc && desc_host && prof32 && m0_32 && factor && v && z, "null pointer argument");
    CV_CHECK(v != z, "v and z must not alias");
    const cv_fdfd_desc dh = desc_host[0];
    CV_CHECK(dh.nx > 0 && dh.ny > 0 && dh.ny <= kBtdMax && dh.cell_off == 0 && dh.nblk == 1,
             "single instance with ny <= %d required", kBtdMax);
    return btd_apply(desc, dh, prof32, m0_32, m1_32, factor, v, z, transpose, as_stream(stream));
}
