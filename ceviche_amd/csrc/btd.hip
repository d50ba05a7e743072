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

#include <algorithm>
#include <vector>

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

// LDS of one Gauss-Jordan inversion (one per workgroup).
struct GjLds {
    cx<float> colbuf[2][kBtdMax];
    cx<float> rowbuf[2][kBtdMax];
    cx<float> swapbuf[kBtdMax];
    int perm[kBtdMax];
    int colmap[kBtdMax];
    int invmap[kBtdMax];
    int bad;
};

// Inverts the n x n matrix held in m (wave w: rows 8w .. 8w+7, lane l: columns
// l and l+64, zero outside n x n) in place: Gauss-Jordan elimination with
// partial pivoting by physical row swaps, then the column interchanges undone
// through stage (UNPERMUTE), or left to the caller: physical column c of m is
// logical column g.invmap[c].  On return m holds the inverse (zero padding),
// logical or physical column order.  A zero pivot
// sets g.bad = code (if g.bad is still 0).  Called by all threads.
template <bool UNPERMUTE>
__device__ void gj_invert(cx<float> (&m)[kFRows][kFCols], int n, int r0, int w, int l, int t, GjLds& g, int code,
                          cx<float>* stage) {
    cx<float> (&colbuf)[2][kBtdMax] = g.colbuf;
    cx<float> (&rowbuf)[2][kBtdMax] = g.rowbuf;
    cx<float>* const swapbuf = g.swapbuf;
    int* const perm = g.perm;
    int* const colmap = g.colmap;
    int& bad = g.bad;
    const int i = code - 1;
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
            const float gmax = fmaxf(fmaxf(lane_f(mx, 0), lane_f(mx, 16)), fmaxf(lane_f(mx, 32), lane_f(mx, 48)));
            const unsigned long long low = __ballot(best == gmax && arg < 64);
            const unsigned long long any = __ballot(best == gmax);
            if (!(gmax > 0.0f)) pr = k;
            else if (low) pr = __ffsll(low) - 1;
            else pr = __ffsll(any) - 1 + 64;
            pr = __builtin_amdgcn_readfirstlane(pr);
            if (t == 0) {
                perm[k] = pr;
                if (!(gmax > 0.0f) && bad == 0) bad = 1 + i;
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
        // Row k becomes the pivot row scaled by 1/piv (with 1/piv in column
        // k); every other row r loses f_r times it, f_r = S[r][k] after the
        // swap.  Since S[r][k] = f_r, subtracting f_r (rj + e_k) leaves
        // -f_r / piv in column k, so one update form serves every entry:
        //   rj[c] = (c == k ? 1 : S[pr][c]) / piv,  rjx = rj + e_k,
        //   S[r][c] -= f_r rjx[c]  (r != k),  S[k][c] = rj[c].
        // Row pr first takes the old row k (its owner wave only).
        const cx<float> ck = colbuf[b][pr], cp = colbuf[b][k];
        const float d2 = ck.re * ck.re + ck.im * ck.im;
        const float inv_d2 = (d2 > 0.0f) ? 1.0f / d2 : 0.0f;
        const cx<float> ipiv = mk<float>(ck.re * inv_d2, -ck.im * inv_d2);
        cx<float> rj[kFCols], rjx[kFCols];
#pragma unroll
        for (int q = 0; q < kFCols; ++q) {
            const int c = l + 64 * q;
            const bool isk = (c == k);
            rj[q] = (isk ? mk<float>(1.0f, 0.0f) : rowbuf[b][c]) * ipiv;
            rjx[q] = mk<float>(rj[q].re + (isk ? 1.0f : 0.0f), rj[q].im);
        }
        const bool own_pr = (pr != k) && ((pr >> 3) == w);
        if (own_pr) {
            const int pp = pr - r0;
#pragma unroll
            for (int q = 0; q < kFCols; ++q) {
                const cx<float> sw = swapbuf[l + 64 * q];
#pragma unroll
                for (int p = 0; p < kFRows; ++p) m[p][q] = (p == pp) ? sw : m[p][q];
            }
        }
        cx<float> f[kFRows];
#pragma unroll
        for (int p = 0; p < kFRows; ++p) f[p] = colbuf[b][r0 + p];
        if (own_pr) {
            const int pp = pr - r0;
#pragma unroll
            for (int p = 0; p < kFRows; ++p) f[p] = (p == pp) ? cp : f[p];
        }
#pragma unroll
        for (int p = 0; p < kFRows; ++p)
#pragma unroll
            for (int q = 0; q < kFCols; ++q) {   // four FMAs per entry
                m[p][q].re = fmaf(-f[p].re, rjx[q].re, fmaf(f[p].im, rjx[q].im, m[p][q].re));
                m[p][q].im = fmaf(-f[p].re, rjx[q].im, fmaf(-f[p].im, rjx[q].re, m[p][q].im));
            }
        if ((k >> 3) == w) {
            const int pk = k - r0;
#pragma unroll
            for (int q = 0; q < kFCols; ++q)
#pragma unroll
                for (int p = 0; p < kFRows; ++p) m[p][q] = (p == pk) ? rj[q] : m[p][q];
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
    if (!UNPERMUTE) {
        __syncthreads();
        if (t < n) g.invmap[colmap[t]] = t;
        __syncthreads();
        return;
    }
#pragma unroll
    for (int p = 0; p < kFRows; ++p)
#pragma unroll
        for (int q = 0; q < kFCols; ++q) stage[(r0 + p) * kBtdMax + l + 64 * q] = m[p][q];
    __syncthreads();
#pragma unroll
    for (int p = 0; p < kFRows; ++p) {
        const int r = r0 + p;
#pragma unroll
        for (int q = 0; q < kFCols; ++q) {
            const int c = l + 64 * q;
            m[p][q] = (r < n && c < n) ? stage[r * kBtdMax + colmap[c]] : czero<float>();
        }
    }
    __syncthreads();
}

__global__ void __launch_bounds__(kBtdThreads) k_btd_factor(BtdFactorArgs a) {
    __shared__ GjLds g;
    __shared__ cx<float> stage[kBtdMax * kBtdMax];
    __shared__ cx<float> upprev[kBtdMax];
    __shared__ cx<float> cdg[kBtdMax], clo[kBtdMax], cym[kBtdMax], cyp[kBtdMax];
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
    if (t == 0) g.bad = 0;
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
        gj_invert<true>(m, n, r0, w, l, t, g, 1 + i, stage);
        cx<float>* out = a.sinv + (int64_t)i * n * n;
#pragma unroll
        for (int p = 0; p < kFRows; ++p) {
            const int r = r0 + p;
#pragma unroll
            for (int q = 0; q < kFCols; ++q) {
                const int c = l + 64 * q;
                if (r < n && c < n) out[(int64_t)r * n + c] = m[p][q];
            }
        }
    }
    __syncthreads();
    if (t == 0) *a.status = g.bad;
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


// --------------------------------------------------------------------------
// Block cyclic reduction (BCR) of the same block-tridiagonal matrix: level l
// has N blocks; the odd ones are inverted (all at once, one workgroup each)
// and eliminated, the even ones form level l+1:
//   alpha_j = L_j D_{j-1}^-1,  beta_j = U_j D_{j+1}^-1,
//   gamma_j = D_{j-1}^-1 U_{j-1},  delta_j = D_{j+1}^-1 L_{j+1},
//   D'_j = D_j - L_j gamma_j - U_j delta_j,  L'_j = -alpha_j L_{j-1},  U'_j = -beta_j U_{j+1}.
// A solve reduces the right-hand side level by level (b'_j = b_j - alpha_j
// b_{j-1} - beta_j b_{j+1}; A^T: gamma_j^T, delta_j^T), inverts the last block
// and substitutes back (x_j = D_j^-1 (b_j - L_j x_{j-1} - U_j x_{j+1}); A^T:
// D_j^-T (b_j - U_{j-1}^T x_{j-1} - L_{j+1}^T x_{j+1})).  The sequential
// factorisation's nx dependent inversions become ceil(log2 nx) levels of
// independent ones; every matrix is an n x n complex64 "slot".

// Level 0's L_i and U_i are diagonal (stored dense); ops mark them so the
// products and solves touch only their diagonals.
struct BcrGemmOp {     // C = X + s1 A1 B1 + s2 A2 B2 (slot -1: absent)
    int c, x, a1, b1, a2, b2;
    float s1, s2;
    int diag;          // diagonal factors: 1 A1, 2 B1, 4 A2, 8 B2
};

struct BcrVecOp {      // t = b - M1 u1 - M2 u2 (A^T: M1t^T, M2t^T); out = M0 t (A^T: M0^T t) or t
    int out, b, u1, u2, m1, m2, m1t, m2t, m0;
    int diag;          // diagonal matrices: 1 M1, 2 M2, 4 M1t, 8 M2t
};

// Several factorisations of the same shape run as one (blockIdx.y = which);
// their op tables are identical, only the slot arrays differ.
constexpr int kBcrBatch = 8;

struct BcrSlots {
    cx<float>* slots[kBcrBatch];
    int* status[kBcrBatch];
};

struct BcrBuild {
    const cv_fdfd_desc* desc[kBcrBatch];
    const cx<float>* prof[kBcrBatch];
    const float* m0[kBcrBatch];
    const float* m1[kBcrBatch];
};

// D_i, L_i = diag(lo_i), U_i = diag(up_i) of level 0 (x-wrap dropped);
// dls[3 i + 0 / 1 / 2] = slots of D_i, L_i, U_i (-1: none)
__global__ void __launch_bounds__(256)
k_bcr_build0(BcrBuild in, BcrSlots out, const int* __restrict__ dls) {
    const int fb = blockIdx.y;
    const cv_fdfd_desc* __restrict__ desc = in.desc[fb];
    const cx<float>* __restrict__ prof = in.prof[fb];
    const float* __restrict__ m0 = in.m0[fb];
    const float* __restrict__ m1 = in.m1[fb];
    cx<float>* __restrict__ slots = out.slots[fb];
    __shared__ cx<float> cdg[kBtdMax], clo[kBtdMax], cup[kBtdMax], cym[kBtdMax], cyp[kBtdMax];
    const cv_fdfd_desc d = load_desc(desc, 0);
    const int n = d.ny, nx = d.nx, i = blockIdx.x;
    const bool xid = d.flags & CV_FLAG_X_IDENTITY, yid = d.flags & CV_FLAG_Y_IDENTITY;
    for (int r = threadIdx.x; r < n; r += blockDim.x) {
        const Row5<float> rw = row_coefs<float>(d, prof, m0, m1, i, r);
        cdg[r] = rw.c;
        clo[r] = (i > 0 && !xid) ? rw.xm : czero<float>();
        cup[r] = (i < nx - 1 && !xid) ? rw.xp : czero<float>();
        cym[r] = yid ? czero<float>() : rw.ym;
        cyp[r] = yid ? czero<float>() : rw.yp;
    }
    __syncthreads();
    const int64_t nn = (int64_t)n * n;
    cx<float>* D = slots + (int64_t)dls[3 * i] * nn;
    cx<float>* L = dls[3 * i + 1] >= 0 ? slots + (int64_t)dls[3 * i + 1] * nn : nullptr;
    cx<float>* U = dls[3 * i + 2] >= 0 ? slots + (int64_t)dls[3 * i + 2] * nn : nullptr;
    for (int64_t e = threadIdx.x; e < nn; e += blockDim.x) {
        const int r = (int)(e / n), c = (int)(e - (int64_t)r * n);
        cx<float> v = czero<float>();
        if (c == r) v = v + cdg[r];
        if (c == ((r == 0) ? n - 1 : r - 1)) v = v + cym[r];
        if (c == ((r == n - 1) ? 0 : r + 1)) v = v + cyp[r];
        D[e] = v;
        if (L) L[e] = (c == r) ? clo[r] : czero<float>();
        if (U) U[e] = (c == r) ? cup[r] : czero<float>();
    }
}

__global__ void k_zero_u32_btd(unsigned int* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// In-place inversion of the listed slots, one workgroup each of 64 ceil(n / 8)
// threads; a zero pivot records 1 + slot in *status.
//
// Gauss-Jordan with partial pivoting and neither row swaps nor row scaling
// during the elimination: wave w owns columns 8w .. 8w+7, lane l rows l and
// l + 64.  Step k (column k, held by one wave in one register slot):
//   the owner wave picks the largest |S[r][k]| among rows not yet used as
//   pivots (pr, 1/piv = ipiv), publishes f_r = S[r][k] ipiv (0 for r = pr)
//   and writes column k itself: -f_r, and 1 in row pr (whose scale ipiv is
//   recorded instead of applied);
//   after one barrier every wave reads the stored pivot row u_c from lane
//   pr & 63 (readlane, scalar) and updates S[r][c] -= f_r u_c (u_k = 0).
// Row pr, scaled by ipiv, is the pivot row of textbook Gauss-Jordan; since
// every later update of a row is linear in it, the scale can wait: the
// inverse is A^-1[i][j] = scale[piv[i]] S[piv[i]][inv[j]] with piv[k] = pr of
// step k and inv its inverse permutation (logical row i lives in physical
// row piv[i], logical column j in physical column inv[j]).
struct GjcLds {
    cx<float> fbuf[2][kBtdMax];   // f_r of the step, double-buffered
    int prbuf[2];
    cx<float> scale[kBtdMax];     // per physical row: 1 / its pivot
    int piv[kBtdMax];             // step -> pivot row
    int inv[kBtdMax];             // pivot row -> step
    int bad;
};

__global__ void __launch_bounds__(kBtdThreads)
k_bcr_inv(BcrSlots B, const int* __restrict__ list, int n) {
    __shared__ GjcLds g;
    cx<float>* __restrict__ slots = B.slots[blockIdx.y];
    int* __restrict__ status = B.status[blockIdx.y];
    const int t = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6), l = t & 63;
    const int nw = (n + 7) / 8;
    const int slot = list[blockIdx.x];
    const int code = 1 + slot;
    cx<float>* S = slots + (int64_t)slot * n * n;
    if (t == 0) g.bad = 0;
    cx<float> m[2][8];           // rows l, l + 64; columns 8w .. 8w+7
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const int r = l + 64 * q, cc = 8 * w + c;
            m[q][c] = (r < n && cc < n) ? S[(int64_t)r * n + cc] : czero<float>();
        }
    bool used0 = false, used1 = false;
    __syncthreads();
    for (int kb = 0; kb < nw; ++kb) {
#pragma unroll
        for (int cc = 0; cc < 8; ++cc) {
            const int k = 8 * kb + cc;
            if (k >= n) break;     // uniform
            const int b = k & 1;
            if (w == kb) {
                const cx<float> c0 = m[0][cc], c1 = m[1][cc];
                const int ra = l, rb = l + 64;
                const float mag0 = (ra < n && !used0) ? cabs2(c0) : -1.0f;
                const float mag1 = (rb < n && !used1) ? cabs2(c1) : -1.0f;
                const float best = fmaxf(mag0, mag1);
                const int arg = (mag1 > mag0) ? rb : ra;
                float mx = best;
                mx = fmaxf(mx, dpp_f<0xB1>(mx));    // quad_perm [1,0,3,2]
                mx = fmaxf(mx, dpp_f<0x4E>(mx));    // quad_perm [2,3,0,1]
                mx = fmaxf(mx, dpp_f<0x141>(mx));   // row_half_mirror
                mx = fmaxf(mx, dpp_f<0x140>(mx));   // row_mirror
                const float gmax = fmaxf(fmaxf(lane_f(mx, 0), lane_f(mx, 16)), fmaxf(lane_f(mx, 32), lane_f(mx, 48)));
                // lowest row attaining the maximum: rows < 64 first
                const unsigned long long low = __ballot(best == gmax && arg < 64);
                const unsigned long long any = __ballot(best == gmax);
                int pr = low ? (__ffsll(low) - 1) : (__ffsll(any) - 1 + 64);
                pr = __builtin_amdgcn_readfirstlane(pr);
                const int pl = pr & 63;
                const bool hi = pr >= 64;
                const cx<float> pv = hi ? c1 : c0;
                const cx<float> piv = mk<float>(lane_f(pv.re, pl), lane_f(pv.im, pl));
                const float d2 = piv.re * piv.re + piv.im * piv.im;
                const float inv_d2 = (d2 > 0.0f) ? 1.0f / d2 : 0.0f;
                const cx<float> ipiv = mk<float>(piv.re * inv_d2, -piv.im * inv_d2);
                const cx<float> f0 = (ra == pr) ? czero<float>() : c0 * ipiv;
                const cx<float> f1 = (rb == pr) ? czero<float>() : c1 * ipiv;
                g.fbuf[b][ra] = f0;
                g.fbuf[b][rb] = f1;
                m[0][cc] = (ra == pr) ? mk<float>(1.0f, 0.0f) : mk<float>(-f0.re, -f0.im);
                m[1][cc] = (rb == pr) ? mk<float>(1.0f, 0.0f) : mk<float>(-f1.re, -f1.im);
                if (l == 0) {
                    g.prbuf[b] = pr;
                    g.scale[pr] = ipiv;
                    g.piv[k] = pr;
                    if (!(d2 > 0.0f) && g.bad == 0) g.bad = code;
                }
            }
            __syncthreads();
            const int pr = __builtin_amdgcn_readfirstlane(g.prbuf[b]);
            const int pl = pr & 63;
            const bool hi = pr >= 64;
            if (l == pl) {
                if (hi) used1 = true;
                else used0 = true;
            }
            const cx<float> f0 = g.fbuf[b][l], f1 = g.fbuf[b][l + 64];
            // the stored pivot row in this wave's columns (column k itself excluded)
            float ur[8], ui[8];
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const cx<float> v = hi ? m[1][c] : m[0][c];
                ur[c] = lane_f(v.re, pl);
                ui[c] = lane_f(v.im, pl);
            }
            if (w == kb) {
                ur[cc] = 0.0f;
                ui[cc] = 0.0f;
            }
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                m[0][c].re = fmaf(-f0.re, ur[c], fmaf(f0.im, ui[c], m[0][c].re));
                m[0][c].im = fmaf(-f0.re, ui[c], fmaf(-f0.im, ur[c], m[0][c].im));
                m[1][c].re = fmaf(-f1.re, ur[c], fmaf(f1.im, ui[c], m[1][c].re));
                m[1][c].im = fmaf(-f1.re, ui[c], fmaf(-f1.im, ur[c], m[1][c].im));
            }
        }
    }
    __syncthreads();
    if (t < n) g.inv[g.piv[t]] = t;
    __syncthreads();
    // physical (r, c) -> logical (inv[r], piv[c]), times the row's scale
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int r = l + 64 * q;
        if (r >= n) continue;
        const int li = g.inv[r];
        const cx<float> sc = g.scale[r];
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const int cc = 8 * w + c;
            if (cc < n) S[(int64_t)li * n + g.piv[cc]] = sc * m[q][c];
        }
    }
    if (t == 0 && g.bad != 0) atomicCAS(status, 0, g.bad);
}

// Batched n x n complex GEMMs (n <= 128).  256 threads (16 x 16) own a
// (16 TM) x (16 TM) output tile, TM x TM outputs per thread: TM = 4 with
// four 64 x 64 tiles for n > 96, one tile of 80 or 96 (TM = 5, 6) for
// 64 < n <= 96, so few padded entries are computed.  K in chunks of 16
// staged through LDS, the next chunk loaded into registers while the
// current one is multiplied.  Ops whose every term has a diagonal factor
// (level 0) are row / column scalings over all ops' workgroups instead.
constexpr int kGemmThreads = 256;
constexpr int kGemmK = 16;

static inline int gemm_tm(int n) { return n > 96 ? 4 : (n > 80 ? 6 : (n > 64 ? 5 : 4)); }
static inline int gemm_tiles(int n) { const int tile = 16 * gemm_tm(n); return (n + tile - 1) / tile; }

template <int TM>
__global__ void __launch_bounds__(kGemmThreads)
k_bcr_gemm(BcrSlots B, const BcrGemmOp* __restrict__ ops, int n, int tiles, int per_op) {
    constexpr int kTile = 16 * TM;
    cx<float>* __restrict__ slots = B.slots[blockIdx.y];
    __shared__ cx<float> As[kGemmK][kTile + 1];   // As[k][row], padded against bank conflicts
    __shared__ cx<float> Bs[kGemmK][kTile];       // Bs[k][col]
    // per_op workgroups per op: tiles^2 output tiles, or slices of a scaling
    const int opi = blockIdx.x / per_op, tq = blockIdx.x - opi * per_op;
    const BcrGemmOp op = ops[opi];
    const int nn = n * n;
    const int t = threadIdx.x, ty = t / 16, tx = t % 16;
    const bool has1 = op.a1 >= 0 && op.b1 >= 0, has2 = op.a2 >= 0 && op.b2 >= 0;
    if ((!has1 || (op.diag & 3)) && (!has2 || (op.diag & 12))) {
        // row / column scaling: the op's workgroups take consecutive slices
        // of the row-major entries
        const int q0 = (int)(((int64_t)nn * tq) / per_op), q1 = (int)(((int64_t)nn * (tq + 1)) / per_op);
        cx<float>* C = slots + (int64_t)op.c * nn;
        const cx<float>* X = op.x >= 0 ? slots + (int64_t)op.x * nn : nullptr;
        for (int e = q0 + t; e < q1; e += kGemmThreads) {
            const int r = e / n, c = e - r * n;
            cx<float> v = X ? X[e] : czero<float>();
#pragma unroll
            for (int term = 0; term < 2; ++term) {
                if (!(term ? has2 : has1)) continue;
                const cx<float>* A = slots + (int64_t)(term ? op.a2 : op.a1) * nn;
                const cx<float>* Bm = slots + (int64_t)(term ? op.b2 : op.b1) * nn;
                const bool adiag = op.diag & (term ? 4 : 1);
                const cx<float> pr = adiag ? A[r * n + r] * Bm[e] : A[e] * Bm[c * n + c];
                const float sg = term ? op.s2 : op.s1;
                v = v + mk<float>(sg * pr.re, sg * pr.im);
            }
            C[e] = v;
        }
        return;
    }
    if (tq >= tiles * tiles) return;   // extra workgroups of a launch sized for scalings
    const int row0 = (tq / tiles) * kTile, col0 = (tq - (tq / tiles) * tiles) * kTile;
    cx<float> acc[TM][TM];
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TM; ++b) acc[a][b] = czero<float>();
    // chunks (term, k0) in sequence; A is read along k (16 lanes per row,
    // contiguous), B along the tile's columns
    const int nk = (n + kGemmK - 1) / kGemmK;
    const int first = has1 ? 0 : 1;
    const int nchunks = ((has1 ? 1 : 0) + (has2 ? 1 : 0)) * nk;
    cx<float> ra[TM], rb[TM];    // kGemmK * kTile entries of each, over 256 threads
    auto load = [&](int ch) {
        const int term = (ch < nk) ? first : 1, k0 = (ch % nk) * kGemmK;
        const cx<float>* A = slots + (int64_t)(term ? op.a2 : op.a1) * nn;
        const cx<float>* Bm = slots + (int64_t)(term ? op.b2 : op.b1) * nn;
        const float sg = term ? op.s2 : op.s1;
#pragma unroll
        for (int u = 0; u < TM; ++u) {
            const int e = t + kGemmThreads * u;
            const int ar = row0 + e / kGemmK, ak = k0 + e % kGemmK;
            const cx<float> av = (ar < n && ak < n) ? A[ar * n + ak] : czero<float>();
            ra[u] = mk<float>(sg * av.re, sg * av.im);
            const int bk = k0 + e / kTile, bc = col0 + e % kTile;
            rb[u] = (bc < n && bk < n) ? Bm[bk * n + bc] : czero<float>();
        }
    };
    if (nchunks > 0) load(0);
    for (int ch = 0; ch < nchunks; ++ch) {
        __syncthreads();
#pragma unroll
        for (int u = 0; u < TM; ++u) {
            const int e = t + kGemmThreads * u;
            As[e % kGemmK][e / kGemmK] = ra[u];
            Bs[e / kTile][e % kTile] = rb[u];
        }
        __syncthreads();
        if (ch + 1 < nchunks) load(ch + 1);
#pragma unroll
        for (int kk = 0; kk < kGemmK; ++kk) {
            cx<float> av[TM], bv[TM];
#pragma unroll
            for (int a = 0; a < TM; ++a) av[a] = As[kk][TM * ty + a];
#pragma unroll
            for (int b = 0; b < TM; ++b) bv[b] = Bs[kk][TM * tx + b];
#pragma unroll
            for (int a = 0; a < TM; ++a)
#pragma unroll
                for (int b = 0; b < TM; ++b) {
                    acc[a][b].re = fmaf(av[a].re, bv[b].re, fmaf(-av[a].im, bv[b].im, acc[a][b].re));
                    acc[a][b].im = fmaf(av[a].re, bv[b].im, fmaf(av[a].im, bv[b].re, acc[a][b].im));
                }
        }
    }
    cx<float>* C = slots + (int64_t)op.c * nn;
    const cx<float>* X = op.x >= 0 ? slots + (int64_t)op.x * nn : nullptr;
#pragma unroll
    for (int a = 0; a < TM; ++a) {
        const int r = row0 + TM * ty + a;
        if (r >= n) continue;
#pragma unroll
        for (int b = 0; b < TM; ++b) {
            const int c = col0 + TM * tx + b;
            if (c >= n) continue;
            cx<float> v = acc[a][b];
            if (X) v = v + X[r * n + c];
            C[r * n + c] = v;
        }
    }
}

// scaling: every op of the launch is a level-0 row / column scaling (eight
// workgroups per op); otherwise every op is a dense product (one per tile)
static void launch_gemm(const BcrSlots& S, const BcrGemmOp* ops, int count, int cnt, int n, bool scaling,
                        hipStream_t s) {
    const int tiles = gemm_tiles(n);
    const int per_op = scaling ? 8 : tiles * tiles;
    const dim3 grid(per_op * count, cnt);
    switch (gemm_tm(n)) {
        case 5: k_bcr_gemm<5><<<grid, kGemmThreads, 0, s>>>(S, ops, n, tiles, per_op); break;
        case 6: k_bcr_gemm<6><<<grid, kGemmThreads, 0, s>>>(S, ops, n, tiles, per_op); break;
        default: k_bcr_gemm<4><<<grid, kGemmThreads, 0, s>>>(S, ops, n, tiles, per_op); break;
    }
}

// One BCR solve op per workgroup of 8 threads per block row (rounded up to
// whole waves), several factorisations of the same shape at once
// (blockIdx.y).  Every matrix entry a thread needs is loaded at the start,
// so each op waits for memory once instead of once per row loop:
//   A:   thread (row r = t / 8, part = t % 8) holds columns part, part + 8, ...
//        and 8 lanes combine by shuffles;
//   A^T: thread (column c = t % n, part = t / n) holds rows part, part + 8,
//        ..., coalesced along c, and the 8 parts combine through LDS.
// Vector index i < nx is the right-hand side v, nx <= i < 2 nx the solution
// z (level 0, read from / written to the caller's buffers directly); the
// others live in the factor's work vectors.
constexpr int kVecPer = kBtdMax / 8;    // entries per thread and matrix

struct BcrVecBatch {
    const cx<float>* slots[kBcrBatch];
    cx<float>* work[kBcrBatch];
    const cx<float>* v[kBcrBatch];
    cx<float>* z[kBcrBatch];
};

// Thread roles, 64 ceil(8 n / 64) threads: A: (row a = t / 8, part p = t % 8);
// A^T: (column a = t % n, part p = t / n), threads with p >= 8 idle.
template <bool T>
struct VecRole {
    int a, p;
    bool on;
    __device__ __forceinline__ VecRole(int n) {
        const int t = threadIdx.x;
        a = T ? t % n : t / 8;
        p = T ? t / n : t % 8;
        on = (p < 8) && (a < n);
    }
};

template <bool T>
__device__ __forceinline__ void bcr_load(const cx<float>* __restrict__ M, int n, bool diag, const VecRole<T>& v,
                                         cx<float> (&e)[kVecPer]) {
#pragma unroll
    for (int i = 0; i < kVecPer; ++i) {
        const int b = v.p + 8 * i;
        const int r = T ? b : v.a, c = T ? v.a : b;
        e[i] = (v.on && b < n && (!diag || r == c)) ? M[(int64_t)r * n + c] : czero<float>();
    }
}

// sum over this thread's held entries of M x (A: along its row; A^T: down its column)
__device__ __forceinline__ cx<float> bcr_dot(const cx<float> (&e)[kVecPer], const cx<float>* x, int p) {
    cx<float> acc = czero<float>();
#pragma unroll
    for (int i = 0; i < kVecPer; ++i) {
        const cx<float> xv = x[p + 8 * i];
        acc.re = fmaf(e[i].re, xv.re, fmaf(-e[i].im, xv.im, acc.re));
        acc.im = fmaf(e[i].re, xv.im, fmaf(e[i].im, xv.re, acc.im));
    }
    return acc;
}

// out[] = base[] + sg * (sum over the 8 parts), ends with a barrier
template <bool T>
__device__ __forceinline__ void bcr_combine(cx<float> acc, const VecRole<T>& v, const cx<float>* base,
                                            cx<float>* out, cx<float> (&red)[8][kBtdMax], int n, float sg) {
    const int t = threadIdx.x;
    if (!T) {
#pragma unroll
        for (int off = 1; off < 8; off <<= 1) {
            acc.re += __shfl_xor(acc.re, off, 64);
            acc.im += __shfl_xor(acc.im, off, 64);
        }
        if (v.p == 0 && v.a < n) out[v.a] = base[v.a] + mk<float>(sg * acc.re, sg * acc.im);
    } else {
        if (v.on) red[v.p][v.a] = acc;
        __syncthreads();
        if (t < n) {
            cx<float> s2 = czero<float>();
#pragma unroll
            for (int g = 0; g < 8; ++g) s2 = s2 + red[g][t];
            out[t] = base[t] + mk<float>(sg * s2.re, sg * s2.im);
        }
    }
    __syncthreads();
}

// Solve-op kinds of a launch: every op's matrices are in registers before
// the first barrier; the kinds bound the registers a launch needs.
constexpr int kVecGeneral = 0;   // M1, M2 and M0 (any present)
constexpr int kVecReduce = 1;    // M1, M2 only: the reduction sweep
constexpr int kVecDiag = 2;      // diagonal M1, M2 and a dense M0: level 0 back-substitution

template <bool TRANS, int KIND>
__global__ void __launch_bounds__(kBtdThreads)
k_bcr_vec(BcrVecBatch B, const BcrVecOp* __restrict__ ops, int n, int nx) {
    __shared__ cx<float> bs[kBtdMax], x1[kBtdMax], x2[kBtdMax], tv[kBtdMax], res[kBtdMax];
    __shared__ cx<float> red[8][kBtdMax];
    const int f = blockIdx.y;
    const cx<float>* __restrict__ slots = B.slots[f];
    cx<float>* const work = B.work[f];
    const BcrVecOp op = ops[blockIdx.x];
    const int64_t nn = (int64_t)n * n;
    const int t = threadIdx.x;
    const VecRole<TRANS> role(n);
    auto vec_in = [&](int i) -> const cx<float>* {
        return i < nx ? B.v[f] + (int64_t)i * n : (i < 2 * nx ? B.z[f] + (int64_t)(i - nx) * n : work + (int64_t)i * n);
    };
    const int m1 = TRANS ? op.m1t : op.m1, m2 = TRANS ? op.m2t : op.m2;
    const bool h1 = op.u1 >= 0 && m1 >= 0, h2 = op.u2 >= 0 && m2 >= 0;
    const bool h0 = (KIND != kVecReduce) && op.m0 >= 0;
    cx<float> e1[kVecPer], e2[kVecPer], e0[kVecPer];
    cx<float> d1 = czero<float>(), d2 = czero<float>();    // kVecDiag: this thread's diagonal entries
    // the thread holding the diagonal entry of its row (A) / column (A^T)
    const bool holder = role.on && (role.p == (role.a & 7));
    if (KIND == kVecDiag) {
        const int64_t dd = (int64_t)role.a * n + role.a;
        if (h1 && holder) d1 = slots[(int64_t)m1 * nn + dd];
        if (h2 && holder) d2 = slots[(int64_t)m2 * nn + dd];
    } else {
        if (h1) bcr_load<TRANS>(slots + (int64_t)m1 * nn, n, op.diag & (TRANS ? 4 : 1), role, e1);
        if (h2) bcr_load<TRANS>(slots + (int64_t)m2 * nn, n, op.diag & (TRANS ? 8 : 2), role, e2);
    }
    if (h0) bcr_load<TRANS>(slots + (int64_t)op.m0 * nn, n, false, role, e0);
    for (int i = t; i < kBtdMax; i += blockDim.x) {
        const bool in = i < n;
        bs[i] = in ? vec_in(op.b)[i] : czero<float>();
        tv[i] = czero<float>();
        res[i] = czero<float>();
        x1[i] = (in && h1) ? vec_in(op.u1)[i] : czero<float>();
        x2[i] = (in && h2) ? vec_in(op.u2)[i] : czero<float>();
    }
    __syncthreads();
    // t = b - M1 u1 - M2 u2
    cx<float> acc = czero<float>();
    if (role.on) {
        if (KIND == kVecDiag) {
            acc = acc + d1 * x1[role.a] + d2 * x2[role.a];
        } else {
            if (h1) acc = acc + bcr_dot(e1, x1, role.p);
            if (h2) acc = acc + bcr_dot(e2, x2, role.p);
        }
    }
    // ops never write a right-hand side (index < nx)
    cx<float>* const outp = (op.out < 2 * nx) ? B.z[f] + (int64_t)(op.out - nx) * n : work + (int64_t)op.out * n;
    if (!h0) {
        if (h1 || h2) {
            bcr_combine<TRANS>(acc, role, bs, tv, red, n, -1.0f);
            if (t < n) outp[t] = tv[t];
        } else if (t < n) {
            outp[t] = bs[t];
        }
        return;
    }
    if (h1 || h2) bcr_combine<TRANS>(acc, role, bs, tv, red, n, -1.0f);
    const cx<float>* const tin = (h1 || h2) ? tv : bs;
    // out = M0 t
    const cx<float> acc0 = role.on ? bcr_dot(e0, tin, role.p) : czero<float>();
    bcr_combine<TRANS>(acc0, role, res, res, red, n, 1.0f);
    if (t < n) outp[t] = res[t];
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
    CV_CHECK(!(dh.flags & CV_FLAG_CPLX_MAT), "complex Hz materials are not supported by the direct factorisation");
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

// ------------------------------------------------------------------ BCR host
namespace cv {

struct BcrPlan {
    int nx = 0, n = 0;
    int n_slots = 0, n_vecs = 0;
    std::vector<int> dls;                         // level-0 D, L, U slots (3 nx)
    std::vector<std::vector<int>> inv;            // per level: slots to invert
    std::vector<std::vector<BcrGemmOp>> g1, g2;   // per level: products, then next-level blocks
    std::vector<std::vector<BcrVecOp>> fwd, bwd;  // per level: RHS reduction, back-substitution
    int final_slot = -1;
    BcrVecOp final_op{};
    // byte offsets inside the factor buffer
    int64_t off_dls = 0, off_final = 0, off_slots = 0, off_vecs = 0, bytes = 0;
    std::vector<int64_t> off_inv, off_g1, off_g2, off_fwd, off_bwd;
};

static int64_t al256b(int64_t x) { return (x + 255) & ~int64_t(255); }

static void bcr_plan(int nx, int n, BcrPlan& P) {
    P = BcrPlan();
    P.nx = nx;
    P.n = n;
    int ns = 0, nv = 0;
    // level 0 vectors: b at 0 .. nx-1, x at nx .. 2nx-1 (contiguous for the copies)
    std::vector<int> b(nx), x(nx), D(nx), L(nx, -1), U(nx, -1);
    for (int j = 0; j < nx; ++j) b[j] = nv++;
    for (int j = 0; j < nx; ++j) x[j] = nv++;
    P.dls.assign(3 * (size_t)nx, -1);
    for (int j = 0; j < nx; ++j) {
        D[j] = ns++;
        if (j >= 1) L[j] = ns++;
        if (j + 1 < nx) U[j] = ns++;
        P.dls[3 * j] = D[j];
        P.dls[3 * j + 1] = L[j];
        P.dls[3 * j + 2] = U[j];
    }
    int N = nx;
    while (N > 1) {
        std::vector<int> inv, D2, L2, U2, b2, x2;
        std::vector<BcrGemmOp> g1, g2;
        std::vector<BcrVecOp> fw, bw;
        for (int j = 1; j < N; j += 2) inv.push_back(D[j]);
        const int N2 = (N + 1) / 2;
        D2.assign(N2, -1);
        L2.assign(N2, -1);
        U2.assign(N2, -1);
        b2.assign(N2, -1);
        x2.assign(N2, -1);
        const int dg = (N == nx) ? 1 : 0;   // level 0: L and U are diagonal
        for (int j = 0; j < N; j += 2) {
            const int jj = j / 2;
            int al = -1, be = -1, ga = -1, de = -1;
            if (j >= 1) {
                al = ns++;
                ga = ns++;
                g1.push_back({al, -1, L[j], D[j - 1], -1, -1, 1.0f, 0.0f, dg * 1});
                g1.push_back({ga, -1, D[j - 1], U[j - 1], -1, -1, 1.0f, 0.0f, dg * 2});
            }
            if (j + 1 < N) {
                be = ns++;
                de = ns++;
                g1.push_back({be, -1, U[j], D[j + 1], -1, -1, 1.0f, 0.0f, dg * 1});
                g1.push_back({de, -1, D[j + 1], L[j + 1], -1, -1, 1.0f, 0.0f, dg * 2});
            }
            D2[jj] = ns++;
            g2.push_back({D2[jj], D[j], (j >= 1) ? L[j] : -1, ga, (j + 1 < N) ? U[j] : -1, de, -1.0f, -1.0f, dg * 5});
            if (j >= 2) {
                L2[jj] = ns++;
                g2.push_back({L2[jj], -1, al, L[j - 1], -1, -1, -1.0f, 0.0f, dg * 2});
            }
            if (j + 2 < N) {
                U2[jj] = ns++;
                g2.push_back({U2[jj], -1, be, U[j + 1], -1, -1, -1.0f, 0.0f, dg * 2});
            }
            b2[jj] = nv++;
            x2[jj] = nv++;
            fw.push_back({b2[jj], b[j], (j >= 1) ? b[j - 1] : -1, (j + 1 < N) ? b[j + 1] : -1, al, be, ga, de, -1, 0});
            bw.push_back({x[j], x2[jj], -1, -1, -1, -1, -1, -1, -1, 0});     // even: copy from the next level
        }
        for (int j = 1; j < N; j += 2) {
            const bool up = (j + 1 < N);
            bw.push_back({x[j], b[j], x2[(j - 1) / 2], up ? x2[(j + 1) / 2] : -1, L[j], up ? U[j] : -1, U[j - 1],
                          up ? L[j + 1] : -1, D[j], dg * 15});
        }
        P.inv.push_back(inv);
        P.g1.push_back(g1);
        P.g2.push_back(g2);
        P.fwd.push_back(fw);
        P.bwd.push_back(bw);
        D = D2;
        L = L2;
        U = U2;
        b = b2;
        x = x2;
        N = N2;
    }
    P.final_slot = D[0];
    P.final_op = {x[0], b[0], -1, -1, -1, -1, -1, -1, D[0], 0};
    P.n_slots = ns;
    P.n_vecs = nv;
    // layout: header | dls | per level inv, g1, g2, fwd, bwd | final op | slots | vecs
    int64_t off = 256;
    P.off_dls = off;
    off = al256b(off + 4 * (int64_t)P.dls.size());
    for (size_t lv = 0; lv < P.inv.size(); ++lv) {
        P.off_inv.push_back(off);
        off = al256b(off + 4 * (int64_t)P.inv[lv].size());
        P.off_g1.push_back(off);
        off = al256b(off + (int64_t)sizeof(BcrGemmOp) * P.g1[lv].size());
        P.off_g2.push_back(off);
        off = al256b(off + (int64_t)sizeof(BcrGemmOp) * P.g2[lv].size());
        P.off_fwd.push_back(off);
        off = al256b(off + (int64_t)sizeof(BcrVecOp) * P.fwd[lv].size());
        P.off_bwd.push_back(off);
        off = al256b(off + (int64_t)sizeof(BcrVecOp) * P.bwd[lv].size());
    }
    P.off_final = off;
    off = al256b(off + sizeof(BcrVecOp) + 4);
    P.off_slots = off;
    off = al256b(off + (int64_t)P.n_slots * n * n * 8);
    P.off_vecs = off;
    off = al256b(off + (int64_t)P.n_vecs * n * 8);
    P.bytes = off;
}

// the plan of the last shape applied on this host thread (applies repeat it)
static const BcrPlan& cached_plan(int nx, int n) {
    thread_local BcrPlan P;
    if (P.nx != nx || P.n != n) bcr_plan(nx, n, P);
    return P;
}

int bcr_apply_batch(const cv_fdfd_desc& dh, int nb, const void* const* factors, const cx<float>* v, cx<float>* z,
                    int64_t stride, int transpose, hipStream_t s) {
    const BcrPlan& P = cached_plan(dh.nx, dh.ny);
    const int n = dh.ny, nx = dh.nx;
    const char* base0 = (const char*)factors[0];
    for (int f0 = 0; f0 < nb; f0 += kBcrBatch) {
        const int cnt = std::min(kBcrBatch, nb - f0);
        BcrVecBatch B{};
        for (int q = 0; q < cnt; ++q) {
            char* base = (char*)factors[f0 + q];
            B.slots[q] = (const cx<float>*)(base + P.off_slots);
            B.work[q] = (cx<float>*)(base + P.off_vecs);
            B.v[q] = v + (int64_t)(f0 + q) * stride;
            B.z[q] = z + (int64_t)(f0 + q) * stride;
        }
        // every factor of this shape holds the same op tables: use the first one's
        const int threads = 64 * ((8 * n + 63) / 64);
        auto launch = [&](int64_t off, int count, int kind) {
            if (count <= 0) return;
            const BcrVecOp* ops = (const BcrVecOp*)(base0 + off);
            const dim3 grid(count, cnt);
            if (transpose) {
                if (kind == kVecReduce) k_bcr_vec<true, kVecReduce><<<grid, threads, 0, s>>>(B, ops, n, nx);
                else if (kind == kVecDiag) k_bcr_vec<true, kVecDiag><<<grid, threads, 0, s>>>(B, ops, n, nx);
                else k_bcr_vec<true, kVecGeneral><<<grid, threads, 0, s>>>(B, ops, n, nx);
            } else {
                if (kind == kVecReduce) k_bcr_vec<false, kVecReduce><<<grid, threads, 0, s>>>(B, ops, n, nx);
                else if (kind == kVecDiag) k_bcr_vec<false, kVecDiag><<<grid, threads, 0, s>>>(B, ops, n, nx);
                else k_bcr_vec<false, kVecGeneral><<<grid, threads, 0, s>>>(B, ops, n, nx);
            }
        };
        for (size_t lv = 0; lv < P.fwd.size(); ++lv) launch(P.off_fwd[lv], (int)P.fwd[lv].size(), kVecReduce);
        launch(P.off_final, 1, kVecGeneral);
        for (size_t lv = P.bwd.size(); lv-- > 0;)
            launch(P.off_bwd[lv], (int)P.bwd[lv].size(), lv == 0 ? kVecDiag : kVecGeneral);
        CV_LAUNCH_CHECK();
    }
    return CV_OK;
}

int bcr_apply_impl(const cv_fdfd_desc& dh, const void* factor, const void* v, void* z, int transpose,
                   hipStream_t s) {
    return bcr_apply_batch(dh, 1, &factor, (const cx<float>*)v, (cx<float>*)z, 0, transpose, s);
}

int direct_apply(int kind, const cv_fdfd_desc* desc, const cv_fdfd_desc& dh, const void* prof32, const void* m0_32,
                 const void* m1_32, const void* factor, const void* v, void* z, int transpose, hipStream_t s) {
    if (kind == CV_FACTOR_BCR) return bcr_apply_impl(dh, factor, v, z, transpose, s);
    return btd_apply(desc, dh, prof32, m0_32, m1_32, factor, v, z, transpose, s);
}

}  // namespace cv

extern "C" int64_t cv_fdfd_bcr_factor_bytes(int nx, int ny) {
    if (nx <= 0 || ny <= 0 || ny > kBtdMax) return 0;
    BcrPlan P;
    bcr_plan(nx, ny, P);
    return P.bytes;
}

namespace cv {
// factorises nb instances of one shape; desc[q] / desc_host[q] describe instance q
static int bcr_factor_batch(int nb, const cv_fdfd_desc* const* desc, const cv_fdfd_desc* desc_host,
                            const void* const* prof32, const void* const* m0_32, const void* const* m1_32,
                            void* const* factor, int64_t factor_bytes, hipStream_t s) {
    CV_CHECK(nb >= 1 && desc && desc_host && prof32 && m0_32 && m1_32 && factor, "null pointer argument");
    const cv_fdfd_desc dh = desc_host[0];
    for (int q = 0; q < nb; ++q) {
        const cv_fdfd_desc& d = desc_host[q];
        CV_CHECK(desc[q] && prof32[q] && m0_32[q] && factor[q], "null pointer argument (instance %d)", q);
        CV_CHECK(d.nx > 0 && d.ny > 0 && d.cell_off == 0, "one instance starting at cell 0 is required");
        CV_CHECK(!(d.flags & CV_FLAG_CPLX_MAT), "complex Hz materials are not supported by the direct factorisation");
        CV_CHECK(d.nx == dh.nx && d.ny == dh.ny && d.pol == dh.pol, "instance %d differs in shape from instance 0", q);
        CV_CHECK(d.ny <= kBtdMax, "block cyclic reduction needs ny <= %d (ny = %d)", kBtdMax, d.ny);
        CV_CHECK(d.nblk == 1 && d.n_cpl == 0, "frequency-coupled systems are not supported");
        CV_CHECK(d.pol == CV_POL_EZ || m1_32[q], "Hz operator needs m1");
    }
    BcrPlan P;
    bcr_plan(dh.nx, dh.ny, P);
    CV_CHECK(factor_bytes >= P.bytes, "factor buffer too small: %lld < %lld", (long long)factor_bytes,
             (long long)P.bytes);
    // tables: staged in one host buffer, uploaded into every factor (each
    // stays self-contained), kept alive until the final synchronisation
    std::vector<char> host((size_t)P.off_slots, 0);
    auto put = [&](int64_t off, const void* src, size_t bytes) { if (bytes) memcpy(host.data() + off, src, bytes); };
    put(P.off_dls, P.dls.data(), 4 * P.dls.size());
    for (size_t lv = 0; lv < P.inv.size(); ++lv) {
        put(P.off_inv[lv], P.inv[lv].data(), 4 * P.inv[lv].size());
        put(P.off_g1[lv], P.g1[lv].data(), sizeof(BcrGemmOp) * P.g1[lv].size());
        put(P.off_g2[lv], P.g2[lv].data(), sizeof(BcrGemmOp) * P.g2[lv].size());
        put(P.off_fwd[lv], P.fwd[lv].data(), sizeof(BcrVecOp) * P.fwd[lv].size());
        put(P.off_bwd[lv], P.bwd[lv].data(), sizeof(BcrVecOp) * P.bwd[lv].size());
    }
    put(P.off_final, &P.final_op, sizeof(BcrVecOp));
    put(P.off_final + sizeof(BcrVecOp), &P.final_slot, 4);
    for (int q = 0; q < nb; ++q) {
        char* base = (char*)factor[q];
        CV_HIP(hipMemcpyAsync(base, host.data(), host.size(), hipMemcpyHostToDevice, s));
        k_zero_u32_btd<<<1, 64, 0, s>>>((unsigned int*)base, 1);   // status: header word 0
    }
    const char* base0 = (const char*)factor[0];
    const int n = dh.ny;
    for (int f0 = 0; f0 < nb; f0 += kBcrBatch) {
        const int cnt = std::min(kBcrBatch, nb - f0);
        BcrSlots S{};
        BcrBuild in{};
        for (int q = 0; q < cnt; ++q) {
            char* base = (char*)factor[f0 + q];
            S.slots[q] = (cx<float>*)(base + P.off_slots);
            S.status[q] = (int*)base;
            in.desc[q] = desc[f0 + q];
            in.prof[q] = (const cx<float>*)prof32[f0 + q];
            in.m0[q] = (const float*)m0_32[f0 + q];
            in.m1[q] = (const float*)m1_32[f0 + q];
        }
        const int inv_threads = 64 * ((n + kFRows - 1) / kFRows);
        k_bcr_build0<<<dim3(dh.nx, cnt), 256, 0, s>>>(in, S, (const int*)(base0 + P.off_dls));
        for (size_t lv = 0; lv < P.inv.size(); ++lv) {
            if (!P.inv[lv].empty())
                k_bcr_inv<<<dim3((int)P.inv[lv].size(), cnt), inv_threads, 0, s>>>(
                    S, (const int*)(base0 + P.off_inv[lv]), n);
            if (!P.g1[lv].empty())
                launch_gemm(S, (const BcrGemmOp*)(base0 + P.off_g1[lv]), (int)P.g1[lv].size(), cnt, n, lv == 0, s);
            if (!P.g2[lv].empty())
                launch_gemm(S, (const BcrGemmOp*)(base0 + P.off_g2[lv]), (int)P.g2[lv].size(), cnt, n, lv == 0, s);
        }
        k_bcr_inv<<<dim3(1, cnt), inv_threads, 0, s>>>(S, (const int*)(base0 + P.off_final + sizeof(BcrVecOp)), n);
        CV_LAUNCH_CHECK();
    }
    std::vector<int> st(nb, 0);
    for (int q = 0; q < nb; ++q)
        CV_HIP(hipMemcpyAsync(&st[q], factor[q], sizeof(int), hipMemcpyDeviceToHost, s));
    CV_HIP(hipStreamSynchronize(s));
    for (int q = 0; q < nb; ++q)
        if (st[q] != 0) {
            set_error("zero pivot in block slot %d of the block cyclic reduction (instance %d)", st[q] - 1, q);
            return CV_ERR_SINGULAR;
        }
    return CV_OK;
}
}  // namespace cv

extern "C" int cv_fdfd_bcr_factor(const cv_fdfd_desc* desc, const cv_fdfd_desc* desc_host, const void* prof32,
                                  const void* m0_32, const void* m1_32, void* factor, int64_t factor_bytes,
                                  void* stream) {
    return bcr_factor_batch(1, &desc, desc_host, &prof32, &m0_32, &m1_32, &factor, factor_bytes, as_stream(stream));
}

extern "C" int cv_fdfd_bcr_factor_batch(int nb, const cv_fdfd_desc* const* desc, const cv_fdfd_desc* desc_host,
                                        const void* const* prof32, const void* const* m0_32,
                                        const void* const* m1_32, void* const* factors, int64_t factor_bytes,
                                        void* stream) {
    return bcr_factor_batch(nb, desc, desc_host, prof32, m0_32, m1_32, factors, factor_bytes, as_stream(stream));
}

extern "C" int cv_fdfd_bcr_apply(const cv_fdfd_desc* desc, const cv_fdfd_desc* desc_host, const void* factor,
                                 const void* v, void* z, int transpose, void* stream) {
    CV_CHECK(desc && desc_host && factor && v && z, "null pointer argument");
    CV_CHECK(v != z, "v and z must not alias");
    const cv_fdfd_desc dh = desc_host[0];
    CV_CHECK(dh.nx > 0 && dh.ny > 0 && dh.ny <= kBtdMax && dh.cell_off == 0 && dh.nblk == 1,
             "single instance with ny <= %d required", kBtdMax);
    return bcr_apply_impl(dh, factor, v, z, transpose, as_stream(stream));
}
