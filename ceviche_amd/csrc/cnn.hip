// Kernels of the convolutional preconditioner: 3x3 convolutions with 32 input
// and 32 output channels on the bf16 matrix cores (fp32 accumulation), 2x
// down/up sampling, and the packing of complex residuals into network
// channels.  No counterpart exists in the reference (ceviche solves with a
// sparse direct factorisation, ceviche/solvers.py:37-62); the network is the
// learned preconditioner inside the flexible GMRES of fgmres.hip.
//
// Activations are NHWC bf16 with C = 32: one pixel is 64 bytes, four 16-byte
// chunks of 8 channels.  A convolution is an implicit GEMM per 3x3 tap:
//     out[co][pixel] += W_t[co][ci] * in[ci][pixel + shift_t]
// mapped onto v_mfma_f32_16x16x32_bf16 with M = 16 output channels,
// N = 16 pixels of one row, K = 32 input channels.  For that instruction lane
// l holds A[m = l%16][k = 8(l/16)..+8], B[k = 8(l/16)..+8][n = l%16] and
// D[m = 4(l/16)+i][n = l%16], i = 0..3, so a B operand is exactly one 16-byte
// chunk of one pixel, and the weights (A operands, 18 per layer) stay in
// registers for the whole workgroup.
#include "fdfd_stencil.h"

#include <algorithm>

namespace cv {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

constexpr int kC = 32;                    // channels of every convolution
constexpr int kTH = 8, kTW = 32;          // output tile per workgroup (rows x cols)
constexpr int kHH = kTH + 2, kHW = kTW + 2;
constexpr int kWElems = 9 * kC * kC;      // weights of one layer, fp32 master layout [tap][ci][co]
constexpr int kWPack = 9 * 2 * 64 * 8;    // bf16 A operands [tap][co half][lane][8]
constexpr int kWgPartial = kWElems + kC;  // per-workgroup partial of dW and db

__device__ __forceinline__ float bf2f(uint32_t v16) { return __uint_as_float(v16 << 16); }

// fp32 -> bf16, round to nearest even (NaN stays NaN)
__device__ __forceinline__ uint32_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (u >> 16) | ((u & 0xffffu) ? 0x40u : 0u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return u >> 16;
}

__device__ __forceinline__ bool bf_pos(uint32_t v16) { return !(v16 & 0x8000u) && (v16 & 0x7fffu); }

// zero the bf16 lanes of v whose mask lane is not > 0
__device__ __forceinline__ uint4 mask_chunk(uint4 v, uint4 m) {
    uint32_t* pv = reinterpret_cast<uint32_t*>(&v);
    const uint32_t* pm = reinterpret_cast<const uint32_t*>(&m);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint32_t lo = bf_pos(pm[q] & 0xffffu) ? 0x0000ffffu : 0u;
        const uint32_t hi = bf_pos(pm[q] >> 16) ? 0xffff0000u : 0u;
        pv[q] &= (lo | hi);
    }
    return v;
}

__device__ __forceinline__ bf16x8 as_bf16x8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }

// ---------------------------------------------------------------- weights
// fp32 master W[tap][ci][co] -> A operands.  flip = 0: the forward layer.
// flip = 1: the layer whose forward pass is the data gradient of W, i.e.
// W'[t][ci'][co'] = W[8 - t][co'][ci'] (taps mirrored, channels swapped).
__global__ void __launch_bounds__(kBlock) k_pack_w(const float* __restrict__ w, uint16_t* __restrict__ pack,
                                                   int flip) {
    const int idx = blockIdx.x * kBlock + threadIdx.x;
    if (idx >= kWPack) return;
    const int e = idx & 7, l = (idx >> 3) & 63, h = (idx >> 9) & 1, t = idx >> 10;
    const int m = 16 * h + (l & 15);      // output channel of the packed layer
    const int k = 8 * (l >> 4) + e;       // input channel of the packed layer
    const float v = flip ? w[((8 - t) * kC + m) * kC + k] : w[(t * kC + k) * kC + m];
    pack[idx] = (uint16_t)f2bf(v);
}

// ------------------------------------------------------------ convolution
struct ConvArgs {
    const uint16_t* in;     // (B, H, W, 32)
    const uint16_t* mask;   // optional (B, H, W, 32): input lanes with mask <= 0 read as zero
    const uint16_t* wpack;  // kWPack A operands
    const float* bias;      // optional (32)
    uint16_t* out;          // (B, H, W, 32)
    int B, H, W, tiles_x, tiles_y;
    int relu, accumulate;
};

// out = act(conv3x3(in (.) [mask > 0], W) + bias (+ out if accumulate)),
// zero padding at the image border.  One workgroup = one 8 x 32 output tile:
// the 10 x 34 halo tile is staged in LDS, each wave computes two rows as four
// 16-pixel segments, 18 MFMAs per segment.
__global__ void __launch_bounds__(kBlock) k_conv3x3(ConvArgs a) {
    __shared__ __attribute__((aligned(16))) uint4 tile[kHH * kHW * 4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int bid = blockIdx.x;
    const int tx = bid % a.tiles_x;
    bid /= a.tiles_x;
    const int ty = bid % a.tiles_y;
    const int b = bid / a.tiles_y;
    const int y0 = ty * kTH, x0 = tx * kTW;
    const size_t img = (size_t)b * a.H * a.W;

    bf16x8 wa[9][2];
    const uint4* wp = reinterpret_cast<const uint4*>(a.wpack);
#pragma unroll
    for (int t = 0; t < 9; ++t)
#pragma unroll
        for (int h = 0; h < 2; ++h) wa[t][h] = as_bf16x8(wp[(t * 2 + h) * 64 + lane]);

    const uint4* in4 = reinterpret_cast<const uint4*>(a.in);
    const uint4* mk4 = reinterpret_cast<const uint4*>(a.mask);
    for (int e = tid; e < kHH * kHW * 4; e += kBlock) {
        const int chunk = e & 3, pix = e >> 2;
        const int hy = pix / kHW, hx = pix - hy * kHW;
        const int gy = y0 + hy - 1, gx = x0 + hx - 1;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (gy >= 0 && gy < a.H && gx >= 0 && gx < a.W) {
            const size_t off = (img + (size_t)gy * a.W + gx) * 4 + chunk;
            v = in4[off];
            if (mk4) v = mask_chunk(v, mk4[off]);
        }
        tile[e] = v;
    }
    __syncthreads();

#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int r = 2 * wave + (s >> 1), c0 = (s & 1) * 16;
        f32x4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            const int dy = t / 3, dx = t - 3 * (t / 3);
            const int hp = (r + dy) * kHW + c0 + (lane & 15) + dx;
            const bf16x8 bv = as_bf16x8(tile[hp * 4 + (lane >> 4)]);
            acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa[t][0], bv, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa[t][1], bv, acc1, 0, 0, 0);
        }
        const int gy = y0 + r, gx = x0 + c0 + (lane & 15);
        if (gy < a.H && gx < a.W) {
            const size_t pix = img + (size_t)gy * a.W + gx;
            uint2* o2 = reinterpret_cast<uint2*>(a.out);
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int co0 = 16 * h + 4 * (lane >> 4);
                const f32x4 acc = h ? acc1 : acc0;
                float v[4] = {acc[0], acc[1], acc[2], acc[3]};
                if (a.bias) {
#pragma unroll
                    for (int i = 0; i < 4; ++i) v[i] += a.bias[co0 + i];
                }
                const size_t o = (pix * kC + co0) >> 2;
                if (a.accumulate) {
                    const uint2 old = o2[o];
                    v[0] += bf2f(old.x & 0xffffu);
                    v[1] += bf2f(old.x >> 16);
                    v[2] += bf2f(old.y & 0xffffu);
                    v[3] += bf2f(old.y >> 16);
                }
                if (a.relu) {
#pragma unroll
                    for (int i = 0; i < 4; ++i) v[i] = fmaxf(v[i], 0.f);
                }
                o2[o] = make_uint2(f2bf(v[0]) | (f2bf(v[1]) << 16), f2bf(v[2]) | (f2bf(v[3]) << 16));
            }
        }
    }
}

// ---------------------------------------------------------- weight gradient
// dW[t][ci][co] = sum_p x[p + shift_t][ci] g[p][co], db[co] = sum_p g[p][co],
// g = gy (.) [ymask > 0].  GEMM per tap with M = 16 output channels, N = 16
// input channels, K = the 32 pixels of one tile row; the operands come from
// channel-major copies of the tile in LDS (three copies of x, one per column
// shift, so every operand read is an aligned 16-byte chunk).  Wave w owns the
// (co half, ci half) block w for all 9 taps; workgroups loop over tiles and
// write one partial each, summed in a fixed order by k_wgrad_reduce.
__global__ void __launch_bounds__(kBlock) k_conv3x3_wgrad(const uint16_t* __restrict__ x,
                                                          const uint16_t* __restrict__ gy,
                                                          const uint16_t* __restrict__ ymask,
                                                          float* __restrict__ partial, int B, int H, int W,
                                                          int tiles_x, int tiles_y) {
    __shared__ __attribute__((aligned(16))) uint16_t xT[3][kC][kHH][kTW];
    __shared__ __attribute__((aligned(16))) uint16_t gT[kC][kTH][kTW];
    __shared__ float bred[kBlock];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int hc = wave >> 1, hi = wave & 1;
    const int n_tiles = tiles_x * tiles_y * B;
    f32x4 acc[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    float bacc = 0.f;
    const uint4* x4 = reinterpret_cast<const uint4*>(x);
    const uint4* g4 = reinterpret_cast<const uint4*>(gy);
    const uint4* m4 = reinterpret_cast<const uint4*>(ymask);

    for (int tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        int q = tile;
        const int tx = q % tiles_x;
        q /= tiles_x;
        const int ty = q % tiles_y;
        const int b = q / tiles_y;
        const int y0 = ty * kTH, x0 = tx * kTW;
        const size_t img = (size_t)b * H * W;
        __syncthreads();  // previous tile's reads are done
        for (int e = tid; e < kHH * kHW * 4; e += kBlock) {
            const int chunk = e & 3, pix = e >> 2;
            const int hy = pix / kHW, hx = pix - hy * kHW;
            const int gyy = y0 + hy - 1, gxx = x0 + hx - 1;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (gyy >= 0 && gyy < H && gxx >= 0 && gxx < W) v = x4[(img + (size_t)gyy * W + gxx) * 4 + chunk];
            const uint32_t* pv = reinterpret_cast<const uint32_t*>(&v);
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) {
                const int c = hx - dx;
                if (c < 0 || c >= kTW) continue;
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    xT[dx][chunk * 8 + k][hy][c] = (uint16_t)((k & 1) ? (pv[k >> 1] >> 16) : (pv[k >> 1] & 0xffffu));
            }
        }
        for (int e = tid; e < kTH * kTW * 4; e += kBlock) {
            const int chunk = e & 3, pix = e >> 2;
            const int ry = pix / kTW, rc = pix - ry * kTW;
            const int gyy = y0 + ry, gxx = x0 + rc;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (gyy < H && gxx < W) {
                const size_t off = (img + (size_t)gyy * W + gxx) * 4 + chunk;
                v = g4[off];
                if (m4) v = mask_chunk(v, m4[off]);
            }
            const uint32_t* pv = reinterpret_cast<const uint32_t*>(&v);
#pragma unroll
            for (int k = 0; k < 8; ++k)
                gT[chunk * 8 + k][ry][rc] = (uint16_t)((k & 1) ? (pv[k >> 1] >> 16) : (pv[k >> 1] & 0xffffu));
        }
        __syncthreads();
        // bias: thread (co = tid/8, row = tid%8) sums one tile row of one channel
        {
            const int co = tid >> 3, ry = tid & 7;
            float s = 0.f;
#pragma unroll 8
            for (int c = 0; c < kTW; ++c) s += bf2f(gT[co][ry][c]);
            bacc += s;
        }
#pragma unroll
        for (int ry = 0; ry < kTH; ++ry) {
            const bf16x8 av = as_bf16x8(*reinterpret_cast<const uint4*>(&gT[16 * hc + (lane & 15)][ry][8 * (lane >> 4)]));
#pragma unroll
            for (int t = 0; t < 9; ++t) {
                const int dy = t / 3, dx = t - 3 * (t / 3);
                const bf16x8 bv =
                    as_bf16x8(*reinterpret_cast<const uint4*>(&xT[dx][16 * hi + (lane & 15)][ry + dy][8 * (lane >> 4)]));
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[t], 0, 0, 0);
            }
        }
    }
    float* pw = partial + (size_t)blockIdx.x * kWgPartial;
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        const int ci = 16 * hi + (lane & 15);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int co = 16 * hc + 4 * (lane >> 4) + i;
            pw[(t * kC + ci) * kC + co] = acc[t][i];
        }
    }
    bred[tid] = bacc;
    __syncthreads();
    if (tid < kC) {
        float s = 0.f;
#pragma unroll
        for (int r = 0; r < 8; ++r) s += bred[tid * 8 + r];
        pw[kWElems + tid] = s;
    }
}

// First stage of the partial sum: segment blockIdx.y adds the partials of
// workgroups [kWgSeg y, kWgSeg (y+1)) in workgroup order, one thread per
// element, into seg[y].  The second stage (k_wgrad_reduce over the segments)
// adds the segments in order, so the total is summed in a fixed order.
constexpr int kWgSeg = 32;

__global__ void __launch_bounds__(kBlock) k_wgrad_segsum(const float* __restrict__ partial, int n_wg,
                                                         float* __restrict__ seg) {
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= kWgPartial) return;
    const int g0 = blockIdx.y * kWgSeg, g1 = min(n_wg, g0 + kWgSeg);
    float s = 0.f;
    for (int g = g0; g < g1; ++g) s += partial[(size_t)g * kWgPartial + e];
    seg[(size_t)blockIdx.y * kWgPartial + e] = s;
}

// dw (+)= sum over workgroups of partial[:, :kWElems], db (+)= ... [kWElems:]
__global__ void __launch_bounds__(kBlock) k_wgrad_reduce(const float* __restrict__ partial, int n_wg,
                                                         float* __restrict__ dw, float* __restrict__ db,
                                                         int accumulate) {
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= kWgPartial) return;
    float s = 0.f;
    for (int g = 0; g < n_wg; ++g) s += partial[(size_t)g * kWgPartial + e];
    float* dst = (e < kWElems) ? (dw + e) : (db + (e - kWElems));
    *dst = accumulate ? (*dst + s) : s;
}

// ------------------------------------------------------ resampling by two
// out (coarse Hc x Wc, Hc = ceil(H/2)) = scale * sum of the 2x2 fine pixels
// that exist
__global__ void __launch_bounds__(kBlock) k_down2(const uint4* __restrict__ in, uint4* __restrict__ out, int B,
                                                  int H, int W, int Hc, int Wc, float scale) {
    const int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t total = (int64_t)B * Hc * Wc * 4;
    if (idx >= total) return;
    const int chunk = (int)(idx & 3);
    int64_t p = idx >> 2;
    const int j = (int)(p % Wc);
    p /= Wc;
    const int i = (int)(p % Hc);
    const int b = (int)(p / Hc);
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int di = 0; di < 2; ++di)
#pragma unroll
        for (int dj = 0; dj < 2; ++dj) {
            const int y = 2 * i + di, xx = 2 * j + dj;
            if (y >= H || xx >= W) continue;
            const uint4 v = in[(((size_t)b * H + y) * W + xx) * 4 + chunk];
            const uint32_t* pv = reinterpret_cast<const uint32_t*>(&v);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                s[2 * k] += bf2f(pv[k] & 0xffffu);
                s[2 * k + 1] += bf2f(pv[k] >> 16);
            }
        }
    uint4 o;
    uint32_t* po = reinterpret_cast<uint32_t*>(&o);
#pragma unroll
    for (int k = 0; k < 4; ++k) po[k] = f2bf(scale * s[2 * k]) | (f2bf(scale * s[2 * k + 1]) << 16);
    out[idx] = o;
}

// out (fine H x W) = scale * in[coarse pixel (y/2, x/2)] + (add ? add : 0);
// out may be the same buffer as add
__global__ void __launch_bounds__(kBlock) k_up2(const uint4* __restrict__ in, const uint4* add, uint4* out, int B,
                                                int H, int W, int Hc, int Wc, float scale) {
    const int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t total = (int64_t)B * H * W * 4;
    if (idx >= total) return;
    const int chunk = (int)(idx & 3);
    int64_t p = idx >> 2;
    const int xx = (int)(p % W);
    p /= W;
    const int y = (int)(p % H);
    const int b = (int)(p / H);
    const uint4 v = in[(((size_t)b * Hc + (y >> 1)) * Wc + (xx >> 1)) * 4 + chunk];
    const uint32_t* pv = reinterpret_cast<const uint32_t*>(&v);
    uint4 a4 = make_uint4(0u, 0u, 0u, 0u);
    if (add) a4 = add[idx];
    const uint32_t* pa = reinterpret_cast<const uint32_t*>(&a4);
    uint4 o;
    uint32_t* po = reinterpret_cast<uint32_t*>(&o);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float lo = scale * bf2f(pv[k] & 0xffffu) + bf2f(pa[k] & 0xffffu);
        const float hi = scale * bf2f(pv[k] >> 16) + bf2f(pa[k] >> 16);
        po[k] = f2bf(lo) | (f2bf(hi) << 16);
    }
    out[idx] = o;
}

// ------------------------------------------------------- network in / out
// out[b][pixel] channels: 0, 1 = rs * rscale[b] * (Re, Im) of r[r_idx[b]][pixel];
// 2 .. 2+nf-1 = feat[f_idx[b]][pixel][:]; the rest 0.  Null index arrays
// mean the identity.
template <typename R>
__global__ void __launch_bounds__(kBlock) k_pack_input(const cx<R>* __restrict__ r, const int* __restrict__ r_idx,
                                                       const float* __restrict__ rscale, float rs,
                                                       const float* __restrict__ feat, const int* __restrict__ f_idx,
                                                       int nf, uint4* __restrict__ out, int B, int64_t n) {
    const int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (idx >= (int64_t)B * n) return;
    const int b = (int)(idx / n);
    const int64_t p = idx - (int64_t)b * n;
    const int rb = r_idx ? r_idx[b] : b;
    const cx<R> z = r[(int64_t)rb * n + p];
    const float s = rs * (rscale ? rscale[b] : 1.f);
    float ch[kC];
#pragma unroll
    for (int c = 0; c < kC; ++c) ch[c] = 0.f;
    ch[0] = s * (float)z.re;
    ch[1] = s * (float)z.im;
    if (feat) {
        const int fb = f_idx ? f_idx[b] : b;
        const float* f = feat + ((int64_t)fb * n + p) * nf;
#pragma unroll
        for (int c = 0; c < kC - 2; ++c)
            if (c < nf) ch[2 + c] = f[c];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        uint4 o;
        uint32_t* po = reinterpret_cast<uint32_t*>(&o);
#pragma unroll
        for (int k = 0; k < 4; ++k) po[k] = f2bf(ch[8 * q + 2 * k]) | (f2bf(ch[8 * q + 2 * k + 1]) << 16);
        out[idx * 4 + q] = o;
    }
}

// z[b][pixel] = os * oscale[b] * (ch0 + i ch1)
template <typename R>
__global__ void __launch_bounds__(kBlock) k_unpack(const uint32_t* __restrict__ net, const float* __restrict__ oscale,
                                                   float os, cx<R>* __restrict__ z, int B, int64_t n) {
    const int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (idx >= (int64_t)B * n) return;
    const int b = (int)(idx / n);
    const uint32_t v = net[idx * (kC / 2)];
    const float s = os * (oscale ? oscale[b] : 1.f);
    z[idx] = mk<R>((R)(s * bf2f(v & 0xffffu)), (R)(s * bf2f(v >> 16)));
}

// Relative squared error of channels 0, 1 against e[e_idx[b]]:
//   loss += gscale * inv_norm[e_idx[b]] * |ch - e|^2
//   g[..][0:2] = 2 gscale inv_norm (ch - e), g[..][2:32] = 0
__global__ void __launch_bounds__(kBlock) k_loss(const uint32_t* __restrict__ net, const cx<float>* __restrict__ e,
                                                 const int* __restrict__ e_idx, const float* __restrict__ inv_norm,
                                                 uint4* __restrict__ g, double* __restrict__ loss, float gscale,
                                                 const cx<float>* __restrict__ extra, float extra_scale, int B,
                                                 int64_t n) {
    __shared__ double lds[kWaves];
    const int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    double l = 0.0;
    if (idx < (int64_t)B * n) {
        const int b = (int)(idx / n);
        const int64_t p = idx - (int64_t)b * n;
        const int eb = e_idx ? e_idx[b] : b;
        const cx<float> t = e[(int64_t)eb * n + p];
        const uint32_t v = net[idx * (kC / 2)];
        const float w = gscale * inv_norm[eb];
        const float d0 = bf2f(v & 0xffffu) - t.re, d1 = bf2f(v >> 16) - t.im;
        l = (double)w * ((double)d0 * d0 + (double)d1 * d1);
        float g0 = 2.f * w * d0, g1 = 2.f * w * d1;
        if (extra) {
            const cx<float> x = extra[idx];
            g0 += extra_scale * x.re;
            g1 += extra_scale * x.im;
        }
        g[idx * 4] = make_uint4(f2bf(g0) | (f2bf(g1) << 16), 0u, 0u, 0u);
        g[idx * 4 + 1] = make_uint4(0u, 0u, 0u, 0u);
        g[idx * 4 + 2] = make_uint4(0u, 0u, 0u, 0u);
        g[idx * 4 + 3] = make_uint4(0u, 0u, 0u, 0u);
    }
    double v[1] = {l};
    block_sum<1>(v, lds);
    if (threadIdx.x == 0) __hip_atomic_fetch_add(loss, v[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Per-sample FDFD operator of a training batch: sample b lives at x + b n
// and uses instance inst[b] of a multi-instance operator (all instances of
// equal size n, materials at desc[inst].cell_off).
//   herm = 0:  y_b = A x_b - (r ? r[r_idx[b]] : 0), loss += |y_b|^2 * lscale
//   herm = 1:  y_b = A^H x_b
__global__ void __launch_bounds__(kBlock)
k_batch_apply(const cv_fdfd_desc* __restrict__ desc, const cx<float>* __restrict__ prof, const float* __restrict__ m0,
              const float* __restrict__ m1, const int* __restrict__ inst, const cx<float>* __restrict__ x,
              const cx<float>* __restrict__ r, const int* __restrict__ r_idx, cx<float>* __restrict__ y,
              double* __restrict__ loss, double lscale, int herm, int B, int64_t n) {
    __shared__ double lds[kWaves];
    const int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    double l = 0.0;
    if (idx < (int64_t)B * n) {
        const int b = (int)(idx / n);
        const int64_t loc = idx - (int64_t)b * n;
        const cv_fdfd_desc d = load_desc(desc, inst[b]);
        const int i = (int)(loc / d.ny), j = (int)(loc - (int64_t)i * d.ny);
        const cx<float>* xb = x + (int64_t)b * n - d.cell_off;   // material index -> sample field
        cx<float> out;
        if (herm) {
            auto fx = [&](int64_t k) { return conj(xb[k]); };
            out = conj(apply_any<float, true>(d, prof, m0, m1, i, j, fx, conj(x[idx])));
        } else {
            auto fx = [&](int64_t k) { return xb[k]; };
            out = apply_any<float, false>(d, prof, m0, m1, i, j, fx, x[idx]);
            if (r) out = out - r[(int64_t)(r_idx ? r_idx[b] : b) * n + loc];
            l = lscale * (double)norm2(out);
        }
        y[idx] = out;
    }
    if (loss) {
        double v[1] = {l};
        block_sum<1>(v, lds);
        if (threadIdx.x == 0) __hip_atomic_fetch_add(loss, v[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

static int grid_of(int64_t n) { return (int)((n + kBlock - 1) / kBlock); }

static int conv_launch(const void* in, const void* mask, const void* wpack, const float* bias, void* out, int B,
                       int H, int W, int relu, int accumulate, hipStream_t s) {
    CV_CHECK(in && wpack && out, "null pointer argument");
    CV_CHECK(B > 0 && H > 0 && W > 0, "empty tensor (%d, %d, %d)", B, H, W);
    ConvArgs a;
    a.in = (const uint16_t*)in;
    a.mask = (const uint16_t*)mask;
    a.wpack = (const uint16_t*)wpack;
    a.bias = bias;
    a.out = (uint16_t*)out;
    a.B = B;
    a.H = H;
    a.W = W;
    a.tiles_x = (W + kTW - 1) / kTW;
    a.tiles_y = (H + kTH - 1) / kTH;
    a.relu = relu;
    a.accumulate = accumulate;
    const int64_t blocks = (int64_t)a.tiles_x * a.tiles_y * B;
    CV_CHECK(blocks < (1ll << 31), "tensor too large");
    k_conv3x3<<<(unsigned)blocks, kBlock, 0, s>>>(a);
    CV_LAUNCH_CHECK();
    return CV_OK;
}

static int wgrad_blocks(int B, int H, int W) {
    const int64_t tiles = (int64_t)((W + kTW - 1) / kTW) * ((H + kTH - 1) / kTH) * B;
    return (int)std::min<int64_t>(tiles, 1024);
}

// Level sizes of the U-Net: level 0 is the grid, each next level halves
// (rounding up) both dimensions.
static void level_dims(int H, int W, int levels, int* h, int* w) {
    h[0] = H;
    w[0] = W;
    for (int l = 1; l < levels; ++l) {
        h[l] = (h[l - 1] + 1) / 2;
        w[l] = (w[l - 1] + 1) / 2;
    }
}

static int64_t act_bytes(int B, int h, int w) { return ((int64_t)B * h * w * kC * 2 + 255) & ~int64_t(255); }

}  // namespace cv

using namespace cv;

extern "C" int cv_cnn_pack_weights(const float* w, void* wpack, int flip, int n_layers, void* stream) {
    CV_CHECK(w && wpack, "null pointer argument");
    CV_CHECK(n_layers > 0, "n_layers must be positive");
    hipStream_t s = as_stream(stream);
    for (int l = 0; l < n_layers; ++l)
        k_pack_w<<<grid_of(kWPack), kBlock, 0, s>>>(w + (size_t)l * kWElems, (uint16_t*)wpack + (size_t)l * kWPack,
                                                    flip ? 1 : 0);
    CV_LAUNCH_CHECK();
    return CV_OK;
}

extern "C" int cv_conv3x3_fwd_bf16(const void* in, const void* wpack, const float* bias, void* out, int B, int H,
                                   int W, int relu, void* stream) {
    return conv_launch(in, nullptr, wpack, bias, out, B, H, W, relu, 0, as_stream(stream));
}

extern "C" int cv_conv3x3_bwd_data_bf16(const void* gy, const void* ymask, const void* wpack_flip, void* gx, int B,
                                        int H, int W, int accumulate, void* stream) {
    return conv_launch(gy, ymask, wpack_flip, nullptr, gx, B, H, W, 0, accumulate, as_stream(stream));
}

// workspace: one partial per workgroup, then one per segment of kWgSeg partials
extern "C" int64_t cv_conv3x3_bwd_weight_workspace_bytes(int B, int H, int W) {
    if (B <= 0 || H <= 0 || W <= 0) return 0;
    const int64_t nwg = wgrad_blocks(B, H, W);
    const int64_t nseg = (nwg + kWgSeg - 1) / kWgSeg;
    return (nwg + nseg) * kWgPartial * (int64_t)sizeof(float);
}

extern "C" int cv_conv3x3_bwd_weight_bf16(const void* x, const void* gy, const void* ymask, float* dw, float* db,
                                          int B, int H, int W, int accumulate, void* workspace,
                                          int64_t workspace_bytes, void* stream) {
    CV_CHECK(x && gy && dw && db && workspace, "null pointer argument");
    CV_CHECK(B > 0 && H > 0 && W > 0, "empty tensor (%d, %d, %d)", B, H, W);
    const int64_t need = cv_conv3x3_bwd_weight_workspace_bytes(B, H, W);
    CV_CHECK(workspace_bytes >= need, "workspace too small: %lld < %lld", (long long)workspace_bytes,
             (long long)need);
    hipStream_t s = as_stream(stream);
    const int nwg = wgrad_blocks(B, H, W);
    k_conv3x3_wgrad<<<nwg, kBlock, 0, s>>>((const uint16_t*)x, (const uint16_t*)gy, (const uint16_t*)ymask,
                                           (float*)workspace, B, H, W, (W + kTW - 1) / kTW, (H + kTH - 1) / kTH);
    CV_LAUNCH_CHECK();
    const int nseg = (nwg + kWgSeg - 1) / kWgSeg;
    float* seg = (float*)workspace + (size_t)nwg * kWgPartial;
    k_wgrad_segsum<<<dim3(grid_of(kWgPartial), nseg), kBlock, 0, s>>>((const float*)workspace, nwg, seg);
    CV_LAUNCH_CHECK();
    k_wgrad_reduce<<<grid_of(kWgPartial), kBlock, 0, s>>>(seg, nseg, dw, db, accumulate);
    CV_LAUNCH_CHECK();
    return CV_OK;
}

extern "C" int cv_cnn_down2(const void* in, void* out, int B, int H, int W, float scale, void* stream) {
    CV_CHECK(in && out, "null pointer argument");
    CV_CHECK(B > 0 && H > 0 && W > 0, "empty tensor");
    const int Hc = (H + 1) / 2, Wc = (W + 1) / 2;
    const int64_t total = (int64_t)B * Hc * Wc * 4;
    k_down2<<<grid_of(total), kBlock, 0, as_stream(stream)>>>((const uint4*)in, (uint4*)out, B, H, W, Hc, Wc, scale);
    CV_LAUNCH_CHECK();
    return CV_OK;
}

extern "C" int cv_cnn_up2(const void* in, const void* add, void* out, int B, int H, int W, float scale,
                          void* stream) {
    CV_CHECK(in && out, "null pointer argument");
    CV_CHECK(B > 0 && H > 0 && W > 0, "empty tensor");
    const int Hc = (H + 1) / 2, Wc = (W + 1) / 2;
    const int64_t total = (int64_t)B * H * W * 4;
    k_up2<<<grid_of(total), kBlock, 0, as_stream(stream)>>>((const uint4*)in, (const uint4*)add, (uint4*)out, B, H,
                                                            W, Hc, Wc, scale);
    CV_LAUNCH_CHECK();
    return CV_OK;
}

extern "C" int cv_cnn_pack_input(int dtype, const void* r, const int* r_idx, const float* rscale, float rs,
                                 const float* feat, const int* f_idx, int nf, void* out, int B, int H, int W,
                                 void* stream) {
    CV_CHECK(r && out, "null pointer argument");
    CV_CHECK(B > 0 && H > 0 && W > 0, "empty tensor");
    CV_CHECK(nf >= 0 && nf <= kC - 2, "nf must be in [0, %d]", kC - 2);
    CV_CHECK(dtype == CV_C64 || dtype == CV_C128, "unknown dtype %d", dtype);
    const int64_t n = (int64_t)H * W;
    hipStream_t s = as_stream(stream);
    if (dtype == CV_C64)
        k_pack_input<float><<<grid_of(B * n), kBlock, 0, s>>>((const cx<float>*)r, r_idx, rscale, rs, feat, f_idx,
                                                              nf, (uint4*)out, B, n);
    else
        k_pack_input<double><<<grid_of(B * n), kBlock, 0, s>>>((const cx<double>*)r, r_idx, rscale, rs, feat, f_idx,
                                                               nf, (uint4*)out, B, n);
    CV_LAUNCH_CHECK();
    return CV_OK;
}

extern "C" int cv_cnn_unpack(int dtype, const void* net, const float* oscale, float os, void* z, int B, int H,
                             int W, void* stream) {
    CV_CHECK(net && z, "null pointer argument");
    CV_CHECK(B > 0 && H > 0 && W > 0, "empty tensor");
    CV_CHECK(dtype == CV_C64 || dtype == CV_C128, "unknown dtype %d", dtype);
    const int64_t n = (int64_t)H * W;
    hipStream_t s = as_stream(stream);
    if (dtype == CV_C64)
        k_unpack<float><<<grid_of(B * n), kBlock, 0, s>>>((const uint32_t*)net, oscale, os, (cx<float>*)z, B, n);
    else
        k_unpack<double><<<grid_of(B * n), kBlock, 0, s>>>((const uint32_t*)net, oscale, os, (cx<double>*)z, B, n);
    CV_LAUNCH_CHECK();
    return CV_OK;
}

extern "C" int cv_cnn_loss(const void* net, const void* e, const int* e_idx, const float* inv_norm, void* g,
                           double* loss, float gscale, const void* extra, float extra_scale, int B, int H, int W,
                           void* stream) {
    CV_CHECK(net && e && inv_norm && g && loss, "null pointer argument");
    CV_CHECK(B > 0 && H > 0 && W > 0, "empty tensor");
    const int64_t n = (int64_t)H * W;
    k_loss<<<grid_of(B * n), kBlock, 0, as_stream(stream)>>>((const uint32_t*)net, (const cx<float>*)e, e_idx,
                                                             inv_norm, (uint4*)g, loss, gscale,
                                                             (const cx<float>*)extra, extra_scale, B, n);
    CV_LAUNCH_CHECK();
    return CV_OK;
}

extern "C" int cv_fdfd_batch_apply(const cv_fdfd_desc* desc, const cv_fdfd_desc* desc_host, int n_inst,
                                   const void* prof, const void* m0, const void* m1, const int* inst, const void* x,
                                   const void* r, const int* r_idx, void* y, double* loss, double lscale, int herm,
                                   int B, void* stream) {
    CV_CHECK(desc && desc_host && prof && m0 && inst && x && y, "null pointer argument");
    CV_CHECK(n_inst > 0 && B > 0, "empty batch");
    const int64_t n = (int64_t)desc_host[0].nx * desc_host[0].ny;
    for (int k = 0; k < n_inst; ++k) {
        CV_CHECK((int64_t)desc_host[k].nx * desc_host[k].ny == n, "instances must have equal sizes");
        CV_CHECK(desc_host[k].pol == CV_POL_EZ || m1, "Hz operator needs m1");
        CV_CHECK(desc_host[k].n_cpl == 0, "frequency-coupled systems are not supported");
    }
    k_batch_apply<<<grid_of(B * n), kBlock, 0, as_stream(stream)>>>(
        desc, (const cx<float>*)prof, (const float*)m0, (const float*)m1, inst, (const cx<float>*)x,
        (const cx<float>*)r, r_idx, (cx<float>*)y, loss, lscale, herm, B, n);
    CV_LAUNCH_CHECK();
    return CV_OK;
}

// ------------------------------------------------------------- the network
// Layers (2L + 1 for L levels): 0 = input conv, 1 .. L-1 = down convs of
// levels 0 .. L-2, L = bottom conv, L+1 .. 2L-1 = up convs of levels L-2 .. 0,
// 2L = output conv.  Forward pass:
//   e0 = relu(conv_in(x0))
//   h_l = relu(conv_d_l(e_l)), e_{l+1} = down2(h_l) / 4        l = 0 .. L-2
//   b_{L-1} = relu(conv_bottom(e_{L-1}))
//   u_l = up2(b_{l+1}) + h_l, b_l = relu(conv_u_l(u_l))         l = L-2 .. 0
//   y = conv_out(b_0)                      (channels 0, 1 carry the output)
extern "C" int64_t cv_unet_workspace_bytes(int levels, int B, int H, int W) {
    if (levels < 1 || levels > CV_UNET_MAX_LEVELS || B <= 0 || H <= 0 || W <= 0) return 0;
    int h[CV_UNET_MAX_LEVELS], w[CV_UNET_MAX_LEVELS];
    level_dims(H, W, levels, h, w);
    int64_t total = 0;
    for (int l = 0; l < levels; ++l) total += 3 * act_bytes(B, h[l], w[l]);
    return total;
}

extern "C" int cv_unet_apply(const cv_unet* net, int dtype, const void* r, float rs, float os, void* z, int H, int W,
                             void* workspace, int64_t workspace_bytes, void* stream) {
    CV_CHECK(net && net->wpack && net->bias && r && z && workspace, "null pointer argument");
    const int L = net->levels;
    CV_CHECK(L >= 1 && L <= CV_UNET_MAX_LEVELS, "levels must be in [1, %d]", CV_UNET_MAX_LEVELS);
    CV_CHECK(net->nf >= 0 && net->nf <= kC - 2, "nf must be in [0, %d]", kC - 2);
    CV_CHECK(net->nf == 0 || net->feat, "features missing");
    const int64_t need = cv_unet_workspace_bytes(L, 1, H, W);
    CV_CHECK(need > 0 && workspace_bytes >= need, "workspace too small: %lld < %lld", (long long)workspace_bytes,
             (long long)need);
    hipStream_t s = as_stream(stream);
    int h[CV_UNET_MAX_LEVELS], w[CV_UNET_MAX_LEVELS];
    level_dims(H, W, L, h, w);
    char* base = (char*)workspace;
    void* A[CV_UNET_MAX_LEVELS];
    void* Bf[CV_UNET_MAX_LEVELS];
    void* S[CV_UNET_MAX_LEVELS];
    for (int l = 0; l < L; ++l) {
        const int64_t nb = act_bytes(1, h[l], w[l]);
        A[l] = base;
        Bf[l] = base + nb;
        S[l] = base + 2 * nb;
        base += 3 * nb;
    }
    const uint16_t* wp = (const uint16_t*)net->wpack;
    auto W_ = [&](int layer) { return (const void*)(wp + (size_t)layer * kWPack); };
    auto b_ = [&](int layer) { return net->bias + (size_t)layer * kC; };
    int rc = cv_cnn_pack_input(dtype, r, nullptr, nullptr, rs, net->feat, nullptr, net->nf, A[0], 1, H, W, stream);
    if (rc != CV_OK) return rc;
    if ((rc = conv_launch(A[0], nullptr, W_(0), b_(0), Bf[0], 1, h[0], w[0], 1, 0, s)) != CV_OK) return rc;
    for (int l = 0; l + 1 < L; ++l) {
        if ((rc = conv_launch(Bf[l], nullptr, W_(1 + l), b_(1 + l), S[l], 1, h[l], w[l], 1, 0, s)) != CV_OK)
            return rc;
        if ((rc = cv_cnn_down2(S[l], Bf[l + 1], 1, h[l], w[l], 0.25f, stream)) != CV_OK) return rc;
    }
    if ((rc = conv_launch(Bf[L - 1], nullptr, W_(L), b_(L), A[L - 1], 1, h[L - 1], w[L - 1], 1, 0, s)) != CV_OK)
        return rc;
    for (int l = L - 2; l >= 0; --l) {
        if ((rc = cv_cnn_up2(A[l + 1], S[l], Bf[l], 1, h[l], w[l], 1.0f, stream)) != CV_OK) return rc;
        const int layer = L + 1 + (L - 2 - l);
        if ((rc = conv_launch(Bf[l], nullptr, W_(layer), b_(layer), A[l], 1, h[l], w[l], 1, 0, s)) != CV_OK)
            return rc;
    }
    if ((rc = conv_launch(A[0], nullptr, W_(2 * L), b_(2 * L), Bf[0], 1, h[0], w[0], 0, 0, s)) != CV_OK) return rc;
    return cv_cnn_unpack(dtype, Bf[0], nullptr, os, z, 1, H, W, stream);
}
