// Shared device/host helpers for the ceviche_amd HIP library (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>
#include <math.h>

#include "ceviche_amd.h"

namespace cv {

// ------------------------------------------------------------------ errors
void set_error(const char* fmt, ...);

#define CV_HIP(expr)                                                          \
    do {                                                                      \
        hipError_t _e = (expr);                                               \
        if (_e != hipSuccess) {                                               \
            ::cv::set_error("%s:%d %s: %s", __FILE__, __LINE__, #expr,       \
                            hipGetErrorString(_e));                           \
            return CV_ERR_HIP;                                                \
        }                                                                     \
    } while (0)

#define CV_CHECK(cond, ...)                                                   \
    do {                                                                      \
        if (!(cond)) {                                                        \
            ::cv::set_error(__VA_ARGS__);                                     \
            return CV_ERR_INVALID;                                            \
        }                                                                     \
    } while (0)

#define CV_LAUNCH_CHECK()                                                     \
    do {                                                                      \
        hipError_t _e = hipGetLastError();                                    \
        if (_e != hipSuccess) {                                               \
            ::cv::set_error("%s:%d kernel launch: %s", __FILE__, __LINE__,   \
                            hipGetErrorString(_e));                           \
            return CV_ERR_HIP;                                                \
        }                                                                     \
    } while (0)

// ----------------------------------------------------------------- complex
template <typename R>
struct alignas(2 * sizeof(R)) cx {
    R re, im;
};

template <typename R>
__host__ __device__ __forceinline__ cx<R> mk(R re, R im) {
    cx<R> z;
    z.re = re;
    z.im = im;
    return z;
}
template <typename R>
__host__ __device__ __forceinline__ cx<R> operator+(cx<R> a, cx<R> b) { return mk(a.re + b.re, a.im + b.im); }
template <typename R>
__host__ __device__ __forceinline__ cx<R> operator-(cx<R> a, cx<R> b) { return mk(a.re - b.re, a.im - b.im); }
template <typename R>
__host__ __device__ __forceinline__ cx<R> operator-(cx<R> a) { return mk(-a.re, -a.im); }
template <typename R>
__host__ __device__ __forceinline__ cx<R> operator*(cx<R> a, cx<R> b) {
    return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <typename R>
__host__ __device__ __forceinline__ cx<R> operator*(R s, cx<R> a) { return mk(s * a.re, s * a.im); }
template <typename R>
__host__ __device__ __forceinline__ cx<R> conj(cx<R> a) { return mk(a.re, -a.im); }
template <typename R>
__host__ __device__ __forceinline__ R norm2(cx<R> a) { return a.re * a.re + a.im * a.im; }
template <typename R>
__host__ __device__ __forceinline__ cx<R> czero() { return mk<R>(R(0), R(0)); }
template <typename R>
__host__ __device__ __forceinline__ cx<R> cone() { return mk<R>(R(1), R(0)); }

template <typename R, typename S>
__host__ __device__ __forceinline__ cx<R> ccast(cx<S> a) { return mk<R>(R(a.re), R(a.im)); }

// complex division in double precision (scalars only)
__host__ __device__ __forceinline__ cx<double> cdiv(cx<double> a, cx<double> b) {
    double d = b.re * b.re + b.im * b.im;
    return mk((a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d);
}

__host__ __device__ __forceinline__ bool cfinite(cx<double> a) { return isfinite(a.re) && isfinite(a.im); }

// ------------------------------------------------------------ reductions
constexpr int kBlock = 256;           // threads per workgroup (4 wave64)
constexpr int kWaves = kBlock / 64;

// Sum NV doubles over the workgroup; the result is valid in thread 0 and the
// summation order is fixed (deterministic).  `lds` holds kWaves*NV doubles.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* lds) {
#pragma unroll
    for (int c = 0; c < NV; ++c) {
        double s = v[c];
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
        v[c] = s;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
#pragma unroll
        for (int c = 0; c < NV; ++c) lds[wave * NV + c] = v[c];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < NV; ++c) {
            double s = 0.0;
            for (int w = 0; w < kWaves; ++w) s += lds[w * NV + c];
            v[c] = s;
        }
    }
    __syncthreads();
}

__device__ __forceinline__ void store_agent(double* p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double load_agent(const double* p) {
    return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Last-workgroup protocol over the tiles of one instance.  Every workgroup
// of the instance calls this once with its block-summed values (valid in
// thread 0); it publishes them to partials[tile*NV..] and returns true in
// the workgroup that arrived last, after that workgroup has summed all the
// instance's partials (fixed order) into `tot` (valid in all threads).
//
// Only the partials and the counter cross workgroups inside the kernel, so
// they are agent-scope atomics (coherent across the XCDs' L2s without a full
// L2 write-back / invalidate); s_waitcnt(0) makes the partial stores complete
// before the counter increment is issued.  The next kernel sees everything
// else through the usual kernel-boundary cache maintenance.
template <int NV>
__device__ __forceinline__ bool last_block_reduce(const double (&mine)[NV], double* partials,
                                                  unsigned int* counter, int tile, int tile_first,
                                                  int tile_count, double* lds, double (&tot)[NV]) {
    __shared__ int s_last;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < NV; ++c) store_agent(&partials[(size_t)tile * NV + c], mine[c]);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_s_waitcnt(0);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        const unsigned int ticket =
            __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_last = (ticket == (unsigned int)(tile_count - 1));
    }
    __syncthreads();
    if (!s_last) return false;
    double acc[NV];
#pragma unroll
    for (int c = 0; c < NV; ++c) acc[c] = 0.0;
    // 4 independent loads in flight per thread (fixed summation order)
    for (int t0 = 0; t0 < tile_count; t0 += 4 * kBlock) {
        double v[4][NV];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int t = t0 + u * kBlock + (int)threadIdx.x;
#pragma unroll
            for (int c = 0; c < NV; ++c)
                v[u][c] = t < tile_count ? load_agent(&partials[(size_t)(tile_first + t) * NV + c]) : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int c = 0; c < NV; ++c) acc[c] += v[u][c];
    }
    block_sum<NV>(acc, lds);
    __shared__ double s_tot[NV];
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < NV; ++c) s_tot[c] = acc[c];
        __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);  // ready for the next kernel
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < NV; ++c) tot[c] = s_tot[c];
    return true;
}

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace cv
