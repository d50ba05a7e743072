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

// 1 / z in the operand's precision
template <typename R>
__host__ __device__ __forceinline__ cx<R> crecip(cx<R> z) {
    const R d = z.re * z.re + z.im * z.im;
    return mk<R>(z.re / d, -z.im / d);
}

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

// Last-workgroup protocol over the tiles of one instance, two levels deep.
// Every workgroup of the instance calls this once with its block-summed
// values (valid in thread 0).  Level 1: the instance's tiles form groups of
// kGroup consecutive tiles; each workgroup publishes its values to
// partials[tile*NV..] and increments its group's counter (gcount at the
// group's first tile index); the workgroup that arrives last in the group
// sums the group's partials (fixed order) into gpartials[group first tile*NV..]
// and increments the instance counter (counter).  Level 2: the workgroup that
// arrives last there sums the group partials (fixed order) into `tot` (valid
// in all threads) and returns true; every other workgroup returns false.
//
// One counter per workgroup of a large instance would serialise thousands of
// atomics on one address (about 12 ns each on MI355X, 48 µs for 4096
// workgroups); with groups of 64 the atomics spread over many addresses.
// Only the partials and the counters cross workgroups inside the kernel, so
// they are agent-scope atomics (coherent across the XCDs' L2s without a full
// L2 write-back / invalidate); s_waitcnt(0) makes the partial stores complete
// before the counter increment is issued.  The next kernel sees everything
// else through the usual kernel-boundary cache maintenance.  The last
// workgroups reset the counters they used, so consecutive kernels can share
// them; they must be zero before the first kernel.
constexpr int kGroup = 64;

template <int NV>
__device__ __forceinline__ bool last_block_reduce(const double (&mine)[NV], double* partials, double* gpartials,
                                                  unsigned int* counter, unsigned int* gcount, int tile,
                                                  int tile_first, int tile_count, double* lds, double (&tot)[NV]) {
    __shared__ int s_last;
    __shared__ double s_tot[NV];
    const int g = (tile - tile_first) / kGroup;
    const int g_first = tile_first + g * kGroup;
    const int g_size = min(kGroup, tile_count - g * kGroup);
    const int n_groups = (tile_count + kGroup - 1) / kGroup;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < NV; ++c) store_agent(&partials[(size_t)tile * NV + c], mine[c]);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_s_waitcnt(0);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        const unsigned int ticket =
            __hip_atomic_fetch_add(&gcount[g_first], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_last = (ticket == (unsigned int)(g_size - 1));
    }
    __syncthreads();
    if (!s_last) return false;
    double acc[NV];
#pragma unroll
    for (int c = 0; c < NV; ++c)
        acc[c] = ((int)threadIdx.x < g_size) ? load_agent(&partials[(size_t)(g_first + threadIdx.x) * NV + c]) : 0.0;
    block_sum<NV>(acc, lds);
    if (n_groups == 1) {   // the group is the whole instance: acc is the total
        if (threadIdx.x == 0) {
#pragma unroll
            for (int c = 0; c < NV; ++c) s_tot[c] = acc[c];
            __hip_atomic_store(&gcount[g_first], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < NV; ++c) tot[c] = s_tot[c];
        return true;
    }
    if (threadIdx.x == 0) {
        __hip_atomic_store(&gcount[g_first], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
        for (int c = 0; c < NV; ++c) store_agent(&gpartials[(size_t)g_first * NV + c], acc[c]);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_s_waitcnt(0);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        const unsigned int ticket = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_last = (ticket == (unsigned int)(n_groups - 1));
    }
    __syncthreads();
    if (!s_last) return false;
#pragma unroll
    for (int c = 0; c < NV; ++c) acc[c] = 0.0;
    for (int gi = threadIdx.x; gi < n_groups; gi += kBlock) {
#pragma unroll
        for (int c = 0; c < NV; ++c) acc[c] += load_agent(&gpartials[(size_t)(tile_first + gi * kGroup) * NV + c]);
    }
    block_sum<NV>(acc, lds);
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
