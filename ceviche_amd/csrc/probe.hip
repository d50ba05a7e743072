// Weighted probe sums over field snapshots: the reduction inside ceviche's
// measure_fields (ceviche/utils.py:316-332, sum(fields[component] * probe)
// per probe and step) with the probe weights in CSR form, so a line or mode
// probe touches only its support.  Accumulates in FP64.
#include "common.h"

#include <algorithm>

namespace cv {

constexpr int kProbeBlock = 256;

// out[t * out_stride + p] (+)= sum_{k in [rowptr[p], rowptr[p+1])} w[k] * field[t * f_stride + col[k]]
template <typename R>
__global__ void __launch_bounds__(kProbeBlock)
k_probe_weights(const R* __restrict__ field, int64_t f_stride, int T, const int64_t* __restrict__ rowptr,
                const int64_t* __restrict__ col, const double* __restrict__ w, int n_rows, R* __restrict__ out,
                int64_t out_stride, int accumulate) {
    __shared__ double lds[kProbeBlock / 64];
    const int p = blockIdx.x;
    const int64_t k0 = rowptr[p], k1 = rowptr[p + 1];
    for (int t = blockIdx.y; t < T; t += gridDim.y) {
        const R* f = field + (int64_t)t * f_stride;
        double acc = 0.0;
        for (int64_t k = k0 + threadIdx.x; k < k1; k += kProbeBlock) acc += w[k] * (double)f[col[k]];
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
        if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = acc;
        __syncthreads();
        if (threadIdx.x == 0) {
            double s = 0.0;
            for (int q = 0; q < kProbeBlock / 64; ++q) s += lds[q];
            R* o = out + (int64_t)t * out_stride + p;
            *o = accumulate ? R((double)*o + s) : R(s);
        }
        __syncthreads();
    }
}

}  // namespace cv

using namespace cv;

extern "C" int cv_probe_weights(int dtype, const void* field, int64_t field_stride, int T, const int64_t* rowptr,
                                const int64_t* col, const double* w, int n_rows, void* out, int64_t out_stride,
                                int accumulate, void* stream) {
    CV_CHECK(T >= 0 && n_rows >= 0, "negative sizes");
    if (T == 0 || n_rows == 0) return CV_OK;
    CV_CHECK(field && rowptr && col && w && out, "null pointer argument");
    CV_CHECK(field_stride >= 0 && out_stride >= n_rows, "bad strides");
    hipStream_t s = as_stream(stream);
    dim3 grid((unsigned)n_rows, (unsigned)std::min(T, 65535));
    if (dtype == CV_F64)
        k_probe_weights<double><<<grid, kProbeBlock, 0, s>>>((const double*)field, field_stride, T, rowptr, col, w,
                                                             n_rows, (double*)out, out_stride, accumulate);
    else if (dtype == CV_F32)
        k_probe_weights<float><<<grid, kProbeBlock, 0, s>>>((const float*)field, field_stride, T, rowptr, col, w,
                                                            n_rows, (float*)out, out_stride, accumulate);
    else
        CV_CHECK(false, "unknown dtype %d", dtype);
    CV_LAUNCH_CHECK();
    return CV_OK;
}
