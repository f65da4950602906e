// k_entropy.hip -- histograms of reconstructions for the entropy metric of
// centre sweeps (find_shift.cpp:2078-2111): bin (int)((v - min) / inv) with
// inv = (max - min) / n_bins, counted when 0 <= bin <= n_bins, over both images
// of a candidate.  Each block counts into an LDS histogram and adds it to the
// candidate's global one with integer atomics: the counts do not depend on
// the order.
#include <algorithm>

#include "gr_fft.hpp"
#include "gr_internal.h"

#pragma clang fp contract(off)

namespace {

__global__ void __launch_bounds__(256) k_hist(const float *__restrict__ img, long cand_stride, long slice_stride,
                                              long row_stride, int on, float vmin, float inv, int nb1,
                                              unsigned *__restrict__ hist) {
    extern __shared__ unsigned sh[];
    const int c = blockIdx.y;
    for (int b = threadIdx.x; b < nb1; b += blockDim.x) sh[b] = 0u;
    __syncthreads();
    const long per = (long)on * on, npix = 2 * per;
    const float *base = img + (long)c * cand_stride;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < npix; i += (long)gridDim.x * blockDim.x) {
        const long sl = i / per, rem = i - sl * per;
        const long r = rem / on, col = rem - r * on;
        const float x = (base[sl * slice_stride + r * row_stride + col] - vmin) / inv;
        // (int)x counts for 0 <= (int)x <= n_bins, i.e. -1 < x < n_bins + 1
        if (x > -1.0f && x < (float)nb1) atomicAdd(&sh[(int)x], 1u);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nb1; b += blockDim.x)
        if (sh[b]) atomicAdd(&hist[(long)c * nb1 + b], sh[b]);
}

__global__ void k_zero_u32(unsigned *p, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

}  // namespace

hipError_t gr_launch_hist(const float *img, long cand_stride, long slice_stride, long row_stride, int on,
                          int n_cand, float vmin, float vmax, int n_bins, unsigned *hist, hipStream_t s) {
    if (n_bins < 1 || n_bins > 16383 || n_cand < 1 || !(vmax > vmin)) return hipErrorInvalidValue;
    const int nb1 = n_bins + 1;
    const long nh = (long)n_cand * nb1;
    hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)((nh + 255) / 256)), dim3(256), 0, s, hist, nh);
    const float inv = (vmax - vmin) / (float)n_bins;
    const long npix = 2L * on * on;
    const unsigned bx = (unsigned)std::min<long>(256, (npix + 255) / 256);
    const size_t lds = sizeof(unsigned) * nb1;
    hipError_t e = gr_allow_lds(k_hist, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_hist, dim3(bx, n_cand), dim3(256), lds, s, img, cand_stride, slice_stride, row_stride, on,
                       vmin, inv, nb1, hist);
    return hipGetLastError();
}
