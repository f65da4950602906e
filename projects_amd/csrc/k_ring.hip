// k_ring.hip -- ring-artefact correction of sinograms in place, the
// preprocessing step tomo_recon runs between flat/dark normalisation and
// gridrec (RingCorrectionSinogram, tomo_recon.cpp:3905-3972).  Per slice and
// iteration:
//   m_i   = mean of row (angle) i,  mt = mean of the m_i
//   d_ij *= mt / m_i                          (rows with m_i == 0 untouched)
//   c_j   = mean over i of column j,  l_j = (c_{j-1} + c_j + c_{j+1}) / 3
//           (l_0 = c_0, l_{n-1} = c_{n-1})
//   d_ij  = d_ij - coeff (c_j - l_j)  if that is > 0, else 0
// Means are sums in a fixed order (per-thread strided partials combined by a
// tree), so results are repeatable; they differ from the reference's
// sequential float sums only by rounding.
#include "gr_internal.h"

#pragma clang fp contract(off)

namespace {

struct RingArgs {
    float *sino;
    long slice_stride, row_stride;
    int n_ang, n_det;
    float coeff;
    float *row_mean;  // [n_slices][n_ang]
    float *col_mean;  // [n_slices][n_det]
};

__device__ __forceinline__ float block_sum(float v, float *red) {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    const float s = red[0];
    __syncthreads();
    return s;
}

// Row means of one slice per block row: grid (n_ang, n_slices).
__global__ void __launch_bounds__(256) k_row_means(RingArgs A) {
    __shared__ float red[256];
    const int i = blockIdx.x, s = blockIdx.y;
    const float *row = A.sino + s * A.slice_stride + (long)i * A.row_stride;
    float v = 0.0f;
    for (int j = threadIdx.x; j < A.n_det; j += blockDim.x) v += row[j];
    const float sum = block_sum(v, red);
    if (threadIdx.x == 0) A.row_mean[(long)s * A.n_ang + i] = sum / (float)A.n_det;
}

// Column means of the row-renormalised data: grid (ceil(n_det/256), n_slices),
// one column per thread, rows in order; the global mean is recomputed per
// block from the row means (same order in every block).
__global__ void __launch_bounds__(256) k_col_means(RingArgs A) {
    __shared__ float red[256];
    const int s = blockIdx.y;
    const float *rm = A.row_mean + (long)s * A.n_ang;
    float v = 0.0f;
    for (int i = threadIdx.x; i < A.n_ang; i += blockDim.x) v += rm[i];
    const float mt = block_sum(v, red) / (float)A.n_ang;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= A.n_det) return;
    const float *col = A.sino + s * A.slice_stride + j;
    float c = 0.0f;
    for (int i = 0; i < A.n_ang; i++) {
        const float m = rm[i];
        const float d = col[(long)i * A.row_stride];
        c += m != 0.0f ? d * mt / m : d;
    }
    A.col_mean[(long)s * A.n_det + j] = c / (float)A.n_ang;
}

// Renormalise and subtract the ring estimate: grid (n_ang, n_slices).
__global__ void __launch_bounds__(256) k_ring_apply(RingArgs A) {
    __shared__ float red[256];
    const int i = blockIdx.x, s = blockIdx.y;
    const float *rm = A.row_mean + (long)s * A.n_ang;
    float v = 0.0f;
    for (int k = threadIdx.x; k < A.n_ang; k += blockDim.x) v += rm[k];
    const float mt = block_sum(v, red) / (float)A.n_ang;
    const float m = rm[i];
    const float *cm = A.col_mean + (long)s * A.n_det;
    float *row = A.sino + s * A.slice_stride + (long)i * A.row_stride;
    const int n = A.n_det;
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        float d = row[j];
        if (m != 0.0f) d = d * mt / m;
        const float l = (j == 0 || j == n - 1) ? cm[j] : (cm[j - 1] + cm[j] + cm[j + 1]) / 3.0f;
        const float t = cm[j] - l;
        row[j] = (d - t * A.coeff) > 0.0f ? d - t * A.coeff : 0.0f;
    }
}

}  // namespace

hipError_t gr_launch_ring(float *sino, long slice_stride, long row_stride, int n_slices, int n_ang, int n_det,
                          float coeff, int iterations, float *row_mean, float *col_mean, hipStream_t s) {
    RingArgs A{sino, slice_stride, row_stride, n_ang, n_det, coeff, row_mean, col_mean};
    for (int it = 0; it < iterations; it++) {
        hipLaunchKernelGGL(k_row_means, dim3(n_ang, n_slices), dim3(256), 0, s, A);
        hipLaunchKernelGGL(k_col_means, dim3((n_det + 255) / 256, n_slices), dim3(256), 0, s, A);
        hipLaunchKernelGGL(k_ring_apply, dim3(n_ang, n_slices), dim3(256), 0, s, A);
    }
    return hipGetLastError();
}
