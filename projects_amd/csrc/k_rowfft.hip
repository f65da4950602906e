// k_rowfft.hip -- K1: one workgroup per (slice pair, angle).
//
// Replaces phase 1a/1b of do_recon (gridrec/GridrecCPP/Gridrec5.2/grid.cpp:240-274):
//   cproj[j] = (G1[n][j], G2[n][j]) for j < n_det, 0 up to pdim   (grid.cpp:240-252)
//   four1(cproj, pdim, +1)                                          (grid.cpp:254)
//   Cdata1[j] = filphase[j] * cproj[j]                              (grid.cpp:272)
//   Cdata2[j] = conj(filphase[j]) * cproj[pdim-j]  for j = 1 .. pdim/2-1
// (filphase times the ROI phase factor when an ROI offset is set, grid.cpp:265-270).
//
// The two sinogram rows stream from HBM with consecutive lanes on consecutive
// detector cells; the half-padding caller convention (tomo_recon.cpp:3784-3807)
// is generated while loading, so padded rows never exist in memory.  The row
// is transformed in LDS (gr_fft.hpp) and the filtered split spectra are written
// as two rows of pdim/2 complex values (entry 0 unused, written as zero).
#include "gr_fft.hpp"
#include "gr_internal.h"

#pragma clang fp contract(off)

namespace {

struct RowFftArgs {
    const float *sino;
    long slice_stride, row_stride;
    int n_slices, pair0, n_ang;
    int n_in, n_eff, pad_off, pad;
    int pdim, half, flag;
    const float2 *filphase;
    const float *roi_offset;
    const float2 *tw;
    float2 *spec1, *spec2;
};

template <int MAXB>
__global__ void __launch_bounds__(512) k_rowfft(RowFftArgs A) {
    extern __shared__ float2 row[];
    const int n = blockIdx.x;
    const int p = blockIdx.y;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const long s1 = 2L * (A.pair0 + p), s2 = s1 + 1;
    const float *r1 = A.sino + s1 * A.slice_stride + (long)n * A.row_stride;
    const float *r2 = s2 < A.n_slices ? A.sino + s2 * A.slice_stride + (long)n * A.row_stride : nullptr;

    for (int j = tid; j < A.pdim; j += nthr) {
        float re = 0.0f, im = 0.0f;
        if (j < A.n_eff) {
            int q = j;
            if (A.pad) {
                q = j - A.pad_off;
                q = q < 0 ? 0 : (q > A.n_in - 1 ? A.n_in - 1 : q);
            }
            re = r1[q];
            if (r2) im = r2[q];
        }
        row[j] = make_float2(re, im);
    }
    __syncthreads();
    gr_lds_fft<+1, MAXB>(row, A.pdim, A.tw, tid, nthr);

    float2 *o1 = A.spec1 + ((long)p * A.n_ang + n) * A.half;
    float2 *o2 = A.spec2 + ((long)p * A.n_ang + n) * A.half;
    const float offset = A.roi_offset[n];
    for (int j = tid; j < A.half; j += nthr) {
        if (j == 0) {
            o1[0] = make_float2(0.0f, 0.0f);
            o2[0] = make_float2(0.0f, 0.0f);
            continue;
        }
        float2 ct = A.filphase[j];
        if (A.flag) {  // grid.cpp:267-269
            const float arg = (float)j * offset;
            const float phr = cosf(arg), phi = -sinf(arg);
            ct = make_float2(ct.x * phr - ct.y * phi, ct.x * phi + ct.y * phr);
        }
        const float2 P = row[j], Q = row[A.pdim - j];
        o1[j] = make_float2(ct.x * P.x - ct.y * P.y, ct.x * P.y + ct.y * P.x);
        const float ci = -ct.y;
        o2[j] = make_float2(ct.x * Q.x - ci * Q.y, ct.x * Q.y + ci * Q.x);
    }
}

}  // namespace

hipError_t gr_launch_rowfft(const GrGeom &g, const GrDevTables &t, int n_ang, const float *sino,
                            long slice_stride, long row_stride, int n_slices, int pair0, int pairs,
                            float2 *spec1, float2 *spec2, hipStream_t s) {
    RowFftArgs A;
    A.sino = sino;
    A.slice_stride = slice_stride;
    A.row_stride = row_stride;
    A.n_slices = n_slices;
    A.pair0 = pair0;
    A.n_ang = n_ang;
    A.n_in = g.n_det_in;
    A.n_eff = g.n_det_eff;
    A.pad_off = g.n_det_in / 2;
    A.pad = g.pad == GR_PAD_HALF_EDGE;
    A.pdim = (int)g.pdim;
    A.half = (int)g.half;
    A.flag = g.flag;
    A.filphase = t.filphase;
    A.roi_offset = t.roi_offset;
    A.tw = t.tw_row;
    A.spec1 = spec1;
    A.spec2 = spec2;
    const int nthr = gr_fft_threads(g.pdim);
    const long per = (g.pdim / 4 + nthr - 1) / nthr;
    const dim3 grid(n_ang, pairs), block(nthr);
    const size_t lds = sizeof(float2) * g.pdim;
#define GR_LAUNCH(MB)                                                   \
    do {                                                                \
        hipError_t e_ = gr_allow_lds(k_rowfft<MB>, lds);                \
        if (e_ != hipSuccess) return e_;                                \
        hipLaunchKernelGGL(k_rowfft<MB>, grid, block, lds, s, A);       \
    } while (0)
    switch (per) {
    case 0:
    case 1: GR_LAUNCH(1); break;
    case 2: GR_LAUNCH(2); break;
    case 4: GR_LAUNCH(4); break;
    case 8: GR_LAUNCH(8); break;
    case 16: GR_LAUNCH(16); break;
    default: return hipErrorInvalidValue;
    }
#undef GR_LAUNCH
    return hipGetLastError();
}
