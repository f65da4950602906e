// k_rowfft.hip -- K1: one workgroup per (two slice pairs, angle).
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
// is generated while loading, so padded rows never exist in memory.  With
// detector counts as input (PRE != 0) the steps the callers run on the host
// before gridrec happen here too, per value: flat/dark normalisation
// (PRE 1: 1 - (p - d)/(w1/2 + w2/2 - d), tomo_recon.cpp:1877; PRE 2: the log
// form log((w1/2 + w2/2 - d)/(p - d)), 0 where undefined,
// find_shift.cpp:1685-1697, in float) and the sub-pixel shift by xoffset with
// linear interpolation and edge extension (tomo_recon.cpp:3708-3780).  The row
// is transformed in LDS (gr_fft.hpp) and the filtered split spectra are written
// as two rows of pdim/2 complex values (entry 0 unused, written as zero).
#include "gr_fft.hpp"
#include "gr_internal.h"

#pragma clang fp contract(off)

namespace {

struct RowFftArgs {
    const float *sino;
    long slice_stride, row_stride;
    GrPre pre;
    int n_slices, pair0, n_ang, pairs;
    int n_in, n_eff, pad_off, pad;
    int pdim, half, flag, radix;
    const float2 *filphase;
    const float *roi_offset;
    const float2 *tw;
    float2 *spec1, *spec2;
};

// Normalised value of detector cell q of one row of counts.
template <int PRE>
__device__ __forceinline__ float normalised(const unsigned short *raw, const float *dark, const float *w1,
                                            const float *w2, int q) {
    const float pj = (float)raw[q], d = dark[q], a = w1[q], b = w2[q];
    if (PRE == 1) return 1.0f - 1.0f * (pj - d) / (0.5f * a + 0.5f * b - d);
    const float dv = pj - d;
    if (dv <= 0.0f) return 0.0f;
    const float r = (0.5f * (a + b) - d) / dv;
    return r > 0.0f ? logf(r) : 0.0f;
}

// Row value at cell k after the shift: linear interpolation at k - xoffset,
// cells beyond either end taking the edge value (tomo_recon.cpp:3713-3778,
// find_shift.cpp:1737-1778); get(q) is the unshifted value of cell q.
template <typename Get>
__device__ __forceinline__ float shift_interp(Get get, int n, float xoffset, int k) {
    const float kk = (float)k - xoffset;
    const int nkk = (int)floorf(kk);
    const int q0 = nkk < 0 ? 0 : (nkk > n - 1 ? n - 1 : nkk);
    const int q1 = nkk + 1 < 0 ? 0 : (nkk + 1 > n - 1 ? n - 1 : nkk + 1);
    float v = 0.0f;
    v += get(q0) * ((float)(nkk + 1) - kk);
    v += get(q1) * (kk - (float)nkk);
    return v;
}

template <int PRE>
__device__ __forceinline__ float shifted(const unsigned short *raw, const float *dark, const float *w1,
                                         const float *w2, int n, float xoffset, int k) {
    return shift_interp([&](int q) { return normalised<PRE>(raw, dark, w1, w2, q); }, n, xoffset, k);
}

// A workgroup transforms CW pairs of one angle side by side (CW groups of
// nthr = blockDim / CW threads, each with its own LDS row; the workgroup's
// barriers are shared, so the groups run the same passes in lockstep) and
// writes the CW neighbouring slots of each interleaved record with one store.
// Workgroup b takes angle (b / 8 / G) * 8 + b % 8 and pair group (b / 8) % G,
// G = ceil(pairs / CW): the pair groups of one angle run back to back on one
// XCD (workgroup b runs on XCD b mod 8), where their stores into the
// interleaved spectra meet in that XCD's L2.
template <int MAXB, int PRE, int CW>
__global__ void __launch_bounds__(1024) k_rowfft(RowFftArgs A) {
    extern __shared__ float2 lds[];
    const int G = (A.pairs + CW - 1) / CW;
    const int rb = blockIdx.x >> 3;
    const int pg = rb % G;
    const int n = (rb / G) * 8 + (blockIdx.x & 7);
    if (n >= A.n_ang) return;
    const int nthr = (int)blockDim.x / CW;
    const int grp = (int)threadIdx.x / nthr, tid = (int)threadIdx.x % nthr;
    const int p = pg * CW + grp;  // this group's pair
    const bool valid = p < A.pairs;
    float2 *row = lds + (long)grp * A.pdim;
    const long s1 = 2L * (A.pair0 + p), s2 = s1 + 1;
    const bool has2 = s2 < A.n_slices;

    // loops unrolled to the per-thread bound (pdim <= 4 MAXB nthr), every
    // load issued before the first LDS store: one memory latency per row
    // instead of one per element
    constexpr int EPT = 4 * MAXB;
    float2 v[EPT];
#pragma unroll
    for (int r = 0; r < EPT; r++) {
        const int j = tid + r * nthr;
        float re = 0.0f, im = 0.0f;
        if (valid && j < A.pdim && j < A.n_eff) {
            int q = j;
            if (A.pad) {
                q = j - A.pad_off;
                q = q < 0 ? 0 : (q > A.n_in - 1 ? A.n_in - 1 : q);
            }
            if (PRE == 0) {
                re = A.sino[s1 * A.slice_stride + (long)n * A.row_stride + q];
                if (has2) im = A.sino[s2 * A.slice_stride + (long)n * A.row_stride + q];
            } else if (PRE == 3) {
                // every pair is the same two slices under its own shift
                const float xo = A.pre.pair_xoffset[A.pair0 + p];
                const float *f1 = A.sino + (long)n * A.row_stride;
                const float *f2 = f1 + A.slice_stride;
                re = shift_interp([&](int c) { return f1[c]; }, A.n_in, xo, q);
                im = shift_interp([&](int c) { return f2[c]; }, A.n_in, xo, q);
            } else {
                const GrPre &P = A.pre;
                const long f1 = s1 * P.field_stride, f2 = s2 * P.field_stride;
                re = shifted<PRE>(P.raw + s1 * P.raw_slice_stride + (long)n * P.raw_row_stride, P.dark + f1,
                                  P.white1 + f1, P.white2 + f1, A.n_in, P.xoffset, q);
                if (has2)
                    im = shifted<PRE>(P.raw + s2 * P.raw_slice_stride + (long)n * P.raw_row_stride, P.dark + f2,
                                      P.white1 + f2, P.white2 + f2, A.n_in, P.xoffset, q);
            }
        }
        v[r] = make_float2(re, im);
    }
#pragma unroll
    for (int r = 0; r < EPT; r++) {
        const int j = tid + r * nthr;
        if (j < A.pdim) row[gr_lds_idx(j)] = v[r];
    }
    __syncthreads();
    gr_lds_fft<+1, MAXB>(row, A.pdim, A.tw, tid, nthr, A.radix);

    // epilogue: every thread of the workgroup takes frequencies j and writes
    // the CW pairs' slots of record j (CW * 8 contiguous bytes; slots of pairs
    // past the launch's last, whose rows were zeros, get zeros)
    const int p0 = pg * CW;
    const long rec = ((long)(p0 / kSpecGroup) * A.n_ang + n) * A.half * kSpecGroup + p0 % kSpecGroup;
    float2 *o1 = A.spec1 + rec;
    float2 *o2 = A.spec2 + rec;
    const float offset = A.roi_offset[n];
    const int nall = (int)blockDim.x;
    constexpr int EPH = (2 * MAXB + CW - 1) / CW;  // half = pdim/2 <= 2 MAXB nthr
    float2 fp[EPH];
#pragma unroll
    for (int r = 0; r < EPH; r++) {
        const int j = (int)threadIdx.x + r * nall;
        if (j < A.half) fp[r] = A.filphase[j];
    }
#pragma unroll
    for (int r = 0; r < EPH; r++) {
        const int j = (int)threadIdx.x + r * nall;
        if (j >= A.half) continue;
        float2 c1[CW], c2[CW];
        if (j == 0) {
#pragma unroll
            for (int g = 0; g < CW; g++) c1[g] = c2[g] = make_float2(0.0f, 0.0f);
        } else {
            float2 ct = fp[r];
            if (A.flag) {  // grid.cpp:267-269
                const float arg = (float)j * offset;
                const float phr = cosf(arg), phi = -sinf(arg);
                ct = make_float2(ct.x * phr - ct.y * phi, ct.x * phi + ct.y * phr);
            }
            const float ci = -ct.y;
#pragma unroll
            for (int g = 0; g < CW; g++) {
                const float2 P = lds[(long)g * A.pdim + gr_lds_idx(j)];
                const float2 Q = lds[(long)g * A.pdim + gr_lds_idx(A.pdim - j)];
                c1[g] = make_float2(ct.x * P.x - ct.y * P.y, ct.x * P.y + ct.y * P.x);
                c2[g] = make_float2(ct.x * Q.x - ci * Q.y, ct.x * Q.y + ci * Q.x);
            }
        }
        const long o = (long)j * kSpecGroup;
        if (CW == 1) {
            o1[o] = c1[0];
            o2[o] = c2[0];
        } else {
#pragma unroll
            for (int g = 0; g < CW; g += 2) {
                reinterpret_cast<float4 *>(o1 + o + g)[0] = make_float4(c1[g].x, c1[g].y, c1[g + 1].x, c1[g + 1].y);
                reinterpret_cast<float4 *>(o2 + o + g)[0] = make_float4(c2[g].x, c2[g].y, c2[g + 1].x, c2[g + 1].y);
            }
        }
    }
}

}  // namespace

// Pairs per workgroup: kRowPairs side by side when they fit a workgroup
// (<= 1024 threads) and the CU's LDS (160 KiB) and the launch has that many
// pairs, else fewer (a power of two that divides the 8 slots of a record).
// At pdim = 4096 on MI355X, per launch of 8 pairs: 0.33 ms with one pair per
// workgroup, 0.20 with two, 0.27 with four.
constexpr int kRowPairs = 2;

static int rowfft_pairs_per_block(long pdim, int pairs) {
    int cw = kRowPairs;
    while (cw > 1 && (cw > pairs || (long)cw * gr_fft_threads(pdim) > 1024 || (long)cw * pdim * 8 > 163840)) cw /= 2;
    return cw;
}

hipError_t gr_launch_rowfft(const GrGeom &g, const GrDevTables &t, int n_ang, const float *sino,
                            long slice_stride, long row_stride, const GrPre *pre, int n_slices, int pair0,
                            int pairs, float2 *spec1, float2 *spec2, hipStream_t s) {
    RowFftArgs A;
    A.sino = sino;
    A.pre = pre ? *pre : GrPre{0, nullptr, 0, 0, nullptr, nullptr, nullptr, 0, 0.0f, nullptr};
    const int mode = pre ? pre->mode : 0;
    A.slice_stride = slice_stride;
    A.row_stride = row_stride;
    A.n_slices = n_slices;
    A.pair0 = pair0;
    A.n_ang = n_ang;
    A.pairs = pairs;
    A.n_in = g.n_det_in;
    A.n_eff = g.n_det_eff;
    A.pad_off = g.n_det_in / 2;
    A.pad = g.pad == GR_PAD_HALF_EDGE;
    A.pdim = (int)g.pdim;
    A.half = (int)g.half;
    A.flag = g.flag;
    A.radix = g.fft_radix;
    A.filphase = t.filphase;
    A.roi_offset = t.roi_offset;
    A.tw = t.tw_row;
    A.spec1 = spec1;
    A.spec2 = spec2;
    const int nthr = gr_fft_threads(g.pdim);
    const long per = (g.pdim / 4 + nthr - 1) / nthr;
    const int cw = rowfft_pairs_per_block(g.pdim, pairs);
    const dim3 grid((unsigned)((n_ang + 7) / 8 * 8 * ((pairs + cw - 1) / cw))), block(cw * nthr);
    const size_t lds = sizeof(float2) * cw * gr_lds_len(g.pdim);
#define GR_LAUNCH_C(MB, PR, CW)                                              \
    do {                                                                     \
        hipError_t e_ = gr_allow_lds(k_rowfft<MB, PR, CW>, lds);             \
        if (e_ != hipSuccess) return e_;                                     \
        hipLaunchKernelGGL((k_rowfft<MB, PR, CW>), grid, block, lds, s, A);  \
    } while (0)
#define GR_LAUNCH_P(MB, PR)                                                  \
    do {                                                                     \
        if (cw == 4) GR_LAUNCH_C(MB, PR, 4);                                 \
        else if (cw == 2) GR_LAUNCH_C(MB, PR, 2);                            \
        else GR_LAUNCH_C(MB, PR, 1);                                         \
    } while (0)
#define GR_LAUNCH(MB)                                                   \
    do {                                                                \
        if (mode == GR_PRE_FLATDARK) GR_LAUNCH_P(MB, 1);                \
        else if (mode == GR_PRE_FLATDARK_LOG) GR_LAUNCH_P(MB, 2);       \
        else if (mode == 3) GR_LAUNCH_P(MB, 3);                         \
        else GR_LAUNCH_P(MB, 0);                                        \
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
#undef GR_LAUNCH_C
#undef GR_LAUNCH
#undef GR_LAUNCH_P
    return hipGetLastError();
}
