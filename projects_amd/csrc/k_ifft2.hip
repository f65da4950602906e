// k_ifft2.hip -- K3 + K4: 2D inverse FFT of the grid, deapodization and crop.
//
// Replaces phases 2 and 3 of do_recon (gridrec/GridrecCPP/Gridrec5.2/grid.cpp:311-411):
//   fourn(H, {M, M}, 2, -1)                      x[p][q] = sum_{u,v} H[u][v] e^{-2 pi i (up+vq)/M}
//   S1[j][k] = winv[j]*winv[k] * Re x[iu(j)][iv(k)]
//   S2[j][k] = winv[j]*winv[k] * Im x[iu(j)][iv(k)]
// where iu(j) = M-M02+j for j < M02 and j-M02 after (the four-corner wrap).
//
// Only the requested window [off, off+n) of rows and columns is produced, so
// both passes are pruned: K3 transforms every row of Ht (one per b, over a)
// but keeps only the n outputs p = iu(j) of the window; K4 transforms those
// n columns down b, keeps the n outputs q = iv(k) and writes the two images
// row by row with the deapodization product fused in.
#include "gr_fft.hpp"
#include "gr_gather.hpp"
#include "gr_internal.h"

#pragma clang fp contract(off)

namespace {

__device__ __forceinline__ long corner_index(long j, long M, long M02) {
    return j < M02 ? M - M02 + j : j - M02;
}

struct RowsArgs {
    const float2 *ht;
    float2 *z;
    const float2 *tw;
    const int2 *row_tiles;  // per tile row the written tiles [lo, hi]; the others are zeros
    int M, radix;
    long M02, out_off, out_n;
};

// Output hook of the row transform: FFT output i lands in window column
// jj = j' - out_off, where corner_index(j') = i (outputs that map to no
// window column are dropped).
struct WindowRow {
    float2 *dst;
    int M, m02, off, out_n;
    static constexpr bool enabled = true;
    __device__ __forceinline__ void operator()(int i, float2 v) const {
        const int jp = i >= M - m02 ? i - (M - m02) : i + m02;
        const int jj = jp - off;
        if (jj >= 0 && jj < out_n) dst[jj] = v;
    }
};

// Loads and LDS stores in loops unrolled to the per-thread bound (M <= 4 MAXB
// nthr), every load issued before the first store: one memory latency per
// row instead of one per element.  CW rows per workgroup, transformed side by
// side by CW thread groups (shared barriers).
template <int MAXB, int CW>
__global__ void __launch_bounds__(1024) k_ifft_rows(RowsArgs A) {
    extern __shared__ float2 lds[];
    constexpr int EPT = 4 * MAXB;
    const int nthr = (int)blockDim.x / CW;
    const int grp = (int)threadIdx.x / nthr, tid = (int)threadIdx.x % nthr;
    const int b = blockIdx.x * CW + grp, p = blockIdx.y;
    const int M = A.M;
    float2 *row = lds + (long)grp * M;
    const float2 *src = A.ht + ((long)p * M + b) * M;
    const int2 lh = A.row_tiles[b >> 3];
    const int lo = lh.x, hi = lh.y;
    float2 v[EPT];
#pragma unroll
    for (int r = 0; r < EPT; r++) {
        const int i = tid + r * nthr, tx = i >> 3;
        v[r] = i < M && tx >= lo && tx <= hi ? src[i] : make_float2(0.0f, 0.0f);
    }
#pragma unroll
    for (int r = 0; r < EPT; r++) {
        const int i = tid + r * nthr;
        if (i < M) row[gr_lds_idx(i)] = v[r];
    }
    __syncthreads();
    float2 *dst = A.z + ((long)p * M + b) * A.out_n;
    const int out_n = (int)A.out_n, off = (int)A.out_off, m02 = (int)A.M02;
    // when the last pass is radix 16, the window's outputs go to HBM straight
    // from its registers: output i is window column j' = corner_index^-1(i)
    if (gr_lds_fft<-1, MAXB>(row, M, A.tw, tid, nthr, A.radix, WindowRow{dst, M, m02, off, out_n})) return;
#pragma unroll
    for (int r = 0; r < EPT; r++) {
        const int jj = tid + r * nthr;
        if (jj < out_n) v[r] = row[gr_lds_idx((int)corner_index(off + jj, M, m02))];
    }
#pragma unroll
    for (int r = 0; r < EPT; r++) {
        const int jj = tid + r * nthr;
        if (jj < out_n) dst[jj] = v[r];
    }
}

struct ColsArgs {
    const float2 *z;
    float *out;
    const float2 *tw;
    const float *winv;
    long out_slice_stride, out_row_stride;
    int M, cw, n_slices, pair0, radix;
    long M02, out_off, out_n;
};

// Output hook of the column transform: FFT output i is image column
// kk = k' - out_off of this row, corner_index(k') = i, scaled by the
// deapodization cu * winv[k'] (grid.cpp:401); rows past the window have
// r1 = nullptr and write nothing.
struct WindowCol {
    float *r1, *r2;
    const float *winv;
    float cu;
    int M, m02, off, out_n;
    static constexpr bool enabled = true;
    __device__ __forceinline__ void operator()(int i, float2 v) const {
        const int kp = i >= M - m02 ? i - (M - m02) : i + m02;
        const int kk = kp - off;
        if (r1 && kk >= 0 && kk < out_n) {
            const float corrn = cu * winv[kp];
            r1[kk] = corrn * v.x;
            if (r2) r2[kk] = corrn * v.y;
        }
    }
};

// Column group of a workgroup.  Workgroups are dealt to the 8 XCDs round
// robin, each XCD with its own L2; numbering the groups so that x, x+8, x+16,
// ... (one XCD) take neighbouring column groups lets the groups that share a
// 128-byte line of Z fetch it through one L2 instead of four.
__device__ __forceinline__ long column_group(unsigned x, unsigned ngroups) {
    if (ngroups % 8u) return x;
    return (long)(x % 8u) * (ngroups / 8u) + x / 8u;
}

// CW > 0: CW neighbouring columns per workgroup, transformed side by side by
// CW groups of nthr/CW threads (the workgroup's barriers are shared, so the
// groups run the same passes in lockstep); the fill reads the CW columns of a
// row from one cache line.  The column gathers -- one line per row and column
// group, strided by a row of Z -- are what bounds this kernel, so wider groups
// fetch fewer lines for the same bytes.  CW = 0: A.cw columns one after
// another on all threads (small M, where CW * M is small).
template <int MAXB, int CW>
__global__ void __launch_bounds__(1024) k_ifft_cols(ColsArgs A) {
    extern __shared__ float2 cols[];  // cw columns of gr_lds_len(M) slots
    const int p = blockIdx.y;
    const int M = A.M, cw = CW ? CW : A.cw;
    const long c0 = column_group(blockIdx.x, gridDim.x) * cw;
    const int nthr = CW ? (int)blockDim.x / CW : (int)blockDim.x;
    const int tid = CW ? (int)threadIdx.x % nthr : (int)threadIdx.x;
    const int grp = CW ? (int)threadIdx.x / nthr : 0;
    const float2 *zp = A.z + (long)p * M * A.out_n;
    const int out_n = (int)A.out_n;
    constexpr int EPT = 4 * MAXB;  // rows per thread and column: M <= 4 MAXB nthr
    if (CW) {
        // unrolled, every load before the first LDS store; the CW columns of
        // a row come from neighbouring lanes
        float2 v[EPT];
#pragma unroll
        for (int r = 0; r < EPT; r++) {
            const int idx = (int)threadIdx.x + r * (int)blockDim.x;
            const int b = idx / CW, c = idx - b * CW;
            v[r] = b < M && c0 + c < out_n ? zp[(long)b * out_n + c0 + c] : make_float2(0.0f, 0.0f);
        }
#pragma unroll
        for (int r = 0; r < EPT; r++) {
            const int idx = (int)threadIdx.x + r * (int)blockDim.x;
            const int b = idx / CW, c = idx - b * CW;
            if (b < M) cols[c * gr_lds_len(M) + gr_lds_idx(b)] = v[r];
        }
    } else if (cw == 1) {
        const float2 *zc = zp + c0;
        for (int b = tid; b < M; b += nthr) cols[gr_lds_idx(b)] = c0 < out_n ? zc[(long)b * out_n] : make_float2(0.0f, 0.0f);
    } else {
        for (int idx = threadIdx.x; idx < cw * M; idx += blockDim.x) {  // CW: cw * M <= 20480; else <= 16384
            const int b = idx / cw, c = idx - b * cw;
            cols[c * gr_lds_len(M) + gr_lds_idx(b)] =
                (c0 + c < out_n) ? zp[(long)b * out_n + c0 + c] : make_float2(0.0f, 0.0f);
        }
    }
    __syncthreads();
    const long s1 = 2L * (A.pair0 + p), s2 = s1 + 1;
    float *o1 = A.out + s1 * A.out_slice_stride;
    float *o2 = s2 < A.n_slices ? A.out + s2 * A.out_slice_stride : nullptr;
    if (CW) {
        // when the last pass is radix 16, each group writes its image row
        // straight from that pass's registers (deapodization on the way)
        const long jj = c0 + grp;
        const bool live = jj < A.out_n;
        WindowCol wc{live ? o1 + jj * A.out_row_stride : nullptr,
                     live && o2 ? o2 + jj * A.out_row_stride : nullptr,
                     A.winv, live ? A.winv[A.out_off + jj] : 0.0f, M, (int)A.M02, (int)A.out_off, out_n};
        if (gr_lds_fft<-1, MAXB>(cols + grp * gr_lds_len(M), M, A.tw, tid, nthr, A.radix, wc)) return;
    } else {
        for (int c = 0; c < cw; c++) gr_lds_fft<-1, MAXB>(cols + c * gr_lds_len(M), M, A.tw, tid, nthr, A.radix);
    }

    for (int c = CW ? grp : 0; c < (CW ? grp + 1 : cw); c++) {
        const long jj = c0 + c;
        if (jj >= A.out_n) break;
        const float cu = A.winv[A.out_off + jj];
        const float2 *col = cols + c * gr_lds_len(M);
        float *r1 = o1 + jj * A.out_row_stride, *r2 = o2 ? o2 + jj * A.out_row_stride : nullptr;
        const int off = (int)A.out_off, m02 = (int)A.M02;
        if (CW) {  // unrolled: the winv loads of all of a thread's outputs in flight at once
            float wk[EPT];
            float2 v[EPT];
#pragma unroll
            for (int r = 0; r < EPT; r++) {
                const int kk = tid + r * nthr, k = off + kk;
                if (kk < out_n) {
                    wk[r] = A.winv[k];
                    v[r] = col[gr_lds_idx(k < m02 ? M - m02 + k : k - m02)];
                }
            }
#pragma unroll
            for (int r = 0; r < EPT; r++) {
                const int kk = tid + r * nthr;
                if (kk < out_n) {
                    const float corrn = cu * wk[r];  // grid.cpp:401
                    r1[kk] = corrn * v[r].x;
                    if (r2) r2[kk] = corrn * v[r].y;
                }
            }
            continue;
        }
        for (int kk = tid; kk < out_n; kk += nthr) {
            const int k = off + kk;
            const float corrn = cu * A.winv[k];  // grid.cpp:401
            const float2 v = col[gr_lds_idx(k < m02 ? M - m02 + k : k - m02)];
            r1[kk] = corrn * v.x;
            if (r2) r2[kk] = corrn * v.y;
        }
    }
}

inline int maxb_of(long N, int nthr) { return (int)((N / 4 + nthr - 1) / nthr); }

__global__ void k_row_tiles(int2 *out, int ntr, float M2, float rmax) {
    const int ty = blockIdx.x * blockDim.x + threadIdx.x;
    if (ty >= ntr) return;
    int lo, hi;
    gr_active_tile_range(ty, ntr, M2, rmax, lo, hi);
    out[ty] = make_int2(lo, hi);
}

}  // namespace

hipError_t gr_launch_row_tiles(const GrGeom &g, int2 *out, hipStream_t s) {
    const int ntr = (int)(g.M >> 3);
    hipLaunchKernelGGL(k_row_tiles, dim3((ntr + 255) / 256), dim3(256), 0, s, out, ntr, (float)(g.M >> 1),
                       gr_grid_rmax(g));
    return hipGetLastError();
}

// Rows of Ht per workgroup of the row transform: one (at M = 4096 on MI355X
// 0.57 ms per launch of 8 pairs, 0.59 with 2 side by side, 0.67 with 4).
constexpr int kRowsPerBlock = 1;

hipError_t gr_launch_ifft_rows(const GrGeom &g, const GrDevTables &t, const float2 *ht, float2 *z,
                               int pairs, hipStream_t s) {
    RowsArgs A{ht, z, t.tw_grid, t.row_tiles, (int)g.M, g.fft_radix, g.M02, g.out_off, g.out_n};
    const int nthr = gr_fft_threads(g.M);
    int cw = kRowsPerBlock;  // rows side by side: a power of two dividing M, within 1024 threads and the LDS
    while (cw > 1 && ((long)cw * nthr > 1024 || (long)cw * g.M * 8 > 163840 || g.M % cw)) cw /= 2;
    const dim3 grid((unsigned)(g.M / cw), pairs), block(cw * nthr);
    const size_t lds = sizeof(float2) * cw * gr_lds_len(g.M);
#define GR_ROWS_K(MB, CW)                                                          \
    do {                                                                           \
        hipError_t e_ = gr_allow_lds(k_ifft_rows<MB, CW>, lds);                    \
        if (e_ != hipSuccess) return e_;                                           \
        hipLaunchKernelGGL((k_ifft_rows<MB, CW>), grid, block, lds, s, A);         \
    } while (0)
#define GR_ROWS(MB)                                                                \
    do {                                                                           \
        if (cw == 4) GR_ROWS_K(MB, 4);                                             \
        else if (cw == 2) GR_ROWS_K(MB, 2);                                        \
        else GR_ROWS_K(MB, 1);                                                     \
    } while (0)
    switch (maxb_of(g.M, nthr)) {
    case 0:
    case 1: GR_ROWS(1); break;
    case 2: GR_ROWS(2); break;
    case 4: GR_ROWS(4); break;
    case 8: GR_ROWS(8); break;
    case 16: GR_ROWS(16); break;
    default: return hipErrorInvalidValue;
    }
#undef GR_ROWS
#undef GR_ROWS_K
    return hipGetLastError();
}

// Columns per workgroup.  Side by side (CW > 0) while CW FFT thread groups
// fit a workgroup (<= 1024 threads) and CW columns fit the CU's LDS (160 KiB);
// beyond that, one after another on all threads while cw * M <= 16384.  The
// default is the widest of 4, 2, 1 side by side that fits: the column gathers
// bound the kernel and a line then serves that many columns (at M = 4096 on
// MI355X: 0.57 ms per launch of 8 pairs with 1, 0.50 with 2, 0.48 with 4).
// Callers may ask for a width (gr_plan.cols_per_block).

static bool cols_side_by_side(long M, int cw) {
    return cw >= 1 && cw <= 8 && (cw & (cw - 1)) == 0 && (long)cw * gr_fft_threads(M) <= 1024 &&
           (long)cw * M * 8 <= 163840;
}

bool gr_cols_width_ok(long M, int cols) {
    return cols_side_by_side(M, cols) || (cols >= 1 && (long)cols * M <= 16384);
}

static int cols_per_block(long M, long out_n, int pairs) {
    (void)out_n;
    (void)pairs;
    return cols_side_by_side(M, 4) ? 4 : cols_side_by_side(M, 2) ? 2 : 1;
}

hipError_t gr_launch_ifft_cols(const GrGeom &g, const GrDevTables &t, const float2 *z, float *out,
                               long out_slice_stride, long out_row_stride, int n_slices, int pair0,
                               int pairs, int cols, hipStream_t s) {
    ColsArgs A;
    A.z = z;
    A.out = out;
    A.tw = t.tw_grid;
    A.winv = t.winv;
    A.out_slice_stride = out_slice_stride;
    A.out_row_stride = out_row_stride;
    A.M = (int)g.M;
    A.cw = cols > 0 && gr_cols_width_ok(g.M, cols) ? cols : cols_per_block(g.M, g.out_n, pairs);
    A.n_slices = n_slices;
    A.pair0 = pair0;
    A.radix = g.fft_radix;
    A.M02 = g.M02;
    A.out_off = g.out_off;
    A.out_n = g.out_n;
    const int nthr = gr_fft_threads(g.M);
    const bool side = cols_side_by_side(g.M, A.cw);
    const dim3 grid((unsigned)((g.out_n + A.cw - 1) / A.cw), pairs), block(side ? A.cw * nthr : nthr);
    const size_t lds = sizeof(float2) * A.cw * gr_lds_len(g.M);
#define GR_COLS_K(MB, CW)                                                                \
    do {                                                                                 \
        hipError_t e_ = gr_allow_lds(k_ifft_cols<MB, CW>, lds);                          \
        if (e_ != hipSuccess) return e_;                                                 \
        hipLaunchKernelGGL((k_ifft_cols<MB, CW>), grid, block, lds, s, A);               \
    } while (0)
#define GR_COLS(MB)                                                                      \
    do {                                                                                 \
        if (!side) GR_COLS_K(MB, 0);                                                     \
        else if (A.cw == 1) GR_COLS_K(MB, 1);                                            \
        else if (A.cw == 2) GR_COLS_K(MB, 2);                                            \
        else if (A.cw == 4) GR_COLS_K(MB, 4);                                            \
        else if (A.cw == 8) GR_COLS_K(MB, 8);                                            \
        else GR_COLS_K(MB, 0);                                                           \
    } while (0)
    switch (maxb_of(g.M, nthr)) {
    case 0:
    case 1: { GR_COLS(1); } break;
    case 2: { GR_COLS(2); } break;
    case 4: { GR_COLS(4); } break;
    case 8: { GR_COLS(8); } break;
    case 16: { GR_COLS(16); } break;
    default: return hipErrorInvalidValue;
    }
#undef GR_COLS_K
#undef GR_COLS
    return hipGetLastError();
}
