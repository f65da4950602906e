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
#include <algorithm>
#include <atomic>

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
// window column are dropped), at Z's column-blocked place for the row (dst =
// the pair's Z + 4 b).
struct WindowRow {
    float2 *dst;
    int M, m02, off, out_n;
    static constexpr bool enabled = true;
    __device__ __forceinline__ void prep(int, int) {}
    __device__ __forceinline__ void operator()(int i, float2 v, int) const {
        const int jp = i >= M - m02 ? i - (M - m02) : i + m02;
        const int jj = jp - off;
        if (jj >= 0 && jj < out_n) dst[(long)(jj >> 2) * M * 4 + (jj & 3)] = v;
    }
};

// Row of Z written by row-transform workgroup x (of nx, CW rows each): the
// workgroups are dealt to the 8 XCDs round robin, so x, x + 8, ... share one
// XCD's L2; numbering them so that the 4 rows whose 32-byte pieces make up
// each 128-byte line of the column-blocked Z come from one XCD, one after
// another, lets its L2 merge the pieces into whole lines.
__device__ __forceinline__ int row_group(int x, int nx, int CW) {
    const int G = 4 / CW;  // workgroups per 4 rows
    if (G <= 1 || nx % (8 * G)) return x;
    return (x / (8 * G)) * (8 * G) + (x % 8) * G + (x / 8) % G;
}

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
    const int b = row_group((int)blockIdx.x, (int)gridDim.x, CW) * CW + grp, p = blockIdx.y;
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
    float2 *dst = A.z + (long)p * M * gr_z_cols(A.out_n) + 4L * b;
    const int out_n = (int)A.out_n, off = (int)A.out_off, m02 = (int)A.M02;
    // when the last pass is radix 16, the window's outputs go to HBM straight
    // from its registers: output i is window column j' = corner_index^-1(i)
    if (gr_lds_fft<-1, MAXB>(row, M, A.tw, tid, nthr, A.radix, WindowRow{dst, M, m02, off, out_n})) return;
    // an array of its own (reusing v would keep the loaded row alive, in
    // registers, through the whole transform)
    float2 o[EPT];
#pragma unroll
    for (int r = 0; r < EPT; r++) {
        const int jj = tid + r * nthr;
        o[r] = row[gr_lds_idx((int)corner_index(jj < out_n ? off + jj : off, M, m02))];
    }
#pragma unroll
    for (int r = 0; r < EPT; r++) {
        const int jj = tid + r * nthr;
        if (jj < out_n) dst[(long)(jj >> 2) * M * 4 + (jj & 3)] = o[r];
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
// r1 = nullptr and write nothing.  prep loads winv[k'] of every output of the
// thread (k' < M always: no bounds test, so the loads are issued together
// rather than each behind the previous output's stores).  NS: output slots.
template <int NS>
struct WindowCol {
    float *r1, *r2;
    const float *winv;
    float cu;
    int M, m02, off, out_n;
    float wv[NS];
    static constexpr bool enabled = true;
    __device__ __forceinline__ int kprime(int i) const { return i >= M - m02 ? i - (M - m02) : i + m02; }
    __device__ __forceinline__ void prep(int i, int s) { wv[s] = winv[kprime(i)]; }
    __device__ __forceinline__ void operator()(int i, float2 v, int s) const {
        const int kk = kprime(i) - off;
        if (r1 && kk >= 0 && kk < out_n) {
            const float corrn = cu * wv[s];
            r1[kk] = corrn * v.x;
            if (r2) r2[kk] = corrn * v.y;
        }
    }
};

// Column group of a workgroup.  Workgroups are dealt to the 8 XCDs round
// robin, each XCD with its own L2; numbering the groups so that x, x+8, x+16,
// ... (one XCD) take neighbouring column groups keeps neighbouring blocks of
// the column-blocked Z on one XCD (narrower groups than four columns share
// each block's lines).
__device__ __forceinline__ long column_group(unsigned x, unsigned ngroups) {
    if (ngroups % 8u) return x;
    return (long)(x % 8u) * (ngroups / 8u) + x / 8u;
}

// CW > 0: CW neighbouring columns per workgroup, transformed side by side by
// CW groups of nthr/CW threads (the workgroup's barriers are shared, so the
// groups run the same passes in lockstep); the fill reads the CW columns of
// a row as neighbouring bytes of the column-blocked Z (four columns: one
// contiguous block).  CW = 0: A.cw columns one after another on all threads
// (small M, where CW * M is small).
template <int MAXB, int CW>
__global__ void __launch_bounds__(1024) k_ifft_cols(ColsArgs A) {
    extern __shared__ float2 cols[];  // cw columns of gr_lds_len(M) slots
    const int p = blockIdx.y;
    const int M = A.M, cw = CW ? CW : A.cw;
    const long c0 = column_group(blockIdx.x, gridDim.x) * cw;
    const int nthr = CW ? (int)blockDim.x / CW : (int)blockDim.x;
    const int tid = CW ? (int)threadIdx.x % nthr : (int)threadIdx.x;
    const int grp = CW ? (int)threadIdx.x / nthr : 0;
    const float2 *zp = A.z + (long)p * M * gr_z_cols(A.out_n);  // column-blocked
    const int out_n = (int)A.out_n;
    constexpr int EPT = 4 * MAXB;  // rows per thread and column: M <= 4 MAXB nthr
    if (CW) {
        // unrolled, every load before the first LDS store (buffer loads: a
        // column past the window reads zeros without a branch); the CW
        // columns of a row come from neighbouring lanes
        const GrBuf zb = gr_buf(zp, (unsigned)((long)M * gr_z_cols(out_n) * sizeof(float2)));
        float2 v[EPT];
#pragma unroll
        for (int r = 0; r < EPT; r++) {
            const int idx = (int)threadIdx.x + r * (int)blockDim.x;
            const int b = idx / CW, c = idx - b * CW;
            const bool in = b < M && c0 + c < out_n;
            v[r] = gr_buf_ld2(zb, in ? (unsigned)(gr_z_index(M, b, c0 + c) * sizeof(float2)) : kGrBufNone);
        }
#pragma unroll
        for (int r = 0; r < EPT; r++) {
            const int idx = (int)threadIdx.x + r * (int)blockDim.x;
            const int b = idx / CW, c = idx - b * CW;
            if (b < M) cols[c * gr_lds_len(M) + gr_lds_idx(b)] = v[r];
        }
    } else if (cw == 1) {
        for (int b = tid; b < M; b += nthr)
            cols[gr_lds_idx(b)] = c0 < out_n ? zp[gr_z_index(M, b, c0)] : make_float2(0.0f, 0.0f);
    } else {
        for (int idx = threadIdx.x; idx < cw * M; idx += blockDim.x) {  // CW: cw * M <= 20480; else <= 16384
            const int b = idx / cw, c = idx - b * cw;
            cols[c * gr_lds_len(M) + gr_lds_idx(b)] =
                (c0 + c < out_n) ? zp[gr_z_index(M, b, c0 + c)] : make_float2(0.0f, 0.0f);
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
        WindowCol<16 * ((MAXB + 3) / 4)> wc{live ? o1 + jj * A.out_row_stride : nullptr,
                                            live && o2 ? o2 + jj * A.out_row_stride : nullptr,
                                            A.winv, live ? A.winv[A.out_off + jj] : 0.0f, M, (int)A.M02,
                                            (int)A.out_off, out_n, {}};
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

// Persistent column transform for M = 4096 with 4 columns per item and
// radix-16 passes (the plain 2048-detector geometry): the columns and the
// 32 KiB of twiddles fill the CU's 160 KiB of LDS, so one workgroup of 1024
// threads runs per CU and loops over (pair, column group) items, each one
// contiguous 128 KB block of the column-blocked Z.  Each thread
// loads rows j, j + 256, ..., j + 3840 of one column -- exactly the inputs of
// the first radix-16 butterfly j -- so that pass runs from the registers that
// received the loads; the next item's loads are issued into the same
// registers right after, and stay in flight while passes 2 and 3 run on
// twiddles from LDS (no global load in between: the loads of all of a wave
// complete in order, so any would wait for them).  The last pass ends in LDS
// and an epilogue scales and stores the image row (an output hook would need
// sixteen more registers for its deapodization factors: spills).  Same
// arithmetic, twiddles and output values as k_ifft_cols: bitwise its images.
// Needs 1 <= out_n <= 2048.
constexpr int kPfM = 4096, kPfCw = 4, kPfThreads = 1024;
constexpr int kPfNthr = kPfThreads / kPfCw;  // threads per column
constexpr int kPfColStride = kPfM + 4;       // padded: the four columns' first-pass stores use different banks
constexpr int kPfTw = 15 * 16 + 15 * 256;    // pass-ordered radix-16 twiddles of 4096 points
constexpr size_t kPfLds = sizeof(float2) * ((size_t)kPfCw * kPfColStride + kPfTw);

__global__ void __launch_bounds__(kPfThreads) k_ifft_cols_pf(ColsArgs A, int n_items) {
    extern __shared__ float2 lds[];
    float2 *tw = lds + kPfCw * kPfColStride;
    const int out_n = (int)A.out_n, off = (int)A.out_off, m02 = (int)A.M02;
    const int ngroups = (out_n + kPfCw - 1) / kPfCw;
    for (int i = (int)threadIdx.x; i < kPfTw; i += kPfThreads) tw[i] = A.tw[i];
    float2 v[16];
    float cu_next = 0.0f;
    // the item's column loads: thread t reads rows t/4 + 256 r of column t % 4
    // (the inputs of first-pass butterfly t/4); a column past the window (the
    // last group only) reads zeros through the buffer bound, no branch
    auto load = [&](int item) {
        const int t = (int)threadIdx.x, lc = t % kPfCw;
        const int p = item / ngroups, x = item - p * ngroups;
        const int c0 = (int)column_group((unsigned)x, (unsigned)ngroups) * kPfCw;
        // the group's column block: 4096 rows x 32 bytes, contiguous
        const GrBuf zb = gr_buf(A.z + (long)p * kPfM * gr_z_cols(out_n) + (long)c0 * kPfM,
                                (unsigned)(kPfM * kPfCw * sizeof(float2)));
        const bool in = c0 + lc < out_n;
        const unsigned o0 = in ? (unsigned)(((t / kPfCw) * kPfCw + lc) * sizeof(float2)) : kGrBufNone;
#pragma unroll
        for (int r = 0; r < 16; r++) v[r] = gr_buf_ld2(zb, o0, (unsigned)(r * kPfNthr * kPfCw * sizeof(float2)));
        const int grp = t / kPfNthr;
        cu_next = A.winv[off + (c0 + grp < out_n ? c0 + grp : out_n - 1)];
    };
    int item = (int)blockIdx.x;
    if (item < n_items) load(item);
    __syncthreads();  // twiddles in place
    while (item < n_items) {  // the same trip count in every thread of the block
        // thread indices recomputed from an opaque copy each iteration: held
        // from outside the loop, their derived addresses would not fit the
        // 128 registers of a 1024-thread workgroup
        int t = (int)threadIdx.x;
        asm volatile("" : "+v"(t));
        gr_fft_first_pass16<-1>(lds + (t % kPfCw) * kPfColStride, v, t / kPfCw);
        const int cur = item;
        const float cu = cu_next;
        item += (int)gridDim.x;
        __syncthreads();
        if (item < n_items) load(item);
        const int grp = t / kPfNthr, tid = t % kPfNthr;
        const int p = cur / ngroups, x = cur - p * ngroups;
        const long jj = column_group((unsigned)x, (unsigned)ngroups) * kPfCw + grp;
        const long s1 = 2L * (A.pair0 + p), s2 = s1 + 1;
        const bool live = jj < out_n;
        float *r1 = live ? A.out + s1 * A.out_slice_stride + jj * A.out_row_stride : nullptr;
        float *r2 = live && s2 < A.n_slices ? A.out + s2 * A.out_slice_stride + jj * A.out_row_stride : nullptr;
        float2 *col = lds + grp * kPfColStride;
        gr_lds_fft<-1, 4, GrFftToLds, true>(col, kPfM, tw, tid, kPfNthr, 16, GrFftToLds(), true);
        // image row jj: output kk = tid + 256 r (r < 8 covers a window of up to 2048)
        float wk[8];
        float2 o[8];
#pragma unroll
        for (int r = 0; r < 8; r++) {
            const int kk = tid + r * kPfNthr, k = off + (kk < out_n ? kk : 0);
            wk[r] = A.winv[k];
            o[r] = col[gr_lds_idx(k < m02 ? kPfM - m02 + k : k - m02)];
        }
#pragma unroll
        for (int r = 0; r < 8; r++) {
            const int kk = tid + r * kPfNthr;
            if (r1 && kk < out_n) {
                const float corrn = cu * wk[r];  // grid.cpp:401
                r1[kk] = corrn * o[r].x;
                if (r2) r2[kk] = corrn * o[r].y;
            }
        }
        __syncthreads();  // the columns are free for the next item's first pass
    }
}

// Persistent row transform for the same geometry (M = 4096, radix 16, window
// <= 2048): an item is four neighbouring rows b0 .. b0+3 of Ht (one 8-row tile
// row: one active-tile interval), one per 256-thread group; each thread loads
// elements j + 256 r of its row (buffer loads: inactive tiles read zeros, no
// branch), runs the first radix-16 butterfly j from those registers, then
// issues the next item's loads, which stay in flight through passes 2 and 3
// (twiddles from LDS).  The epilogue writes Z column-blocked in whole lines:
// sixteen consecutive lanes store one 128-byte line -- the four rows' four
// columns of one column block.  Same arithmetic and output values as
// k_ifft_rows: bitwise.
__global__ void __launch_bounds__(kPfThreads) k_ifft_rows_pf(RowsArgs A, int n_items) {
    extern __shared__ float2 lds[];
    float2 *tw = lds + kPfCw * kPfColStride;
    const int out_n = (int)A.out_n, off = (int)A.out_off, m02 = (int)A.M02;
    const int groups = kPfM / kPfCw;  // items per pair
    for (int i = (int)threadIdx.x; i < kPfTw; i += kPfThreads) tw[i] = A.tw[i];
    float2 v[16];
    // the active-tile interval of an item's tile row (its four rows share
    // one), fetched an item ahead: waited for only where the loads issued
    // before it are waited for anyway
    auto interval = [&](int item) {
        const int p = item / groups;
        return item < n_items ? A.row_tiles[((item - p * groups) * kPfCw) >> 3] : make_int2(1, 0);
    };
    auto load = [&](int item, int2 lh, int t) {  // t: threadIdx.x (an opaque copy inside the loop)
        const int grp = t / kPfNthr, tid = t % kPfNthr;
        const int p = item / groups, b0 = (item - p * groups) * kPfCw;
        const GrBuf hb = gr_buf(A.ht + ((long)p * kPfM + b0) * kPfM, (unsigned)(kPfCw * kPfM * sizeof(float2)));
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int i = tid + r * kPfNthr, tx = i >> 3;
            const bool in = tx >= lh.x && tx <= lh.y;
            v[r] = gr_buf_ld2(hb, in ? (unsigned)((grp * kPfM + i) * sizeof(float2)) : kGrBufNone);
        }
    };
    int item = (int)blockIdx.x;
    if (item < n_items) load(item, interval(item), (int)threadIdx.x);
    int2 lh_next = interval(item + (int)gridDim.x);
    __syncthreads();  // twiddles in place
    while (item < n_items) {  // the same trip count in every thread of the block
        int t = (int)threadIdx.x;
        asm volatile("" : "+v"(t));  // indices recomputed per item (as k_ifft_cols_pf)
        const int grp = t / kPfNthr, tid = t % kPfNthr;
        float2 *row = lds + grp * kPfColStride;
        gr_fft_first_pass16<-1>(row, v, tid);
        const int cur = item;
        item += (int)gridDim.x;
        __syncthreads();
        if (item < n_items) {
            load(item, lh_next, t);
            lh_next = interval(item + (int)gridDim.x);
        }
        gr_lds_fft<-1, 4, GrFftToLds, true>(row, kPfM, tw, tid, kPfNthr, 16, GrFftToLds(), true);
        const int p = cur / groups, b0 = (cur - p * groups) * kPfCw;
        float2 *zp = A.z + (long)p * kPfM * gr_z_cols(out_n);
#pragma unroll
        for (int r = 0; r < 8; r++) {  // 4 rows x up to 2048 columns = 8192 outputs
            const int L = t + r * kPfThreads, blk = L >> 4, rr = (L >> 2) & 3, c = L & 3;
            const int jj = 4 * blk + c;
            if (jj < out_n) {
                const int k = off + jj;
                zp[((long)blk * kPfM + b0 + rr) * 4 + c] =
                    lds[rr * kPfColStride + gr_lds_idx(k < m02 ? kPfM - m02 + k : k - m02)];
            }
        }
        __syncthreads();  // the rows are free for the next item's first pass
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

// A figure of the current device, queried once per device (0 when the query
// fails).
static int device_attr(hipDeviceAttribute_t attr, std::atomic<int> (&cache)[64]) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
    int v = cache[dev].load(std::memory_order_relaxed);
    if (v == 0) {
        if (hipDeviceGetAttribute(&v, attr, dev) != hipSuccess || v <= 0) v = -1;
        cache[dev].store(v, std::memory_order_relaxed);
    }
    return v > 0 ? v : 0;
}

int gr_cu_count() {
    static std::atomic<int> cache[64];
    return device_attr(hipDeviceAttributeMultiprocessorCount, cache);
}

int gr_max_lds_per_block() {
    static std::atomic<int> optin[64], plain[64];
    return std::max(device_attr(hipDeviceAttributeSharedMemPerBlockOptin, optin),
                    device_attr(hipDeviceAttributeMaxSharedMemoryPerBlock, plain));
}

// The persistent transforms need the whole 160 KiB of a CU's LDS.
static bool pf_fits() { return gr_cu_count() > 0 && (size_t)gr_max_lds_per_block() >= kPfLds; }

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
    if (g.M == kPfM && nthr == kPfNthr && g.fft_radix == 16 && g.out_n <= 8 * kPfNthr && pf_fits()) {
        const int n_items = (int)(g.M / kPfCw) * pairs;
        const int nblk = std::min(n_items, gr_cu_count());
        hipError_t e = gr_allow_lds(k_ifft_rows_pf, kPfLds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_ifft_rows_pf, dim3(nblk), dim3(kPfThreads), kPfLds, s, A, n_items);
        return hipGetLastError();
    }
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
    if (side && A.cw == kPfCw && g.M == kPfM && nthr == kPfNthr && g.fft_radix == 16 && g.out_n <= 8 * kPfNthr &&
        pf_fits()) {
        const int n_items = (int)((g.out_n + kPfCw - 1) / kPfCw) * pairs;
        const int nblk = std::min(n_items, gr_cu_count());
        hipError_t e = gr_allow_lds(k_ifft_cols_pf, kPfLds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_ifft_cols_pf, dim3(nblk), dim3(kPfThreads), kPfLds, s, A, n_items);
        return hipGetLastError();
    }
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
