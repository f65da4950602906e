// k_cnn.hip -- the sparse-angle residual CNN: forward, backward and Adam.
//
// Network (no counterpart in the reference; it corrects gridrec's streaks on
// sparse-angle scans): five 3x3 convolutions 1 -> 32 -> 32 -> 32 -> 32 -> 1,
// ReLU after the first four, y = x/scale + conv5(a4) + b5.  BF16 weights and
// activations, FP32 accumulation, MSE loss, Adam on FP32 master weights.
//
// Layouts: images [n][H][W] f32; activations and their gradients [n][H][W][32]
// bf16 (channels last: one pixel is one 64-byte record); parameters one flat
// f32 vector W1[32][9] b1[32] W2[32][32][9] b2 W3 b3 W4 b4 W5[32][9] b5[1]
// (taps ky*3+kx, correlation, zero padding).
//
// The three 32 -> 32 layers run on MFMA (v_mfma_f32_16x16x32_bf16): a block
// owns an 8 x 32 pixel tile; its input with a one-pixel halo is staged in
// LDS with an 80-byte pixel stride (the 16 lanes of a 16-byte read then hit
// 16 disjoint 4-bank groups).  Forward: D[co][pixel] = W_t[co][ci] x
// X_t[ci][pixel] summed over the 9 taps; each lane keeps its weight fragments
// (2 x 9 x 8 bf16) in registers for the whole block.  The input gradient is
// the same kernel with the taps flipped and the channel roles swapped
// (packed by k_pack), the ReLU mask of the layer input fused into its
// epilogue.  The weight gradient is D[co][ci] = dZ[co][pixel] x A_t[pixel][ci]
// with K running over the pixels of a tile row; dZ and three pre-shifted
// copies of A (one per kx, so every 16-byte read is aligned) are staged
// transposed in LDS.  Blocks loop over tiles and write per-block partial sums
// that k_reduce adds in a fixed order: results are bitwise repeatable.
//
// The 1 -> 32 and 32 -> 1 layers run on MFMA as well: K over the 9 taps of
// the single input channel, and W5 in row 0 of A.
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <utility>

#include "gr_fft.hpp"
#include "gr_internal.h"
#include "../../include/gridrec_cnn.h"

#pragma clang fp contract(off)

int gr_set_error(int code, const char *msg);

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int CH = 32;
constexpr int TH = 8, TW = 32;            // output tile (rows x columns)
constexpr int HS = TH + 2, WS = TW + 2;   // with the halo
constexpr int PSTR = 40;                  // LDS pixel stride in bf16 (80 bytes)
constexpr int OFF_W1 = 0, OFF_B1 = 288, OFF_W2 = 320, OFF_B2 = OFF_W2 + 9216, OFF_W3 = OFF_B2 + 32,
              OFF_B3 = OFF_W3 + 9216, OFF_W4 = OFF_B3 + 32, OFF_B4 = OFF_W4 + 9216, OFF_W5 = OFF_B4 + 32,
              OFF_B5 = OFF_W5 + 288, N_PARAMS = OFF_B5 + 1;
constexpr int PACK_LAYER = 2 * 9 * CH * CH;  // forward + flipped, bf16, per 32->32 layer
constexpr int WG_REC = 9 * CH * CH + CH;     // weight-gradient partial record of a 32->32 layer
constexpr int IN_REC = 9 * CH + CH;          // of the 1->32 layer
constexpr int OUT_REC = 9 * CH + 1;          // of the 32->1 layer
constexpr int MAX_WG_BLOCKS = 512;
constexpr int MAX_CONV_BLOCKS = 2048;  // 8 per CU: each loops over tiles with its weights in registers

static_assert(N_PARAMS == 28353, "parameter layout");

__device__ __forceinline__ uint16_t to_bf16(float f) {
    // round to nearest, ties to even: one v_cvt_pk_bf16_f32 (for two values)
    // instead of four integer operations per value.  On every non-NaN float
    // it equals the integer rounding (u + 0x7FFF + ((u >> 16) & 1)) >> 16 of
    // the bit pattern u, the oracle's (checked over all 2^32 patterns on
    // MI355X: no difference).
    return __builtin_bit_cast(uint16_t, (__bf16)f);
}
__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ float round_bf16(float f) { return from_bf16(to_bf16(f)); }
__device__ __forceinline__ bool positive_bf16(uint16_t h) { return !(h & 0x8000u) && h != 0; }

__device__ __forceinline__ __bf16 bf16_of(float f) { return __builtin_bit_cast(__bf16, to_bf16(f)); }

// Stage a tile's activations with their one-pixel halo (HS x WS pixels of 32
// bf16 channels, four 16-byte pieces each; zeros outside the image) from HBM
// into LDS at pixel stride PSTR, by a workgroup of 256 threads: a thread's
// loads go out in batches of BATCH, each batch issued before its LDS stores,
// so the tile costs ceil(6 / BATCH) memory latencies instead of one per loop
// trip (BATCH bounds the registers the staging holds).
template <int BATCH>
__device__ __forceinline__ void stage_tile(const uint16_t *__restrict__ x, uint16_t *sx, long ibase, int H, int W,
                                           int y0, int x0, int tid) {
    constexpr int N = HS * WS * 4, PER = (N + 255) / 256;
#pragma unroll
    for (int k0 = 0; k0 < PER; k0 += BATCH) {
        uint4 v[BATCH];
#pragma unroll
        for (int k = 0; k < BATCH; k++) {
            const int i = tid + (k0 + k) * 256;
            const int pix = i >> 2, ch = i & 3;
            const int r = pix / WS, c = pix - r * WS;
            const int gy = y0 + r - 1, gx = x0 + c - 1;
            v[k] = make_uint4(0u, 0u, 0u, 0u);
            if (k0 + k < PER && i < N && gy >= 0 && gy < H && gx >= 0 && gx < W)
                v[k] = *reinterpret_cast<const uint4 *>(x + ((ibase + (long)gy * W + gx) * CH + ch * 8));
        }
#pragma unroll
        for (int k = 0; k < BATCH; k++) {
            const int i = tid + (k0 + k) * 256;
            if (k0 + k < PER && i < N) *reinterpret_cast<uint4 *>(sx + (i >> 2) * PSTR + (i & 3) * 8) = v[k];
        }
    }
}

// The 1 -> 32 and 32 -> 1 layers on MFMA too (v_mfma_f32_16x16x32_bf16, the
// lane layout of the 32 -> 32 layers: A row = lane & 15, B column = lane & 15,
// k = 8 * (lane >> 4) + e; D column = lane & 15, rows 4 * (lane >> 4) + e).
//
// 1 -> 32: K runs over the 9 taps (k >= 9 zero), so one MFMA per 16 output
// channels gives D[co][pixel] = sum_t W1[co][t] x(pixel + t).  w1 holds the
// bf16 values of W1 as floats, [co * 9 + t].
__device__ __forceinline__ void conv_in_frag(const float *w1, int lane, bf16x8 (&A)[2]) {
#pragma unroll
    for (int nb = 0; nb < 2; nb++)
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const int k = 8 * (lane >> 4) + e;
            A[nb][e] = bf16_of(k < 9 ? w1[(nb * 16 + (lane & 15)) * 9 + k] : 0.0f);
        }
}

// B of the 1 -> 32 layer for the pixel whose 3 x 3 neighbourhood starts at
// xs[r * pitch + c] (bf16 values held as floats).
__device__ __forceinline__ bf16x8 conv_in_b(const float *xs, int pitch, int r, int c, int lane) {
    bf16x8 B;
#pragma unroll
    for (int e = 0; e < 8; e++) {
        const int k = 8 * (lane >> 4) + e;
        B[e] = bf16_of(k < 9 ? xs[(r + k / 3) * pitch + c + k % 3] : 0.0f);
    }
    return B;
}

// 32 -> 1: nine MFMAs (one per tap, K over the 32 input channels) with W5 in
// row 0 of A and zeros in rows 1..15; the result D[0][pixel] sits in element 0
// of lanes 0..15.  w5 holds the bf16 values of W5 as floats, [ci * 9 + t].
__device__ __forceinline__ void conv_out_frag(const float *w5, int lane, bf16x8 (&A)[9]) {
#pragma unroll
    for (int t = 0; t < 9; t++)
#pragma unroll
        for (int e = 0; e < 8; e++)
            A[t][e] = bf16_of((lane & 15) == 0 ? w5[(8 * (lane >> 4) + e) * 9 + t] : 0.0f);
}

// D of the 32 -> 1 layer for the pixel whose neighbourhood starts at record
// (r, c) of an LDS region of `pitch` records (PSTR bf16 each).
__device__ __forceinline__ floatx4 conv_out_mfma(const uint16_t *in, int pitch, int r, int c, int lane,
                                                 const bf16x8 (&A)[9]) {
    floatx4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int t = 0; t < 9; t++) {
        const bf16x8 B = *reinterpret_cast<const bf16x8 *>(in + ((r + t / 3) * pitch + c + t % 3) * PSTR +
                                                           8 * (lane >> 4));
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[t], B, acc, 0, 0, 0);
    }
    return acc;
}

struct TileMap {
    int tiles_x, tiles_y;
    __device__ void of(int tile, int &img, int &y0, int &x0) const { of_rows(tile, TH, img, y0, x0); }
    // tiles of th rows (tiles_y counts them)
    __device__ void of_rows(int tile, int th, int &img, int &y0, int &x0) const {
        const int per = tiles_x * tiles_y;
        img = tile / per;
        const int r = tile - img * per;
        y0 = (r / tiles_x) * th;
        x0 = (r % tiles_x) * TW;
    }
};

// A tile of TR rows: its activations with their halo (TR + 2 rows of WS
// pixels, stage_tile's pieces) loaded into registers through a buffer bounded
// to the tile's image: pieces outside the image read zeros with no branch,
// so all the loads are in flight together and none is waited for until the
// pieces are stored (needs H W 64 < 2^31).
template <int TR>
struct Stage {
    static constexpr int rows = TR + 2, n = rows * WS * 4, per = (n + 255) / 256;
};
constexpr int STAGE_PER = Stage<TH>::per;
template <int TR>
__device__ __forceinline__ void prefetch_tile(const uint16_t *__restrict__ x, int H, int W, int img, int y0, int x0,
                                              int tid, uint4 (&v)[Stage<TR>::per]) {
    const GrBuf b = gr_buf(x + (long)img * H * W * CH, (unsigned)((long)H * W * CH * sizeof(uint16_t)));
#pragma unroll
    for (int k = 0; k < Stage<TR>::per; k++) {
        const int i = tid + k * 256;
        const int pix = i >> 2, ch = i & 3;
        const int r = pix / WS, c = pix - r * WS;
        const int gy = y0 + r - 1, gx = x0 + c - 1;
        const bool in = i < Stage<TR>::n && gy >= 0 && gy < H && gx >= 0 && gx < W;
        v[k] = gr_buf_ld16(b, in ? (unsigned)(((gy * W + gx) * CH + ch * 8) * sizeof(uint16_t)) : kGrBufNone);
    }
}
template <int TR>
__device__ __forceinline__ void put_tile(uint16_t *sx, int tid, const uint4 (&v)[Stage<TR>::per]) {
#pragma unroll
    for (int k = 0; k < Stage<TR>::per; k++) {
        const int i = tid + k * 256;
        if (i < Stage<TR>::n) *reinterpret_cast<uint4 *>(sx + (i >> 2) * PSTR + (i & 3) * 8) = v[k];
    }
}

// ---------------------------------------------------------------------------
// 32 -> 32 convolution on MFMA.  MODE 0: y = bf16(relu(acc + bias));
// MODE 1 (input gradient): y = bf16(acc * (mask > 0)).  PF (MODE 0, an
// image under 2^31 bytes): the next tile's activations are loaded into
// registers while this tile's MFMAs run (the bias is held in registers, so
// no global load sits between them and the wait), on tiles of TR = 16 rows
// (the halo 1.2x the output rows instead of 1.33x, half the barriers per
// pixel); same values, same arithmetic.
template <int MODE, bool PF = false, int TR = TH>
__global__ void __launch_bounds__(256) k_conv32(const uint16_t *__restrict__ x, const uint16_t *__restrict__ wp,
                                                const float *__restrict__ bias, const uint16_t *__restrict__ mask,
                                                uint16_t *__restrict__ y, int H, int W, int ntiles, TileMap tm) {
    static_assert(!PF || MODE == 0, "prefetch: forward only");
    static_assert(PF || TR == TH, "staged tiles are TH rows");
    __shared__ __align__(16) uint16_t sx[(TR + 2) * WS * PSTR];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // weight fragments, loaded once per wave for all its tiles: lane holds
    // row co = nb*16 + (lane & 15), k = ci = 8*(lane >> 4) + e
    bf16x8 A[2][9];
#pragma unroll
    for (int nb = 0; nb < 2; nb++)
#pragma unroll
        for (int t = 0; t < 9; t++)
            A[nb][t] = *reinterpret_cast<const bf16x8 *>(wp + ((t * CH + nb * 16 + (lane & 15)) * CH + 8 * (lane >> 4)));
    float bz[2][4];
    if (PF) {
#pragma unroll
        for (int nb = 0; nb < 2; nb++)
#pragma unroll
            for (int e = 0; e < 4; e++) bz[nb][e] = bias[nb * 16 + 4 * (lane >> 4) + e];
        if ((int)blockIdx.x < ntiles) {
            int img, y0, x0;
            tm.of_rows(blockIdx.x, TR, img, y0, x0);
            uint4 pv[Stage<TR>::per];
            prefetch_tile<TR>(x, H, W, img, y0, x0, tid, pv);
            put_tile<TR>(sx, tid, pv);
        }
    }
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    int img, y0, x0;
    tm.of_rows(tile, TR, img, y0, x0);
    const long ibase = (long)img * H * W;
    uint4 pv[Stage<TR>::per];
    const int next = tile + (int)gridDim.x;
    if (PF) {
        __syncthreads();  // this tile's pieces are in sx
        if (next < ntiles) {
            int ni, ny, nx;
            tm.of_rows(next, TR, ni, ny, nx);
            prefetch_tile<TR>(x, H, W, ni, ny, nx, tid, pv);
        }
    } else {
        __syncthreads();  // the previous tile's reads of sx are done
        stage_tile<3>(x, sx, ibase, H, W, y0, x0, tid);
        __syncthreads();
    }

    const int j = lane & 15, g = lane >> 4;
#pragma unroll 1
    for (int q = 0; q < TR / 2; q++) {
        const int r = wave * (TR / 4) + (q >> 1), cb = (q & 1) * 16;
        floatx4 acc0 = {0.0f, 0.0f, 0.0f, 0.0f}, acc1 = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int t = 0; t < 9; t++) {
            const int ky = t / 3, kx = t - ky * 3;
            const bf16x8 B = *reinterpret_cast<const bf16x8 *>(sx + ((r + ky) * WS + cb + j + kx) * PSTR + 8 * g);
            acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[0][t], B, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[1][t], B, acc1, 0, 0, 0);
        }
        // lane holds D[co = nb*16 + 4g + e][pixel j]
        const int gy = y0 + r, gx = x0 + cb + j;
        if (gy < H && gx < W) {
            const long p = ibase + (long)gy * W + gx;
#pragma unroll
            for (int nb = 0; nb < 2; nb++) {
                const floatx4 a = nb ? acc1 : acc0;
                const int co0 = nb * 16 + 4 * g;
                float v[4] = {a[0], a[1], a[2], a[3]};
                if (MODE == 0) {
#pragma unroll
                    for (int e = 0; e < 4; e++) {
                        const float z = v[e] + (PF ? bz[nb][e] : bias[co0 + e]);
                        v[e] = z > 0.0f ? z : 0.0f;
                    }
                } else {
                    const uint2 m = *reinterpret_cast<const uint2 *>(mask + p * CH + co0);
                    const uint16_t mm[4] = {(uint16_t)(m.x & 0xffff), (uint16_t)(m.x >> 16), (uint16_t)(m.y & 0xffff),
                                            (uint16_t)(m.y >> 16)};
#pragma unroll
                    for (int e = 0; e < 4; e++) v[e] = positive_bf16(mm[e]) ? v[e] : 0.0f;
                }
                uint2 o;
                o.x = (uint32_t)to_bf16(v[0]) | ((uint32_t)to_bf16(v[1]) << 16);
                o.y = (uint32_t)to_bf16(v[2]) | ((uint32_t)to_bf16(v[3]) << 16);
                *reinterpret_cast<uint2 *>(y + p * CH + co0) = o;
            }
        }
    }
    if (PF && next < ntiles) {
        __syncthreads();  // this tile's reads of sx are done
        put_tile<TR>(sx, tid, pv);
    }
  }
}

// ---------------------------------------------------------------------------
// Inference with the first layer fused into the second (a1 never reaches
// HBM), on 16 x 32 output tiles: the 1 -> 32 layer runs on the tile grown
// by one pixel -- R12 x RC = 18 x 34 pixels in 39 groups of 16 (1.2x the
// output pixels; 8 x 32 tiles grew 1.33x and had twice the barriers per
// pixel) -- into LDS, zeros outside the image (the padding the second
// layer's staging gives), then the second layer on the 16 x 32 tile from
// there.
// Every value is computed by the same MFMA sequence and rounding as in the
// layer-by-layer kernels: bitwise.  (The last layer fused into the fourth the
// same way was slower: recomputing the 32 -> 32 layer on the grown tile costs
// 1.9x its MFMA work and the 62 KB workgroup halved the occupancy.)
constexpr int TH12 = 16;                     // output tile rows of the fused layers 1 + 2
constexpr int R12 = TH12 + 2, RC = TW + 2;   // the grown tile
constexpr int XR12 = TH12 + 4, XC2 = TW + 4; // its input (x) region, two-pixel halo

// Store lane's four channels co0..co0+3 of grown-tile pixel (r, c) as bf16 of
// relu(acc + bias) -- or zeros for a pixel outside the image.
__device__ __forceinline__ void put_grown(uint16_t *dst, int r, int c, int co0, floatx4 a, const float *bias,
                                          bool inside) {
    float v[4];
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const float z = a[e] + bias[co0 + e];
        v[e] = inside && z > 0.0f ? z : 0.0f;
    }
    uint2 o;
    o.x = (uint32_t)to_bf16(v[0]) | ((uint32_t)to_bf16(v[1]) << 16);
    o.y = (uint32_t)to_bf16(v[2]) | ((uint32_t)to_bf16(v[3]) << 16);
    *reinterpret_cast<uint2 *>(dst + (r * RC + c) * PSTR + co0) = o;
}

// Layers 1 and 2: a2 = layer2(layer1(x)) on 8 x 32 tiles, persistent as
// k_conv32: bf16(x/scale) with a two-pixel halo in LDS, a1 on the grown tile
// in LDS (k_conv_in's arithmetic), then k_conv32<0>'s.
__global__ void __launch_bounds__(256) k_conv12(const float *__restrict__ x, float scale,
                                                const float *__restrict__ prm, const uint16_t *__restrict__ wp,
                                                uint16_t *__restrict__ y, int H, int W, int ntiles, TileMap tm) {
    __shared__ float sxs[XR12 * XC2];
    __shared__ __align__(16) uint16_t sx[R12 * RC * PSTR];
    __shared__ float sw[9 * CH + 2 * CH];  // W1 (bf16 values), b1, b2
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int i = tid; i < 9 * CH; i += 256) sw[i] = round_bf16(prm[OFF_W1 + i]);
    if (tid < CH) {
        sw[9 * CH + tid] = prm[OFF_B1 + tid];
        sw[10 * CH + tid] = prm[OFF_B2 + tid];
    }
    __syncthreads();
    bf16x8 A1[2];
    conv_in_frag(sw, lane, A1);
    bf16x8 A[2][9];
#pragma unroll
    for (int nb = 0; nb < 2; nb++)
#pragma unroll
        for (int t = 0; t < 9; t++)
            A[nb][t] = *reinterpret_cast<const bf16x8 *>(wp + ((t * CH + nb * 16 + (lane & 15)) * CH + 8 * (lane >> 4)));
    const int j = lane & 15, g = lane >> 4;
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        int img, y0, x0;
        tm.of_rows(tile, TH12, img, y0, x0);
        const long ibase = (long)img * H * W;
        __syncthreads();  // the previous tile's reads of sxs and sx are done
        {
            constexpr int PER = (XR12 * XC2 + 255) / 256;
            float v[PER];
#pragma unroll
            for (int k = 0; k < PER; k++) {
                const int i = tid + k * 256;
                const int r = i / XC2, c = i - r * XC2, gy = y0 + r - 2, gx = x0 + c - 2;
                v[k] = i < XR12 * XC2 && gy >= 0 && gy < H && gx >= 0 && gx < W ? x[ibase + (long)gy * W + gx] : 0.0f;
            }
#pragma unroll
            for (int k = 0; k < PER; k++) {
                const int i = tid + k * 256;
                const int r = i / XC2, c = i - r * XC2, gy = y0 + r - 2, gx = x0 + c - 2;
                if (i < XR12 * XC2) sxs[i] = (gy >= 0 && gy < H && gx >= 0 && gx < W) ? round_bf16(v[k] / scale) : 0.0f;
            }
        }
        __syncthreads();
        // a1 on the grown tile: its 612 pixels packed row after row into 39
        // groups of 16 over the 4 waves (a pixel's MFMA column depends on its
        // own inputs only, so the packing changes no value)
        constexpr int NG1 = (R12 * RC + 15) / 16;
#pragma unroll 1
        for (int q = wave; q < NG1; q += 4) {
            const int pix = q * 16 + j, pp = pix < R12 * RC ? pix : R12 * RC - 1;
            const int r = pp / RC, c = pp - r * RC;
            const bf16x8 B = conv_in_b(sxs, XC2, r, c, lane);
            const floatx4 z = {0.0f, 0.0f, 0.0f, 0.0f};
            const floatx4 acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A1[0], B, z, 0, 0, 0);
            const floatx4 acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A1[1], B, z, 0, 0, 0);
            if (pix < R12 * RC) {
                const int gy = y0 + r - 1, gx = x0 + c - 1;
                const bool inside = gy >= 0 && gy < H && gx >= 0 && gx < W;
                put_grown(sx, r, c, 4 * g, acc0, sw + 9 * CH, inside);
                put_grown(sx, r, c, 16 + 4 * g, acc1, sw + 9 * CH, inside);
            }
        }
        __syncthreads();
#pragma unroll 1
        for (int q = 0; q < TH12 / 2; q++) {
            const int r = wave * (TH12 / 4) + (q >> 1), cb = (q & 1) * 16;
            floatx4 acc0 = {0.0f, 0.0f, 0.0f, 0.0f}, acc1 = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int t = 0; t < 9; t++) {
                const int ky = t / 3, kx = t - ky * 3;
                const bf16x8 B = *reinterpret_cast<const bf16x8 *>(sx + ((r + ky) * WS + cb + j + kx) * PSTR + 8 * g);
                acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[0][t], B, acc0, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[1][t], B, acc1, 0, 0, 0);
            }
            const int gy = y0 + r, gx = x0 + cb + j;
            if (gy < H && gx < W) {
                const long p = ibase + (long)gy * W + gx;
#pragma unroll
                for (int nb = 0; nb < 2; nb++) {
                    const floatx4 a = nb ? acc1 : acc0;
                    const int co0 = nb * 16 + 4 * g;
                    float v[4];
#pragma unroll
                    for (int e = 0; e < 4; e++) {
                        const float zz = a[e] + sw[10 * CH + co0 + e];
                        v[e] = zz > 0.0f ? zz : 0.0f;
                    }
                    uint2 o;
                    o.x = (uint32_t)to_bf16(v[0]) | ((uint32_t)to_bf16(v[1]) << 16);
                    o.y = (uint32_t)to_bf16(v[2]) | ((uint32_t)to_bf16(v[3]) << 16);
                    *reinterpret_cast<uint2 *>(y + p * CH + co0) = o;
                }
            }
        }
    }
}

// Weight gradient of a 32 -> 32 layer: partial[block] = (dW[t][co][ci], db[co])
// over the tiles blockIdx.x, blockIdx.x + gridDim.x, ...
constexpr int DZ_CO = TH * TW + 8;     // sdz channel stride (bf16)
constexpr int SA_CI = HS * TW + 8;     // sa channel stride (bf16)
constexpr size_t WGRAD_LDS = sizeof(uint16_t) * (CH * DZ_CO + 3 * CH * SA_CI) + sizeof(float) * 256 * 8;

__global__ void __launch_bounds__(256) k_conv32_wgrad(const uint16_t *__restrict__ a, const uint16_t *__restrict__ dz,
                                                      float *__restrict__ part, int H, int W, int ntiles, TileMap tm) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    uint16_t *sdz = reinterpret_cast<uint16_t *>(smem_raw);      // [co][r*TW + c]
    uint16_t *sa = sdz + CH * DZ_CO;                             // [kx][ci][r*TW + c]
    float *sdb = reinterpret_cast<float *>(sa + 3 * CH * SA_CI); // [256][8]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int j = lane & 15, g = lane >> 4;
    floatx4 acc[9];
#pragma unroll
    for (int i = 0; i < 9; i++) acc[i] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    float db[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const int ch = tid & 3;  // the channel chunk this thread stages (256 % 4 == 0)
    // The next tile's records are loaded into registers while this tile's are
    // added (its loads are in flight during the MFMAs): DZ_N pieces of dz and
    // A_N of a per thread.
    constexpr int DZ_N = TH * TW * 4 / 256, A_N = (HS * WS * 4 + 255) / 256;
    uint4 pdz[DZ_N], pa[A_N];
    auto fetch = [&](int tile) {
        int img, y0, x0;
        tm.of(tile, img, y0, x0);
        const long ibase = (long)img * H * W;
#pragma unroll
        for (int k = 0; k < DZ_N; k++) {
            const int pix = (tid + 256 * k) >> 2, r = pix / TW, c = pix - r * TW, gy = y0 + r, gx = x0 + c;
            pdz[k] = make_uint4(0u, 0u, 0u, 0u);
            if (gy < H && gx < W) pdz[k] = *reinterpret_cast<const uint4 *>(dz + ((ibase + (long)gy * W + gx) * CH + ch * 8));
        }
#pragma unroll
        for (int k = 0; k < A_N; k++) {
            const int i = tid + 256 * k, pix = i >> 2, r = pix / WS, cc = pix - r * WS;
            const int gy = y0 + r - 1, gx = x0 + cc - 1;
            pa[k] = make_uint4(0u, 0u, 0u, 0u);
            if (i < HS * WS * 4 && gy >= 0 && gy < H && gx >= 0 && gx < W)
                pa[k] = *reinterpret_cast<const uint4 *>(a + ((ibase + (long)gy * W + gx) * CH + ch * 8));
        }
    };
    if (blockIdx.x < ntiles) fetch(blockIdx.x);

    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        __syncthreads();  // previous tile's reads done
#pragma unroll
        for (int k = 0; k < DZ_N; k++) {
            const int pix = (tid + 256 * k) >> 2;
            const uint32_t w[4] = {pdz[k].x, pdz[k].y, pdz[k].z, pdz[k].w};
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const uint16_t h = (uint16_t)(w[e >> 1] >> (16 * (e & 1)));
                sdz[(ch * 8 + e) * DZ_CO + pix] = h;
                db[e] += from_bf16(h);
            }
        }
#pragma unroll
        for (int k = 0; k < A_N; k++) {
            const int i = tid + 256 * k, pix = i >> 2;
            if (i >= HS * WS * 4) break;
            const int r = pix / WS, cc = pix - r * WS;
            const uint32_t w[4] = {pa[k].x, pa[k].y, pa[k].z, pa[k].w};
#pragma unroll
            for (int kx = 0; kx < 3; kx++) {
                const int c = cc - kx;  // sa[kx] column c holds input column x0 + c + kx - 1
                if (c < 0 || c >= TW) continue;
#pragma unroll
                for (int e = 0; e < 8; e++)
                    sa[(kx * CH + ch * 8 + e) * SA_CI + r * TW + c] = (uint16_t)(w[e >> 1] >> (16 * (e & 1)));
            }
        }
        if (tile + (int)gridDim.x < ntiles) fetch(tile + gridDim.x);
        __syncthreads();
#pragma unroll 1
        for (int r = 0; r < TH; r++) {
#pragma unroll
            for (int i = 0; i < 9; i++) {
                const int b = wave * 9 + i;  // 36 blocks: tap t, co half, ci half
                const int t = b >> 2, nbo = (b >> 1) & 1, nbi = b & 1;
                const int ky = t / 3, kx = t - ky * 3;
                const bf16x8 Av = *reinterpret_cast<const bf16x8 *>(sdz + (nbo * 16 + j) * DZ_CO + r * TW + 8 * g);
                const bf16x8 Bv =
                    *reinterpret_cast<const bf16x8 *>(sa + (kx * CH + nbi * 16 + j) * SA_CI + (r + ky) * TW + 8 * g);
                acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Av, Bv, acc[i], 0, 0, 0);
            }
        }
    }
    float *out = part + (long)blockIdx.x * WG_REC;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        const int b = wave * 9 + i;
        const int t = b >> 2, nbo = (b >> 1) & 1, nbi = b & 1;
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int co = nbo * 16 + 4 * g + e, ci = nbi * 16 + j;
            out[(t * CH + co) * CH + ci] = acc[i][e];
        }
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 8; e++) sdb[tid * 8 + e] = db[e];
    __syncthreads();
    if (tid < CH) {  // channel tid: chunk tid/8, element tid%8, summed over the threads of that chunk in order
        float s = 0.0f;
        for (int k = tid >> 3; k < 256; k += 4) s += sdb[k * 8 + (tid & 7)];
        out[9 * CH * CH + tid] = s;
    }
}

// ---------------------------------------------------------------------------
// 1 -> 32 layer: a1 = bf16(relu(W1 * bf16(x/scale) + b1)) on MFMA: a block
// owns an 8 x 32 pixel tile (persistent over tiles), stages bf16(x/scale)
// with a one-pixel halo in LDS, and each wave computes 4 groups of 16 pixels.
__global__ void __launch_bounds__(256) k_conv_in(const float *__restrict__ x, float scale,
                                                 const float *__restrict__ prm, uint16_t *__restrict__ a1, int H,
                                                 int W, int ntiles, TileMap tm) {
    __shared__ float sxs[HS * WS];
    __shared__ float sw[9 * CH + CH];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int i = tid; i < 9 * CH; i += 256) sw[i] = round_bf16(prm[OFF_W1 + i]);
    if (tid < CH) sw[9 * CH + tid] = prm[OFF_B1 + tid];
    __syncthreads();
    bf16x8 A[2];
    conv_in_frag(sw, lane, A);
    const int j = lane & 15, g = lane >> 4;
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        int img, y0, x0;
        tm.of(tile, img, y0, x0);
        const long ibase = (long)img * H * W;
        __syncthreads();  // the previous tile's reads of sxs are done
        {  // both loads of a thread in flight before the stores
            constexpr int PER = (HS * WS + 255) / 256;
            float v[PER];
#pragma unroll
            for (int k = 0; k < PER; k++) {
                const int i = tid + k * 256;
                const int r = i / WS, c = i - r * WS, gy = y0 + r - 1, gx = x0 + c - 1;
                v[k] = i < HS * WS && gy >= 0 && gy < H && gx >= 0 && gx < W ? x[ibase + (long)gy * W + gx] : 0.0f;
            }
#pragma unroll
            for (int k = 0; k < PER; k++) {
                const int i = tid + k * 256;
                const int r = i / WS, c = i - r * WS, gy = y0 + r - 1, gx = x0 + c - 1;
                if (i < HS * WS) sxs[i] = (gy >= 0 && gy < H && gx >= 0 && gx < W) ? round_bf16(v[k] / scale) : 0.0f;
            }
        }
        __syncthreads();
#pragma unroll 1
        for (int q = 0; q < 4; q++) {
            const int r = wave * 2 + (q >> 1), cb = (q & 1) * 16;
            const bf16x8 B = conv_in_b(sxs, WS, r, cb + j, lane);
            const floatx4 z = {0.0f, 0.0f, 0.0f, 0.0f};
            const floatx4 acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[0], B, z, 0, 0, 0);
            const floatx4 acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[1], B, z, 0, 0, 0);
            const int gy = y0 + r, gx = x0 + cb + j;
            if (gy < H && gx < W) {
                const long p = ibase + (long)gy * W + gx;
#pragma unroll
                for (int nb = 0; nb < 2; nb++) {
                    const floatx4 a = nb ? acc1 : acc0;
                    const int co0 = nb * 16 + 4 * g;
                    float v[4];
#pragma unroll
                    for (int e = 0; e < 4; e++) {
                        const float zz = a[e] + sw[9 * CH + co0 + e];
                        v[e] = zz > 0.0f ? zz : 0.0f;
                    }
                    uint2 o;
                    o.x = (uint32_t)to_bf16(v[0]) | ((uint32_t)to_bf16(v[1]) << 16);
                    o.y = (uint32_t)to_bf16(v[2]) | ((uint32_t)to_bf16(v[3]) << 16);
                    *reinterpret_cast<uint2 *>(a1 + p * CH + co0) = o;
                }
            }
        }
    }
}

// Sum acc[9][8] (tap, channel of the thread's group of 8) and extra[8] over
// the threads of the block that share a channel group (tid mod 4), in a fixed
// order: lanes of a wave by a butterfly over lane bits 2..5, then the four
// waves in order through LDS; the block's record gets W[ch][t] at
// ch * 9 + t (parameter order) and the extras at 9 * CH + ch.
__device__ __forceinline__ void block_sum_groups(float (&acc)[9][8], float (&extra)[8], float *red, float *out,
                                                 bool with_extra) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, cg = tid & 3;
#pragma unroll
    for (int o = 4; o < 64; o <<= 1) {
#pragma unroll
        for (int t = 0; t < 9; t++)
#pragma unroll
            for (int e = 0; e < 8; e++) acc[t][e] += __shfl_xor(acc[t][e], o);
#pragma unroll
        for (int e = 0; e < 8; e++) extra[e] += __shfl_xor(extra[e], o);
    }
    // lanes 0..3 of each wave hold the wave's sums of groups 0..3
    if (lane < 4) {
#pragma unroll
        for (int t = 0; t < 9; t++)
#pragma unroll
            for (int e = 0; e < 8; e++) red[(wave * 4 + cg) * 80 + t * 8 + e] = acc[t][e];
#pragma unroll
        for (int e = 0; e < 8; e++) red[(wave * 4 + cg) * 80 + 72 + e] = extra[e];
    }
    __syncthreads();
    for (int i = tid; i < 4 * 80; i += blockDim.x) {  // i = group * 80 + slot
        const int grp = i / 80, sl = i - grp * 80;
        if (sl >= 72 && !with_extra) continue;
        float v = 0.0f;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) v += red[(w * 4 + grp) * 80 + sl];
        if (sl < 72) {
            const int t = sl >> 3, ch = grp * 8 + (sl & 7);
            out[ch * 9 + t] = v;
        } else {
            out[9 * CH + grp * 8 + (sl - 72)] = v;
        }
    }
}

// Weight gradient of the 1 -> 32 layer: partial[block] = (dW1[co][t], db1[co])
// over the tiles blockIdx.x, blockIdx.x + gridDim.x, ...  Thread tid takes the
// 8 channels (tid mod 4) * 8.. of the output pixels tid / 4 + 64 k of a tile:
// one 16-byte load of dz per pixel (four threads read a pixel's 64-byte
// record, a wave 16 whole records), all four issued before use, and the nine
// taps of bf16(x/scale) from the tile staged with its halo in LDS; the sums
// stay in registers across the block's tiles.
__global__ void __launch_bounds__(256) k_conv_in_wgrad(const float *__restrict__ x, float scale,
                                                       const uint16_t *__restrict__ dz, float *__restrict__ part,
                                                       int H, int W, int ntiles, TileMap tm) {
    __shared__ float sx[HS * WS];
    __shared__ float red[16 * 80];
    const int tid = threadIdx.x, cg = tid & 3, ps = tid >> 2;
    float acc[9][8], db[8];
#pragma unroll
    for (int e = 0; e < 8; e++) {
        db[e] = 0.0f;
#pragma unroll
        for (int t = 0; t < 9; t++) acc[t][e] = 0.0f;
    }
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        int img, y0, x0;
        tm.of(tile, img, y0, x0);
        const long ibase = (long)img * H * W;
        uint4 v[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int p = ps + 64 * k, r = p / TW, c = p - r * TW, gy = y0 + r, gx = x0 + c;
            v[k] = make_uint4(0u, 0u, 0u, 0u);
            if (gy < H && gx < W) v[k] = *reinterpret_cast<const uint4 *>(dz + (ibase + (long)gy * W + gx) * CH + cg * 8);
        }
        __syncthreads();  // the previous tile's reads of sx are done
        for (int i = tid; i < HS * WS; i += 256) {
            const int r = i / WS, c = i - r * WS, gy = y0 + r - 1, gx = x0 + c - 1;
            sx[i] = (gy >= 0 && gy < H && gx >= 0 && gx < W) ? round_bf16(x[ibase + (long)gy * W + gx] / scale) : 0.0f;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int p = ps + 64 * k, r = p / TW, c = p - r * TW;
            const uint32_t w4[4] = {v[k].x, v[k].y, v[k].z, v[k].w};
            float d[8];
#pragma unroll
            for (int e = 0; e < 8; e++) {
                d[e] = from_bf16((uint16_t)(w4[e >> 1] >> (16 * (e & 1))));
                db[e] += d[e];
            }
#pragma unroll
            for (int t = 0; t < 9; t++) {
                const float xv = sx[(r + t / 3) * WS + c + t % 3];
#pragma unroll
                for (int e = 0; e < 8; e++) acc[t][e] = __builtin_fmaf(d[e], xv, acc[t][e]);
            }
        }
    }
    block_sum_groups(acc, db, red, part + (long)blockIdx.x * IN_REC, true);
}

// ---------------------------------------------------------------------------
// 32 -> 1 layer with the residual: r = W5 * a4 + b5, yn = x/scale + r,
// y = yn * scale; per-tile sums of r^2 and (x/scale)^2 for the fallback test.
// An 8 x 32 tile at a time: a4 with a one-pixel halo staged in LDS, each wave
// 4 groups of 16 pixels on MFMA (conv_out_mfma), lanes 0..15 finishing them;
// the tile's sums by a pairwise tree (sred[i] += sred[i + h], h = 128 .. 1:
// the two steps across waves through LDS, the rest inside wave 0 by lane
// shuffles, the same additions in the same order).  PF (an image under 2^31
// bytes of activations): persistent workgroups, the next tile's a4 pieces and
// x values loaded into registers while this tile is computed; otherwise one
// tile per workgroup.  The same values either way.
template <bool PF>
__global__ void __launch_bounds__(256) k_conv_out(const uint16_t *__restrict__ a4, const float *__restrict__ x,
                                                  float scale, const float *__restrict__ prm, float *__restrict__ y,
                                                  float *__restrict__ yn, float *__restrict__ stats, int H, int W,
                                                  int ntiles, TileMap tm) {
    __shared__ __align__(16) uint16_t sx[HS * WS * PSTR];
    __shared__ float sw[9 * CH + 1];
    __shared__ float sred[2][TH * TW];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int j = lane & 15;
    for (int i = tid; i < 9 * CH; i += 256) sw[i] = round_bf16(prm[OFF_W5 + i]);
    if (tid == 0) sw[9 * CH] = prm[OFF_B5];
    // x of the lane's four pixels (lanes 0..15; zero outside the image)
    auto load_x = [&](int img, int y0, int x0, float (&xv)[4]) {
        const GrBuf b = gr_buf(x + (long)img * H * W, (unsigned)((long)H * W * sizeof(float)));
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int gy = y0 + wave * 2 + (q >> 1), gx = x0 + (q & 1) * 16 + j;
            const bool in = lane < 16 && gy < H && gx < W;
            xv[q] = gr_buf_ldf(b, in ? (unsigned)((gy * W + gx) * sizeof(float)) : kGrBufNone);
        }
    };
    float xv[4];
    int tile = PF ? (int)blockIdx.x : 0;
    if (PF) {
        if (tile < ntiles) {
            int img, y0, x0;
            tm.of(tile, img, y0, x0);
            uint4 pv[STAGE_PER];
            prefetch_tile<TH>(a4, H, W, img, y0, x0, tid, pv);
            load_x(img, y0, x0, xv);
            put_tile<TH>(sx, tid, pv);
        }
    } else {
        tile = (int)blockIdx.x;
        int img, y0, x0;
        tm.of(tile, img, y0, x0);
        stage_tile<6>(a4, sx, (long)img * H * W, H, W, y0, x0, tid);
    }
    __syncthreads();
    bf16x8 A[9];
    conv_out_frag(sw, lane, A);
    const float b5 = sw[9 * CH];
    for (; tile < ntiles; tile += PF ? (int)gridDim.x : ntiles) {
        int img, y0, x0;
        tm.of(tile, img, y0, x0);
        const long ibase = (long)img * H * W;
        __syncthreads();  // this tile's pieces in sx; the previous tile's sums read
        const int next = tile + (int)gridDim.x;
        uint4 pv[STAGE_PER];
        float xn_next[4];
        if (PF && next < ntiles) {
            int ni, ny, nx;
            tm.of(next, ni, ny, nx);
            prefetch_tile<TH>(a4, H, W, ni, ny, nx, tid, pv);
            load_x(ni, ny, nx, xn_next);
        }
#pragma unroll 1
        for (int q = 0; q < 4; q++) {
            const int r = wave * 2 + (q >> 1), c = (q & 1) * 16 + j;
            const floatx4 acc = conv_out_mfma(sx, WS, r, c, lane, A);
            const int gy = y0 + r, gx = x0 + c;
            if (lane < 16) {  // every pixel of the tile writes its sums: zeros outside the image
                float s2 = 0.0f, x2 = 0.0f;
                if (gy < H && gx < W) {
                    const float s = acc[0] + b5;
                    const long p = ibase + (long)gy * W + gx;
                    const float xn = (PF ? xv[q] : x[p]) / scale;
                    const float v = xn + s;
                    if (yn) yn[p] = v;
                    if (y) y[p] = v * scale;
                    s2 = s * s;
                    x2 = xn * xn;
                }
                sred[0][r * TW + c] = s2;
                sred[1][r * TW + c] = x2;
            }
        }
        __syncthreads();  // sx reads and sred writes done
        if (PF && next < ntiles) put_tile<TH>(sx, tid, pv);
        if (tid < 128) {
            sred[0][tid] += sred[0][tid + 128];
            sred[1][tid] += sred[1][tid + 128];
        }
        __syncthreads();
        if (tid < 64) {
            float r0 = sred[0][tid] + sred[0][tid + 64], r1 = sred[1][tid] + sred[1][tid + 64];
#pragma unroll
            for (int h = 32; h > 0; h >>= 1) {
                r0 += __shfl_down(r0, h, 64);
                r1 += __shfl_down(r1, h, 64);
            }
            if (tid == 0 && stats) {
                stats[2L * tile] = r0;
                stats[2L * tile + 1] = r1;
            }
        }
        if (PF) {
#pragma unroll
            for (int q = 0; q < 4; q++) xv[q] = xn_next[q];
        }
    }
}

// ---------------------------------------------------------------------------
// Inference, all five layers fused: a block owns an FH x FW output tile and
// computes each layer on the tile grown by the halo the later layers need
// (x: 5 pixels, a1: 4, a2: 3, a3: 2, a4: 1), keeping every activation in LDS
// (a1/a3 and a2/a4 share two buffers), so HBM sees x once and y once instead
// of four 64-byte activation records per pixel each way.  Per pixel the
// arithmetic and its order are those of k_conv_in, k_conv32<0> and
// k_conv_out, so y is bitwise the unfused forward's; activations of pixels
// outside the image are zero, the unfused path's zero padding.
constexpr int FH = 16, FW = 32;
constexpr int XR = FH + 10, XC = FW + 10;
constexpr int R1 = FH + 8, C1 = FW + 8;
constexpr int R2 = FH + 6, C2 = FW + 6;
constexpr size_t FUSED_LDS = sizeof(float) * XR * XC + sizeof(uint16_t) * (size_t)(R1 * C1 + R2 * C2) * PSTR +
                             sizeof(float) * (IN_REC + OUT_REC) + sizeof(float) * 2 * FH * FW;

// One 32 -> 32 layer from an (Ri x Ci) input region to the (Ri-2 x Ci-2)
// output region, LDS to LDS; (oy, ox) is the image position of output (0, 0).
__device__ __forceinline__ void fused_conv32(const uint16_t *in, int Ci, uint16_t *out, int Ro, int Co,
                                             const uint16_t *__restrict__ wp, const float *__restrict__ bias,
                                             int oy, int ox, int H, int W, int lane, int wave, int nwaves) {
    bf16x8 A[2][9];
#pragma unroll
    for (int nb = 0; nb < 2; nb++)
#pragma unroll
        for (int t = 0; t < 9; t++)
            A[nb][t] = *reinterpret_cast<const bf16x8 *>(wp + ((t * CH + nb * 16 + (lane & 15)) * CH + 8 * (lane >> 4)));
    const int j = lane & 15, g = lane >> 4;
    const int npix = Ro * Co, ngroups = (npix + 15) / 16;
    // two 16-pixel groups per trip: four independent accumulation chains
    for (int grp0 = 2 * wave; grp0 < ngroups; grp0 += 2 * nwaves) {
        int pp[2], rr[2], cc[2];
#pragma unroll
        for (int h = 0; h < 2; h++) {
            pp[h] = (grp0 + h) * 16 + j;
            const int pc = pp[h] < npix ? pp[h] : npix - 1;
            rr[h] = pc / Co;
            cc[h] = pc - rr[h] * Co;
        }
        floatx4 acc[2][2];
#pragma unroll
        for (int h = 0; h < 2; h++) acc[h][0] = acc[h][1] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int t = 0; t < 9; t++) {
            const int ky = t / 3, kx = t - ky * 3;
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const bf16x8 B =
                    *reinterpret_cast<const bf16x8 *>(in + ((rr[h] + ky) * Ci + cc[h] + kx) * PSTR + 8 * g);
                acc[h][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[0][t], B, acc[h][0], 0, 0, 0);
                acc[h][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[1][t], B, acc[h][1], 0, 0, 0);
            }
        }
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const int p = pp[h], r = rr[h], c = cc[h];
        if (p < npix) {
            const int gy = oy + r, gx = ox + c;
            const bool inside = gy >= 0 && gy < H && gx >= 0 && gx < W;
#pragma unroll
            for (int nb = 0; nb < 2; nb++) {
                const floatx4 a = acc[h][nb];
                const int co0 = nb * 16 + 4 * g;
                float v[4];
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const float z = a[e] + bias[co0 + e];
                    v[e] = inside && z > 0.0f ? z : 0.0f;
                }
                uint2 o;
                o.x = (uint32_t)to_bf16(v[0]) | ((uint32_t)to_bf16(v[1]) << 16);
                o.y = (uint32_t)to_bf16(v[2]) | ((uint32_t)to_bf16(v[3]) << 16);
                *reinterpret_cast<uint2 *>(out + p * PSTR + co0) = o;
            }
        }
      }
    }
}

__global__ void __launch_bounds__(512) k_cnn_fused(const float *__restrict__ x, float scale,
                                                   const float *__restrict__ prm, const uint16_t *__restrict__ packed,
                                                   float *__restrict__ y, float *__restrict__ stats, int H, int W,
                                                   int tiles_x, int tiles_y) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float *sx = reinterpret_cast<float *>(smem_raw);                 // [XR][XC] bf16(x/scale) as float
    uint16_t *bufA = reinterpret_cast<uint16_t *>(sx + XR * XC);     // a1 [R1][C1], then a3
    uint16_t *bufB = bufA + R1 * C1 * PSTR;                          // a2 [R2][C2], then a4
    float *sw1 = reinterpret_cast<float *>(bufB + R2 * C2 * PSTR);   // W1 (bf16 values), b1
    float *sw5 = sw1 + IN_REC;                                       // W5 (bf16 values), b5
    float *sred = sw5 + OUT_REC;                                     // [2][FH*FW]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwaves = blockDim.x >> 6;
    const int per = tiles_x * tiles_y;
    const int img = blockIdx.x / per, rt = blockIdx.x - img * per;
    const int y0 = (rt / tiles_x) * FH, x0 = (rt % tiles_x) * FW;
    const long ibase = (long)img * H * W;

    for (int i = tid; i < 9 * CH; i += blockDim.x) sw1[i] = round_bf16(prm[OFF_W1 + i]);
    for (int i = tid; i < CH; i += blockDim.x) sw1[9 * CH + i] = prm[OFF_B1 + i];
    for (int i = tid; i < 9 * CH; i += blockDim.x) sw5[i] = round_bf16(prm[OFF_W5 + i]);
    if (tid == 0) sw5[9 * CH] = prm[OFF_B5];
    for (int i = tid; i < XR * XC; i += blockDim.x) {
        const int r = i / XC, c = i - r * XC, gy = y0 - 5 + r, gx = x0 - 5 + c;
        sx[i] = (gy >= 0 && gy < H && gx >= 0 && gx < W) ? round_bf16(x[ibase + (long)gy * W + gx] / scale) : 0.0f;
    }
    __syncthreads();
    // layer 1 on the a1 region, 16 pixels per MFMA group (as k_conv_in)
    {
        bf16x8 A1[2];
        conv_in_frag(sw1, lane, A1);
        const int j = lane & 15, g = lane >> 4, npix = R1 * C1;
        for (int grp = wave; grp * 16 < npix; grp += nwaves) {
            const int p = grp * 16 + j, pc = p < npix ? p : npix - 1;
            const int r = pc / C1, c = pc - r * C1;
            const bf16x8 B = conv_in_b(sx, XC, r, c, lane);
            const floatx4 z = {0.0f, 0.0f, 0.0f, 0.0f};
            const floatx4 acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A1[0], B, z, 0, 0, 0);
            const floatx4 acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A1[1], B, z, 0, 0, 0);
            if (p < npix) {
                const int gy = y0 - 4 + r, gx = x0 - 4 + c;
                const bool inside = gy >= 0 && gy < H && gx >= 0 && gx < W;
#pragma unroll
                for (int nb = 0; nb < 2; nb++) {
                    const floatx4 a = nb ? acc1 : acc0;
                    const int co0 = nb * 16 + 4 * g;
                    float v[4];
#pragma unroll
                    for (int e = 0; e < 4; e++) {
                        const float zz = a[e] + sw1[9 * CH + co0 + e];
                        v[e] = inside && zz > 0.0f ? zz : 0.0f;
                    }
                    uint2 o;
                    o.x = (uint32_t)to_bf16(v[0]) | ((uint32_t)to_bf16(v[1]) << 16);
                    o.y = (uint32_t)to_bf16(v[2]) | ((uint32_t)to_bf16(v[3]) << 16);
                    *reinterpret_cast<uint2 *>(bufA + p * PSTR + co0) = o;
                }
            }
        }
    }
    __syncthreads();
    fused_conv32(bufA, C1, bufB, R2, C2, packed, prm + OFF_B2, y0 - 3, x0 - 3, H, W, lane, wave, nwaves);
    __syncthreads();
    fused_conv32(bufB, C2, bufA, R2 - 2, C2 - 2, packed + PACK_LAYER, prm + OFF_B3, y0 - 2, x0 - 2, H, W, lane, wave,
                 nwaves);
    __syncthreads();
    fused_conv32(bufA, C2 - 2, bufB, R2 - 4, C2 - 4, packed + 2 * PACK_LAYER, prm + OFF_B4, y0 - 1, x0 - 1, H, W,
                 lane, wave, nwaves);
    __syncthreads();
    // layer 5 + residual, 16 pixels per MFMA group (as k_conv_out)
    const int C4 = C2 - 4;
    for (int i = tid; i < 2 * FH * FW; i += blockDim.x) sred[i] = 0.0f;
    __syncthreads();
    {
        bf16x8 A5[9];
        conv_out_frag(sw5, lane, A5);
        const int j = lane & 15;
        for (int grp = wave; grp * 16 < FH * FW; grp += nwaves) {
            const int q = grp * 16 + j, r = q / FW, c = q - r * FW, gy = y0 + r, gx = x0 + c;
            const floatx4 acc = conv_out_mfma(bufB, C4, r, c, lane, A5);
            if (lane < 16 && gy < H && gx < W) {
                const float s = acc[0] + sw5[9 * CH];
                const long p = ibase + (long)gy * W + gx;
                const float xn = x[p] / scale;
                y[p] = (xn + s) * scale;
                sred[q] = s * s;
                sred[FH * FW + q] = xn * xn;
            }
        }
    }
    __syncthreads();
    for (int h = FH * FW / 2; h > 0; h >>= 1) {
        for (int i = tid; i < h; i += blockDim.x) {
            sred[i] += sred[i + h];
            sred[FH * FW + i] += sred[FH * FW + i + h];
        }
        __syncthreads();
    }
    if (tid == 0 && stats) {
        stats[2 * blockIdx.x] = sred[0];
        stats[2 * blockIdx.x + 1] = sred[FH * FW];
    }
}

// Loss: dy = 2 (yn - t/scale) / N; per-block partial sums of (yn - t/scale)^2.
__global__ void __launch_bounds__(256) k_loss(const float *__restrict__ yn, const float *__restrict__ target,
                                              float scale, float *__restrict__ dy, float *__restrict__ part, long N) {
    __shared__ float sred[256];
    float s = 0.0f;
    for (long p = (long)blockIdx.x * 256 + threadIdx.x; p < N; p += (long)gridDim.x * 256) {
        const float e = yn[p] - target[p] / scale;
        dy[p] = 2.0f * e / (float)N;
        s += e * e;
    }
    sred[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) sred[threadIdx.x] += sred[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sred[0];
}

// Input gradient of the 32 -> 1 layer, masked: dz4 = bf16(sum_t dy[p - off_t] W5[ci][t] * (a4 > 0)).
__global__ void __launch_bounds__(256) k_conv_out_dgrad(const float *__restrict__ dy, const uint16_t *__restrict__ a4,
                                                        const float *__restrict__ prm, uint16_t *__restrict__ dz4,
                                                        int n, int H, int W) {
    __shared__ float sw[9 * CH];
    for (int i = threadIdx.x; i < 9 * CH; i += blockDim.x) sw[i] = round_bf16(prm[OFF_W5 + i]);
    __syncthreads();
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (long)n * H * W) return;
    const long img = p / ((long)H * W);
    const int rem = (int)(p - img * H * W), py = rem / W, px = rem - py * W;
    float d[9];
#pragma unroll
    for (int t = 0; t < 9; t++) {
        const int yy = py - (t / 3 - 1), xx = px - (t % 3 - 1);
        d[t] = (yy >= 0 && yy < H && xx >= 0 && xx < W) ? dy[(img * H + yy) * W + xx] : 0.0f;
    }
    const uint4 *m4 = reinterpret_cast<const uint4 *>(a4 + p * CH);
    uint4 *o4 = reinterpret_cast<uint4 *>(dz4 + p * CH);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint4 m = m4[k];
        const uint32_t mw[4] = {m.x, m.y, m.z, m.w};
        uint32_t ow[4];
#pragma unroll
        for (int e2 = 0; e2 < 4; e2++) {
            uint32_t word = 0;
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int ci = k * 8 + e2 * 2 + h;
                float s = 0.0f;
#pragma unroll
                for (int t = 0; t < 9; t++) s += d[t] * sw[ci * 9 + t];
                const uint16_t mk = (uint16_t)(mw[e2] >> (16 * h));
                word |= (uint32_t)to_bf16(positive_bf16(mk) ? s : 0.0f) << (16 * h);
            }
            ow[e2] = word;
        }
        o4[k] = make_uint4(ow[0], ow[1], ow[2], ow[3]);
    }
}

// Weight gradient of the 32 -> 1 layer: partial[block] = (dW5[ci][t], db5)
// over the tiles blockIdx.x, blockIdx.x + gridDim.x, ...  dW5[ci][t] = sum_p
// dy[p] a4[p + off_t][ci] is taken as a sum over the input pixels q of the
// tile with its halo, dy[q - off_t] the output pixel each tap reaches: thread
// tid takes the 8 channels (tid mod 4) * 8.. of the halo pixels tid / 4 +
// 64 k (one 16-byte load per pixel, a wave 16 whole records, all six issued
// before use) and dy of the tile from LDS; sums in registers across tiles.
__global__ void __launch_bounds__(256) k_conv_out_wgrad(const float *__restrict__ dy, const uint16_t *__restrict__ a4,
                                                        float *__restrict__ part, int H, int W, int ntiles,
                                                        TileMap tm) {
    __shared__ float sdy[TH * TW];
    __shared__ float red[16 * 80];
    constexpr int NQ = (HS * WS + 63) / 64;  // halo pixels per thread: 6
    const int tid = threadIdx.x, cg = tid & 3, ps = tid >> 2;
    float acc[9][8], db[8], dsum = 0.0f;
#pragma unroll
    for (int e = 0; e < 8; e++) {
        db[e] = 0.0f;
#pragma unroll
        for (int t = 0; t < 9; t++) acc[t][e] = 0.0f;
    }
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        int img, y0, x0;
        tm.of(tile, img, y0, x0);
        const long ibase = (long)img * H * W;
        uint4 v[NQ];
#pragma unroll
        for (int k = 0; k < NQ; k++) {
            const int q = ps + 64 * k, r = q / WS, c = q - r * WS, gy = y0 + r - 1, gx = x0 + c - 1;
            v[k] = make_uint4(0u, 0u, 0u, 0u);
            if (q < HS * WS && gy >= 0 && gy < H && gx >= 0 && gx < W)
                v[k] = *reinterpret_cast<const uint4 *>(a4 + (ibase + (long)gy * W + gx) * CH + cg * 8);
        }
        __syncthreads();  // the previous tile's reads of sdy are done
        {
            const int r = tid / TW, c = tid - r * TW, gy = y0 + r, gx = x0 + c;
            const float d = (gy < H && gx < W) ? dy[ibase + (long)gy * W + gx] : 0.0f;
            sdy[tid] = d;
            dsum += d;  // db5: thread tid takes pixel tid of every tile
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < NQ; k++) {
            const int q = ps + 64 * k, qr = q / WS, qc = q - qr * WS;
            const uint32_t w4[4] = {v[k].x, v[k].y, v[k].z, v[k].w};
            float a[8];
#pragma unroll
            for (int e = 0; e < 8; e++) a[e] = from_bf16((uint16_t)(w4[e >> 1] >> (16 * (e & 1))));
#pragma unroll
            for (int t = 0; t < 9; t++) {
                const int pr = qr - t / 3, pc = qc - t % 3;  // output pixel of tap t
                const float d = (q < HS * WS && pr >= 0 && pr < TH && pc >= 0 && pc < TW) ? sdy[pr * TW + pc] : 0.0f;
#pragma unroll
                for (int e = 0; e < 8; e++) acc[t][e] = __builtin_fmaf(d, a[e], acc[t][e]);
            }
        }
    }
    float *out = part + (long)blockIdx.x * OUT_REC;
    block_sum_groups(acc, db, red, out, false);
    // db5: the threads' pixel sums in thread order
    __syncthreads();
    red[tid] = dsum;
    __syncthreads();
    if (tid == 0) {
        float s5 = 0.0f;
        for (int i = 0; i < 256; i++) s5 += red[i];
        out[9 * CH] = s5;
    }
}

// Sum of nparts partial records in block order (double) into the gradient
// vector.  kind 0: 32->32 record (t,co,ci)+db -> W[co][ci][t], b; kind 1 and 2
// records are already in parameter order.
// A block of 1024 threads takes 64 consecutive elements: wave w sums the
// records w, w + 16, ... for its lane's element (coalesced 256-byte rows, eight
// loads in flight per batch), then lane e of wave 0 adds the 16 wave sums in
// wave order -- a fixed order, so results are bitwise repeatable.
constexpr int RED_WAVES = 16;
__global__ void __launch_bounds__(1024) k_reduce(const float *__restrict__ part, int nparts, int rec, int kind,
                                                 float *__restrict__ gw, float *__restrict__ gb) {
    __shared__ double sw[RED_WAVES][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + lane;
    double s = 0.0;
    if (i < rec) {
        for (int k0 = wave; k0 < nparts; k0 += 8 * RED_WAVES) {
            float v[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int k = k0 + u * RED_WAVES;
                v[u] = k < nparts ? part[(long)k * rec + i] : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < 8; u++) s += (double)v[u];
        }
    }
    sw[wave][lane] = s;
    __syncthreads();
    if (wave != 0 || i >= rec) return;
    s = 0.0;
    for (int w = 0; w < RED_WAVES; w++) s += sw[w][lane];
    const float v = (float)s;
    if (kind == 0) {
        if (i < 9 * CH * CH) {
            const int t = i / (CH * CH), co = (i / CH) % CH, ci = i % CH;
            gw[(co * CH + ci) * 9 + t] = v;
        } else {
            gb[i - 9 * CH * CH] = v;
        }
    } else {
        const int nw = 9 * CH;
        if (i < nw) gw[i] = v;
        else gb[i - nw] = v;
    }
}

// Sum of per-block partials in order, times f: out[grp*stride + comp] =
// f * sum_i part[(grp*per + i)*stride + comp].
// One block of 256 threads per output: thread t adds the partials t, t + 256,
// ... in order (double, eight loads in flight per batch), then the 256 sums
// are added by a fixed pairwise tree -- bitwise repeatable.
__global__ void __launch_bounds__(256) k_sum_groups(const float *__restrict__ part, int groups, int per, int stride,
                                                    double f, float *__restrict__ out) {
    __shared__ double sd[256];
    const int k = blockIdx.x, tid = threadIdx.x;
    if (k >= groups * stride) return;
    const int grp = k / stride, comp = k - grp * stride;
    const float *src = part + (long)grp * per * stride + comp;
    double s = 0.0;
    for (int i0 = tid; i0 < per; i0 += 8 * 256) {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int i = i0 + u * 256;
            v[u] = i < per ? src[(long)i * stride] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < 8; u++) s += (double)v[u];
    }
    sd[tid] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (tid < h) sd[tid] += sd[tid + h];
        __syncthreads();
    }
    if (tid == 0) out[k] = (float)(sd[0] * f);
}

// Packed bf16 weights of the three 32->32 layers: forward [t][co][ci] and
// input-gradient [t][ci][co] = W[co][ci][8 - t].
__global__ void k_pack(const float *__restrict__ prm, uint16_t *__restrict__ packed) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * 9 * CH * CH) return;
    const int l = i / (9 * CH * CH), k = i - l * 9 * CH * CH;
    const int t = k / (CH * CH), a = (k / CH) % CH, b = k % CH;
    const float *w = prm + OFF_W2 + l * (9 * CH * CH + CH);
    uint16_t *fw = packed + (long)l * PACK_LAYER, *bw = fw + 9 * CH * CH;
    fw[k] = to_bf16(w[(a * CH + b) * 9 + t]);      // a = co, b = ci
    bw[k] = to_bf16(w[(b * CH + a) * 9 + 8 - t]);  // a = ci (output), b = co (input)
}

__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, int n, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] = p[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

int cnn_fail(int code, const char *msg) { return gr_set_error(code, msg); }
#define CNN_HIP(call)                                                       \
    do {                                                                    \
        hipError_t e_ = (call);                                             \
        if (e_ != hipSuccess) return cnn_fail(GR_EHIP, hipGetErrorString(e_)); \
    } while (0)

TileMap tiles_of(int H, int W) { return TileMap{(W + TW - 1) / TW, (H + TH - 1) / TH}; }

int check_dims(int n, int H, int W) {
    if (n < 1 || H < 1 || W < 1 || (long)n * H * W > (1L << 31) - 1) return cnn_fail(GR_EINVAL, "invalid image batch");
    const TileMap tm = tiles_of(H, W);
    if ((long)tm.tiles_x * tm.tiles_y * n > (1L << 31) - 1) return cnn_fail(GR_EINVAL, "too many tiles");
    return GR_OK;
}

}  // namespace

extern "C" {

long gr_cnn_param_count(void) { return N_PARAMS; }
long gr_cnn_packed_count(void) { return 3L * PACK_LAYER; }

size_t gr_cnn_work_bytes(int n, int H, int W) {
    const TileMap tm = tiles_of(H, W);
    const long ntiles = (long)tm.tiles_x * tm.tiles_y * n;
    const long npx = (long)n * H * W;
    const size_t up = 256;
    auto al = [&](size_t b) { return (b + up - 1) / up * up; };
    return al(sizeof(float) * npx)                         // yn
           + al(sizeof(float) * npx)                       // dy
           + 2 * al(sizeof(uint16_t) * npx * CH)           // dz ping-pong
           + al(sizeof(float) * (size_t)MAX_WG_BLOCKS * WG_REC)  // weight-gradient partials
           + al(sizeof(float) * 2 * (size_t)ntiles)        // forward statistics partials
           + al(sizeof(float) * 1024);                     // loss partials
}

int gr_cnn_pack(const float *params, unsigned short *packed, void *stream) {
    if (!params || !packed) return cnn_fail(GR_EINVAL, "null argument");
    const int n = 3 * 9 * CH * CH;
    hipLaunchKernelGGL(k_pack, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, params, packed);
    CNN_HIP(hipGetLastError());
    return GR_OK;
}

int gr_cnn_forward(const float *params, const unsigned short *packed, const float *x, int n, int H, int W, float scale,
                   unsigned short *a1, unsigned short *a2, unsigned short *a3, unsigned short *a4, float *y,
                   float *yn, float *slice_stats, void *work, void *stream) {
    if (!params || !packed || !x || !a1 || !a2 || !a3 || !a4 || !(scale > 0.0f)) return cnn_fail(GR_EINVAL, "null argument or scale <= 0");
    if (a1 == a2 || a2 == a3 || a3 == a4) return cnn_fail(GR_EINVAL, "consecutive activations must not alias");
    int rc = check_dims(n, H, W);
    if (rc) return rc;
    if (slice_stats && !work) return cnn_fail(GR_EINVAL, "slice statistics need the work buffer");
    hipStream_t s = (hipStream_t)stream;
    const TileMap tm = tiles_of(H, W);
    const int ntiles = tm.tiles_x * tm.tiles_y * n;
    const long npx = (long)n * H * W;
    const uint16_t *pk = reinterpret_cast<const uint16_t *>(packed);
    const int nconv = ntiles < MAX_CONV_BLOCKS ? ntiles : MAX_CONV_BLOCKS;
    // a 32 -> 32 forward layer, with the next-tile prefetch when an image's
    // activations fit a buffer view (< 2^31 bytes)
    // (its workgroups: exactly those resident at once, each looping over tiles)
    const bool pf = (long)H * W * CH * (long)sizeof(uint16_t) < (1L << 31);
    const TileMap tm16{tm.tiles_x, (H + 15) / 16};
    const int nt16 = tm16.tiles_x * tm16.tiles_y * n;
    int npf = 0;
    if (pf) {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_conv32<0, true, 16>, 256, 0) !=
                hipSuccess ||
            per_cu < 1)
            per_cu = 1;
        const long want = (long)per_cu * (gr_cu_count() > 0 ? gr_cu_count() : 256);
        npf = (int)std::min<long>(nt16, want);
    }
    auto conv32 = [&](const uint16_t *in, const uint16_t *wl, const float *b, uint16_t *out) {
        if (pf)
            hipLaunchKernelGGL((k_conv32<0, true, 16>), dim3(npf), dim3(256), 0, s, in, wl, b,
                               (const uint16_t *)nullptr, out, H, W, nt16, tm16);
        else
            hipLaunchKernelGGL((k_conv32<0, false>), dim3(nconv), dim3(256), 0, s, in, wl, b,
                               (const uint16_t *)nullptr, out, H, W, ntiles, tm);
    };
    // inference (a3 aliasing a1 and a4 aliasing a2: the caller keeps no
    // activation): the first layer fused into the second
    const bool fused = a3 == a1 && a4 == a2;
    if (fused) {
        const TileMap tm12{tm.tiles_x, (H + TH12 - 1) / TH12};
        const int nt12 = tm12.tiles_x * tm12.tiles_y * n;
        hipLaunchKernelGGL(k_conv12, dim3(nt12 < MAX_CONV_BLOCKS ? nt12 : MAX_CONV_BLOCKS), dim3(256), 0, s, x, scale,
                           params, pk, a2, H, W, nt12, tm12);
        conv32(a2, pk + PACK_LAYER, params + OFF_B3, a3);
        conv32(a3, pk + 2L * PACK_LAYER, params + OFF_B4, a4);
    } else {
        hipLaunchKernelGGL(k_conv_in, dim3(nconv), dim3(256), 0, s, x, scale, params, a1, H, W, ntiles, tm);
        uint16_t *acts[4] = {a1, a2, a3, a4};
        const int boff[3] = {OFF_B2, OFF_B3, OFF_B4};
        for (int l = 0; l < 3; l++) conv32(acts[l], pk + (long)l * PACK_LAYER, params + boff[l], acts[l + 1]);
    }
    float *stats = nullptr;
    if (slice_stats) {
        char *wb = (char *)work;
        auto al = [](size_t b) { return (b + 255) / 256 * 256; };
        wb += 2 * al(sizeof(float) * npx) + 2 * al(sizeof(uint16_t) * npx * CH) +
              al(sizeof(float) * (size_t)MAX_WG_BLOCKS * WG_REC);
        stats = (float *)wb;
    }
    if (pf) {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_conv_out<true>, 256, 0) !=
                hipSuccess ||
            per_cu < 1)
            per_cu = 1;
        const int nout = (int)std::min<long>(ntiles, (long)per_cu * (gr_cu_count() > 0 ? gr_cu_count() : 256));
        hipLaunchKernelGGL(k_conv_out<true>, dim3(nout), dim3(256), 0, s, a4, x, scale, params, y, yn, stats, H, W,
                           ntiles, tm);
    } else {
        hipLaunchKernelGGL(k_conv_out<false>, dim3(ntiles), dim3(256), 0, s, a4, x, scale, params, y, yn, stats, H, W,
                           ntiles, tm);
    }
    if (slice_stats) {
        const int per = tm.tiles_x * tm.tiles_y;
        hipLaunchKernelGGL(k_sum_groups, dim3(2 * n), dim3(256), 0, s, stats, n, per, 2, 1.0,
                           slice_stats);
    }
    CNN_HIP(hipGetLastError());
    return GR_OK;
}

size_t gr_cnn_infer_work_bytes(int n, int H, int W) {
    return sizeof(float) * 2 * (size_t)((W + FW - 1) / FW) * ((H + FH - 1) / FH) * (n > 0 ? n : 0);
}

int gr_cnn_infer(const float *params, const unsigned short *packed, const float *x, int n, int H, int W, float scale,
                 float *y, float *slice_stats, void *work, void *stream) {
    if (!params || !packed || !x || !y || !(scale > 0.0f)) return cnn_fail(GR_EINVAL, "null argument or scale <= 0");
    int rc = check_dims(n, H, W);
    if (rc) return rc;
    if (slice_stats && !work) return cnn_fail(GR_EINVAL, "slice statistics need the work buffer");
    hipStream_t s = (hipStream_t)stream;
    const int tx = (W + FW - 1) / FW, ty = (H + FH - 1) / FH;
    const long nblocks = (long)tx * ty * n;
    const long npx = (long)n * H * W;
    (void)npx;
    float *stats = slice_stats ? (float *)work : nullptr;
    CNN_HIP(gr_allow_lds(k_cnn_fused, FUSED_LDS));
    hipLaunchKernelGGL(k_cnn_fused, dim3((unsigned)nblocks), dim3(FH * FW), FUSED_LDS, s, x, scale, params,
                       reinterpret_cast<const uint16_t *>(packed), y, stats, H, W, tx, ty);
    if (slice_stats)
        hipLaunchKernelGGL(k_sum_groups, dim3(2 * n), dim3(256), 0, s, stats, n, tx * ty, 2, 1.0,
                           slice_stats);
    CNN_HIP(hipGetLastError());
    return GR_OK;
}

int gr_cnn_backward(const float *params, const unsigned short *packed, const float *x, const float *target, int n,
                    int H, int W, float scale, const unsigned short *a1, const unsigned short *a2,
                    const unsigned short *a3, const unsigned short *a4, const float *yn, float *grads, float *loss,
                    void *work, void *stream) {
    if (!params || !packed || !x || !target || !a1 || !a2 || !a3 || !a4 || !yn || !grads || !work || !(scale > 0.0f))
        return cnn_fail(GR_EINVAL, "null argument or scale <= 0");
    int rc = check_dims(n, H, W);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    const TileMap tm = tiles_of(H, W);
    const int ntiles = tm.tiles_x * tm.tiles_y * n;
    const long npx = (long)n * H * W;
    auto al = [](size_t b) { return (b + 255) / 256 * 256; };
    char *wb = (char *)work;
    wb += al(sizeof(float) * npx);  // yn slot (the caller's yn may live there)
    float *dy = (float *)wb;
    wb += al(sizeof(float) * npx);
    uint16_t *dzA = (uint16_t *)wb;
    wb += al(sizeof(uint16_t) * npx * CH);
    uint16_t *dzB = (uint16_t *)wb;
    wb += al(sizeof(uint16_t) * npx * CH);
    float *part = (float *)wb;
    wb += al(sizeof(float) * (size_t)MAX_WG_BLOCKS * WG_REC);
    wb += al(sizeof(float) * 2 * (size_t)ntiles);
    float *lpart = (float *)wb;
    const uint16_t *pk = reinterpret_cast<const uint16_t *>(packed);
    const int nwg = ntiles < MAX_WG_BLOCKS ? ntiles : MAX_WG_BLOCKS;
    const int nloss = (int)std::min<long>(1024, (npx + 255) / 256);

    hipLaunchKernelGGL(k_loss, dim3(nloss), dim3(256), 0, s, yn, target, scale, dy, lpart, npx);
    if (loss) hipLaunchKernelGGL(k_sum_groups, dim3(1), dim3(256), 0, s, lpart, 1, nloss, 1, 1.0 / (double)npx, loss);
    // layer 5
    hipLaunchKernelGGL(k_conv_out_wgrad, dim3(nwg), dim3(256), 0, s, dy, a4, part, H, W, ntiles, tm);
    hipLaunchKernelGGL(k_reduce, dim3((OUT_REC + 63) / 64), dim3(64 * RED_WAVES), 0, s, part, nwg, OUT_REC, 1,
                       grads + OFF_W5, grads + OFF_B5);
    hipLaunchKernelGGL(k_conv_out_dgrad, dim3((unsigned)((npx + 255) / 256)), dim3(256), 0, s, dy, a4, params, dzA,
                       n, H, W);
    // layers 4, 3, 2
    const unsigned short *acts[4] = {a1, a2, a3, a4};
    const int woff[3] = {OFF_W2, OFF_W3, OFF_W4}, boff[3] = {OFF_B2, OFF_B3, OFF_B4};
    CNN_HIP(gr_allow_lds(k_conv32_wgrad, WGRAD_LDS));
    uint16_t *dz = dzA, *dzn = dzB;
    for (int l = 2; l >= 0; l--) {  // layer l+2: input acts[l], output gradient dz
        const uint16_t *ain = reinterpret_cast<const uint16_t *>(acts[l]);
        hipLaunchKernelGGL(k_conv32_wgrad, dim3(nwg), dim3(256), WGRAD_LDS, s, ain, dz, part, H, W, ntiles, tm);
        hipLaunchKernelGGL(k_reduce, dim3((WG_REC + 63) / 64), dim3(64 * RED_WAVES), 0, s, part, nwg, WG_REC, 0,
                           grads + woff[l], grads + boff[l]);
        hipLaunchKernelGGL(k_conv32<1>, dim3(ntiles < MAX_CONV_BLOCKS ? ntiles : MAX_CONV_BLOCKS), dim3(256), 0, s,
                           dz, pk + (long)l * PACK_LAYER + 9 * CH * CH, (const float *)nullptr, ain, dzn, H, W,
                           ntiles, tm);
        std::swap(dz, dzn);
    }
    // layer 1
    hipLaunchKernelGGL(k_conv_in_wgrad, dim3(nwg), dim3(256), 0, s, x, scale, dz, part, H, W, ntiles, tm);
    hipLaunchKernelGGL(k_reduce, dim3((IN_REC + 63) / 64), dim3(64 * RED_WAVES), 0, s, part, nwg, IN_REC, 1,
                       grads + OFF_W1, grads + OFF_B1);
    CNN_HIP(hipGetLastError());
    return GR_OK;
}

int gr_cnn_adam(float *params, const float *grads, float *m, float *v, long count, int step, float lr, float beta1,
                float beta2, float eps, void *stream) {
    if (!params || !grads || !m || !v || count < 1 || count > (1L << 30) || step < 1)
        return cnn_fail(GR_EINVAL, "invalid Adam arguments");
    const float c1 = (float)(1.0 / (1.0 - pow((double)beta1, step)));
    const float c2 = (float)(1.0 / (1.0 - pow((double)beta2, step)));
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, (hipStream_t)stream, params,
                       grads, m, v, (int)count, lr, beta1, beta2, eps, c1, c2);
    CNN_HIP(hipGetLastError());
    return GR_OK;
}

}  // extern "C"
