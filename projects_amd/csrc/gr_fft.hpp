// gr_fft.hpp -- power-of-two complex FFT of one row held in LDS.
//
// Stockham autosort formulation (N. K. Govindaraju et al., "High performance
// discrete Fourier transforms on graphics processors", SC'08): passes of
// radix 16 (a 4 x 4 decomposition in registers) while 16 divides what is left,
// then radix 4, then one radix-2 pass when needed (max_radix 4 skips the
// radix-16 passes: a launch-tuning choice); output in natural order.
// Each pass reads all of a thread's butterfly inputs into registers, crosses a
// barrier, and writes the outputs back into the same LDS row, so one buffer of
// N complex values suffices.  SIGN = +1 computes sum_k x_k e^{+2 pi i jk/N},
// SIGN = -1 the e^{-...} transform; neither is normalised (the conventions of
// the reference's four1 with isign=+1 and fourn with isign=-1,
// gridrec/GridrecCPP/Gridrec5.2/fftgr.cpp:8-120).
//
// Twiddles come pass-ordered (gr_twp_len, built on the host from the table
// w[m] = e^{SIGN 2 pi i m / N} rounded from double): for each pass with
// Ns > 1, in pass order, the entries [r-1][k] = w[r k N/(Ns R)] for
// r = 1 .. R-1 and k < Ns (R the pass's radix; the radix-2 pass: [k] = w[k]),
// first the passes of max_radix 16, then those of max_radix 4.  A pass's
// lanes then read consecutive entries -- one or a few cache lines per load
// instead of up to one per lane -- and multiply by the very same values.
// MAXB bounds the radix-4 butterflies per thread: N/4 <= MAXB * nthreads.
//
// LDS layout: element i of the row lives at gr_lds_idx(i) = i XOR ((i >> 4)
// mod 32), a permutation of each aligned block of 32 slots (the identity
// below 32 points).  The first radix-16 pass writes 16 consecutive outputs
// per thread, so in natural order the lanes of a store sit 128 bytes apart,
// all in the same banks; the swizzle spreads every pass's loads and stores
// over the banks (checked against 32- and 64-bank models for 2048 .. 8192
// points) without extra LDS.  Callers that fill the row or read the result
// use the same mapping.
#pragma once

#include <hip/hip_runtime.h>

__host__ __device__ __forceinline__ int gr_lds_idx(int i) { return i ^ ((i >> 4) & 31); }

// Raw buffer view of the `bytes` bytes at base (<= 2^31): a load at a byte
// offset at or past the end returns zeros and touches no memory.  Loads with
// a lane-dependent bound thus need no branch: every one of a thread's loads
// is issued before the first wait (a guarded load whose result is copied
// inside its branch waits for it there).
struct GrBuf {
#if defined(__HIP_DEVICE_COMPILE__)
    __amdgpu_buffer_rsrc_t r;
#endif
};
constexpr unsigned kGrBufNone = 0x80000000u;  // an offset past every view's end
__device__ __forceinline__ GrBuf gr_buf(const void *base, unsigned bytes) {
    GrBuf b;
#if defined(__HIP_DEVICE_COMPILE__)
    b.r = __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), 0, (int)bytes, 0x00020000);
#else
    (void)base;
    (void)bytes;
#endif
    return b;
}
// Four bytes at byte offset off (zero past the end).
__device__ __forceinline__ float gr_buf_ldf(GrBuf b, unsigned off) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(b.r, (int)off, 0, 0));
#else
    (void)b;
    (void)off;
    return 0.0f;
#endif
}
// Sixteen bytes at byte offset off (zeros past the end).
__device__ __forceinline__ uint4 gr_buf_ld16(GrBuf b, unsigned off) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(b.r, (int)off, 0, 0));
#else
    (void)b;
    (void)off;
    return make_uint4(0u, 0u, 0u, 0u);
#endif
}
// soff: a wave-uniform part of the offset (a scalar register), added to off.
__device__ __forceinline__ float2 gr_buf_ld2(GrBuf b, unsigned off, unsigned soff = 0) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(b.r, (int)off, (int)soff, 0));
#else
    (void)b;
    (void)off;
    (void)soff;
    return make_float2(0.0f, 0.0f);
#endif
}

// Entries of the pass-ordered twiddles of an N-point transform whose passes
// go up to max_radix (16 or 4).
__host__ __device__ __forceinline__ long gr_twp_len(long N, int max_radix) {
    long Ns = 1, len = 0;
    while (max_radix >= 16 && Ns * 16 <= N) {
        if (Ns > 1) len += 15 * Ns;
        Ns *= 16;
    }
    while (Ns * 4 <= N) {
        if (Ns > 1) len += 3 * Ns;
        Ns *= 4;
    }
    if (Ns < N) len += Ns;
    return len;
}
__host__ __device__ __forceinline__ long gr_lds_len(long N) { return N; }

// Complex product with one fused multiply-add per component (one rounding
// fewer than product-then-add).  The fusion is written out rather than left to
// the compiler's contraction so that every instantiation of a transform (any
// column width, thread count or radix bound) rounds alike: the same input
// gives the same bits whichever kernel shape computes it.  It takes a third of
// the complex products' VALU instructions away from kernels that the VALU
// bounds.
__device__ __forceinline__ float2 gr_cmul(float2 a, float2 b) {
    const float yy = a.y * b.y, yx = a.y * b.x;
    return make_float2(__builtin_fmaf(a.x, b.x, -yy), __builtin_fmaf(a.x, b.y, yx));
}

// 4-point DFT in place with root SIGN*i.
template <int SIGN>
__device__ __forceinline__ void gr_dft4(float2 &v0, float2 &v1, float2 &v2, float2 &v3) {
    const float2 a0 = make_float2(v0.x + v2.x, v0.y + v2.y);
    const float2 a1 = make_float2(v0.x - v2.x, v0.y - v2.y);
    const float2 a2 = make_float2(v1.x + v3.x, v1.y + v3.y);
    const float2 d = make_float2(v1.x - v3.x, v1.y - v3.y);
    // a3 = d * (SIGN i)
    const float2 a3 = SIGN > 0 ? make_float2(-d.y, d.x) : make_float2(d.y, -d.x);
    v0 = make_float2(a0.x + a2.x, a0.y + a2.y);
    v2 = make_float2(a0.x - a2.x, a0.y - a2.y);
    v1 = make_float2(a1.x + a3.x, a1.y + a3.y);
    v3 = make_float2(a1.x - a3.x, a1.y - a3.y);
}

// 16-point DFT in place with root w = e^{SIGN 2 pi i/16}: slots r = 4a + b
// hold x[r]; a 4-point DFT over a for each b, the twiddle w^(bc), a 4-point
// DFT over b for each c; X[c + 4d] ends in slot 4c + d.
template <int SIGN>
__device__ __forceinline__ void gr_dft16(float2 (&v)[16]) {
    constexpr float C8 = 0.923879532511286756f, S8 = 0.382683432365089772f, R2 = 0.707106781186547524f;
    const float s = (float)SIGN;
#pragma unroll
    for (int b = 0; b < 4; b++) gr_dft4<SIGN>(v[b], v[4 + b], v[8 + b], v[12 + b]);
    v[5] = gr_cmul(v[5], make_float2(C8, s * S8));      // w^1
    v[6] = gr_cmul(v[6], make_float2(R2, s * R2));      // w^2
    v[7] = gr_cmul(v[7], make_float2(S8, s * C8));      // w^3
    v[9] = gr_cmul(v[9], make_float2(R2, s * R2));      // w^2
    v[10] = make_float2(-s * v[10].y, s * v[10].x);     // w^4 = SIGN i
    v[11] = gr_cmul(v[11], make_float2(-R2, s * R2));   // w^6
    v[13] = gr_cmul(v[13], make_float2(S8, s * C8));    // w^3
    v[14] = gr_cmul(v[14], make_float2(-R2, s * R2));   // w^6
    v[15] = gr_cmul(v[15], make_float2(-C8, -s * S8));  // w^9
#pragma unroll
    for (int c = 0; c < 4; c++) gr_dft4<SIGN>(v[4 * c], v[4 * c + 1], v[4 * c + 2], v[4 * c + 3]);
}

// No output hook: the transform ends in buf.
struct GrFftToLds {
    static constexpr bool enabled = false;
    __device__ void prep(int, int) {}
    __device__ void operator()(int, float2, int) const {}
};

// The first radix-16 pass of butterfly j (Ns = 1, no twiddles) on values
// already in registers, v[r] = x[j + r N/16], and its outputs stored to buf
// where gr_lds_fft's own first pass puts them.
template <int SIGN>
__device__ __forceinline__ void gr_fft_first_pass16(float2 *buf, float2 (&v)[16], int j) {
    gr_dft16<SIGN>(v);
#pragma unroll
    for (int q = 0; q < 16; q++) buf[gr_lds_idx((j << 4) + q)] = v[4 * (q & 3) + (q >> 2)];
}

// Returns true when an enabled output hook `out` received the result instead
// of buf: when the last pass is a radix-16 one, out(i, X[i], s) is called for
// every output i straight from the registers of that pass (no LDS store, no
// barrier), s < 16 MAXB16 the output's slot among the thread's outputs; before
// any of these calls, out.prep(i, s) is called for every one of them (the
// hook issues the loads its outputs need there, all in flight together).
// Otherwise the result is in buf as usual and false is returned.  Every output
// i is handed over exactly once; for a fixed pass slot the threads hand over
// consecutive i.  first_done: the caller ran the first radix-16 pass
// (gr_fft_first_pass16 for every butterfly, then a barrier); only with
// max_radix 16 and N >= 16.  TW_LDS: tw is a copy of the radix-16 part of
// the table in LDS (only with max_radix 16 and N a power of 16).
template <int SIGN, int MAXB, class Out = GrFftToLds, bool TW_LDS = false>
__device__ __forceinline__ bool gr_lds_fft(float2 *buf, int N, const float2 *__restrict__ tw, int tid, int nthr,
                                           int max_radix = 16, Out out = Out(), bool first_done = false) {
    int Ns = first_done ? 16 : 1;  // the Ns = 1 pass reads no twiddles: tofs stays at the table's start
    // twiddles by buffer loads: one lane offset (8 k) for all of a pass's
    // loads, the entry's row in the scalar offset, tofs = the pass's table
    const GrBuf twb = gr_buf(tw, (unsigned)((gr_twp_len(N, 16) + gr_twp_len(N, 4)) * sizeof(float2)));
    unsigned tofs = max_radix >= 16 ? 0u : (unsigned)(gr_twp_len(N, 16) * sizeof(float2));
    constexpr int MAXB16 = (MAXB + 3) / 4;  // radix-16 butterflies per thread
    const int nb16 = N >> 4;
    while (max_radix >= 16 && Ns * 16 <= N) {
        float2 v[MAXB16][16];
#pragma unroll
        for (int c = 0; c < MAXB16; c++) {
            const int j = tid + c * nthr;
            if (j < nb16) {
                const int k = j & (Ns - 1);
                // the twiddles into registers of their own first: all fifteen
                // loads in flight together (loaded next to their products the
                // compiler reuses one register pair, one round trip each)
                float2 w[15];
                if (Ns > 1) {
#pragma unroll
                    for (int r = 1; r < 16; r++)
                        w[r - 1] = TW_LDS ? tw[tofs / 8u + (unsigned)((r - 1) * Ns + k)]
                                          : gr_buf_ld2(twb, 8u * (unsigned)k, tofs + 8u * (unsigned)((r - 1) * Ns));
                }
#pragma unroll
                for (int r = 0; r < 16; r++) v[c][r] = buf[gr_lds_idx(j + r * nb16)];
                if (Ns > 1) {
#pragma unroll
                    for (int r = 1; r < 16; r++) v[c][r] = gr_cmul(v[c][r], w[r - 1]);
                }
                gr_dft16<SIGN>(v[c]);
            }
        }
        if (Out::enabled && Ns * 16 == N) {  // the last pass: outputs straight from registers
#pragma unroll
            for (int c = 0; c < MAXB16; c++) {
                const int j = tid + c * nthr, k = j & (Ns - 1);
                const int base = ((j - k) << 4) + k;
#pragma unroll
                for (int q = 0; q < 16; q++) out.prep(j < nb16 ? base + q * Ns : 0, 16 * c + q);
            }
#pragma unroll
            for (int c = 0; c < MAXB16; c++) {
                const int j = tid + c * nthr;
                if (j < nb16) {
                    const int k = j & (Ns - 1);
                    const int base = ((j - k) << 4) + k;
#pragma unroll
                    for (int q = 0; q < 16; q++) out(base + q * Ns, v[c][4 * (q & 3) + (q >> 2)], 16 * c + q);
                }
            }
            return true;
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < MAXB16; c++) {
            const int j = tid + c * nthr;
            if (j < nb16) {
                const int k = j & (Ns - 1);
                const int base = ((j - k) << 4) + k;
#pragma unroll
                for (int q = 0; q < 16; q++) buf[gr_lds_idx(base + q * Ns)] = v[c][4 * (q & 3) + (q >> 2)];
            }
        }
        __syncthreads();
        if (Ns > 1) tofs += 8u * 15u * (unsigned)Ns;
        Ns <<= 4;
    }
    const int nb4 = N >> 2;
    while (Ns * 4 <= N) {
        float2 v[MAXB][4];
#pragma unroll
        for (int c = 0; c < MAXB; c++) {
            const int j = tid + c * nthr;
            if (j < nb4) {
                const int k = j & (Ns - 1);
                float2 w[3];
                if (Ns > 1) {
#pragma unroll
                    for (int r = 0; r < 3; r++) w[r] = gr_buf_ld2(twb, 8u * (unsigned)k, tofs + 8u * (unsigned)(r * Ns));
                }
#pragma unroll
                for (int r = 0; r < 4; r++) v[c][r] = buf[gr_lds_idx(j + r * nb4)];
                if (Ns > 1) {
#pragma unroll
                    for (int r = 1; r < 4; r++) v[c][r] = gr_cmul(v[c][r], w[r - 1]);
                }
                gr_dft4<SIGN>(v[c][0], v[c][1], v[c][2], v[c][3]);
            }
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < MAXB; c++) {
            const int j = tid + c * nthr;
            if (j < nb4) {
                const int k = j & (Ns - 1);
                const int base = ((j - k) << 2) + k;
#pragma unroll
                for (int r = 0; r < 4; r++) buf[gr_lds_idx(base + r * Ns)] = v[c][r];
            }
        }
        __syncthreads();
        if (Ns > 1) tofs += 8u * 3u * (unsigned)Ns;
        Ns <<= 2;
    }
    if (Ns < N) {  // N = 2 Ns: one radix-2 pass, twiddle e^{SIGN 2 pi i k/N}
        const int nb2 = N >> 1;
        float2 a[2 * MAXB], b[2 * MAXB];
#pragma unroll
        for (int c = 0; c < 2 * MAXB; c++) {
            const int j = tid + c * nthr;
            if (j < nb2) {
                const int k = j & (Ns - 1);
                const float2 x0 = buf[gr_lds_idx(j)];
                const float2 x1 = gr_cmul(buf[gr_lds_idx(j + nb2)], gr_buf_ld2(twb, 8u * (unsigned)k, tofs));
                a[c] = make_float2(x0.x + x1.x, x0.y + x1.y);
                b[c] = make_float2(x0.x - x1.x, x0.y - x1.y);
            }
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < 2 * MAXB; c++) {
            const int j = tid + c * nthr;
            if (j < nb2) {
                const int k = j & (Ns - 1);
                const int base = ((j - k) << 1) + k;
                buf[gr_lds_idx(base)] = a[c];
                buf[gr_lds_idx(base + Ns)] = b[c];
            }
        }
        __syncthreads();
    }
    return false;
}

// Threads per FFT block and the matching MAXB for a row of N points.
static inline int gr_fft_threads(long N) { return N >= 8192 ? 512 : (N >= 1024 ? 256 : (N >= 256 ? (int)(N / 4) : 64)); }

// Allow a kernel more than 64 KiB of dynamic LDS (gfx950 has 160 KiB per CU).
template <typename K>
static inline hipError_t gr_allow_lds(K kernel, size_t bytes) {
    if (bytes <= 65536) return hipSuccess;
    return hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}
