// k_grid.hip -- K2: PSWF convolution gridding as a gather, no atomics.
//
// Replaces phase 1c of do_recon (gridrec/GridrecCPP/Gridrec5.2/grid.cpp:276-302),
// which scatters float `+=` into H.  Here every wavefront owns an 8 x 8 tile
// of grid cells, one cell per lane, and pulls the contributions of all
// samples whose convolution box covers its cell (gr_gather.hpp), so each
// cell is written once and the result does not depend on scheduling.
//
// The geometry of a tap -- box test, table indices, weight -- depends only on
// the plan, not on the slices: a wavefront therefore grids its tile for P
// slice pairs at once, evaluating each tap's weight once and applying it to
// the P pairs' spectra (P accumulators per lane).
//
// Which angles can reach a tile: the samples of angle n lie on the line
// through the grid centre with direction (COSE[n], SINE[n]); a sample reaches
// a cell only if it is within L/2 of it in both axes, so the line passes
// within dmax = 3.5*sqrt2 + (L/2)*sqrt2 (+ margin) of the tile centre.  For a
// tile at radius R and polar angle phi that is the set of line angles within
// asin(dmax/R) of phi modulo pi; angle keys atan2(SINE, COSE) mod pi are sorted
// on the host and the window is found by binary search, wave-uniform.  Each
// angle of the window is then tested exactly (perpendicular distance against
// the tile's half-width in the line's normal direction) and skipped when its
// line misses the tile.  Central tiles scan every angle; tiles beyond the
// largest sample radius are written as zeros.  Tiles are listed nearest to the
// centre first, so the heaviest work is dispatched first.
//
// Per angle, the samples that can reach the tile form one contiguous range of
// j for the direct term and one for the mirrored term; the wave stages the P
// pairs' spectra of that range in its own LDS slot (64 samples at a time) and
// each lane then runs its own exact candidate range against LDS.
//
// Accumulation: GR_ACCUM_PLAIN adds each tap's float product in float, in
// the reference's order.  GR_ACCUM_COMPENSATED sums the taps of up to
// kMergeAngles consecutive sample lines (a few taps each) with fused
// multiply-adds into a partial, and adds each partial into the cell's running
// sum with an error-free TwoSum (Knuth), carrying the rounding errors in a
// second float: the cell's value is the sum plus the carried error, rounded
// once.
//
// Output: Ht[pair][b][a] = H[a][b] (transposed, so K3's rows run over a).
#include "gr_fft.hpp"
#include "gr_gather.hpp"
#include "gr_internal.h"

#pragma clang fp contract(off)

namespace {

constexpr int kTile = 8;        // cells per tile side; one tile per wavefront
constexpr int kMergeAngles = 4;  // sample lines per compensated partial

struct GridArgs {
    const float2 *spec1, *spec2;
    float2 *ht;
    const float4 *ang;   // sorted: (COSE, SINE, 1/COSE or 0, 1/SINE or 0)
    const int *aord;     // sorted -> angle index
    const float *akey;   // sorted keys
    const unsigned *tiles;
    const float *wtbl;
    int ntiles, n_ang, half, lds_wtbl, pair_base, pair_end;
    GrGatherGeom g;
    float L2m, dmax, rmax;
};

// s + c += x with s the float sum and c the carried rounding errors.
__device__ __forceinline__ void two_sum_add(float &s, float &c, float x) {
    const float t = s + x;
    const float bp = t - s;
    c += (s - (t - bp)) + (x - bp);
    s = t;
}

// Close the open partial of every pair into its compensated sum.
template <int P>
__device__ __forceinline__ void merge(float4 (&acc)[P], float2 (&part)[P]) {
#pragma unroll
    for (int p = 0; p < P; p++) {
        two_sum_add(acc[p].x, acc[p].y, part[p].x);
        two_sum_add(acc[p].z, acc[p].w, part[p].y);
        part[p] = make_float2(0.0f, 0.0f);
    }
}

// Stage samples c0 .. c0+nj-1 of angle n for the P pairs, then add the taps
// of samples [j0, j1] into cell (qa, qb).  In a partial group (nvalid < P)
// the spare slots repeat the last valid pair: their accumulators are never
// stored, and unconditional loads keep the P loads issued back to back.
template <int ACCUM, int P>
__device__ __forceinline__ void chunk(const float2 *spec, long pair_stride, int nvalid, long row_off, int c0, int nj, int j0,
                                      int j1, float C, float S, float qa, float qb, const GrGatherGeom &g,
                                      const float *wtbl, float2 (*stage)[64], int lane, float4 (&acc)[P],
                                      float2 (&part)[P]) {
#pragma unroll
    for (int p = 0; p < P; p++)
        if (lane < nj) stage[p][lane] = spec[(p < nvalid ? p : nvalid - 1) * pair_stride + row_off + c0 + lane];
    __builtin_amdgcn_wave_barrier();
    for (int j = j0; j <= j1; j++) {
        float w;
        if (gr_tap(j, C, S, qa, qb, g, wtbl, w)) {
#pragma unroll
            for (int p = 0; p < P; p++) {
                const float2 v = stage[p][j - c0];
                if (ACCUM == GR_ACCUM_COMPENSATED) {
                    part[p].x = __builtin_fmaf(w, v.x, part[p].x);
                    part[p].y = __builtin_fmaf(w, v.y, part[p].y);
                } else {
                    acc[p].x += w * v.x;
                    acc[p].z += w * v.y;
                }
            }
        }
    }
    __builtin_amdgcn_wave_barrier();
}

template <int ACCUM, int P>
__global__ void __launch_bounds__(256) k_grid(GridArgs A) {
    extern __shared__ float smem[];
    float *s_wtbl = smem;
    for (int i = threadIdx.x; i <= A.g.ltbl; i += blockDim.x) s_wtbl[i] = A.wtbl[i];
    __syncthreads();

    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ti = blockIdx.x * (blockDim.x >> 6) + wave;
    if (ti >= A.ntiles) return;
    float2(*stage)[64] = reinterpret_cast<float2(*)[64]>(smem + A.lds_wtbl) + wave * P;
    const int pair0 = A.pair_base + blockIdx.y * P;
    const int nvalid = min(P, A.pair_end - pair0);
    const unsigned tt = A.tiles[ti];
    const int tx = (int)(tt & 0xffffu), ty = (int)(tt >> 16);
    const int a = tx * kTile + (lane & 7), b = ty * kTile + (lane >> 3);
    const int M = A.g.M;
    const float M2 = A.g.fM2;
    const float cu = (float)(tx * kTile) + 3.5f - M2, cv = (float)(ty * kTile) + 3.5f - M2;

    const bool valid = a >= 1 && b >= 1 && a < M && b < M;
    const float pa = (float)a, pb = (float)b, ma = (float)(M - a), mb = (float)(M - b);
    const float du = pa - M2, dv = pb - M2;
    const float reach = 3.5f + A.L2m;  // tile half-width plus box half-width
    float4 acc[P];   // (sum re, error re, sum im, error im)
    float2 part[P];  // compensated mode: this angle's partial
#pragma unroll
    for (int p = 0; p < P; p++) {
        acc[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        part[p] = make_float2(0.0f, 0.0f);
    }

    if (sqrtf(cu * cu + cv * cv) <= A.rmax) {
        int lines = 0;  // sample lines in the open partial
        int r0, r1, r2, r3;
        gr_angle_window(tx, ty, M2, A.dmax, A.akey, A.n_ang, r0, r1, r2, r3);
        const long pair_stride = (long)A.n_ang * A.half;
        const float2 *base1 = A.spec1 + pair0 * pair_stride;
        const float2 *base2 = A.spec2 + pair0 * pair_stride;
        for (int side = 0; side < 2; side++) {
            const int k0 = side ? r2 : r0, k1 = side ? r3 : r1;
            for (int k = k0; k < k1; k++) {
                const float4 cs = A.ang[k];
                // does the line pass through the tile grown by the box half-width?
                if (fabsf(cu * cs.y - cv * cs.x) > reach * (fabsf(cs.x) + fabsf(cs.y)) + 0.05f) continue;
                int td0, td1, tm0, tm1;
                gr_candidates_r(cu, cv, cs.z, cs.w, reach, A.g, td0, td1, tm0, tm1);
                if (td0 > td1 && tm0 > tm1) continue;
                int jd0, jd1, jm0, jm1;
                gr_candidates_r(du, dv, cs.z, cs.w, A.L2m, A.g, jd0, jd1, jm0, jm1);
                if (!valid) {
                    jd1 = jd0 - 1;
                    jm1 = jm0 - 1;
                }
                const long row_off = (long)A.aord[k] * A.half;
                for (int c0 = td0; c0 <= td1; c0 += 64)
                    chunk<ACCUM, P>(base1, pair_stride, nvalid, row_off, c0, min(64, td1 - c0 + 1), max(jd0, c0),
                                    min(min(jd1, td1), c0 + 63), cs.x, cs.y, pa, pb, A.g, s_wtbl, stage, lane, acc,
                                    part);
                for (int c0 = tm0; c0 <= tm1; c0 += 64)
                    chunk<ACCUM, P>(base2, pair_stride, nvalid, row_off, c0, min(64, tm1 - c0 + 1), max(jm0, c0),
                                    min(min(jm1, tm1), c0 + 63), cs.x, cs.y, ma, mb, A.g, s_wtbl, stage, lane, acc,
                                    part);
                if (ACCUM == GR_ACCUM_COMPENSATED && ++lines == kMergeAngles) {
                    merge<P>(acc, part);
                    lines = 0;
                }
            }
        }
        if (ACCUM == GR_ACCUM_COMPENSATED && lines) merge<P>(acc, part);
    }
    if (a < M && b < M) {
#pragma unroll
        for (int p = 0; p < P; p++) {
            if (p >= nvalid) break;
            const float4 s = acc[p];
            const float2 h = ACCUM == GR_ACCUM_COMPENSATED ? make_float2(s.x + s.y, s.z + s.w) : make_float2(s.x, s.z);
            A.ht[((long)(pair0 + p) * M + b) * M + a] = h;
        }
    }
}

template <int ACCUM, int P>
hipError_t launch(GridArgs A, int groups, int ntiles, size_t lds_base, int wpb, hipStream_t s) {
    const size_t lds = lds_base + sizeof(float2) * 64 * wpb * P;
    hipError_t e = gr_allow_lds(k_grid<ACCUM, P>, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_grid<ACCUM, P>), dim3((ntiles + wpb - 1) / wpb, groups), dim3(64 * wpb), lds, s, A);
    return hipGetLastError();
}

template <int ACCUM>
hipError_t launch_all(GridArgs A, int pairs, int ntiles, size_t lds_base, int wpb, hipStream_t s) {
    // groups of 8 pairs; a remainder of 5-7 rides in one more (partial) group
    // of 8, one of 3-4 in a group of 4, 2 and 1 in their own widths: a tap's
    // geometry costs the same for any width, so one partial wide group beats
    // several narrow launches
    const int full = pairs / 8, rem = pairs % 8;
    hipError_t e = hipSuccess;
    A.pair_end = pairs;
    A.pair_base = 0;
    if (full + (rem >= 5) > 0) e = launch<ACCUM, 8>(A, full + (rem >= 5), ntiles, lds_base, wpb, s);
    if (e != hipSuccess || rem == 0 || rem >= 5) return e;
    A.pair_base = full * 8;
    if (rem >= 3) return launch<ACCUM, 4>(A, 1, ntiles, lds_base, wpb, s);
    if (rem == 2) return launch<ACCUM, 2>(A, 1, ntiles, lds_base, wpb, s);
    return launch<ACCUM, 1>(A, 1, ntiles, lds_base, wpb, s);
}

}  // namespace

int gr_grid_tile_size() { return kTile; }

hipError_t gr_launch_grid(const GrGeom &g, const GrDevTables &t, int n_ang, long ltbl,
                          const float2 *spec1, const float2 *spec2, float2 *ht, int pairs,
                          int accum, hipStream_t s) {
    GridArgs A;
    A.spec1 = spec1;
    A.spec2 = spec2;
    A.ht = ht;
    A.ang = t.ang;
    A.aord = t.aord;
    A.akey = t.akey;
    A.tiles = t.tiles;
    A.wtbl = t.wtbl;
    A.ntiles = t.ntiles;
    A.n_ang = n_ang;
    A.half = (int)g.half;
    A.lds_wtbl = (int)((ltbl + 1 + 1) & ~1L);  // keep the staging area 8-byte aligned
    A.pair_base = 0;
    A.g.scale = g.scale;
    A.g.L2 = g.L2;
    A.g.tblspcg = g.tblspcg;
    A.g.fM2 = (float)(g.M >> 1);
    A.g.M = (int)g.M;
    A.g.half = (int)g.half;
    A.g.ltbl = (int)ltbl;
    A.L2m = g.L2 + 0.01f;
    A.dmax = 3.5f * 1.41421356f + A.L2m * 1.41421356f + 0.05f;
    // largest sample radius (|(COSE,SINE)| <= 1 + 1e-4 for the float recurrence)
    A.rmax = g.scale * (float)(g.half - 1) * 1.0001f + A.dmax + 0.5f;
    const size_t lds_base = sizeof(float) * A.lds_wtbl;
    const int wpb = g.grid_wpb == 1 || g.grid_wpb == 2 ? g.grid_wpb : 4;
    if (accum == GR_ACCUM_PLAIN) return launch_all<GR_ACCUM_PLAIN>(A, pairs, t.ntiles, lds_base, wpb, s);
    return launch_all<GR_ACCUM_COMPENSATED>(A, pairs, t.ntiles, lds_base, wpb, s);
}
