// k_grid.hip -- K2: PSWF convolution gridding as a gather, no atomics.
//
// Replaces phase 1c of do_recon (gridrec/GridrecCPP/Gridrec5.2/grid.cpp:276-302),
// which scatters float `+=` into H.  Here every wavefront owns an 8 x 8 tile
// of grid cells, one cell per lane, and pulls the contributions of all
// samples whose convolution box covers its cell (gr_gather.hpp), so each
// cell is written once and the result does not depend on scheduling.
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
// j for the direct term and one for the mirrored term.  The box test and the
// table weight separate by axis, so the wave evaluates them once per (sample,
// column) and once per (sample, row) -- 2 x 8 values per sample instead of 64
// -- and stages them with the sample spectra in its own LDS slot (64 samples
// at a time).  Each lane then runs its own candidate range against LDS: a tap
// is one float product wu*wv, exactly the reference's convolv, a multiply by
// the spectrum and the compensated add.
//
// Accumulation: GR_ACCUM_COMPENSATED keeps a Kahan correction per component;
// GR_ACCUM_PLAIN adds in float like the reference.
//
// Output: Ht[pair][b][a] = H[a][b] (transposed, so K3's rows run over a).
#include "gr_fft.hpp"
#include "gr_gather.hpp"
#include "gr_internal.h"

#pragma clang fp contract(off)

namespace {

struct GridArgs {
    const float2 *spec1, *spec2;
    float2 *ht;
    const float4 *ang;   // sorted: (COSE, SINE, 1/COSE or 0, 1/SINE or 0)
    const int *aord;     // sorted -> angle index
    const float *akey;   // sorted keys
    const unsigned *tiles;
    const float *wtbl;
    int ntiles, n_ang, half, lds_wtbl;
    GrGatherGeom g;
    float L2m, dmax, rmax;
};

template <int ACCUM>
__device__ __forceinline__ void accumulate(float &sr, float &cr, float &si, float &ci, float xr, float xi) {
    if (ACCUM == GR_ACCUM_COMPENSATED) {
        gr_kahan(sr, cr, xr);
        gr_kahan(si, ci, xi);
    } else {
        sr += xr;
        si += xi;
    }
}

// Per-wave LDS of one chunk of <= 64 samples of one angle: their spectra, the
// U-axis weight factor of each sample for the tile's 8 columns, the V-axis
// factor for its 8 rows.
struct WaveStage {
    float2 v[64];
    float wu[64 * 8];
    float wv[64 * 8];
};

// Samples c0 .. c0+nj-1 of one angle, for the tile whose cells (columns qa0 +
// da*i, rows qb0 + da*i, i < 8; da = +1 direct, -1 mirrored) the lane's cell
// (qa0 + da*ai, qb0 + da*bi) belongs to: stage the spectra and the per-axis
// weight factors, then accumulate the lane's taps over its candidate range
// [j0, j1].
template <int ACCUM>
__device__ __forceinline__ void chunk_taps(const float2 *row, int c0, int nj, int j0, int j1, float C, float S,
                                           float qa0, float qb0, float da, int ai, int bi, int lane,
                                           const GrGatherGeom &g, const float *wtbl, WaveStage *st, float &sr,
                                           float &cr, float &si, float &ci) {
    if (lane < nj) st->v[lane] = row[c0 + lane];
    for (int e = lane; e < nj * 8; e += 64) {
        float U, V;
        gr_sample_uv(c0 + (e >> 3), C, S, g, U, V);
        const float q = (float)(e & 7) * da;
        st->wu[e] = gr_axis_weight(U, qa0 + q, g, wtbl);
        st->wv[e] = gr_axis_weight(V, qb0 + q, g, wtbl);
    }
    __builtin_amdgcn_wave_barrier();
    const int lo = max(j0, c0), hi = min(j1, c0 + nj - 1);
    for (int j = lo; j <= hi; j++) {
        const int e = j - c0;
        const float w = st->wu[e * 8 + ai] * st->wv[e * 8 + bi];
        if (w != 0.0f) {
            const float2 v = st->v[e];
            accumulate<ACCUM>(sr, cr, si, ci, w * v.x, w * v.y);
        }
    }
    __builtin_amdgcn_wave_barrier();
}

template <int ACCUM>
__global__ void __launch_bounds__(256) k_grid(GridArgs A) {
    extern __shared__ float smem[];
    float *s_wtbl = smem;
    for (int i = threadIdx.x; i <= A.g.ltbl; i += blockDim.x) s_wtbl[i] = A.wtbl[i];
    __syncthreads();

    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ti = blockIdx.x * 4 + wave;
    if (ti >= A.ntiles) return;
    WaveStage *st = reinterpret_cast<WaveStage *>(smem + A.lds_wtbl) + wave;
    const int p = blockIdx.y;
    const unsigned tt = A.tiles[ti];
    const int tx = (int)(tt & 0xffffu), ty = (int)(tt >> 16);
    const int a = tx * 8 + (lane & 7), b = ty * 8 + (lane >> 3);
    const int M = A.g.M;
    const float M2 = A.g.fM2;
    const float cu = (float)(tx * 8) + 3.5f - M2, cv = (float)(ty * 8) + 3.5f - M2;

    const int ai = lane & 7, bi = lane >> 3;
    const float du = (float)a - M2, dv = (float)b - M2;
    const float ta = (float)(tx * 8), tb = (float)(ty * 8);   // first column / row of the tile
    const float reach = 3.5f + A.L2m;  // tile half-width plus box half-width
    float sr = 0.0f, cr = 0.0f, si = 0.0f, ci = 0.0f;

    if (sqrtf(cu * cu + cv * cv) <= A.rmax) {
        int r0, r1, r2, r3;
        gr_angle_window(tx, ty, M2, A.dmax, A.akey, A.n_ang, r0, r1, r2, r3);
        const float2 *base1 = A.spec1 + (long)p * A.n_ang * A.half;
        const float2 *base2 = A.spec2 + (long)p * A.n_ang * A.half;
        for (int part = 0; part < 2; part++) {
            const int k0 = part ? r2 : r0, k1 = part ? r3 : r1;
            for (int k = k0; k < k1; k++) {
                const float4 cs = A.ang[k];
                // does the line pass through the tile grown by the box half-width?
                if (fabsf(cu * cs.y - cv * cs.x) > reach * (fabsf(cs.x) + fabsf(cs.y)) + 0.05f) continue;
                int td0, td1, tm0, tm1;
                gr_candidates_r(cu, cv, cs.z, cs.w, reach, A.g, td0, td1, tm0, tm1);
                if (td0 > td1 && tm0 > tm1) continue;
                int jd0, jd1, jm0, jm1;
                gr_candidates_r(du, dv, cs.z, cs.w, A.L2m, A.g, jd0, jd1, jm0, jm1);
                const int n = A.aord[k];
                const float2 *row1 = base1 + (long)n * A.half;
                const float2 *row2 = base2 + (long)n * A.half;
                // direct term: cell (a, b) itself
                for (int c0 = td0; c0 <= td1; c0 += 64)
                    chunk_taps<ACCUM>(row1, c0, min(64, td1 - c0 + 1), jd0, jd1, cs.x, cs.y, ta, tb, 1.0f, ai, bi,
                                      lane, A.g, s_wtbl, st, sr, cr, si, ci);
                // mirrored term: the deposit at (M-a, M-b) of grid.cpp:299-300
                for (int c0 = tm0; c0 <= tm1; c0 += 64)
                    chunk_taps<ACCUM>(row2, c0, min(64, tm1 - c0 + 1), jm0, jm1, cs.x, cs.y, (float)M - ta,
                                      (float)M - tb, -1.0f, ai, bi, lane, A.g, s_wtbl, st, sr, cr, si, ci);
            }
        }
    }
    const float hr = ACCUM == GR_ACCUM_COMPENSATED ? sr - cr : sr;
    const float hi = ACCUM == GR_ACCUM_COMPENSATED ? si - ci : si;
    if (a < M && b < M) A.ht[((long)p * M + b) * M + a] = make_float2(hr, hi);
}

}  // namespace

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
    const dim3 grid((t.ntiles + 3) / 4, pairs), block(256);
    const size_t lds = sizeof(float) * A.lds_wtbl + 4 * sizeof(WaveStage);
    hipError_t e;
    if (accum == GR_ACCUM_PLAIN) {
        if ((e = gr_allow_lds(k_grid<GR_ACCUM_PLAIN>, lds)) != hipSuccess) return e;
        hipLaunchKernelGGL(k_grid<GR_ACCUM_PLAIN>, grid, block, lds, s, A);
    } else {
        if ((e = gr_allow_lds(k_grid<GR_ACCUM_COMPENSATED>, lds)) != hipSuccess) return e;
        hipLaunchKernelGGL(k_grid<GR_ACCUM_COMPENSATED>, grid, block, lds, s, A);
    }
    return hipGetLastError();
}
