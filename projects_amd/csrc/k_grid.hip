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
// on the host and the window is found by binary search, wave-uniform.
// Central tiles (R <= dmax + 1) scan every angle.  Tiles are listed nearest to
// the centre first, so the heaviest work is dispatched first.
//
// Accumulation: GR_ACCUM_COMPENSATED keeps a TwoSum error term per component
// (the sum of float products is then exact to within one rounding of the
// final add); GR_ACCUM_PLAIN adds in float like the reference.
//
// Output: Ht[pair][b][a] = H[a][b] (transposed, so K3's rows run over a).
#include "gr_gather.hpp"
#include "gr_internal.h"

#pragma clang fp contract(off)

namespace {

struct GridArgs {
    const float2 *spec1, *spec2;
    float2 *ht;
    const float *cose, *sine, *akey;
    const int *aord;
    const unsigned *tiles;
    const float *wtbl;
    int ntiles, n_ang, half;
    GrGatherGeom g;
    float L2m, dmax;
};

template <int ACCUM>
__global__ void __launch_bounds__(256) k_grid(GridArgs A) {
    extern __shared__ float s_wtbl[];
    for (int i = threadIdx.x; i <= A.g.ltbl; i += blockDim.x) s_wtbl[i] = A.wtbl[i];
    __syncthreads();

    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ti = blockIdx.x * 4 + wave;
    if (ti >= A.ntiles) return;
    const int p = blockIdx.y;
    const unsigned tt = A.tiles[ti];
    const int tx = (int)(tt & 0xffffu), ty = (int)(tt >> 16);
    const int a = tx * 8 + (lane & 7), b = ty * 8 + (lane >> 3);
    const int M = A.g.M;
    const float M2 = A.g.fM2;

    // angle window of the tile (wave-uniform)
    int r0, r1, r2, r3;
    gr_angle_window(tx, ty, M2, A.dmax, A.akey, A.n_ang, r0, r1, r2, r3);

    const bool valid = a >= 1 && b >= 1 && a < M && b < M;
    const float pa = (float)a, pb = (float)b, ma = (float)(M - a), mb = (float)(M - b);
    const float du = pa - M2, dv = pb - M2;
    float sr = 0.0f, cr = 0.0f, si = 0.0f, ci = 0.0f;
    const float2 *sp1 = A.spec1 + (long)p * A.n_ang * A.half;
    const float2 *sp2 = A.spec2 + (long)p * A.n_ang * A.half;

    for (int part = 0; part < 2; part++) {
        const int k0 = part ? r2 : r0, k1 = part ? r3 : r1;
        for (int k = k0; k < k1; k++) {
            const int n = A.aord[k];
            const float C = A.cose[n], S = A.sine[n];
            int jd0, jd1, jm0, jm1;
            gr_candidates(du, dv, C, S, A.L2m, A.g, jd0, jd1, jm0, jm1);
            const int cd = valid ? max(0, jd1 - jd0 + 1) : 0;
            const int cm = valid ? max(0, jm1 - jm0 + 1) : 0;
            const float2 *row1 = sp1 + (long)n * A.half;
            const float2 *row2 = sp2 + (long)n * A.half;
            for (int i = 0; i < cd + cm; i++) {
                const bool mir = i >= cd;
                const int j = mir ? jm0 + (i - cd) : jd0 + i;
                float w;
                if (gr_tap(j, C, S, mir ? ma : pa, mir ? mb : pb, A.g, s_wtbl, w)) {
                    const float2 v = mir ? row2[j] : row1[j];
                    const float xr = w * v.x, xi = w * v.y;
                    if (ACCUM == GR_ACCUM_COMPENSATED) {
                        gr_two_sum(sr, cr, xr);
                        gr_two_sum(si, ci, xi);
                    } else {
                        sr += xr;
                        si += xi;
                    }
                }
            }
        }
    }
    if (a < M && b < M) A.ht[((long)p * M + b) * M + a] = make_float2(sr + cr, si + ci);
}

}  // namespace

hipError_t gr_launch_grid(const GrGeom &g, const GrDevTables &t, int n_ang, long ltbl,
                          const float2 *spec1, const float2 *spec2, float2 *ht, int pairs,
                          int accum, hipStream_t s) {
    GridArgs A;
    A.spec1 = spec1;
    A.spec2 = spec2;
    A.ht = ht;
    A.cose = t.cose;
    A.sine = t.sine;
    A.akey = t.akey;
    A.aord = t.aord;
    A.tiles = t.tiles;
    A.wtbl = t.wtbl;
    A.ntiles = t.ntiles;
    A.n_ang = n_ang;
    A.half = (int)g.half;
    A.g.scale = g.scale;
    A.g.L2 = g.L2;
    A.g.tblspcg = g.tblspcg;
    A.g.fM2 = (float)(g.M >> 1);
    A.g.M = (int)g.M;
    A.g.half = (int)g.half;
    A.g.ltbl = (int)ltbl;
    A.L2m = g.L2 + 0.01f;
    A.dmax = 3.5f * 1.41421356f + A.L2m * 1.41421356f + 0.05f;
    const dim3 grid((t.ntiles + 3) / 4, pairs), block(256);
    const size_t lds = sizeof(float) * (ltbl + 1);
    if (accum == GR_ACCUM_PLAIN)
        hipLaunchKernelGGL(k_grid<GR_ACCUM_PLAIN>, grid, block, lds, s, A);
    else
        hipLaunchKernelGGL(k_grid<GR_ACCUM_COMPENSATED>, grid, block, lds, s, A);
    return hipGetLastError();
}
