// k_grid.hip -- K2: PSWF convolution gridding, tile-local, no atomics.
//
// Replaces phase 1c of do_recon (gridrec/GridrecCPP/Gridrec5.2/grid.cpp:276-302),
// which scatters float `+=` into the whole grid H.  Here every wavefront owns
// a 16 x 16 tile of H as a private accumulator in LDS and performs exactly the
// reference's scatter restricted to that tile; each cell of H is written to
// HBM once, and the result does not depend on scheduling.
//
// Which angles can reach a tile: the samples of angle n lie on the line
// through the grid centre with direction (COSE[n], SINE[n]); a sample touches a
// cell only if it is within L/2 of it in both axes, so the line passes within
// dmax = (7.5 + L/2) * sqrt2 (+ margin) of the tile centre.  For a tile at
// radius R and polar angle phi that is the set of line angles within
// asin(dmax/R) of phi modulo pi; angle keys atan2(SINE, COSE) mod pi are sorted
// on the host and the window is found by binary search.  Each angle of the
// window is then tested exactly (perpendicular distance against the tile's
// half-width in the line's normal direction) and skipped when its line misses
// the tile; its samples that can touch the tile form one contiguous range of
// j for the direct deposit and one for the mirrored deposit
// H[M-iu][M-iv] (grid.cpp:299-300).  Central tiles scan every angle; tiles
// beyond the largest sample radius are written as zeros.  Tiles are listed
// nearest to the centre first, so the heaviest work is dispatched first.
//
// Lanes take (sample, box cell) pairs: G lanes per sample (G >= bw^2, bw =
// floor(L)+1 cells per box side), 64/G samples per pass.  The samples of one
// pass are K steps of j apart, K chosen so their boxes cannot overlap, so no
// two lanes of a pass update the same cell; passes run in a fixed order
// (angle, direct/mirrored, residue class of j mod K, j), so the sum order is
// fixed too.  Each tap evaluates the reference's float box bounds and table
// indices (gr_box_tap) and adds w * Cdata into the tile's LDS cell:
// GR_ACCUM_COMPENSATED keeps a Kahan correction per component,
// GR_ACCUM_PLAIN adds in float like the reference.
//
// Output: Ht[pair][b][a] = H[a][b] (transposed, so K3's rows run over a).
#include "gr_fft.hpp"
#include "gr_gather.hpp"
#include "gr_internal.h"

#pragma clang fp contract(off)

namespace {

constexpr int kTile = 16;  // cells per tile side; one tile per wavefront

struct GridArgs {
    const float2 *spec1, *spec2;
    float2 *ht;
    const float4 *ang;   // sorted: (COSE, SINE, 1/COSE or 0, 1/SINE or 0)
    const int *aord;     // sorted -> angle index
    const float *akey;   // sorted keys
    const unsigned *tiles;
    const float *wtbl;
    int ntiles, n_ang, half, lds_wtbl;
    int K;               // conflict-free step of j
    int bw, G, SP;       // box width, lanes per sample, samples per pass
    GrGatherGeom g;
    float L2m, dmax, rmax, reach;
};

// LDS of one wavefront: its tile's accumulators (sum and Kahan correction of
// the real and imaginary parts, cell = (b-b0)*16 + (a-a0)) and the spectra of
// up to 64 samples of the current angle.
struct WaveTile {
    float4 acc[kTile * kTile];
    float2 v[64];
};

template <int ACCUM>
__device__ __forceinline__ void add_tap(float4 &c, float xr, float xi) {
    if (ACCUM == GR_ACCUM_COMPENSATED) {
        gr_kahan(c.x, c.y, xr);
        gr_kahan(c.z, c.w, xi);
    } else {
        c.x += xr;
        c.z += xi;
    }
}

// Deposit samples c0 .. c0+nj-1 (spectra staged in wt->v) into the tile.
// mirror = 0: cell (iu, iv); mirror = 1: cell (M-iu, M-iv).
template <int ACCUM>
__device__ __forceinline__ void deposit_chunk(const GridArgs &A, WaveTile *wt, const float *wtbl, int c0, int nj,
                                              float C, float S, int mirror, int a0, int b0, int lane) {
    const int M = A.g.M;
    const int slot = lane / A.G, sub = lane - slot * A.G;
    for (int r = 0; r < A.K && r < nj; r++) {
        const int cnt = (nj - r + A.K - 1) / A.K;  // samples r, r+K, ... of the chunk
        for (int q0 = 0; q0 < cnt; q0 += A.SP) {
            const int q = q0 + slot;
            if (q < cnt) {
                const int jl = r + A.K * q;
                const float2 v = wt->v[jl];
                for (int t = sub; t < A.bw * A.bw; t += A.G) {
                    const int dv = t / A.bw, du = t - dv * A.bw;
                    int iu, iv;
                    float w;
                    if (gr_box_tap(c0 + jl, C, S, du, dv, A.g, wtbl, iu, iv, w)) {
                        const int ca = (mirror ? M - iu : iu) - a0, cb = (mirror ? M - iv : iv) - b0;
                        if ((unsigned)ca < (unsigned)kTile && (unsigned)cb < (unsigned)kTile) {
                            float4 &cell = wt->acc[cb * kTile + ca];
                            float4 c = cell;
                            add_tap<ACCUM>(c, w * v.x, w * v.y);
                            cell = c;
                        }
                    }
                }
            }
            __builtin_amdgcn_wave_barrier();  // next pass may touch the same cells
        }
    }
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
    WaveTile *wt = reinterpret_cast<WaveTile *>(smem + A.lds_wtbl) + wave;
    const int p = blockIdx.y;
    const unsigned tt = A.tiles[ti];
    const int a0 = (int)(tt & 0xffffu) * kTile, b0 = (int)(tt >> 16) * kTile;
    const int M = A.g.M;
    const float M2 = A.g.fM2;
    const float half_tile = 0.5f * (float)(kTile - 1);
    const float cu = (float)a0 + half_tile - M2, cv = (float)b0 + half_tile - M2;

    for (int c = lane; c < kTile * kTile; c += 64) wt->acc[c] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    __builtin_amdgcn_wave_barrier();

    if (sqrtf(cu * cu + cv * cv) <= A.rmax) {
        int r0, r1, r2, r3;
        gr_angle_window_at(cu, cv, A.dmax, A.akey, A.n_ang, r0, r1, r2, r3);
        const float2 *base1 = A.spec1 + (long)p * A.n_ang * A.half;
        const float2 *base2 = A.spec2 + (long)p * A.n_ang * A.half;
        for (int part = 0; part < 2; part++) {
            const int k0 = part ? r2 : r0, k1 = part ? r3 : r1;
            for (int k = k0; k < k1; k++) {
                const float4 cs = A.ang[k];
                // does the line pass through the tile grown by the box half-width?
                if (fabsf(cu * cs.y - cv * cs.x) > A.reach * (fabsf(cs.x) + fabsf(cs.y)) + 0.05f) continue;
                int td0, td1, tm0, tm1;
                gr_candidates_r(cu, cv, cs.z, cs.w, A.reach, A.g, td0, td1, tm0, tm1);
                if (td0 > td1 && tm0 > tm1) continue;
                const int n = A.aord[k];
                for (int mirror = 0; mirror < 2; mirror++) {
                    const int t0 = mirror ? tm0 : td0, t1 = mirror ? tm1 : td1;
                    const float2 *row = (mirror ? base2 : base1) + (long)n * A.half;
                    for (int c0 = t0; c0 <= t1; c0 += 64) {
                        const int nj = min(64, t1 - c0 + 1);
                        if (lane < nj) wt->v[lane] = row[c0 + lane];
                        __builtin_amdgcn_wave_barrier();
                        deposit_chunk<ACCUM>(A, wt, s_wtbl, c0, nj, cs.x, cs.y, mirror, a0, b0, lane);
                    }
                }
            }
        }
    }
    __builtin_amdgcn_wave_barrier();
    float2 *dst = A.ht + (long)p * M * M;
    for (int c = lane; c < kTile * kTile; c += 64) {
        const int ca = c & (kTile - 1), cb = c / kTile;
        const float4 s = wt->acc[c];
        const float2 h = ACCUM == GR_ACCUM_COMPENSATED ? make_float2(s.x - s.y, s.z - s.w) : make_float2(s.x, s.z);
        if (a0 + ca < M && b0 + cb < M) dst[(long)(b0 + cb) * M + a0 + ca] = h;
    }
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
    A.lds_wtbl = (int)((ltbl + 1 + 3) & ~3L);  // keep the wave tiles 16-byte aligned
    A.g.scale = g.scale;
    A.g.L2 = g.L2;
    A.g.tblspcg = g.tblspcg;
    A.g.fM2 = (float)(g.M >> 1);
    A.g.M = (int)g.M;
    A.g.half = (int)g.half;
    A.g.ltbl = (int)ltbl;
    A.K = gr_conflict_free_step(g.L, g.scale);
    A.bw = (int)floorf(g.L + 0.01f) + 1;
    A.G = A.bw * A.bw <= 16 ? 16 : (A.bw * A.bw <= 32 ? 32 : 64);
    A.SP = 64 / A.G;
    A.L2m = g.L2 + 0.01f;
    const float half_tile = 0.5f * (float)(kTile - 1);
    A.reach = half_tile + A.L2m;
    A.dmax = A.reach * 1.41421356f + 0.05f;
    // largest sample radius (|(COSE,SINE)| <= 1 + 1e-4 for the float recurrence)
    A.rmax = g.scale * (float)(g.half - 1) * 1.0001f + A.dmax + 0.5f;
    const dim3 grid((t.ntiles + 3) / 4, pairs), block(256);
    const size_t lds = sizeof(float) * A.lds_wtbl + 4 * sizeof(WaveTile);
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
