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
// largest sample radius hold only zeros and are not written (the row
// transform reads them as zero, gr_tile_active).  The tile list is cut into
// eight sectors of polar angle, one per XCD, each nearest to the centre
// first, so the heaviest work is dispatched first and each XCD's L2 holds the
// samples of its own sector.
//
// Per angle, the samples that can reach the tile form one contiguous range of
// j for the direct term and one for the mirrored term; the wave stages the P
// pairs' spectra of that range in its own LDS slot (64 samples at a time) and
// each lane then runs its own exact candidate range against LDS.
//
// Accumulation: GR_ACCUM_PLAIN adds each tap's float product in float, in a
// fixed order (the tile's sample lines by sorted angle key, per line the
// direct samples, then the mirrored ones).  GR_ACCUM_COMPENSATED sums the taps of up to
// kMergeAngles consecutive sample lines (a few taps each) with fused
// multiply-adds into a partial, and adds each partial into the cell's running
// sum with an error-free TwoSum (Knuth), carrying the rounding errors in a
// second float: the cell's value is the sum plus the carried error, rounded
// once.
//
// Two kernels: k_grid derives the taps from the geometry per launch;
// k_grid_op streams them from the plan's precomputed gather operator
// (k_gop_build, below), bitwise the same.
//
// Output: Ht[pair][b][a] = H[a][b] (transposed, so K3's rows run over a).
#include "gr_fft.hpp"
#include "gr_gather.hpp"
#include "gr_internal.h"

#include <algorithm>

#pragma clang fp contract(off)

namespace {

constexpr int kTile = 8;        // cells per tile side; one tile per wavefront
constexpr int kMergeAngles = 16;  // sample lines per compensated partial

struct GridArgs {
    const float2 *spec1, *spec2;
    float2 *ht;
    const float4 *ang;   // sorted: (COSE, SINE, 1/COSE or 0, 1/SINE or 0)
    const int *aord;     // sorted -> angle index
    const float *akey;   // sorted keys
    const unsigned *tiles;
    const float *wtbl;
    const float *dwtbl;  // INTERP convolvent: backward differences of wtbl (null: nearest neighbour)
    int ntiles, sector_len, n_ang, half, lds_wtbl, pair_base, pair_end;
    GrGatherGeom g;
    float L2m, dmax, rmax;
};

// Slot of the tile list for wave `wave` of workgroup `block` in a launch of
// `wpb` waves per workgroup: workgroup b runs on XCD b mod 8 and takes sector
// b mod 8 of the list, its slots first + (b / 8) * wpb + wave in order.
// Returns -1 past the sector's end.
__device__ __forceinline__ int sector_slot(const GridArgs &A, int block, int wpb, int wave, int first) {
    const int slot = first + (block / kGridSectors) * wpb + wave;
    return slot < A.sector_len ? (block % kGridSectors) * A.sector_len + slot : -1;
}

// Offset of pair `pair`'s first value in the interleaved spectra: its group's
// block of n_ang * half records, then its slot in the record.
__device__ __forceinline__ long spec_group_offset(const GridArgs &A, int pair) {
    return (long)(pair / kSpecGroup) * A.n_ang * A.half * kSpecGroup + pair % kSpecGroup;
}

// Element i of a table that no kernel writes, read through the constant
// address space: at a wave-uniform index that is a scalar load, which does not
// wait behind the wave's vector loads in the vector memory counter.
template <class T>
__device__ __forceinline__ T table_load(const T *p, long i) {
#if defined(__HIP_DEVICE_COMPILE__)
    return ((const __attribute__((address_space(4))) T *)p)[i];
#else
    return p[i];  // the host pass of a device function: never called
#endif
}

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

// The P pairs' values of one sample: P consecutive float2 of its interleaved
// record (16-byte loads from P = 2 up; the record is 16-byte aligned).
template <int P>
__device__ __forceinline__ void load_pairs(const float2 *src, float2 (&v)[P]) {
    if (P == 1) {
        v[0] = src[0];
    } else {
        const float4 *q = reinterpret_cast<const float4 *>(src);
#pragma unroll
        for (int i = 0; i < P / 2; i++) {
            const float4 t = q[i];
            v[2 * i] = make_float2(t.x, t.y);
            v[2 * i + 1] = make_float2(t.z, t.w);
        }
    }
}

template <int P>
__device__ __forceinline__ void store_pairs(float2 *dst, const float2 (&v)[P]) {
    if (P == 1) {
        dst[0] = v[0];
    } else {
#pragma unroll
        for (int i = 0; i < P / 2; i++)
            reinterpret_cast<float4 *>(dst)[i] = make_float4(v[2 * i].x, v[2 * i].y, v[2 * i + 1].x, v[2 * i + 1].y);
    }
}

// Add the taps of samples [j0, j1] into cell (qa, qb) from one chunk of a
// sample row staged in the wave's LDS buffer stage (slot l: the P pairs'
// values of sample c0 + l).  In a partial group the spare slots hold other
// data: their accumulators are never stored.
template <int ACCUM, int P>
__device__ __forceinline__ void chunk_taps(const float2 (*stage)[P], int c0, int j0, int j1, float C, float S,
                                           float qa, float qb, const GrGatherGeom &g, const float *wtbl,
                                           const float *dwtbl, float4 (&acc)[P], float2 (&part)[P]) {
    for (int j = j0; j <= j1; j++) {
        float w;
        if (dwtbl ? gr_tap_interp(j, C, S, qa, qb, g, wtbl, dwtbl, w) : gr_tap(j, C, S, qa, qb, g, wtbl, w)) {
            float2 u[P];
            load_pairs<P>(stage[j - c0], u);
#pragma unroll
            for (int p = 0; p < P; p++) {
                if (ACCUM == GR_ACCUM_COMPENSATED) {
                    part[p].x = __builtin_fmaf(w, u[p].x, part[p].x);
                    part[p].y = __builtin_fmaf(w, u[p].y, part[p].y);
                } else {
                    acc[p].x += w * u[p].x;
                    acc[p].z += w * u[p].y;
                }
            }
        }
    }
}

// Walk the sample lines that can reach the wave's tile, in the kernel's fixed
// order (sorted angle keys, the window's two ranges in turn; per line the
// direct samples, then the mirrored ones, in chunks of at most 64 ascending
// j), calling on_chunk(mirror, row_off, c0, nj, j0, j1, C, S) per chunk --
// [c0, c0+nj) the tile's samples of the chunk, [j0, j1] the lane's own
// candidates among them -- and on_line() after each line that passed the
// tile tests.  Tiles beyond the largest sample radius visit nothing.
template <class ChunkFn, class LineFn>
__device__ __forceinline__ void walk_tile(const GridArgs &A, int tx, int ty, bool valid, float du, float dv,
                                          ChunkFn &&on_chunk, LineFn &&on_line) {
    const float M2 = A.g.fM2;
    if (!gr_tile_active(tx, ty, M2, A.rmax)) return;
    const float cu = (float)(tx * kTile) + 3.5f - M2, cv = (float)(ty * kTile) + 3.5f - M2;
    const float reach = 3.5f + A.L2m;  // tile half-width plus box half-width
    int r0, r1, r2, r3;
    gr_angle_window(tx, ty, M2, A.dmax, A.akey, A.n_ang, r0, r1, r2, r3);
    for (int side = 0; side < 2; side++) {
        const int k0 = side ? r2 : r0, k1 = side ? r3 : r1;
        for (int k = k0; k < k1; k++) {
            const float4 cs = table_load(A.ang, k);
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
            const long row_off = (long)table_load(A.aord, k) * A.half;
            for (int c0 = td0; c0 <= td1; c0 += 64)
                on_chunk(0, row_off, c0, min(64, td1 - c0 + 1), max(jd0, c0), min(min(jd1, td1), c0 + 63), cs.x,
                         cs.y);
            for (int c0 = tm0; c0 <= tm1; c0 += 64)
                on_chunk(1, row_off, c0, min(64, tm1 - c0 + 1), max(jm0, c0), min(min(jm1, tm1), c0 + 63), cs.x,
                         cs.y);
            on_line();
        }
    }
}

// Write the wave's 64 cells for the P pairs -- only for tiles that can hold
// a nonzero cell: the row transform reads the other tiles as zeros.
template <int ACCUM, int P>
__device__ __forceinline__ void store_cells(const GridArgs &A, int pair0, int nvalid, int a, int b,
                                            const float4 (&acc)[P]) {
    const int M = A.g.M;
    if (!gr_tile_active(a / kTile, b / kTile, A.g.fM2, A.rmax)) return;
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
__global__ void __launch_bounds__(256) k_grid(GridArgs A) {
    extern __shared__ float smem[];
    // the convolvent table (and its differences for INTERP) in LDS, then the
    // waves' staging areas from lds_wtbl on
    float *s_wtbl = smem;
    float *s_dwtbl = A.dwtbl ? smem + (A.g.ltbl + 1) : nullptr;
    for (int i = threadIdx.x; i <= A.g.ltbl; i += blockDim.x) {
        s_wtbl[i] = A.wtbl[i];
        if (s_dwtbl) s_dwtbl[i] = A.dwtbl[i];
    }
    __syncthreads();

    // the tile slot as a wave-uniform value: the tile, its angle window and
    // each line's angle record then come through scalar loads, which do not
    // wait behind the spectra loads in the vector memory counter
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ti = __builtin_amdgcn_readfirstlane(sector_slot(A, blockIdx.x, blockDim.x >> 6, wave, 0));
    if (ti < 0) return;
    const unsigned tt = table_load(A.tiles, ti);
    if (tt == kGridNoTile) return;
    float2(*stage)[P] = reinterpret_cast<float2(*)[P]>(smem + A.lds_wtbl) + wave * 128;  // two buffers of 64 slots
    const int pair0 = A.pair_base + blockIdx.y * P;
    const int nvalid = min(P, A.pair_end - pair0);
    const int tx = (int)(tt & 0xffffu), ty = (int)(tt >> 16);
    const int a = tx * kTile + (lane & 7), b = ty * kTile + (lane >> 3);
    const int M = A.g.M;
    const float M2 = A.g.fM2;

    const bool valid = a >= 1 && b >= 1 && a < M && b < M;
    const float pa = (float)a, pb = (float)b, ma = (float)(M - a), mb = (float)(M - b);
    float4 acc[P];   // (sum re, error re, sum im, error im)
    float2 part[P];  // compensated mode: the open partial
#pragma unroll
    for (int p = 0; p < P; p++) {
        acc[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        part[p] = make_float2(0.0f, 0.0f);
    }

    int lines = 0;  // sample lines in the open partial
    const long group = spec_group_offset(A, pair0);
    const float2 *base1 = A.spec1 + group, *base2 = A.spec2 + group;
    // one chunk in flight: a chunk's samples are loaded when walk_tile hands
    // it over, the taps of the chunk before are added from the other LDS
    // buffer while the loads travel, and only then do the new values go to
    // LDS -- the loaded registers never live past their chunk.  Every lane
    // loads (past the chunk's end it repeats its last sample, never read), so
    // no load sits in a branch.  The taps and the merges after each line
    // keep their order.
    bool have = false, p_eol = false;
    int cur = 0, p_mirror = 0, p_c0 = 0, p_j0 = 0, p_j1 = -1;
    float p_C = 0.0f, p_S = 0.0f;
    auto add_pending = [&]() {
        chunk_taps<ACCUM, P>(stage + cur * 64, p_c0, p_j0, p_j1, p_C, p_S, p_mirror ? ma : pa, p_mirror ? mb : pb,
                             A.g, s_wtbl, s_dwtbl, acc, part);
        if (p_eol && ACCUM == GR_ACCUM_COMPENSATED && ++lines == kMergeAngles) {
            merge<P>(acc, part);
            lines = 0;
        }
    };
    walk_tile(
        A, tx, ty, valid, pa - M2, pb - M2,
        [&](int mirror, long row_off, int c0, int nj, int j0, int j1, float C, float S) {
            float2 v[P];
            load_pairs<P>((mirror ? base2 : base1) + (row_off + c0 + min(lane, nj - 1)) * kSpecGroup, v);
            if (have) add_pending();
            cur ^= 1;
            __builtin_amdgcn_wave_barrier();
            store_pairs<P>(stage[cur * 64 + lane], v);
            __builtin_amdgcn_wave_barrier();
            have = true;
            p_eol = false;
            p_mirror = mirror;
            p_c0 = c0;
            p_j0 = j0;
            p_j1 = j1;
            p_C = C;
            p_S = S;
        },
        [&]() { p_eol = true; });  // the pending chunk ends its line
    if (have) add_pending();
    if (ACCUM == GR_ACCUM_COMPENSATED && lines) merge<P>(acc, part);
    store_cells<ACCUM, P>(A, pair0, nvalid, a, b, acc);
}

// ---- the precomputed gather operator ----------------------------------------
//
// The taps of a cell -- which sample, direct or mirrored, and its weight --
// depend only on the plan.  A plan that reconstructs many slices lists them
// once (k_gop_build) and then grids every group of pairs by streaming the
// list (k_grid_op) instead of re-evaluating the geometry.
//
// Layout: per tile (wavefront) n_t rows of 64 entries, entry k of lane l at
// gop[off_t + 64 k + l] (one coalesced 512-byte row per k).  Entry =
// (idx | mirror << 30 | merge << 31, weight bits): idx = angle*half + j
// indexes the pair's spec1 (mirror 0) or spec2 (mirror 1); merge asks for
// k_grid's merge of the compensated partial after this entry.  A lane's taps
// come in k_grid's order, so k_grid_op performs, per cell, the same float
// operations in the same order: its images are bitwise k_grid's.  Padding
// entries are (0, 0) or (merge, 0): sample 0 of the pair's first row, which K1
// always writes as zero, times weight 0 -- adding it changes no bit, and
// merging a zero partial changes no bit either.
//
// Two layouts, chosen per tile by the host from the count pass:
// * ragged: each lane's taps back to back, n_t = the largest count in the
//   wave, the merge mark on the lane's last tap before each merge point;
// * aligned: the taps between two merge points (kMergeAngles sample lines)
//   occupy the same rows in every lane, padded to the longest lane, with the
//   merge mark on the group's last row in every lane -- so the merge is
//   wave-uniform.  Near the centre of the grid every line crosses the tile
//   and the lanes' groups are nearly equal: there it costs few extra rows and
//   turns a merge in almost every row (some lane always closes a group) into
//   one per group.
constexpr unsigned kGopMirror = 1u << 30, kGopMerge = 1u << 31, kGopIdx = kGopMirror - 1;
constexpr unsigned kGopAligned = kGopAlignedBit, kGopRows = kGopAligned - 1;  // tile_n: layout bit, rows

__device__ __forceinline__ unsigned wave_max(unsigned v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = max(v, (unsigned)__shfl_xor((int)v, o));
    return v;
}

__device__ __forceinline__ unsigned wave_min(unsigned v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = min(v, (unsigned)__shfl_xor((int)v, o));
    return v;
}

// MODE 0 (count): rows of the ragged layout into tile_n, of the aligned one
// into tile_na.  MODE 1 (fill): the entries at tile_off in the layout that
// tile_n's top bit names (its rows in the low bits).
template <int MODE>
__global__ void __launch_bounds__(256) k_gop_build(GridArgs A, unsigned *tile_n, unsigned *tile_na,
                                                   const unsigned *tile_off, uint2 *gop, unsigned *tile_g,
                                                   const uint4 *hmeta, uint2 *hseg, int gps) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ti = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + wave);  // scalar table loads
    if (ti >= A.ntiles) return;
    const unsigned tt = table_load(A.tiles, ti);
    if (tt == kGridNoTile) {  // padding slot: no rows
        if (MODE == 0 && lane == 0) tile_n[ti] = tile_na[ti] = tile_g[ti] = 0;
        return;
    }
    // tiles left without rows (no taps, or listed in the staged operator) are
    // not filled
    if (MODE == 1 && (tile_n[ti] & kGopRows) == 0) return;
    const int tx = (int)(tt & 0xffffu), ty = (int)(tt >> 16);
    const int a = tx * kTile + (lane & 7), b = ty * kTile + (lane >> 3);
    const int M = A.g.M;
    const float M2 = A.g.fM2;
    const bool valid = a >= 1 && b >= 1 && a < M && b < M;
    const float pa = (float)a, pb = (float)b, ma = (float)(M - a), mb = (float)(M - b);
    const unsigned long base = MODE == 1 ? (unsigned long)tile_off[ti] + lane : 0;
    const bool aligned = MODE == 1 && (tile_n[ti] & kGopAligned);

    unsigned n = 0;   // ragged: taps of this lane so far
    unsigned g0 = 0;  // aligned: first row of the open group
    unsigned ng = 0;  // taps of this lane in the open group
    unsigned na = 0;  // count pass: rows of the aligned layout
    unsigned gc = 0;  // merge points passed
    unsigned nm = 0;  // merge marks in this lane's entries so far
    uint2 pend = make_uint2(0u, 0u);
    bool has_pend = false;  // the lane's latest tap, not yet written (it may get the merge mark)
    auto put = [&](unsigned row, uint2 e) { gop[base + 64ul * row] = e; };
    // a tile cut into segments of gps merge groups (hmeta.x >= 1): per
    // segment boundary the lane's row and merge count
    const uint4 hm = MODE == 1 && hseg ? hmeta[ti] : make_uint4(0u, 0u, 0u, 0u);
    auto mark = [&](unsigned seg) {
        if (hm.x >= 1 && seg <= hm.x) hseg[hm.y + 64ul * seg + lane] = make_uint2(aligned ? g0 : n, nm);
    };
    mark(0);
    auto close_group = [&]() {  // a merge point of k_grid (wave-uniform)
        if (MODE == 0) {
            na += wave_max(ng);
        } else if (aligned) {
            const unsigned len = wave_max(ng);
            if (has_pend) put(g0 + ng - 1, ng == len ? make_uint2(pend.x | kGopMerge, pend.y) : pend);
            for (unsigned r = ng; r < len; r++) put(g0 + r, make_uint2(r + 1 == len ? kGopMerge : 0u, 0u));
            g0 += len;
            nm += len ? 1u : 0u;
        } else if (has_pend) {
            put(n - 1, make_uint2(pend.x | kGopMerge, pend.y));
            nm++;
        }
        has_pend = false;
        ng = 0;
        if (++gc % (unsigned)max(gps, 1) == 0) mark(gc / (unsigned)max(gps, 1));
    };
    int lines = 0;
    walk_tile(
        A, tx, ty, valid, pa - M2, pb - M2,
        [&](int mirror, long row_off, int c0, int nj, int j0, int j1, float C, float S) {
            for (int j = j0; j <= j1; j++) {
                float w;
                const float qa = mirror ? ma : pa, qb = mirror ? mb : pb;
                if (A.dwtbl ? gr_tap_interp(j, C, S, qa, qb, A.g, A.wtbl, A.dwtbl, w)
                            : gr_tap(j, C, S, qa, qb, A.g, A.wtbl, w)) {
                    if (MODE == 1 && has_pend) put(aligned ? g0 + ng - 1 : n - 1, pend);
                    pend = make_uint2((unsigned)(row_off + j) | (mirror ? kGopMirror : 0u), __float_as_uint(w));
                    has_pend = true;
                    n++;
                    ng++;
                }
            }
        },
        [&]() {
            if (++lines == kMergeAngles) {
                lines = 0;
                close_group();
            }
        });
    if (lines) close_group();  // k_grid's final merge
    if (MODE == 0) {
        n = wave_max(n);
        if (lane == 0) {
            tile_n[ti] = n;
            tile_na[ti] = na;
            tile_g[ti] = gc;
        }
    } else {
        mark(hm.x);  // the end of the last segment
        const unsigned rows = tile_n[ti] & kGopRows;
        for (unsigned r = aligned ? g0 : n; r < rows; r++) put(r, make_uint2(0u, 0u));
    }
}

// ---- the staged operator --------------------------------------------------------
//
// Where the operator above hands each lane a sample record per tap -- a
// 64-byte gather per lane and row, 64 records per wave instruction, which
// the texture addresser takes one cache line at a time -- a tile's records
// are reused by ~8 of its cells each.  The staged form loads every record a
// tile needs once into the wave's LDS and lets the taps read it there.
//
// A tile's taps are cut into chunks at k_grid's merge points (every
// kMergeAngles sample lines, and the last line) and wherever the next run of
// samples (of one sample line, from the first to the last sample some lane
// taps) would overflow the chunk's kSopSlots LDS slots.  Per chunk: a
// header {rows | merge bit, slots}, its slots' records (kSopSlots keys
// idx | mirror << 30 as the operator's entries, the unused ones zero; slot
// 0 is record 0 of spec1, which K1 writes as zero), and its rows: per lane
// its taps within the chunk in k_grid's order, the lanes with fewer taps
// (and the rows up to a multiple of kSopRowBatch) padded with slot 0 (zero
// times a weight adds nothing to a partial that never holds -0), each entry slot | iu << 8 | iv << 20 with the weight
// wtbl[iu] * wtbl[iv] read from LDS -- k_grid's product.  A chunk that ends
// at a merge point carries the merge bit: the merge is wave-uniform.  So per
// cell the float operations and their order are k_grid's: bitwise its images.
//
// Chunk 0 of tile slot ti is chunk ti (header ch[ti], keys at 128 ti), so a
// wave can fetch it with the tile's offsets before it knows them; a tile's
// further chunks are ntiles + its first extra chunk + k - 1.
constexpr int kSopSlots = kSopKeysPerChunk;
constexpr unsigned kSopMerge = 1u << 31;
constexpr int kSopRowBatch = 4;  // a chunk's rows: a multiple of this (padded), added a batch at a time

// MODE 0 (count): {0, chunks, 0, rows} of each tile past the first n_heavy
// slots of its sector (zeros elsewhere).  MODE 1 (fill): headers, keys and
// entries; tile holds {first extra chunk, chunks, 0, first row}.
template <int MODE>
__global__ void __launch_bounds__(256) k_sop_build(GridArgs A, int n_heavy, uint4 *tile, uint2 *ch, unsigned *keys,
                                                   unsigned *ent) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ti = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + wave);  // scalar table loads
    if (ti >= A.ntiles) return;
    const unsigned tt = table_load(A.tiles, ti);
    if (tt == kGridNoTile || ti % A.sector_len < n_heavy) {
        if (MODE == 0 && lane == 0) tile[ti] = make_uint4(0u, 0u, 0u, 0u);
        return;
    }
    const int tx = (int)(tt & 0xffffu), ty = (int)(tt >> 16);
    const int a = tx * kTile + (lane & 7), b = ty * kTile + (lane >> 3);
    const int M = A.g.M;
    const float M2 = A.g.fM2;
    const bool valid = a >= 1 && b >= 1 && a < M && b < M;
    const float pa = (float)a, pb = (float)b, ma = (float)(M - a), mb = (float)(M - b);
    const uint4 to = MODE == 1 ? tile[ti] : make_uint4(0u, 0u, 0u, 0u);
    unsigned nch = 0, nrows = 0;  // closed chunks and their rows (wave-uniform)
    unsigned slots = 1;           // the open chunk's slots (slot 0: the zero record)
    unsigned ntc = 0;             // this lane's taps in the open chunk
    auto chunk_index = [&]() { return nch ? (unsigned)A.ntiles + to.x + nch - 1u : (unsigned)ti; };
    auto close = [&](bool merge_after) {  // wave-uniform
        const unsigned rows = (wave_max(ntc) + kSopRowBatch - 1) & ~(unsigned)(kSopRowBatch - 1);
        if (MODE == 1) {
            const unsigned long e0 = 64ul * (to.w + nrows) + lane;
            for (unsigned r = ntc; r < rows; r++) ent[e0 + 64ul * r] = 0u;
            const unsigned ci = chunk_index();
            for (unsigned k = slots + lane; k < (unsigned)kSopSlots; k += 64) keys[(unsigned long)ci * kSopSlots + k] = 0u;
            if (lane == 0) {
                ch[ci] = make_uint2(rows | (merge_after ? kSopMerge : 0u), slots);
                keys[(unsigned long)ci * kSopSlots] = 0u;
            }
        }
        nch++;
        nrows += rows;
        slots = 1;
        ntc = 0;
    };
    int lines = 0;
    walk_tile(
        A, tx, ty, valid, pa - M2, pb - M2,
        [&](int mirror, long row_off, int c0, int nj, int j0, int j1, float C, float S) {
            // stage only the samples some lane taps: [lo, hi] of the run
            const float qa = mirror ? ma : pa, qb = mirror ? mb : pb;
            unsigned lo = 0xffffffffu, hi = 0u;
            for (int j = j0; j <= j1; j++) {
                int iu, iv;
                if (gr_tap_index(j, C, S, qa, qb, A.g, iu, iv)) {
                    lo = min(lo, (unsigned)j);
                    hi = max(hi, (unsigned)j);
                }
            }
            lo = wave_min(lo);
            hi = wave_max(hi);
            if (hi < lo) return;  // no lane taps this run (j >= 1, so hi = 0 means none)
            (void)c0;
            (void)nj;
            const unsigned n = hi - lo + 1u;  // <= nj <= 64
            if (slots + n > (unsigned)kSopSlots) close(false);
            const unsigned sb = slots - lo;  // slot of sample j: sb + j
            if (MODE == 1 && (unsigned)lane < n)
                keys[(unsigned long)chunk_index() * kSopSlots + slots + lane] =
                    (unsigned)(row_off + lo + lane) | (mirror ? kGopMirror : 0u);
            slots += n;
            for (int j = j0; j <= j1; j++) {
                int iu, iv;
                if (gr_tap_index(j, C, S, qa, qb, A.g, iu, iv)) {
                    if (MODE == 1)
                        ent[64ul * (to.w + nrows + ntc) + lane] =
                            (sb + (unsigned)j) | (unsigned)iu << 8 | (unsigned)iv << 20;
                    ntc++;
                }
            }
        },
        [&]() {
            if (++lines == kMergeAngles) {
                lines = 0;
                close(true);
            }
        });
    if (lines) close(true);  // k_grid's final merge
    if (MODE == 0 && lane == 0) tile[ti] = make_uint4(0u, nch, 0u, nrows);
}

struct GopArgs {
    const uint2 *gop;
    const unsigned *tile_n, *tile_off;
    int n_heavy;  // slots [0, n_heavy) of every sector take the one-pair-per-wave path
    // the staged operator of the other slots (null: they stream the entries)
    const uint4 *sop_tile;
    const uint2 *sop_ch;
    const unsigned *sop_keys, *sop_ent;
    int sop_waves;  // persistent light waves per sector (and group of pairs)
    // groups of P pairs whose workgroups are interleaved so that the groups'
    // waves walk the same tiles side by side on one XCD and read each entry
    // through its L2 once (1: a launch of up to P pairs)
    int sop_groups;
    // the heavy work items (tile, segment), heavy tiles cut into segments
    // (GrDevTables), and their partials
    const uint2 *hitems;
    int n_hitems;
    const uint4 *hmeta;
    const uint2 *hseg;
    const unsigned *hsplit;
    int n_split;
    float2 *part;
    int part_pairs;
};

// An entry of the streamed operator, by an ordinary (cached) load: the
// interleaved groups of pairs read each entry through L2 once, and even one
// group ran 15% faster than with non-temporal loads.
__device__ __forceinline__ uint2 gop_load(const uint2 *p) {
    const unsigned long long u = *reinterpret_cast<const unsigned long long *>(p);
    return make_uint2((unsigned)u, (unsigned)(u >> 32));
}

// The P pairs' samples of one entry (spare slots of a partial group read
// other data; their accumulators are never stored).
template <int P>
__device__ __forceinline__ void gop_gather(uint2 t, const float2 *base1, const float2 *base2, float2 (&v)[P]) {
    load_pairs<P>(((t.x & kGopMirror) ? base2 : base1) + (long)(t.x & kGopIdx) * kSpecGroup, v);
}

// k_grid's accumulation of one tap (and its merge, when marked).
template <int ACCUM, int P>
__device__ __forceinline__ void gop_consume(uint2 t, const float2 (&v)[P], float4 (&acc)[P], float2 (&part)[P]) {
    const float w = __uint_as_float(t.y);
#pragma unroll
    for (int p = 0; p < P; p++) {
        if (ACCUM == GR_ACCUM_COMPENSATED) {
            part[p].x = __builtin_fmaf(w, v[p].x, part[p].x);
            part[p].y = __builtin_fmaf(w, v[p].y, part[p].y);
        } else {
            acc[p].x += w * v[p].x;
            acc[p].z += w * v[p].y;
        }
    }
    if (ACCUM == GR_ACCUM_COMPENSATED && (t.x & kGopMerge)) merge<P>(acc, part);
}

constexpr int kHeavyBatch = 16;  // rows loaded at once on the one-pair path

// The one-pair-per-wave path for the tiles near the centre of the grid whose
// cells take hundreds to thousands of taps: the P pairs of a group go to P
// wavefronts, which stream the tile's rows kHeavyBatch at a time (the batch's
// sample loads, then the next batch's entry loads, then the batch's adds), so
// each wave has many loads in flight instead of a chain of dependent ones.
// Per cell and pair the float operations are those of the wide path.
template <int ACCUM>
__device__ __forceinline__ void grid_op_heavy(const GridArgs &A, const GopArgs &O, int ti, int pair, int lane) {
    const unsigned tt = A.tiles[ti];
    const int tx = (int)(tt & 0xffffu), ty = (int)(tt >> 16);
    const int a = tx * kTile + (lane & 7), b = ty * kTile + (lane >> 3);
    const long group = spec_group_offset(A, pair);
    const float2 *base1 = A.spec1 + group, *base2 = A.spec2 + group;
    const uint2 *e = O.gop + (unsigned long)O.tile_off[ti] + lane;
    const unsigned n = O.tile_n[ti] & kGopRows;
    float4 acc[1] = {make_float4(0.0f, 0.0f, 0.0f, 0.0f)};
    float2 part[1] = {make_float2(0.0f, 0.0f)};
    if (n == 0) {
        // a slot no sample reaches: no entry may be read (n - 1 would wrap and
        // the clamped prefetch below would run past the tile's entries); the
        // cells are still stored as zeros, which the row transform reads
        store_cells<ACCUM, 1>(A, pair, 1, a, b, acc);
        return;
    }
    auto entry = [&](unsigned k) {
        const uint2 t = gop_load(e + 64ul * min(k, n - 1));
        return k < n ? t : make_uint2(0u, 0u);
    };
    // software-pipelined: the next batch's entries are in flight while this
    // batch's samples are gathered and added, so a batch costs one memory
    // round trip instead of two
    uint2 t[kHeavyBatch];
#pragma unroll
    for (int i = 0; i < kHeavyBatch; i++) t[i] = entry(i);
    for (unsigned k0 = 0; k0 < n; k0 += kHeavyBatch) {
        float2 v[kHeavyBatch][1];
#pragma unroll
        for (int i = 0; i < kHeavyBatch; i++) gop_gather<1>(t[i], base1, base2, v[i]);
        uint2 tn[kHeavyBatch];
#pragma unroll
        for (int i = 0; i < kHeavyBatch; i++) tn[i] = entry(k0 + kHeavyBatch + i);
#pragma unroll
        for (int i = 0; i < kHeavyBatch; i++) gop_consume<ACCUM, 1>(t[i], v[i], acc, part);
#pragma unroll
        for (int i = 0; i < kHeavyBatch; i++) t[i] = tn[i];
    }
    store_cells<ACCUM, 1>(A, pair, 1, a, b, acc);
}

// A wave's record slots in LDS: plane q holds piece q of every slot (16
// bytes of the P pairs' values for P >= 2, the real and the imaginary float
// for P = 1), so the 16-byte reads of 64 lanes at random slots spread over
// all banks (a 64-byte record stride would leave them a quarter of them).
// Records reach the slots by global-to-LDS loads, which write lane l's piece
// at the load's LDS base + l * piece: slot l (or l + 64) of each plane, no
// registers in between.  (Loading whole records per lane group instead --
// coalesced reads, pieces rotated within a slot against bank conflicts --
// took fewer texture-addresser cycles but more VALU work and was slower.)
template <int P>
struct SopStage {
    static constexpr int piece = P >= 2 ? 16 : 4;  // bytes per lane and load
    static constexpr int planes = P * 8 / piece;
    static constexpr int plane_bytes = kSopSlots * piece;
};

template <int P>
__device__ __forceinline__ void sop_stage(unsigned key, const float2 *base1, const float2 *base2, char *stg,
                                          int slot0) {
    const char *rec = reinterpret_cast<const char *>(((key & kGopMirror) ? base2 : base1) +
                                                     (long)(key & kGopIdx) * kSpecGroup);
#if defined(__HIP_DEVICE_COMPILE__)  // a device builtin: the host pass sees no body
    using St = SopStage<P>;
#pragma unroll
    for (int q = 0; q < St::planes; q++)
        __builtin_amdgcn_global_load_lds(
            (const void *)(rec + q * St::piece),
            (__attribute__((address_space(3))) void *)(stg + q * St::plane_bytes + slot0 * St::piece), St::piece,
            0, 0);
#else
    (void)rec;
    (void)stg;
    (void)slot0;
#endif
}

template <int P>
__device__ __forceinline__ void sop_read(const char *stg, unsigned slot, float2 (&v)[P]) {
    using St = SopStage<P>;
    if (P >= 2) {
#pragma unroll
        for (int q = 0; q < P / 2; q++) {
            const float4 x = *reinterpret_cast<const float4 *>(stg + q * St::plane_bytes + 16 * slot);
            v[2 * q] = make_float2(x.x, x.y);
            v[2 * q + 1] = make_float2(x.z, x.w);
        }
    } else {
        v[0] = make_float2(*reinterpret_cast<const float *>(stg + 4 * slot),
                           *reinterpret_cast<const float *>(stg + St::plane_bytes + 4 * slot));
    }
}

// The light tiles of one sector for the P pairs of a group from the staged
// operator, persistent: the wave takes slots slot0, slot0 + stride, ... and
// loads the next slot's tile, offsets, first chunk header and keys while it
// works on the current one (they sit at positions fixed by the slot), so a
// tile costs one round trip -- its records into the slots and its first
// entries -- plus its arithmetic.  Within a tile the next chunk's header and
// keys load while a chunk's rows are added.  The entries array ends in a
// spare row, so the clamped entry loads of a chunk without rows stay inside.
template <int ACCUM, int P>
__device__ __forceinline__ void grid_sop_light(const GridArgs &A, const GopArgs &O, int sector, int slot0,
                                               int stride, int pair0, int nvalid, int lane, char *stg,
                                               const float *s_w) {
    const long group = spec_group_offset(A, pair0);
    const float2 *base1 = A.spec1 + group, *base2 = A.spec2 + group;
    const int first = sector * A.sector_len;
    struct Next {
        unsigned tt;
        uint4 td;
        uint2 h;
        unsigned k0, k1;
    };
    auto fetch = [&](int slot) {  // slot < sector_len
        const int ti = first + slot;
        Next n;
        n.tt = table_load(A.tiles, ti);  // wave-uniform: scalar loads
        n.td = table_load(O.sop_tile, ti);
        n.h = table_load(O.sop_ch, ti);
        n.k0 = O.sop_keys[(unsigned long)ti * kSopSlots + lane];
        n.k1 = O.sop_keys[(unsigned long)ti * kSopSlots + 64 + lane];
        return n;
    };
    if (slot0 >= A.sector_len) return;
    Next nx = fetch(slot0);
    for (int slot = slot0; slot < A.sector_len; slot += stride) {
        const Next cur = nx;
        const unsigned cur_ti = (unsigned)(first + slot);
        nx = fetch(min(slot + stride, A.sector_len - 1));
        if (cur.tt == kGridNoTile) continue;
        const int tx = (int)(cur.tt & 0xffffu), ty = (int)(cur.tt >> 16);
        const int a = tx * kTile + (lane & 7), b = ty * kTile + (lane >> 3);
        float4 acc[P];
        float2 part[P];
#pragma unroll
        for (int p = 0; p < P; p++) {
            acc[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            part[p] = make_float2(0.0f, 0.0f);
        }
        auto consume = [&](unsigned e) {
            const float w = s_w[(e >> 8) & 0xfffu] * s_w[e >> 20];
            float2 v[P];
            sop_read<P>(stg, e & 0xffu, v);
#pragma unroll
            for (int p = 0; p < P; p++) {
                if (ACCUM == GR_ACCUM_COMPENSATED) {
                    part[p].x = __builtin_fmaf(w, v[p].x, part[p].x);
                    part[p].y = __builtin_fmaf(w, v[p].y, part[p].y);
                } else {
                    acc[p].x += w * v[p].x;
                    acc[p].z += w * v[p].y;
                }
            }
        };
        const unsigned nch = cur.td.y, extra = (unsigned)A.ntiles + cur.td.x;
        const unsigned *ep = O.sop_ent + 64ul * cur.td.w + lane;
        uint2 h = cur.h;
        unsigned k0 = cur.k0, k1 = cur.k1;
        for (unsigned c = 0; c < nch; c++) {
            const unsigned rows = h.x & ~kSopMerge, ns = h.y;
            const unsigned last = rows ? rows - 1u : 0u;
            auto eload = [&](unsigned k) { return ep[64ul * min(k, last)]; };  // cached: see gop_load
            sop_stage<P>(k0, base1, base2, stg, 0);
            if (ns > 64u) sop_stage<P>(k1, base1, base2, stg, 64);
            // the next chunk (or this one again past the last)
            const unsigned cn = c + 1 < nch ? extra + c : (c ? extra + c - 1 : cur_ti);
            const uint2 hn = table_load(O.sop_ch, cn);
            const unsigned kn0 = O.sop_keys[(unsigned long)cn * kSopSlots + lane];
            const unsigned kn1 = O.sop_keys[(unsigned long)cn * kSopSlots + 64 + lane];
            // rows come in whole batches of kSopRowBatch: no per-row guard,
            // so a batch's LDS reads overlap
            constexpr int RB = kSopRowBatch;
            unsigned t[RB], u[RB];
#pragma unroll
            for (int i = 0; i < RB; i++) t[i] = eload(i);
            __builtin_amdgcn_wave_barrier();
            for (unsigned r0 = 0; r0 < rows; r0 += 2 * RB) {
#pragma unroll
                for (int i = 0; i < RB; i++) u[i] = eload(r0 + RB + i);
#pragma unroll
                for (int i = 0; i < RB; i++) consume(t[i]);
                if (r0 + RB >= rows) break;
#pragma unroll
                for (int i = 0; i < RB; i++) t[i] = eload(r0 + 2 * RB + i);
#pragma unroll
                for (int i = 0; i < RB; i++) consume(u[i]);
            }
            ep += 64ul * rows;
            if (ACCUM == GR_ACCUM_COMPENSATED && (h.x & kSopMerge)) merge<P>(acc, part);
            h = hn;
            k0 = kn0;
            k1 = kn1;
            __builtin_amdgcn_wave_barrier();  // this chunk's reads before the next chunk's records
        }
        store_cells<ACCUM, P>(A, pair0, nvalid, a, b, acc);
    }
}

// One tile (seg < 0) or one segment of a cut heavy tile for the P pairs of a
// group from the streamed entries, software-pipelined: entries are loaded
// four rows ahead, the samples of row k+1 are gathered while row k is added.
// A whole tile runs its rows to the end, merging into its sums, and stores
// its cells.  A segment runs each lane's rows between its marks at the
// segment's ends (segments start at merge points, so they hold whole merge
// groups), rows past the lane's end reading as zero entries (sample 0,
// weight 0, no merge; loads stay inside the tile's rows), and each merge
// stores the finished partials to partial slot (first partial group + the
// lane's merge count) instead of adding them to sums: k_gop_merge sums
// them in order -- the same float operations, cut at merges.
template <int ACCUM, int P>
__device__ __forceinline__ void grid_op_tile(const GridArgs &A, const GopArgs &O, int ti, int seg, int pair0,
                                             int nvalid, int lane) {
    const unsigned tt = A.tiles[ti];
    const int tx = (int)(tt & 0xffffu), ty = (int)(tt >> 16);
    const int a = tx * kTile + (lane & 7), b = ty * kTile + (lane >> 3);
    float4 acc[P];
    float2 part[P];
#pragma unroll
    for (int p = 0; p < P; p++) {
        acc[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        part[p] = make_float2(0.0f, 0.0f);
    }
    const long group = spec_group_offset(A, pair0);
    const float2 *base1 = A.spec1 + group, *base2 = A.spec2 + group;
    const uint2 *e = O.gop + (unsigned long)O.tile_off[ti] + lane;
    const unsigned rows = O.tile_n[ti] & kGopRows;
    unsigned r0 = 0, len = rows, m = 0;
    float2 *out = nullptr;
    const size_t gstride = (size_t)O.part_pairs * 64;
    if (seg >= 0) {
        const uint4 hm = O.hmeta[ti];
        const uint2 m0 = O.hseg[hm.y + 64ul * seg + lane], m1 = O.hseg[hm.y + 64ul * (seg + 1) + lane];
        r0 = m0.x;
        len = m1.x - m0.x;
        m = hm.z + m0.y;
        out = O.part + (size_t)pair0 * 64 + lane;
    }
    const unsigned n = seg >= 0 ? wave_max(len) : rows;
    if (n) {
        const uint2 zero = make_uint2(0u, 0u);
        auto row = [&](unsigned k) {
            const uint2 t = gop_load(e + 64ul * min(r0 + k, rows - 1u));
            return k < len ? t : zero;
        };
        auto consume = [&](uint2 t, const float2 (&v)[P]) {
            if (seg < 0) {
                gop_consume<ACCUM, P>(t, v, acc, part);
                return;
            }
            const float w = __uint_as_float(t.y);
#pragma unroll
            for (int p = 0; p < P; p++) {
                part[p].x = __builtin_fmaf(w, v[p].x, part[p].x);
                part[p].y = __builtin_fmaf(w, v[p].y, part[p].y);
            }
            if (t.x & kGopMerge) {
#pragma unroll
                for (int p = 0; p < P; p++) {
                    if (p < nvalid) out[m * gstride + 64 * p] = part[p];
                    part[p] = make_float2(0.0f, 0.0f);
                }
                m++;
            }
        };
        uint2 t0 = row(0), t1 = row(1), t2 = row(2), t3 = row(3);
        float2 va[P], vb[P];
        gop_gather<P>(t0, base1, base2, va);
        for (unsigned k = 0; k < n; k += 2) {
            // rows k (t0, samples in va) and k+1 (t1); t2, t3 hold rows k+2, k+3
            const uint2 c1 = k + 1 < n ? t1 : zero;
            gop_gather<P>(c1, base1, base2, vb);
            const uint2 c0 = t0;
            t0 = t2;
            t2 = row(k + 4);
            consume(c0, va);
            gop_gather<P>(k + 2 < n ? t0 : zero, base1, base2, va);
            t1 = t3;
            t3 = row(k + 5);
            consume(c1, vb);
        }
    }
    if (seg < 0) store_cells<ACCUM, P>(A, pair0, nvalid, a, b, acc);
}

// The cut heavy tiles' sums: per (tile, pair) wave, each lane merges its
// partials in order, as grid_op_heavy's merges would have (the merges of
// groups where the lane had no tap add zero partials, which change no bit),
// and stores its cell.  Partials load sixteen at a time.
__global__ void __launch_bounds__(256) k_gop_merge(GridArgs A, GopArgs O, int pairs) {
    const int w = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (w >= O.n_split * pairs) return;
    const int h = w / pairs, pair = w % pairs;
    const unsigned ti = O.hsplit[h];
    const unsigned tt = A.tiles[ti];
    const int tx = (int)(tt & 0xffffu), ty = (int)(tt >> 16);
    const int a = tx * kTile + (lane & 7), b = ty * kTile + (lane >> 3);
    const uint4 hm = O.hmeta[ti];
    const unsigned total = O.hseg[hm.y + 64ul * hm.x + lane].y, n = wave_max(total);
    const float2 *in = O.part + (size_t)hm.z * O.part_pairs * 64 + (size_t)pair * 64 + lane;
    const size_t gstride = (size_t)O.part_pairs * 64;
    float4 acc[1] = {make_float4(0.0f, 0.0f, 0.0f, 0.0f)};
    for (unsigned k0 = 0; k0 < n; k0 += 16) {
        float2 q[16];
#pragma unroll
        for (int i = 0; i < 16; i++) q[i] = in[min(k0 + i, total ? total - 1u : 0u) * gstride];
#pragma unroll
        for (int i = 0; i < 16; i++) {
            float2 part[1] = {k0 + i < total ? q[i] : make_float2(0.0f, 0.0f)};
            merge<1>(acc, part);
        }
    }
    store_cells<GR_ACCUM_COMPENSATED, 1>(A, pair, 1, a, b, acc);
}
// Workgroups of k_grid_op given to the one-pair-per-wave path (a multiple of
// the sector count, so the other workgroups keep their XCD-to-sector map).
__host__ __device__ __forceinline__ int heavy_blocks(int n_items) {
    return (n_items + 4 * kGridSectors - 1) / (4 * kGridSectors) * kGridSectors;
}

// k_grid's accumulation, fed by the operator.  The first heavy_blocks
// workgroups take the heavy work items (slots [0, n_heavy) of every sector)
// from the streamed entries: with compensated sums one segment of a cut tile
// per wave for the P pairs of the group (grid_op_tile; k_gop_merge sums the
// partials afterwards), otherwise a tile one pair per wave (grid_op_heavy).
// The others take one tile per wave for the P pairs -- SOP: from the staged
// operator (grid_sop_light), otherwise from the streamed entries
// (grid_op_tile).
template <int ACCUM, int P, bool SOP>
__global__ void __launch_bounds__(256, 4) k_grid_op(GridArgs A, GopArgs O) {
    extern __shared__ float smem[];  // staged operator: wtbl, then the waves' record slots
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (SOP) {
        for (int i = threadIdx.x; i <= A.g.ltbl; i += blockDim.x) smem[i] = A.wtbl[i];
        __syncthreads();
    }
    // SOP with G = sop_groups > 1 groups of pairs in one launch: workgroup x
    // is workgroup bx of group gy, numbered so that the G groups' workgroups
    // of one slot are x, x + 8, ..., x + 8 (G - 1): dealt to the same XCD one
    // after another (light ones first, then the heavy ones, interleaved alike)
    int bx = (int)blockIdx.x, gy = (int)blockIdx.y;
    if (SOP && O.sop_groups > 1) {
        const int G = O.sop_groups, nl = 2 * O.sop_waves;  // nl and the heavy count: multiples of 8
        const int x = bx < nl * G ? bx : bx - nl * G;
        gy = (x >> 3) % G;
        bx = (x / (8 * G)) * 8 + (x & 7) + (bx < nl * G ? 0 : nl);
    }
    const int pair0 = A.pair_base + gy * P;
    const int nvalid = min(P, A.pair_end - pair0);
    // SOP: the persistent light workgroups first (they start at once and
    // hold most of the machine), the heavy items after them in the slots
    // left; otherwise the heavy items first (their long chains start early)
    const int hb = heavy_blocks(O.n_hitems);
    const int nlight = SOP ? 2 * O.sop_waves : 0;  // kGridSectors * sop_waves / 4
    const int hbx = bx - nlight;
    if (hbx >= 0 && hbx < hb) {
        // wave: work item hbx * 4 + wave -- a cut tile's segment (the P pairs
        // into partials), or a heavy tile one pair per wave
        const int item = hbx * 4 + wave;
        const uint2 it = O.hitems[min(item, O.n_hitems - 1)];
        const int ti = (int)it.x;
        if (item >= O.n_hitems || A.tiles[ti] == kGridNoTile) return;
        if (O.hmeta) {
            grid_op_tile<ACCUM, P>(A, O, ti, (int)it.y, pair0, nvalid, lane);
        } else {
            const int p = (int)it.y;  // tiles not cut: one item per pair
            if (p < nvalid) grid_op_heavy<ACCUM>(A, O, ti, pair0 + p, lane);
        }
        return;
    }
    if (SOP) {
        // persistent: workgroup b (XCD b mod 8) serves sector b mod 8 with
        // sop_waves waves, each taking every sop_waves-th light slot
        const int wv = __builtin_amdgcn_readfirstlane(wave);
        const int lb = bx;
        char *stg = reinterpret_cast<char *>(smem + A.lds_wtbl) + (size_t)wv * kSopSlots * P * sizeof(float2);
        grid_sop_light<ACCUM, P>(A, O, lb % kGridSectors, O.n_heavy + (lb / kGridSectors) * 4 + wv, O.sop_waves,
                                 pair0, nvalid, lane, stg, smem);
        return;
    }
    const int ti = sector_slot(A, blockIdx.x - hb, 4, wave, O.n_heavy);
    if (ti < 0) return;
    const unsigned tt = A.tiles[ti];
    if (tt == kGridNoTile) return;
    grid_op_tile<ACCUM, P>(A, O, ti, -1, pair0, nvalid, lane);
}

// Persistent workgroups per sector (and group of pairs) of the staged light
// path: 8 x 96 take three of the four workgroups an MI355X CU holds, the
// fourth serves the heavy items.
constexpr int kSopBlocksPerSector = 96;

template <int ACCUM, int P>
hipError_t launch(const GridArgs &A, const GopArgs *O, int groups, size_t lds_base, int wpb, hipStream_t s) {
    if (O) {
        const int light = (A.sector_len - O->n_heavy + 3) / 4;
        const int blocks = heavy_blocks(O->n_hitems) + kGridSectors * light;
        if (!O->sop_tile) {
            hipLaunchKernelGGL((k_grid_op<ACCUM, P, false>), dim3(blocks, groups), dim3(256), 0, s, A, *O);
            return hipGetLastError();
        }
        // the staged light tiles: persistent workgroups, kSopBlocksPerSector
        // per sector shared by the launch's groups of pairs (up to 8 groups
        // interleaved), fewer when the sector has fewer slots
        const int G = groups <= 8 ? groups : 1;
        const int lb = std::max(1, std::min(light, kSopBlocksPerSector / G));
        const int sblocks = heavy_blocks(O->n_hitems) + kGridSectors * lb;
        GopArgs o = *O;
        o.sop_waves = 4 * lb;
        o.sop_groups = G;

        const size_t lds = lds_base + sizeof(float2) * 4 * kSopSlots * P;
        hipError_t e = gr_allow_lds(k_grid_op<ACCUM, P, true>, lds);
        if (e != hipSuccess) return e;
        if (G > 1)
            hipLaunchKernelGGL((k_grid_op<ACCUM, P, true>), dim3(G * sblocks, 1), dim3(256), lds, s, A, o);
        else
            hipLaunchKernelGGL((k_grid_op<ACCUM, P, true>), dim3(sblocks, groups), dim3(256), lds, s, A, o);
        return hipGetLastError();
    }
    const size_t lds = lds_base + sizeof(float2) * 128 * wpb * P;  // two staging buffers per wave
    hipError_t e = gr_allow_lds(k_grid<ACCUM, P>, lds);
    if (e != hipSuccess) return e;
    const int blocks = kGridSectors * ((A.sector_len + wpb - 1) / wpb);
    hipLaunchKernelGGL((k_grid<ACCUM, P>), dim3(blocks, groups), dim3(64 * wpb), lds, s, A);
    return hipGetLastError();
}

template <int ACCUM>
hipError_t launch_all(GridArgs A, const GopArgs *O, int pairs, size_t lds_base, int wpb, hipStream_t s) {
    // groups of 8 pairs; a remainder of 5-7 rides in one more (partial) group
    // of 8, one of 3-4 in a group of 4, 2 and 1 in their own widths: a tap's
    // geometry costs the same for any width, so one partial wide group beats
    // several narrow launches
    const int full = pairs / 8, rem = pairs % 8;
    hipError_t e = hipSuccess;
    A.pair_end = pairs;
    A.pair_base = 0;
    if (full + (rem >= 5) > 0) e = launch<ACCUM, 8>(A, O, full + (rem >= 5), lds_base, wpb, s);
    if (e != hipSuccess || rem == 0 || rem >= 5) return e;
    A.pair_base = full * 8;
    if (rem >= 3) return launch<ACCUM, 4>(A, O, 1, lds_base, wpb, s);
    if (rem == 2) return launch<ACCUM, 2>(A, O, 1, lds_base, wpb, s);
    return launch<ACCUM, 1>(A, O, 1, lds_base, wpb, s);
}

}  // namespace

float gr_grid_rmax(const GrGeom &g) {
    const float L2m = g.L2 + 0.01f;
    const float dmax = 3.5f * 1.41421356f + L2m * 1.41421356f + 0.05f;
    // largest sample radius (|(COSE,SINE)| <= 1 + 1e-4 for the float recurrence)
    return g.scale * (float)(g.half - 1) * 1.0001f + dmax + 0.5f;
}

namespace {

GridArgs make_args(const GrGeom &g, const GrDevTables &t, int n_ang, long ltbl) {
    GridArgs A;
    A.spec1 = A.spec2 = nullptr;
    A.ht = nullptr;
    A.ang = t.ang;
    A.aord = t.aord;
    A.akey = t.akey;
    A.tiles = t.tiles;
    A.wtbl = t.wtbl;
    A.dwtbl = t.dwtbl;
    A.ntiles = t.ntiles;
    A.sector_len = t.sector_len;
    A.n_ang = n_ang;
    A.half = (int)g.half;
    // wtbl (and dwtbl) in LDS; keep the staging area 16-byte aligned
    A.lds_wtbl = (int)(((t.dwtbl ? 2 : 1) * (ltbl + 1) + 3) & ~3L);
    A.pair_base = 0;
    A.pair_end = 0;
    A.g.scale = g.scale;
    A.g.L2 = g.L2;
    A.g.tblspcg = g.tblspcg;
    A.g.fM2 = (float)(g.M >> 1);
    A.g.M = (int)g.M;
    A.g.half = (int)g.half;
    A.g.ltbl = (int)ltbl;
    A.L2m = g.L2 + 0.01f;
    A.dmax = 3.5f * 1.41421356f + A.L2m * 1.41421356f + 0.05f;
    A.rmax = gr_grid_rmax(g);
    return A;
}

}  // namespace

int gr_grid_tile_size() { return kTile; }

hipError_t gr_launch_grid(const GrGeom &g, const GrDevTables &t, int n_ang, long ltbl,
                          const float2 *spec1, const float2 *spec2, float2 *ht, float2 *scratch, int pairs,
                          int accum, hipStream_t s) {
    GridArgs A = make_args(g, t, n_ang, ltbl);
    A.spec1 = spec1;
    A.spec2 = spec2;
    A.ht = ht;
    // segments only with compensated sums (their merges are the cut points)
    const bool cut = t.hmeta && t.n_split > 0 && accum == GR_ACCUM_COMPENSATED;
    GopArgs O{};
    O.sop_groups = 1;
    O.gop = t.gop;
    O.tile_n = t.gop_n;
    O.tile_off = t.gop_off;
    O.n_heavy = t.gop_heavy;
    O.sop_tile = t.sop_tile;
    O.sop_ch = t.sop_ch;
    O.sop_keys = t.sop_keys;
    O.sop_ent = t.sop_ent;
    O.hitems = t.hitems;
    O.n_hitems = t.n_hitems;
    O.hmeta = cut ? t.hmeta : nullptr;
    O.hseg = t.hseg;
    O.hsplit = t.hsplit;
    O.n_split = cut ? t.n_split : 0;
    O.part = scratch;
    O.part_pairs = pairs;
    const GopArgs *op = t.gop ? &O : nullptr;
    const size_t lds_base = sizeof(float) * A.lds_wtbl;
    const int wpb = g.grid_wpb == 1 || g.grid_wpb == 2 ? g.grid_wpb : 4;
    hipError_t e = accum == GR_ACCUM_PLAIN ? launch_all<GR_ACCUM_PLAIN>(A, op, pairs, lds_base, wpb, s)
                                           : launch_all<GR_ACCUM_COMPENSATED>(A, op, pairs, lds_base, wpb, s);
    if (e != hipSuccess || !op || !cut) return e;
    A.pair_base = 0;
    A.pair_end = pairs;
    const int waves = O.n_split * pairs;
    hipLaunchKernelGGL(k_gop_merge, dim3((waves + 3) / 4), dim3(256), 0, s, A, O, pairs);
    return hipGetLastError();
}

bool gr_gop_fits(const GrGeom &g, int n_ang) { return (long)n_ang * g.half <= (long)kGopIdx + 1; }

hipError_t gr_launch_gop_count(const GrGeom &g, const GrDevTables &t, int n_ang, long ltbl, unsigned *tile_n,
                               unsigned *tile_na, unsigned *tile_g, hipStream_t s) {
    const GridArgs A = make_args(g, t, n_ang, ltbl);
    hipLaunchKernelGGL(k_gop_build<0>, dim3((A.ntiles + 3) / 4), dim3(256), 0, s, A, tile_n, tile_na,
                       (const unsigned *)nullptr, (uint2 *)nullptr, tile_g, (const uint4 *)nullptr,
                       (uint2 *)nullptr, 0);
    return hipGetLastError();
}

hipError_t gr_launch_gop_fill(const GrGeom &g, const GrDevTables &t, int n_ang, long ltbl, unsigned *tile_n,
                              const unsigned *tile_off, uint2 *gop, const uint4 *hmeta, uint2 *hseg, int gps,
                              hipStream_t s) {
    const GridArgs A = make_args(g, t, n_ang, ltbl);
    hipLaunchKernelGGL(k_gop_build<1>, dim3((A.ntiles + 3) / 4), dim3(256), 0, s, A, tile_n, (unsigned *)nullptr,
                       tile_off, gop, (unsigned *)nullptr, hmeta, hseg, gps);
    return hipGetLastError();
}

bool gr_sop_fits(const GrGeom &g, const GrDevTables &t, long ltbl) {
    (void)g;
    return !t.dwtbl && ltbl >= 0 && ltbl <= 0xfff;
}

hipError_t gr_launch_sop_count(const GrGeom &g, const GrDevTables &t, int n_ang, long ltbl, int n_heavy,
                               uint4 *tile, hipStream_t s) {
    const GridArgs A = make_args(g, t, n_ang, ltbl);
    hipLaunchKernelGGL(k_sop_build<0>, dim3((A.ntiles + 3) / 4), dim3(256), 0, s, A, n_heavy, tile, (uint2 *)nullptr,
                       (unsigned *)nullptr, (unsigned *)nullptr);
    return hipGetLastError();
}

hipError_t gr_launch_sop_fill(const GrGeom &g, const GrDevTables &t, int n_ang, long ltbl, int n_heavy,
                              const uint4 *tile, uint2 *ch, unsigned *keys, unsigned *ent, hipStream_t s) {
    const GridArgs A = make_args(g, t, n_ang, ltbl);
    hipLaunchKernelGGL(k_sop_build<1>, dim3((A.ntiles + 3) / 4), dim3(256), 0, s, A, n_heavy, (uint4 *)tile, ch,
                       keys, ent);
    return hipGetLastError();
}
