// gr_internal.h -- types shared by the host C-ABI and the HIP kernels of the
// MI355X gridrec library.  Not installed; the public interface is
// include/gridrec_hip.h (C-ABI) and include/grid.h (reference C++ API).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/gridrec_hip.h"

// Sizes and scalars of one geometry, as recon_init derives them
// (gridrec/GridrecCPP/Gridrec5.2/grid.cpp:87-120, 223-226).
struct GrGeom {
    long pdim, M, M0, M02, half;  // half = pdim/2; samples j = 1 .. half-1 are gridded
    float sampl;                  // adjusted M/M0
    float L, scale, L2, tblspcg;
    int flag;                     // ROI offset present
    int n_det_in;                 // width of the caller's rows
    int n_det_eff;                // width after padding (what grid.cpp calls n_det)
    float center_eff;             // centre after padding
    int pad;                      // GR_PAD_*
    int filter_id;
    gr_pswf pswf;
    long out_off, out_n;          // output window inside M0 x M0
    // launch tuning (no effect on results): largest FFT radix (16 or 4) and
    // wavefronts (8x8 tiles) per gridding workgroup (1, 2 or 4)
    int fft_radix = 16, grid_wpb = 2;
};

// Z, the row-transformed grid between K3 and K4, is held column-blocked:
// per pair, [window column / 4][row b < M][window column % 4] float2, so the
// four columns a K4 workgroup transforms are one contiguous M x 32-byte block
// (the window's width rounded up to a multiple of four; the padding columns
// are never written nor stored).
__host__ __device__ __forceinline__ long gr_z_cols(long out_n) { return (out_n + 3) / 4 * 4; }
__host__ __device__ __forceinline__ long gr_z_index(long M, long b, long jj) { return ((jj >> 2) * M + b) * 4 + (jj & 3); }

// Host side (gr_tables.cpp): sizes and the reference's float tables.
int gr_geom_from_params(const gr_params *p, GrGeom *g, const char **why);
int gr_find_pswf(float C, gr_pswf *out);
int gr_filter_id_of(const char *name);
// sine/cose: n_ang; filphase: 2*half floats; wtbl: ltbl+1; winv: M0;
// roi_offset: n_ang (phase offset per angle, grid.cpp:237; zeros without ROI);
// dwtbl: ltbl+1 backward differences of wtbl (filled with p->interp; may be null).
void gr_host_tables(const gr_params *p, const GrGeom &g, float *sine, float *cose,
                    float *filphase, float *wtbl, float *winv, float *roi_offset, float *dwtbl);

// Device tables and workspace of a plan.
struct GrDevTables {
    const float *cose, *sine;      // n_ang
    const float *roi_offset;       // n_ang
    const float2 *filphase;        // half
    const float *wtbl;             // ltbl+1
    const float *dwtbl = nullptr;  // ltbl+1 backward differences: the INTERP convolvent (null: nearest)
    const float *winv;             // M0
    const float2 *tw_row;          // pdim: e^{+2 pi i m/pdim}, pass-ordered (gr_twp_len)
    const float2 *tw_grid;         // M:    e^{-2 pi i m/M}, pass-ordered
    const float *akey;             // n_ang angle keys in [0, pi), sorted
    const int *aord;               // n_ang angle index of each sorted key
    const float4 *ang;             // n_ang in sorted order: COSE, SINE, 1/COSE, 1/SINE (0 if |.| <= 1e-6)
    const unsigned *tiles;         // tile list (tx | ty << 16): kGridSectors sectors of sector_len
                                   // slots, nearest to the centre first, kGridNoTile padding
    int ntiles, sector_len;        // ntiles = kGridSectors * sector_len
    // precomputed gather operator of K2 (null when the plan has none):
    // entries, taps per tile, first entry of each tile
    const uint2 *gop = nullptr;
    const unsigned *gop_n = nullptr, *gop_off = nullptr;
    int gop_heavy = 0;  // leading slots of every sector gridded one pair per wave
    // the heavy work: items {tile slot, segment}, the centre-most slots first;
    // heavy tiles cut into segments of kGopSegGroups merge groups, gridded by
    // one wave per (segment, pair) into partials that k_gop_merge sums:
    // per tile {segments (0/1: not cut), first hseg entry, first partial
    // group, index in hsplit}; per segment boundary and lane {row, merges};
    // the cut tiles' slots
    const uint2 *hitems = nullptr;
    int n_hitems = 0;
    const uint4 *hmeta = nullptr;
    const uint2 *hseg = nullptr;
    const unsigned *hsplit = nullptr;
    int n_split = 0;       // cut tiles
    int split_groups = 0;  // partial groups of all cut tiles
    // the staged form of the operator for the other tiles (null: they use the
    // entries above): per tile {first extra chunk, chunks, 0, first row}; per
    // chunk {rows | merge bit, slots} and 128 keys (the records its LDS slots
    // take), chunk 0 of tile slot t at index t; per chunk row 64 entries
    // (slot | iu << 8 | iv << 20)
    const uint4 *sop_tile = nullptr;
    const uint2 *sop_ch = nullptr;
    const unsigned *sop_keys = nullptr, *sop_ent = nullptr;
    // per tile row (M / 8 rows) the interval [lo, hi] of tiles that can hold
    // a nonzero cell (gr_active_tile_range), evaluated once on the device
    const int2 *row_tiles = nullptr;
};

// Compute units of the current device, and the most LDS one workgroup may
// take (queried once per device; 0 when the query fails).
int gr_cu_count();
int gr_max_lds_per_block();

// ---- kernel launchers (one per .hip file) ----------------------------------

// The spectra of a chunk's pairs are interleaved by groups of kSpecGroup
// pairs: pair p's sample j of angle n is element
// ((p / kSpecGroup) * n_ang + n) * half + j) * kSpecGroup + p % kSpecGroup
// of spec1 (spec2), so the gridding step reads one tap of all the pairs of a
// group as one contiguous 64-byte record.
constexpr int kSpecGroup = 8;

// K1: row FFT (+1), filter/phase, Hermitian split into spec1/spec2 (layout
// above); pair p of the chunk reads slices 2*(pair0+p) and +1.
// Detector-count input of K1: flat/dark normalisation and sub-pixel shift
// applied while loading (mode GR_PRE_FLATDARK or GR_PRE_FLATDARK_LOG).
struct GrPre {
    int mode;
    const unsigned short *raw;
    long raw_slice_stride, raw_row_stride;
    const float *dark, *white1, *white2;
    long field_stride;
    float xoffset;
    // mode 3 (centre sweeps): float rows of `sino`, shifted by
    // pair_xoffset[pair] (device) instead of normalised counts
    const float *pair_xoffset;
};
// sino is read when pre is null or in mode 3, pre->raw otherwise.
hipError_t gr_launch_rowfft(const GrGeom &g, const GrDevTables &t, int n_ang, const float *sino,
                            long slice_stride, long row_stride, const GrPre *pre, int n_slices, int pair0,
                            int pairs, float2 *spec1, float2 *spec2, hipStream_t s);

// K2: tile-local gridding into the transposed grid Ht[pairs][M][M] (Ht[b][a] = H[a][b]).
// The plan lists the kernel's tiles (gr_grid_tile_size() cells per side).
int gr_grid_tile_size();
// Radius (cells, from the grid centre) beyond which tiles hold no nonzero
// cell: gridding leaves them unwritten, the row transform reads them as zero
// (gr_tile_active).
float gr_grid_rmax(const GrGeom &g);
// Sectors of the tile list: one per XCD of the MI355X (workgroup b runs on
// XCD b mod 8), and the padding value of a sector's unused slots.
constexpr int kGridSectors = 8;
constexpr unsigned kGridNoTile = 0xffffffffu;
// With t.gop set, K2 streams the precomputed operator instead of evaluating
// the geometry; the images are bitwise the same.
// `scratch` (pairs * M * out_n complex, free while K2 runs: the row
// transform's output) holds the partials of heavy tiles gridded in segments.
hipError_t gr_launch_grid(const GrGeom &g, const GrDevTables &t, int n_ang, long ltbl,
                          const float2 *spec1, const float2 *spec2, float2 *ht, float2 *scratch, int pairs,
                          int accum, hipStream_t s);

// The precomputed operator: whether the geometry's sample indices fit its
// entries; the count pass (rows per tile of the ragged layout into tile_n,
// of the aligned layout into tile_na); the fill pass (entries at
// tile_off[tile]; tile_n then holds the chosen rows, with kGopAlignedBit set
// for the aligned layout, and tile_off the exclusive sum of 64 * rows).
constexpr unsigned kGopAlignedBit = 1u << 31;
bool gr_gop_fits(const GrGeom &g, int n_ang);
hipError_t gr_launch_gop_count(const GrGeom &g, const GrDevTables &t, int n_ang, long ltbl, unsigned *tile_n,
                               unsigned *tile_na, unsigned *tile_g, hipStream_t s);
// hmeta/hseg (may be null): the tiles gridded in segments (below) get their
// per-lane segment boundaries in hseg.
hipError_t gr_launch_gop_fill(const GrGeom &g, const GrDevTables &t, int n_ang, long ltbl, unsigned *tile_n,
                              const unsigned *tile_off, uint2 *gop, const uint4 *hmeta, uint2 *hseg, int gps,
                              hipStream_t s);
// Merge groups per segment of a heavy tile gridded in segments.
constexpr int kGopSegGroups = 1;
// The staged operator of the tiles past the first n_heavy slots of every
// sector: whether the plan admits it (nearest-neighbour convolvent, table
// indices within its entries); the count pass ({0, chunks, 0, rows} per tile
// into tile); the fill pass (tile then holds {first extra chunk, chunks, 0,
// first row}; ch and keys have ntiles + the extra chunks' entries).
bool gr_sop_fits(const GrGeom &g, const GrDevTables &t, long ltbl);
hipError_t gr_launch_sop_count(const GrGeom &g, const GrDevTables &t, int n_ang, long ltbl, int n_heavy,
                               uint4 *tile, hipStream_t s);
hipError_t gr_launch_sop_fill(const GrGeom &g, const GrDevTables &t, int n_ang, long ltbl, int n_heavy,
                              const uint4 *tile, uint2 *ch, unsigned *keys, unsigned *ent, hipStream_t s);
constexpr int kSopKeysPerChunk = 128;

// K3: FFT (-1) of each row of Ht (over a), output rows of the window only:
// Z[pairs][M][out_n].
// The active-tile interval of every tile row of the M x M grid into
// out[M / 8] (gr_active_tile_range on the device, the very arithmetic of the
// gridding kernels' own test).
hipError_t gr_launch_row_tiles(const GrGeom &g, int2 *out, hipStream_t s);
hipError_t gr_launch_ifft_rows(const GrGeom &g, const GrDevTables &t, const float2 *ht, float2 *z,
                               int pairs, hipStream_t s);

// K4: FFT (-1) down the columns of Z (over b), deapodization and crop into the
// caller's images.  Pair p writes slices 2*(pair0+p) (real part) and +1 (imag)
// unless that slice index is >= n_slices.
hipError_t gr_launch_ifft_cols(const GrGeom &g, const GrDevTables &t, const float2 *z, float *out,
                               long out_slice_stride, long out_row_stride, int n_slices, int pair0,
                               int pairs, int cols_per_block, hipStream_t s);
// Whether gr_launch_ifft_cols takes cols columns per workgroup for M (else it
// falls back to its automatic width).
bool gr_cols_width_ok(long M, int cols);

// FP64 tables on the device, rounded to float: filphase (half), wtbl (ltbl+1),
// winv (M0).
hipError_t gr_launch_tables_fp64(const GrGeom &g, const gr_params *p, float2 *filphase, float *wtbl,
                                 float *winv, hipStream_t s);

// Synthetic sinograms.
struct GrEllipse { float a, b, rho, phi, x0, y0; };
struct GrWave { float amp, fx, fy, phase; };
hipError_t gr_launch_phantom(float *sino, long slice_stride, int n_slices, long slice0,
                             unsigned long long seed, int n_ang, int n_det, const float *centers_dev,
                             int n_ell, float photons, int density, float gauss_rel, int n_tex, float tex_amp,
                             float tex_fmax, hipStream_t s);

// Histograms for the entropy metric: n_cand candidates, each two on x on
// images (slice_stride / row_stride floats apart, candidates cand_stride
// apart); hist receives n_cand x (n_bins + 1) counts.
hipError_t gr_launch_hist(const float *img, long cand_stride, long slice_stride, long row_stride, int on,
                          int n_cand, float vmin, float vmax, int n_bins, unsigned *hist, hipStream_t s);

// Ring-artefact correction in place (tomo_recon.cpp:3905-3972); row_mean and
// col_mean are n_slices x n_ang and n_slices x n_det scratch floats.
hipError_t gr_launch_ring(float *sino, long slice_stride, long row_stride, int n_slices, int n_ang, int n_det,
                          float coeff, int iterations, float *row_mean, float *col_mean, hipStream_t s);
