/*
 * gridrec_hip.h -- C-ABI of the MI355X gridrec library (libgridrec_hip.so).
 *
 * Plain C types only: pointers, sizes, status codes.  A HIP stream is passed
 * as an opaque `void *` (a hipStream_t; NULL = the null stream, which orders
 * the work with the device's other blocking streams).
 *
 * Each entry point names the reference interface it replaces
 * (gridrec/GridrecCPP/Gridrec5.2 unless stated):
 *
 *   gr_get_pswf      <- get_pswf          pswf.cpp:36-47, grid.h:83
 *   gr_filter_lookup <- get_filter        filters.cpp:36-48, grid.h:86
 *   gr_plan_create   <- recon_init        grid.cpp:54-147, grid.h:89
 *   gr_recon_host    <- do_recon          grid.cpp:152-415 (5.x globals form),
 *                                         Gridrec4.9/grid.cpp:146 (explicit form)
 *   gr_recon_device  <- do_recon over a batch of slice pairs already in HBM
 *                       (the pair loop of Multitomo_unit.txt:823-856 and
 *                       tomo_recon.cpp:3700-3863 moved onto the GPU)
 *   gr_recon_device_raw <- the same from detector counts, with the flat/dark
 *                       normalisation and sub-pixel shift of tomo_recon.cpp:1877,
 *                       3708-3780 (or find_shift.cpp:1685-1697) fused in
 *   gr_center_entropy <- the entropy sweep of find_shift.cpp:1727-1905
 *   gr_ring_correct  <- RingCorrectionSinogram, tomo_recon.cpp:3905-3972
 *   gr_plan_destroy  <- recon_free        grid.cpp:418-434
 *
 * All functions return 0 on success and a negative GR_E* code on failure;
 * gr_last_error() returns the message of the calling thread's last failure.
 * Unlike the reference (exit(2) in trig_su/legendre/get_pswf) nothing here
 * terminates the process.
 */
#ifndef GRIDREC_HIP_H
#define GRIDREC_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif
#if defined(__GNUC__)
#pragma GCC visibility push(default)
#endif

enum {
    GR_OK = 0,
    GR_EINVAL = -1,     /* invalid argument or geometry */
    GR_ENOMEM = -2,     /* device or host allocation failed */
    GR_EHIP = -3,       /* a HIP runtime call failed */
    GR_ENOPSWF = -4,    /* no PSWF for this C (pswf.cpp:39-44) */
    GR_ENODEV = -5      /* no usable GPU */
};

/* Padding applied to each sinogram row before the row FFT. */
enum {
    GR_PAD_NONE = 0,       /* the reference: rows of n_det, zero-filled to pdim (grid.cpp:240-252) */
    GR_PAD_HALF_EDGE = 1   /* tomo_recon.cpp:3691-3807 with xoffset 0: width 2*n_det, first/last
                              value replicated into n_det/2 cells each side, centre + n_det/2 */
};

/* Accumulation of the gridding taps into each H cell. */
enum {
    GR_ACCUM_COMPENSATED = 0,  /* fused multiply-add partials of a few sample lines, added with an
                                  error-free 2Sum; sum + compensation at the end */
    GR_ACCUM_PLAIN = 1         /* plain float adds, as grid.cpp:297-300 */
};

/* Where the filter/PSWF/deapodization tables come from. */
enum {
    GR_TABLES_REFERENCE = 0,   /* bit-identical to recon_init's float tables (host) */
    GR_TABLES_FP64 = 1         /* evaluated in FP64 on the GPU, then rounded to float */
};

/* The gridding step's precomputed gather operator: the list of every grid
 * cell's taps (sample, weight), built once per plan on the GPU, after which
 * each group of slice pairs streams it instead of evaluating the geometry.
 * Images are bitwise the same with or without it. */
enum {
    GR_GRIDOP_AUTO = 0,  /* built on the first call that brings the plan's total to 16 pairs, when
                            its size is at most a third of the free device memory */
    GR_GRIDOP_OFF = 1,   /* never built */
    GR_GRIDOP_ON = 2     /* built at plan creation; creation fails if it cannot be */
};

/* PSWF record, layout of pswf_struct (grid.h:37-44). */
typedef struct gr_pswf {
    float C;
    int nt;
    float lmbda;
    float coefs[15];
} gr_pswf;

typedef struct gr_params {
    /* sg_struct (grid.h:26-35) */
    int n_ang;             /* angles per sinogram */
    int n_det;             /* detector cells per row of the input */
    int geom;              /* 0: angle list in degrees; 1: uniform on [0,pi); 2: on [0,2pi) */
    const float *angles;   /* host array of n_ang degrees when geom == 0 (copied at create) */
    float center;          /* rotation axis, 0-based detector coordinate of the input rows */
    /* grid_struct (grid.h:46-58) */
    float C;               /* PSWF parameter: 4, 4.2, 5, 6 or 7 */
    float sampl;           /* oversampling ratio (the plan reports the adjusted M/M0) */
    float MaxPixSiz;
    float R;
    float X0, Y0;          /* ROI offset (units of centre-to-edge distance) */
    char filter[16];       /* shlo|shepp|hann|hamm|hamming|ramp|ramlak; unknown -> shlo */
    long ltbl;             /* PSWF lookup table length (reference default 512) */
    /* extensions of this library */
    int pad;               /* GR_PAD_* */
    long crop_off;         /* output window origin in the M0 x M0 image, rows and columns */
    long crop_n;           /* output window size; 0 means the whole M0 x M0 image */
    int accum;             /* GR_ACCUM_* */
    int tables;            /* GR_TABLES_* */
    int device;            /* HIP device ordinal */
    int max_pairs;         /* slice pairs per launch (workspace sizing); 0 = automatic */
    int grid_op;           /* GR_GRIDOP_* */
    int interp;            /* 1: the convolvent of the reference's compile-time INTERP variant
                              (grid.cpp:18-24, 487-489): linear table interpolation with the
                              backward difference wtbl[i] - wtbl[i-1], dwtbl[0] = 0 (the
                              reference never sets it; its macro does not compile as written);
                              0 (default): nearest neighbour, as the reference ships.
                              GR_TABLES_REFERENCE only */
    /* optional overrides, as recon_init honours grid_struct.filter and
       grid_struct.pswf (grid.cpp:78-81): a filter function evaluated on the
       host for the filter table (only with GR_TABLES_REFERENCE), and a PSWF
       record used instead of the database entry for C.  NULL = by name / C. */
    float (*filter_fn)(float);
    const gr_pswf *pswf;
} gr_params;

typedef struct gr_info {
    long pdim, M, M0, M02;  /* sizes of recon_init (grid.cpp:87-113) */
    float sampl;            /* adjusted sampl = M/M0 (grid.cpp:117) */
    float L, scale;         /* grid.cpp:119-120 */
    int n_det_eff;          /* row width seen by gridrec after padding */
    float center_eff;       /* centre seen by gridrec after padding */
    long out_off, out_n;    /* output window */
    int max_pairs;
    size_t device_bytes;    /* device memory held by the plan */
    size_t grid_op_bytes;   /* of which the gather operator (0 while not built) */
    double grid_op_build_ms; /* host time its build took */
} gr_info;

/* Per-kernel timing of the last gr_recon_device call(s) (ms, summed over
 * launches) when timing is enabled. */
typedef struct gr_timing {
    double rowfft_ms, grid_ms, ifft_rows_ms, ifft_cols_ms;
    long rowfft_launches, grid_launches, ifft_rows_launches, ifft_cols_launches;
    long pairs;
} gr_timing;

typedef struct gr_plan gr_plan;

/* get_pswf: copy of the database record whose C is within 0.01 of C. */
int gr_get_pswf(float C, gr_pswf *out);

/* get_filter: canonical filter id (0 shlo, 1 hann, 2 hamm, 3 ramp) for name;
 * an unknown name selects shlo and, as the reference does, writes "shlo" into
 * name_io (which must hold 16 bytes) -- pass NULL to skip the write-back. */
int gr_filter_lookup(const char *name, char *name_io, int *filter_id);

/* recon_init: sizes, device tables and workspace for one geometry. */
int gr_plan_create(const gr_params *params, gr_plan **plan);
int gr_plan_info(const gr_plan *plan, gr_info *info);

/* Copies of the plan's device tables (the globals SINE, COSE, filphase, wtbl,
 * winv of grid.cpp:32-33): n_ang, n_ang, pdim (pdim/2 complex), ltbl+1 and M0
 * floats.  Any pointer may be NULL. */
int gr_plan_tables(const gr_plan *plan, float *sine, float *cose, float *filphase, float *wtbl,
                   float *winv);

/* do_recon on host rows: G1/G2 are n_ang row pointers of n_det floats (any
 * stride), S1/S2 are out_n row pointers of out_n floats.  G2 may equal G1.
 * Synchronous. */
int gr_recon_host(gr_plan *plan, const float *const *G1, const float *const *G2,
                  float *const *S1, float *const *S2);

/* do_recon for npairs slice pairs on host rows -- the reference's own
 * host-buffer boundary for a whole batch: pair p's rows are G1[p*n_ang + n]
 * and G2[p*n_ang + n] (n < n_ang, n_det floats each, any addresses), its
 * images S1[p*out_n + j] and S2[p*out_n + j] (j < out_n, out_n floats each).
 * Pipelined through two pinned staging slots of max_pairs pairs (allocated on
 * the first host-row call): uploads and downloads run on two streams of the
 * plan's while `stream` reconstructs, and host threads pack and unpack rows
 * meanwhile.  Synchronous on return.  Replaces the per-pair do_recon loop of
 * the callers (Multitomo_unit.txt:823-856). */
int gr_exec_pairs(gr_plan *plan, const float *const *G1, const float *const *G2,
                  float *const *S1, float *const *S2, int npairs, void *stream);

/* do_recon for n_slices slices resident in device memory.  Slice s starts at
 * sino + s*sino_slice_stride and holds n_ang rows of n_det floats spaced
 * sino_row_stride floats apart.  Slices (2p, 2p+1) form pair p and share one
 * complex transform (grid.cpp:240-246); an odd last slice is paired with zeros.
 * Image s is written to out + s*out_slice_stride as out_n rows of out_n floats
 * spaced out_row_stride apart.  Asynchronous on `stream`. */
int gr_recon_device(gr_plan *plan, const float *sino, long sino_slice_stride,
                    long sino_row_stride, float *out, long out_slice_stride,
                    long out_row_stride, int n_slices, void *stream);

/* gr_recon_device without any allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, layouts and images (bitwise) as gr_recon_device; the kernels'
 * scratch (spectra, gridded and row-transformed planes) is one static
 * per-GPU array of gr_noalloc_scratch_bytes() bytes (128 MiB) instead of the
 * plan's workspace.  GR_EINVAL when one slice pair's scratch exceeds it
 * (e.g. a half-padded 1024-detector geometry beyond about 600 angles).  Calls
 * that overlap on one GPU race on that array: serialise them.  A plan whose
 * gather operator is not built yet evaluates the geometry per launch here;
 * no per-kernel timing is recorded. */
int gr_recon_device_noalloc(gr_plan *plan, const float *sino, long sino_slice_stride,
                            long sino_row_stride, float *out, long out_slice_stride,
                            long out_row_stride, int n_slices, void *stream);
size_t gr_noalloc_scratch_bytes(void);

/* do_recon from detector counts: the preprocessing the callers run on the
 * host before gridrec, fused into the row load of the same kernels.
 *   GR_PRE_FLATDARK:     v = 1 - (p - d) / (w1/2 + w2/2 - d)   (tomo_recon.cpp:1877)
 *   GR_PRE_FLATDARK_LOG: v = log((w1/2 + w2/2 - d) / (p - d)), 0 where p <= d or
 *                        the ratio is <= 0 (find_shift.cpp:1685-1697; float here)
 * then each row is shifted by xoffset cells with linear interpolation, cells
 * beyond the ends taking the edge value (tomo_recon.cpp:3708-3780; 0 = no
 * shift), then padded as the plan says.  raw: counts [slice][n_ang][row_stride];
 * dark, white1, white2: one row of n_det floats per slice, field_stride apart
 * (white2 may be NULL: white1 alone).  Slices pair and images land as in
 * gr_recon_device. */
enum { GR_PRE_FLATDARK = 1, GR_PRE_FLATDARK_LOG = 2 };
typedef struct gr_preproc {
    int mode;
    const unsigned short *raw;
    long raw_slice_stride, raw_row_stride;
    const float *dark, *white1, *white2;
    long field_stride;
    float xoffset;
} gr_preproc;
int gr_recon_device_raw(gr_plan *plan, const gr_preproc *pre, float *out, long out_slice_stride,
                        long out_row_stride, int n_slices, void *stream);

/* Centre sweep with the entropy metric (find_shift.cpp:1727-1905,
 * 2078-2111): for each of n_shifts shifts, the two slices of sino (float
 * rows, sino_slice_stride apart) are shifted by that many cells (linear
 * interpolation, edge extension, as gr_preproc.xoffset), reconstructed with
 * the plan (typically GR_PAD_HALF_EDGE and the n_det window, centre n_det/2),
 * and the entropy in bits of the histogram of both images is written to
 * entropy[i] (host): n_bins bins of width (hist_max - hist_min)/n_bins from
 * hist_min, values in bin n_bins (the upper edge) counted, values outside
 * ignored.  The candidates are reconstructed as one batch.  Synchronous. */
int gr_center_entropy(gr_plan *plan, const float *sino, long sino_slice_stride, long sino_row_stride,
                      const float *shifts, int n_shifts, float hist_min, float hist_max, int n_bins,
                      float *entropy, void *stream);

/* Ring-artefact correction of n_slices float sinograms in place, the step
 * tomo_recon runs between normalisation and gridrec (RingCorrectionSinogram,
 * tomo_recon.cpp:3905-3972; the reference runs 20 iterations with
 * RING_COEFF): per iteration each row is rescaled to the global mean row
 * level, and the column means minus their 3-tap low pass, times ring_coeff,
 * are subtracted (results clamped at 0).  Asynchronous on `stream`. */
int gr_ring_correct(float *sino, long slice_stride, long row_stride, int n_slices, int n_ang, int n_det,
                    float ring_coeff, int iterations, void *stream);

/* recon_free. */
int gr_plan_destroy(gr_plan *plan);

/* Build the plan's gather operator now, on `stream` (NULL: the null stream),
 * as GR_GRIDOP_ON does at creation; synchronises that stream.  No effect on a
 * plan that has one or whose GR_GRIDOP_AUTO build declined; GR_ENOMEM when it
 * cannot be built.  No counterpart in the reference (grid.h): an extension for
 * callers that choose when the build's few milliseconds are spent. */
int gr_plan_build_grid_operator(gr_plan *plan, void *stream);

/* Workgroup width of the plan's column transform: cols columns per workgroup
 * (1, 2, 4 or 8 side by side while they fit a workgroup and the CU's LDS, or
 * one after another while cols * M <= 16384); 0 restores the automatic choice.
 * Results do not depend on it (bitwise).  GR_EINVAL for a width neither form
 * takes.  No counterpart in the reference (grid.h): a launch-tuning knob, the
 * one the batch queue's autotuner sets per dataset. */
int gr_plan_set_cols_per_block(gr_plan *plan, int cols);

/* Per-kernel HIP-event timing of gr_recon_device: an event pair around each
 * launch, resolved (waited for) by gr_get_timing. */
int gr_set_timing(gr_plan *plan, int enable);
int gr_get_timing(gr_plan *plan, gr_timing *out, int reset);

/* Synthetic sinograms on the device (the generator of projects_amd/phantom.py):
 * n_slices slices [slice0, slice0+n_slices) of seed, each n_ang x n_det floats,
 * n_ell ellipses; photons > 0 adds Poisson noise and -log; density 0 is
 * U[-0.3,1], 1 is log-uniform over 4 decades; gauss_rel > 0 multiplies each
 * value by (1 + gauss_rel * z), z ~ N(0,1); n_tex > 0 adds a band-limited
 * random texture inside the unit disk, n_tex plane waves with frequencies
 * uniform over |f| <= tex_fmax cycles per unit length and Gaussian amplitudes
 * of rms tex_amp in total.  centers[s] is the axis of slice s (n_slices
 * floats, host).  Synchronous. */
int gr_phantom_sinograms(float *sino, long slice_stride, int n_slices, long slice0,
                         unsigned long long seed, int n_ang, int n_det,
                         const float *centers, int n_ell, float photons, int density,
                         float gauss_rel, int n_tex, float tex_amp, float tex_fmax, void *stream);

/* Batched reconstruction of many datasets of uneven geometry, all resident in
 * device memory -- the dataset loop of the reference's drivers
 * (Multitomo_unit.txt:823-856 runs recon_init/do_recon/recon_free once per
 * dataset; tomo_recon.cpp:3700-3863 once per slice batch) made one launch.
 * Each dataset has its own gr_params and buffers (arguments as in
 * gr_recon_device).  gr_queue_create builds one plan per dataset, autotunes
 * slice pairs per launch, the column-FFT width, the largest FFT radix and
 * the gridding workgroup size per size class (when autotune != 0), spreads the datasets over n_streams streams (1..8) and
 * captures all their launches into one HIP graph; it runs the batch while
 * tuning, so the output buffers are written.  gr_queue_run replays the graph
 * on `stream`, asynchronously; the images
 * are bitwise those of gr_recon_device on each dataset.  The buffers named at
 * create are the ones every run reads and writes. */
typedef struct gr_dataset {
    gr_params params;
    const float *sino;
    long sino_slice_stride, sino_row_stride;
    float *out;
    long out_slice_stride, out_row_stride;
    int n_slices;
} gr_dataset;

typedef struct gr_queue_info {
    int n_datasets, n_streams, size_classes, tuned_candidates;
    long slices;            /* total slices of the batch */
    long graph_nodes;
    size_t workspace_bytes; /* shared per-stream workspaces */
    int datasets_with_operator; /* datasets gridded from a precomputed gather operator */
    size_t operator_bytes;      /* device memory those operators hold */
} gr_queue_info;

typedef struct gr_queue gr_queue;

int gr_queue_create(const gr_dataset *datasets, int n_datasets, int device, int n_streams,
                    int autotune, gr_queue **queue);
int gr_queue_run(gr_queue *queue, void *stream);
/* The same launches without the graph, every dataset in turn on one stream. */
int gr_queue_run_sequential(gr_queue *queue, void *stream);
int gr_queue_get_info(const gr_queue *queue, gr_queue_info *info);
/* Tuned settings of dataset d: pairs per launch, column-FFT width (0 =
 * automatic), largest FFT radix (16 or 4), gridding wavefronts per workgroup,
 * stream index.  Any pointer may be NULL. */
int gr_queue_dataset_settings(const gr_queue *queue, int d, int *max_pairs, int *cols_per_block,
                              int *fft_radix, int *grid_wpb, int *stream);
int gr_queue_destroy(gr_queue *queue);

const char *gr_last_error(void);
const char *gr_version(void);

#if defined(__GNUC__)
#pragma GCC visibility pop
#endif
#ifdef __cplusplus
}
#endif
#endif
