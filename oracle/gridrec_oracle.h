/*
 * gridrec_oracle.h -- CPU oracle for the gridrec hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code.  It is used by tests/, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py, always as the checker, never as the thing
 * measured or shipped.
 *
 * It restates the reference gridrec (gridrec/GridrecCPP/Gridrec5.2/grid.cpp,
 * filters.cpp, pswf.cpp) in plain C with the reference's float/double mix
 * reproduced expression by expression, and adds an FP64 variant that makes the
 * same discrete decisions (box bounds, table indices) with every value held in
 * double.  The FFTs are this file's own radix-2 transforms with the reference's
 * conventions (fftgr.cpp:8-52 four1 with isign=+1 is sum x_k e^{+2 pi i jk/N},
 * fftgr.cpp:56-120 fourn with isign=-1 is e^{-2 pi i ...}; both unnormalised).
 */
#ifndef GRIDREC_ORACLE_H
#define GRIDREC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Filter ids in the order of the reference's filter table (filters.cpp:26-34). */
enum { OR_SHLO = 0, OR_HANN = 1, OR_HAMM = 2, OR_RAMP = 3 };

typedef struct or_params {
    int n_ang;            /* sg_struct.n_ang  (grid.h:28) */
    int n_det;            /* sg_struct.n_det  (grid.h:29), width of the rows handed to do_recon */
    int geom;             /* sg_struct.geom   (grid.h:30): 0 angle list, 1 half circle, 2 full circle */
    int filter_id;        /* OR_SHLO ... OR_RAMP (grid_struct.filter, grid.h:55) */
    float center;         /* sg_struct.center (grid.h:33) */
    float C;              /* pswf->C (grid.h:39) */
    float sampl;          /* grid_struct.sampl (grid.h:49) */
    float MaxPixSiz;      /* grid.h:50 */
    float R;              /* grid.h:51 */
    float X0, Y0;         /* grid.h:52-53 */
    long ltbl;            /* grid.h:56 */
    const float *angles;  /* degrees, used when geom == 0 (grid.h:32) */
    int interp;           /* 1: the convolvent of the reference's compile-time INTERP variant
                             (grid.cpp:18-24, 487-489): wtbl[i] + (X - i) * dwtbl[i], i = (int)X,
                             with the backward difference dwtbl[i] = wtbl[i] - wtbl[i-1] and
                             dwtbl[0] (never set there) taken as 0.  The reference's macro does
                             not compile as written (its argument is not parenthesised, so
                             Cnvlvnt(abs(V-iv)*tblspcg) subscripts wtbl with a float); this is
                             its evident reading.  PARITY UNPINNED against the reference (no
                             build of it exists); pinned by the convolvent's second-order
                             accuracy against the exact PSWF.  or_tables, or_recon_f32 and
                             or_recon_batch_f32 only (the FP64 variants return -1) */
} or_params;

typedef struct or_sizes {
    long pdim, M, M0, M02;
    float sampl;          /* adjusted sampl = M/M0 (grid.cpp:117) */
    float L, scale, L2, tblspcg;
    int flag;             /* ROI offset present (grid.cpp:84-85) */
} or_sizes;

/* Sizes computed exactly as recon_init (grid.cpp:87-120). Returns 0, or -1 on bad input. */
int or_sizes_get(const or_params *p, or_sizes *s);

/* Tables exactly as recon_init builds them (trig_su, filphase_su, pswf_su).
 * sine/cose: n_ang floats; filphase: pdim/2 complex as 2*pdim/2 floats;
 * wtbl: ltbl+1 floats; winv: M0 floats.  Any output pointer may be NULL. */
int or_tables(const or_params *p, float *sine, float *cose, float *filphase,
              float *wtbl, float *winv);

/* FP32 restatement of do_recon (grid.cpp:152-415): rows g1/g2 are n_ang x n_det
 * contiguous; outputs s1/s2 are M0 x M0 contiguous.  Returns 0 or -1. */
int or_recon_f32(const or_params *p, const float *g1, const float *g2,
                 float *s1, float *s2);

/* FP64 oracle: same float SINE/COSE and the same float-evaluated U, V for every
 * discrete decision; every value and every sum in double. */
int or_recon_f64(const or_params *p, const float *g1, const float *g2,
                 double *s1, double *s2);

/* Exact-math FP64 variant: exact angles (geom 1/2: n*geom*pi/n_ang; geom 0 the
 * list), U, V and all decisions in double. */
int or_recon_exact64(const or_params *p, const float *g1, const float *g2,
                     double *s1, double *s2);

/* Mixed-precision oracle for the precision study.  The pipeline runs in double
 * and every operation of a stage whose bit is clear in f64_mask is rounded to
 * float, so each stage can be switched between FP32 and FP64 on its own:
 *   OR_MX_ROWFFT  row FFT and filter/phase multiply (grid.cpp:240-274)
 *   OR_MX_ACCUM   gridding weights and the H accumulation (grid.cpp:276-302)
 *   OR_MX_FFT2    2D inverse FFT (grid.cpp:326-328)
 *   OR_MX_TABLES  filphase, wtbl, winv in double instead of the float tables
 *   OR_MX_DEAPOD  phase-3 products (grid.cpp:401-403)
 * Discrete decisions always use the reference's float U, V.  With mask 0 the
 * result equals or_recon_f32 up to the rare double-rounding of a sum; with all
 * bits set it equals or_recon_f64. */
enum { OR_MX_ROWFFT = 1, OR_MX_ACCUM = 2, OR_MX_FFT2 = 4, OR_MX_TABLES = 8, OR_MX_DEAPOD = 16 };
int or_recon_mixed(const or_params *p, const float *g1, const float *g2,
                   double *s1, double *s2, int f64_mask);

/* Caller convention of gridrec/GPU/tomo_recon.cpp:3691-3807 with xoffset = 0:
 * each row of n_det values is widened to 2*n_det with the first and last
 * detector value replicated into n_det/2 (integer division) cells on each side.
 * dst must hold n_ang * 2*n_det floats. */
void or_pad_half(int n_ang, int n_det, const float *src, float *dst);

/* The callers' preprocessing of detector counts before gridrec, n_ang rows of
 * n_det counts, one row each of dark, white1, white2:
 *   mode 1 -- 1 - (p - d)/(0.5 w1 + 0.5 w2 - d) in float (tomo_recon.cpp:1877);
 *   mode 2 -- log((0.5 (w1 + w2) - d)/(p - d)) in double, 0 where p - d <= 0 or
 *             the ratio is <= 0 (find_shift.cpp:1685-1697);
 * then each row is shifted by xoffset with linear interpolation and edge
 * extension (tomo_recon.cpp:3708-3780).  Returns 0, or -1 on a bad mode. */
int or_preprocess(int mode, int n_ang, int n_det, const unsigned short *raw, const float *dark,
                  const float *white1, const float *white2, float xoffset, float *out);

/* Entropy in bits of the histogram of n values (find_shift.cpp:2078-2111):
 * bin (int)((v - min) / ((max - min) / n_bins)), counted when 0 <= bin <= n_bins;
 * p in float, -p log p with the float log, divided by log 2 in double, summed
 * into a float. */
float or_entropy(const float *data, long n, float vmin, float vmax, int n_bins);

/* RingCorrectionSinogram (tomo_recon.cpp:3905-3972) on one n_ang x n_det
 * sinogram in place, `iterations` rounds (the reference: 20), in the
 * reference's float arithmetic and summation order. */
void or_ring_correct(float *data, int n_ang, int n_det, float ring_coeff, int iterations);

/* Batch driver used for the CPU baseline: n_pairs slice pairs of the raw
 * (unpadded) sinograms sino[2*pair + {0,1}][n_ang][n_det]; when pad_half is set
 * each pair is padded with or_pad_half and reconstructed with center + n_det/2;
 * the crop window [crop_off, crop_off+crop_n)^2 of each M0 x M0 image is stored
 * in out[2*pair + {0,1}][crop_n][crop_n].  p->n_det is the unpadded width.
 * Pairs are spread over n_threads POSIX threads.  Returns 0 or -1. */
int or_recon_batch_f32(const or_params *p, int pad_half, const float *sino,
                       int n_pairs, long crop_off, long crop_n, float *out,
                       int n_threads);

/* The oracle's 1D FFTs (interleaved re/im, n a power of two), for tests. */
int or_fft_f32(float *x, long n, int sign);
int or_fft_f64(double *x, long n, int sign);

#ifdef __cplusplus
}
#endif
#endif
