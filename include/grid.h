/*
 * grid.h -- drop-in replacement for the GridrecCPP gridrec header
 * (gridrec/GridrecCPP/Gridrec5.2/grid.h, and the do_recon form of
 * Gridrec4.9/grid.h:90-91), backed by the MI355X library libgridrec_hip.so.
 *
 * Same types, same function names and signatures (C++ linkage, as the
 * reference is compiled C++), same global variables of the 5.x form, same
 * ownership rules:
 *   - get_pswf(C, &P) points P at a database record (pswf.cpp:36-47);
 *   - get_filter(name) returns a filter function; an unknown name selects shlo
 *     and rewrites `name` (filters.cpp:36-48);
 *   - recon_init(GP, SGP, &imgsiz) sets up one geometry, overwrites GP->sampl
 *     with M/M0 and returns the image size M0 (grid.cpp:54-147);
 *   - do_recon() reads the globals G1, G2 (n_ang row pointers of n_det floats)
 *     and writes the caller-allocated S1, S2 (M0 row pointers of M0 floats)
 *     (grid.cpp:152-415); do_recon(G1, G2, &S1, &S2) is the 4.9 form that
 *     allocates both images with malloc_matrix_f (Gridrec4.9/grid.cpp:333-334);
 *   - recon_free() releases what recon_init set up (grid.cpp:418-434).
 * Where the reference exits (bad geometry, unknown C) these functions print
 * the reason to stderr and leave the outputs untouched; gridrec_last_error()
 * returns it (recon_init and do_recon clear it on entry, so after a call it
 * is empty exactly when that call succeeded).  All reconstruction runs on the GPU (HIP device 0, or the
 * ordinal in the environment variable GRIDREC_DEVICE).
 *
 * Library: libgridrec_grid.so (link with -lgridrec_grid; it loads
 * libgridrec_hip.so from its own directory).  The C-ABI library itself
 * exports none of these C++ names or globals.
 *
 * Accounting of Gridrec5.2/grid.h line by line:
 *   3-13   ANSI and system includes: kept (time.h and sys/stat.h are not
 *          needed by any declaration here and are left to the caller)
 *   16-22  macros max, min, free_matrix, abs, pi, TOLERANCE, LTBL_DEF: kept;
 *          abs is the reference's float-safe macro (define
 *          GRIDREC_NO_ABS_MACRO before including to keep std::abs usable)
 *   24-58  complex, sg_struct, pswf_struct, grid_struct: same layout
 *   61     verbose: defined by the library
 *   68-73  get_parm, data_setup, get_sgram, put_image, usage: NOT provided --
 *          the command-line driver's netCDF file I/O (grid_io.cpp), not part
 *          of the reconstruction API; rel_sgram: provided
 *   74-77  malloc_vector_f/_c, malloc_matrix_f/_c: provided
 *   78-80  prn_gparm, prn_pswf, prn_sgparm: provided
 *   83-91  get_pswf, get_filter, recon_init, do_recon (both forms),
 *          recon_free: provided
 *   94-95  four1, fourn: NOT provided -- the Numerical Recipes transforms
 *          grid.cpp calls internally; every transform of do_recon runs in the
 *          library's own GPU kernels
 *   99-107 the 5.x globals: defined by the library
 *   108-109 fshift, global_index: declared as in the reference, which also
 *          only declares them (the application defines them)
 */
#ifndef GRIDREC_DROPIN_GRID_H
#define GRIDREC_DROPIN_GRID_H

#include <math.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* Macros of grid.h:16-22 (guarded, since C++ headers may define some). */
#ifndef max
#define max(A, B) ((A) > (B) ? (A) : (B))
#endif
#ifndef min
#define min(A, B) ((A) < (B) ? (A) : (B))
#endif
#define free_matrix(A) (free(*(A)), free(A))
#if !defined(abs) && !defined(GRIDREC_NO_ABS_MACRO)
#define abs(A) ((A) > 0 ? (A) : -(A))
#endif
#ifndef pi
#define pi 3.14159265359
#endif
#define TOLERANCE 0.1
#define LTBL_DEF 512

typedef struct { float r; float i; } complex;

typedef struct SG_STRUCT {
    int n_ang;
    int n_det;
    int geom;       /* 0: angle list; 1, 2: uniform in half, full circle */
    float *angles;  /* degrees, when geom == 0 */
    float center;   /* rotation axis location */
} sg_struct;

typedef struct PSWF_STRUCT {
    float C;
    int nt;
    float lmbda;
    float coefs[15];
} pswf_struct;

typedef struct GRID_STRUCT {
    pswf_struct *pswf;
    float sampl;
    float MaxPixSiz;
    float R;
    float X0;
    float Y0;
    char fname[16];
    float (*filter)(float);
    long ltbl;
} grid_struct;

#if defined(__GNUC__)
#pragma GCC visibility push(default)
#endif

extern int verbose;

/* grid_math.cpp:19-107 */
void rel_sgram(float **S);
float *malloc_vector_f(long n);
complex *malloc_vector_c(long n);
float **malloc_matrix_f(long nr, long nc);
complex **malloc_matrix_c(long nr, long nc);
void prn_gparm(grid_struct *G);
void prn_pswf(pswf_struct *P);
void prn_sgparm(sg_struct *S);

/* pswf.cpp */
void get_pswf(float C, pswf_struct **P);

/* filters.cpp */
float (*get_filter(char *name))(float);
float shlo(float x);
float hann(float x);
float hamm(float x);
float ramp(float x);

/* grid.cpp */
void recon_init(grid_struct *gparm, sg_struct *sgparm, long *im_size);
void do_recon();
void do_recon(float **sgram1, float **sgram2, float ***image1, float ***image2);
void recon_free(void);

/* this library */
const char *gridrec_last_error(void);

/* Globals of the 5.x form (grid.h:99-107).  The sizes and the float tables
 * (SINE, COSE, filphase, wtbl, winv) are host copies of what the device uses;
 * the scratch arrays of the CPU algorithm (cproj, H, work, dwtbl) do not
 * exist and stay NULL. */
extern int flag;
extern int n_det, n_ang;
extern long pdim, M, M0, M02, ltbl;
extern float sampl, scale, L, X0, Y0;
extern float *SINE, *COSE, *wtbl, *dwtbl, *work, *winv;
extern complex *cproj, *filphase, **H;
extern float **G1, **G2;
extern float **S1, **S2;
extern long imgsiz;
extern FILE *fshift;     /* declared only, as in the reference */
extern int global_index; /* declared only, as in the reference */

#if defined(__GNUC__)
#pragma GCC visibility pop
#endif

#endif
