/*
 * ref_driver.cpp -- C entry points over the reference's own gridrec, compiled
 * from gridrec/GridrecCPP/Gridrec5.2/{grid,fftgr,filters,pswf,grid_io}.cpp
 * where they lie (see oracle/Makefile, target `ref`).
 *
 * TEST INFRASTRUCTURE ONLY: the resulting oracle/_ref/libgridrec_ref.so pins the
 * C oracle and, where present, is the parity target of the GPU tests.  Nothing
 * in the product links or loads it.
 *
 * Drives the reference exactly as its GUI caller does
 * (Multitomo_unit.txt:793-856): get_pswf, get_filter, recon_init, row pointers
 * for G1/G2/S1/S2, do_recon, recon_free.  grid_io.cpp is linked only for the
 * `verbose` flag it defines (grid_io.cpp:22); its netCDF readers are never
 * referenced and are dropped by --gc-sections.
 */
#include <stdlib.h>
#include <string.h>

#include "grid.h"

extern "C" {

typedef struct ref_args {
    int n_ang, n_det, geom;
    float center, C, sampl, MaxPixSiz, R, X0, Y0;
    long ltbl;
    const float *angles;
    char filter[16];
} ref_args;

typedef struct ref_out {
    long pdim, M, M0, M02;
    float sampl, L, scale;
    float *sine, *cose, *filphase, *wtbl, *winv; /* optional copies of the tables */
} ref_out;

static int setup(const ref_args *a, grid_struct *gp, sg_struct *sg, char *fname, long *isz) {
    pswf_struct *ps = NULL;
    get_pswf(a->C, &ps);
    memset(gp, 0, sizeof(*gp));
    memset(sg, 0, sizeof(*sg));
    sg->n_ang = a->n_ang;
    sg->n_det = a->n_det;
    sg->geom = a->geom;
    sg->angles = (float *)a->angles;
    sg->center = a->center;
    gp->pswf = ps;
    gp->sampl = a->sampl;
    gp->MaxPixSiz = a->MaxPixSiz;
    gp->R = a->R;
    gp->X0 = a->X0;
    gp->Y0 = a->Y0;
    strncpy(fname, a->filter, 15);
    fname[15] = 0;
    gp->filter = get_filter(fname);
    gp->ltbl = a->ltbl;
    recon_init(gp, sg, isz);
    return 0;
}

static void copy_out(const grid_struct *gp, ref_out *o) {
    if (!o) return;
    o->pdim = pdim; o->M = M; o->M0 = M0; o->M02 = M02;
    o->sampl = gp->sampl; o->L = L; o->scale = scale;
    if (o->sine) memcpy(o->sine, SINE, sizeof(float) * n_ang);
    if (o->cose) memcpy(o->cose, COSE, sizeof(float) * n_ang);
    if (o->filphase) memcpy(o->filphase, filphase, sizeof(complex) * (pdim / 2));
    if (o->wtbl) memcpy(o->wtbl, wtbl, sizeof(float) * (ltbl + 1));
    if (o->winv) memcpy(o->winv, winv, sizeof(float) * M0);
}

/* Sizes and tables only (recon_init + recon_free). */
__attribute__((visibility("default"))) int ref_tables(const ref_args *a, ref_out *o) {
    grid_struct gp; sg_struct sg; char fname[32]; long isz = 0;
    setup(a, &gp, &sg, fname, &isz);
    copy_out(&gp, o);
    recon_free();
    return 0;
}

/* One do_recon on rows g1/g2 (n_ang x n_det contiguous) into s1/s2 (M0 x M0). */
__attribute__((visibility("default"))) int ref_recon(const ref_args *a, const float *g1,
                                                     const float *g2, float *s1, float *s2,
                                                     ref_out *o) {
    grid_struct gp; sg_struct sg; char fname[32]; long isz = 0;
    setup(a, &gp, &sg, fname, &isz);
    G1 = (float **)malloc(sizeof(float *) * a->n_ang);
    G2 = (float **)malloc(sizeof(float *) * a->n_ang);
    S1 = (float **)malloc(sizeof(float *) * isz);
    S2 = (float **)malloc(sizeof(float *) * isz);
    for (int n = 0; n < a->n_ang; n++) {
        G1[n] = (float *)g1 + (long)n * a->n_det;
        G2[n] = (float *)g2 + (long)n * a->n_det;
    }
    for (long j = 0; j < isz; j++) { S1[j] = s1 + j * isz; S2[j] = s2 + j * isz; }
    do_recon();
    copy_out(&gp, o);
    recon_free();
    free(G1); free(G2); free(S1); free(S2);
    G1 = G2 = S1 = S2 = NULL;
    return 0;
}

} /* extern "C" */
