// A caller of the drop-in C++ API that sees nothing but include/grid.h -- test
// infrastructure, built by __graft_entry__.build() into libgrid_caller.so and
// called in-process through ctypes by tests/test_gpu_dropin.py.
//
// It drives the API in the order of the reference's own caller, the GUI's
// reconstruction loop (gridrec/GridrecCPP/Gridrec5.2/Multitomo_unit.txt:786-856):
// grid_struct / sg_struct from malloc, get_pswf(C, &GP->pswf), a filter name
// through get_filter, recon_init(GP, SGP, &imgsiz), then one do_recon per pair
// of slices (k += 2; a lone last slice is paired with itself) with n_ang row
// pointers per slice, and recon_free at the end.  The GUI copies each
// projection row with imgsiz as its length; this caller copies n_det values
// (with imgsiz = M0 < n_det the GUI's last row ends in memory it never wrote).
//
//   form 0: do_recon(G1, G2, &S1, &S2), the 4.9 form that allocates both
//           images (Gridrec4.9/grid.h:90-91, grid.cpp:333-334), freed with
//           free_matrix after each pair;
//   form 1: do_recon() on the 5.x globals G1, G2, S1, S2 (Gridrec5.2/grid.h:
//           90, 105-106), images allocated by the caller with malloc_matrix_f.
#include <string.h>

#include "../../include/grid.h"

extern "C" {

const char *dropin_last_error(void) { return gridrec_last_error(); }

// slices: n_slices x n_ang x n_det floats; out: n_slices x M0 x M0 floats.
// Returns 0, or -1 when the library reported an error (dropin_last_error).
int dropin_recon(int form, int n_ang_, int n_det_, int geom, float *angles, float center, const char *filter,
                 float C, float sampl_, float R, float MaxPixSiz, float X0_, float Y0_, long ltbl_, int n_slices,
                 const float *slices, float *out, long *imgsiz_out, float *sampl_out) {
    grid_struct *GP = (grid_struct *)malloc(sizeof(grid_struct));
    sg_struct *SGP = (sg_struct *)malloc(sizeof(sg_struct));
    char name[16];
    memset(GP, 0, sizeof *GP);
    SGP->n_ang = n_ang_;
    SGP->geom = geom;
    SGP->n_det = n_det_;
    SGP->center = center;
    SGP->angles = angles;
    GP->pswf = NULL;
    get_pswf(C, &GP->pswf);
    GP->sampl = sampl_;
    GP->MaxPixSiz = MaxPixSiz;
    GP->R = R;
    GP->X0 = X0_;
    GP->Y0 = Y0_;
    strncpy(name, filter, sizeof name - 1);
    name[sizeof name - 1] = 0;
    GP->filter = get_filter(name);
    strcpy(GP->fname, name);
    GP->ltbl = ltbl_;
    long size = 0;
    recon_init(GP, SGP, &size);
    *imgsiz_out = size;
    *sampl_out = GP->sampl;  // recon_init overwrites it with M/M0
    int rc = size > 0 ? 0 : -1;
    const long plane = (long)n_ang_ * n_det_, img = size * size;
    float **rows1 = (float **)malloc(sizeof(float *) * n_ang_);
    float **rows2 = (float **)malloc(sizeof(float *) * n_ang_);
    float *sino1 = (float *)malloc(sizeof(float) * plane);
    float *sino2 = (float *)malloc(sizeof(float) * plane);
    for (int k = 0; k < n_slices && rc == 0; k += 2) {
        const int k2 = k + 1 < n_slices ? k + 1 : k;
        memcpy(sino1, slices + k * plane, sizeof(float) * plane);
        memcpy(sino2, slices + k2 * plane, sizeof(float) * plane);
        rows1[0] = sino1;
        rows2[0] = sino2;
        for (int i = 1; i < n_ang_; i++) {
            rows1[i] = rows1[i - 1] + n_det_;
            rows2[i] = rows2[i - 1] + n_det_;
        }
        float **img1 = NULL, **img2 = NULL;
        if (form == 0) {
            do_recon(rows1, rows2, &img1, &img2);
        } else {
            G1 = rows1;
            G2 = rows2;
            S1 = img1 = malloc_matrix_f(size, size);
            S2 = img2 = malloc_matrix_f(size, size);
            do_recon();
        }
        if (!img1 || !img2 || gridrec_last_error()[0]) {
            rc = -1;
        } else {
            memcpy(out + k * img, img1[0], sizeof(float) * img);
            if (k2 != k) memcpy(out + k2 * img, img2[0], sizeof(float) * img);
        }
        if (img1) free_matrix(img1);
        if (img2) free_matrix(img2);
    }
    free(sino1);
    free(sino2);
    free(rows1);
    free(rows2);
    recon_free();
    free(GP);
    free(SGP);
    return rc;
}

}  // extern "C"
