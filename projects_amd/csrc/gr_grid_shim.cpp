// gr_grid_shim.cpp -- the reference's C++ API (include/grid.h) over the C-ABI,
// built into its own library (libgridrec_grid.so) so that the C-ABI library
// exports none of the reference's generic global names (M, L, H, flag, ...).
//
// One process-wide plan, as the reference keeps one set of globals
// (gridrec/GridrecCPP/Gridrec5.2/grid.cpp:28-37): recon_init creates it,
// do_recon runs it on the GPU, recon_free destroys it.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>

#include "../../include/gridrec_hip.h"

#undef max
#undef min
#include "../../include/grid.h"
#undef max
#undef min
#undef abs

// ---- globals of the 5.x API ------------------------------------------------
int verbose = 0;
int flag = 0;
int n_det = 0, n_ang = 0;
long pdim = 0, M = 0, M0 = 0, M02 = 0, ltbl = 0;
float sampl = 0, scale = 0, L = 0, X0 = 0, Y0 = 0;
float *SINE = nullptr, *COSE = nullptr, *wtbl = nullptr, *dwtbl = nullptr, *work = nullptr, *winv = nullptr;
complex *cproj = nullptr, *filphase = nullptr, **H = nullptr;
float **G1 = nullptr, **G2 = nullptr;
float **S1 = nullptr, **S2 = nullptr;
long imgsiz = 0;

namespace {
gr_plan *g_plan = nullptr;
gr_pswf g_pswf_copy;
std::string g_last;

void report(const char *where) {
    g_last = std::string(where) + ": " + gr_last_error();
    fprintf(stderr, "gridrec: %s\n", g_last.c_str());
}

// PSWF records handed out by get_pswf, with the layout of pswf_struct.
pswf_struct g_db[5];
bool g_db_ready = false;
void load_db() {
    if (g_db_ready) return;
    const float Cs[5] = {4.0f, 4.2f, 5.0f, 6.0f, 7.0f};
    for (int i = 0; i < 5; i++) {
        gr_pswf r;
        if (gr_get_pswf(Cs[i], &r) == GR_OK) {
            g_db[i].C = r.C;
            g_db[i].nt = r.nt;
            g_db[i].lmbda = r.lmbda;
            memcpy(g_db[i].coefs, r.coefs, sizeof r.coefs);
        }
    }
    g_db_ready = true;
}

void free_host_tables() {
    free(SINE);
    free(COSE);
    free(wtbl);
    free(winv);
    free(filphase);
    SINE = COSE = wtbl = winv = nullptr;
    filphase = nullptr;
}
}  // namespace

const char *gridrec_last_error(void) { return g_last.c_str(); }

// ---- grid_math.cpp:19-107 ---------------------------------------------------
void rel_sgram(float **S) { free_matrix(S); }

float *malloc_vector_f(long n) {
    float *v = (float *)malloc((size_t)n * sizeof(float));
    if (!v) fprintf(stderr, "malloc error in malloc_vector_f for length %ld.\n", n);
    return v;
}

complex *malloc_vector_c(long n) {
    complex *v = (complex *)malloc((size_t)n * sizeof(complex));
    if (!v) fprintf(stderr, "malloc error in malloc_vector_c for length %ld.\n", n);
    return v;
}

template <typename T>
static T **matrix_rows(long nr, long nc) {
    T **m = (T **)malloc((size_t)nr * sizeof(T *));
    if (!m) return nullptr;
    m[0] = (T *)malloc((size_t)nr * (size_t)nc * sizeof(T));
    if (!m[0]) {
        free(m);
        return nullptr;
    }
    for (long i = 1; i < nr; i++) m[i] = m[i - 1] + nc;
    return m;
}

float **malloc_matrix_f(long nr, long nc) {
    float **m = matrix_rows<float>(nr, nc);
    if (!m) fprintf(stderr, "malloc error in malloc_matrix_f for %ld x %ld.\n", nr, nc);
    return m;
}

complex **malloc_matrix_c(long nr, long nc) {
    complex **m = matrix_rows<complex>(nr, nc);
    if (!m) fprintf(stderr, "malloc error in malloc_matrix_c for %ld x %ld.\n", nr, nc);
    return m;
}

// ---- grid_io.cpp:405-430: parameter listings --------------------------------
void prn_gparm(grid_struct *G) {
    printf("Contents of grid_struct:\n\t");
    prn_pswf(G->pswf);
    printf("\n  sampl=%f MaxPixSiz=%f R=%f X0=%f Y0=%f\n  filter=%s ltbl=%ld\n", G->sampl, G->MaxPixSiz, G->R,
           G->X0, G->Y0, G->fname, G->ltbl);
}

void prn_pswf(pswf_struct *P) {
    printf("Contents of pswf: C=%f nt=%d lmbda=%f\n\tCoefficients:\n", P->C, P->nt, P->lmbda);
    for (int i = 0; i < P->nt / 2 + 1 && i < 15; i++) printf("\t\t%f\n", P->coefs[i]);
}

void prn_sgparm(sg_struct *S) {
    printf("Contents of sg_struct:\n\tn_ang=%d n_det=%d geom=%d center=%f\n", S->n_ang, S->n_det, S->geom,
           S->center);
}

// ---- pswf.cpp / filters.cpp ---------------------------------------------------
void get_pswf(float C, pswf_struct **P) {
    load_db();
    for (int i = 0; i < 5; i++) {
        float d = C - g_db[i].C;
        if (d < 0) d = -d;
        if (!((double)d > 0.01)) {
            *P = &g_db[i];
            return;
        }
    }
    g_last = "get_pswf: C not in the database";
    fprintf(stderr, "Prolate parameter, C = %f not in data base\n", C);
}

// The four filters of filters.cpp:10-24; the device tables are built from
// these very functions through gr_params.filter_fn.
float shlo(float x) {
    const double v = sin(3.14159265359 * x) / 3.14159265359;
    return (float)(v > 0 ? v : -v);
}
float hann(float x) {
    const float ax = x > 0 ? x : -x;
    return (float)(((double)ax * 0.5) * (1. + cos(2 * 3.14159265359 * x)));
}
float hamm(float x) {
    const float ax = x > 0 ? x : -x;
    return (float)((double)ax * (0.54 + 0.46 * cos(2 * 3.14159265359 * x)));
}
float ramp(float x) { return x > 0 ? x : -x; }

float (*get_filter(char *name))(float) {
    char before[256];
    strncpy(before, name, sizeof before - 1);
    before[sizeof before - 1] = 0;
    int id = 0;
    gr_filter_lookup(name, name, &id);  // rewrites an unknown name to "shlo"
    if (strcmp(before, name)) printf(" Using default filter: %s.\n", name);
    switch (id) {
    case 1: return hann;
    case 2: return hamm;
    case 3: return ramp;
    default: return shlo;
    }
}

// ---- grid.cpp ------------------------------------------------------------------
void recon_free(void) {
    if (g_plan) gr_plan_destroy(g_plan);
    g_plan = nullptr;
    free_host_tables();
}

void recon_init(grid_struct *GP, sg_struct *SGP, long *im_size) {
    recon_free();
    g_last.clear();
    if (!GP || !SGP || !GP->pswf) {
        // grid.cpp:80 reads GP->pswf without a check; get_pswf exits on an
        // unknown C (pswf.cpp:39-44), so the reference never gets here
        g_last = "recon_init: no grid_struct, sg_struct or PSWF record (get_pswf failed?)";
        fprintf(stderr, "gridrec: %s\n", g_last.c_str());
        return;
    }
    gr_params p;
    memset(&p, 0, sizeof p);
    p.n_ang = SGP->n_ang;
    p.n_det = SGP->n_det;
    p.geom = SGP->geom;
    p.angles = SGP->angles;
    p.center = SGP->center;
    p.C = GP->pswf->C;
    p.sampl = GP->sampl;
    p.MaxPixSiz = GP->MaxPixSiz;
    p.R = GP->R;
    p.X0 = GP->X0;
    p.Y0 = GP->Y0;
    snprintf(p.filter, sizeof p.filter, "%.15s", GP->fname);
    p.ltbl = GP->ltbl;
    p.pad = GR_PAD_NONE;
    p.accum = GR_ACCUM_COMPENSATED;
    p.tables = GR_TABLES_REFERENCE;
    const char *dev = getenv("GRIDREC_DEVICE");
    p.device = dev ? atoi(dev) : 0;
    p.max_pairs = 1;
    p.filter_fn = GP->filter;
    g_pswf_copy.C = GP->pswf->C;
    g_pswf_copy.nt = GP->pswf->nt;
    g_pswf_copy.lmbda = GP->pswf->lmbda;
    memcpy(g_pswf_copy.coefs, GP->pswf->coefs, sizeof g_pswf_copy.coefs);
    p.pswf = &g_pswf_copy;
    if (gr_plan_create(&p, &g_plan) != GR_OK) {
        g_plan = nullptr;
        report("recon_init");
        return;
    }
    gr_info info;
    gr_plan_info(g_plan, &info);
    // globals of grid.cpp:28-37
    n_ang = p.n_ang;
    n_det = p.n_det;
    pdim = info.pdim;
    M = info.M;
    M0 = info.M0;
    M02 = info.M02;
    ltbl = p.ltbl;
    sampl = info.sampl;
    scale = info.scale;
    L = info.L;
    X0 = p.X0;
    Y0 = p.Y0;
    flag = (X0 != 0.0f || Y0 != 0.0f);
    GP->sampl = info.sampl;  // grid.cpp:117
    SINE = malloc_vector_f(n_ang);
    COSE = malloc_vector_f(n_ang);
    wtbl = malloc_vector_f(ltbl + 1);
    winv = malloc_vector_f(M0);
    filphase = malloc_vector_c(pdim / 2);
    if (SINE && COSE && wtbl && winv && filphase)
        gr_plan_tables(g_plan, SINE, COSE, (float *)filphase, wtbl, winv);
    imgsiz = info.M0;
    *im_size = info.M0;
}

void do_recon() {
    g_last.clear();
    if (!g_plan) {
        g_last = "do_recon: recon_init did not succeed";
        fprintf(stderr, "gridrec: %s\n", g_last.c_str());
        return;
    }
    if (gr_recon_host(g_plan, (const float *const *)G1, (const float *const *)G2, S1, S2) != GR_OK)
        report("do_recon");
}

void do_recon(float **sgram1, float **sgram2, float ***image1, float ***image2) {
    g_last.clear();
    if (!g_plan) {
        g_last = "do_recon: recon_init did not succeed";
        fprintf(stderr, "gridrec: %s\n", g_last.c_str());
        return;
    }
    *image1 = malloc_matrix_f(M0, M0);  // Gridrec4.9/grid.cpp:333-334
    *image2 = malloc_matrix_f(M0, M0);
    if (!*image1 || !*image2) return;
    if (gr_recon_host(g_plan, (const float *const *)sgram1, (const float *const *)sgram2, *image1, *image2) != GR_OK)
        report("do_recon");
}
