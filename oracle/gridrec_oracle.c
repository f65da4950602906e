/*
 * gridrec_oracle.c -- CPU oracle for the gridrec hot path.
 *
 * TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker; never linked into or called by
 * the product library.
 *
 * Restates gridrec/GridrecCPP/Gridrec5.2/grid.cpp (arithmetic identical to
 * Gridrec4.9/grid.cpp), filters.cpp and pswf.cpp.  Every expression whose
 * float/double rounding can change a discrete decision (box bounds, table
 * indices, angle tables) is evaluated with the reference's types: on x86-64
 * with SSE, float operands give IEEE float results, a double literal such as
 * the reference's `pi` (grid.h:20) promotes the expression to double, and --
 * the reference being C++ -- sin/cos of a float argument are the float
 * overloads (sinf/cosf).
 * Build with -ffp-contract=off so no multiply-add is fused.
 *
 * Parity pins (tests/test_oracle.py): bit-exact tables and <=1e-6 rel-L2
 * images against oracle/_ref (the reference compiled from its own sources),
 * committed golden vectors, DFT known answers and brute-force FP64 sums on tiny
 * shapes.
 */
#include "gridrec_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define OR_PI 3.14159265359 /* the reference's pi macro, grid.h:20 */

typedef struct { float r, i; } cf;
typedef struct { double r, i; } cd;

/* ---------------------------------------------------------------------------
 * PSWF database: data of gridrec/GridrecCPP/Gridrec5.2/pswf.cpp:5-33
 * (C, Legendre degree nt, eigenvalue lambda, coefficients of P_0, P_2, ...).
 * ------------------------------------------------------------------------- */
typedef struct { float C; int nt; float lmbda; float coefs[15]; } or_pswf;

static const or_pswf OR_PSWF_DB[5] = {
    {4.0f, 16, 0.99588549f,
     {0.5239891E+01f, -0.5308499E+01f, 0.1184591E+01f, -0.1230763E-00f, 0.7371623E-02f,
      -0.2864074E-03f, 0.7789983E-05f, -0.1564700E-06f, 0.2414647E-08f}},
    {4.2f, 16, 0.99657887f,
     {0.6062942E+01f, -0.6450252E+01f, 0.1551875E+01f, -0.1755960E-01f, 0.1150712E-01f,
      -0.4903653E-03f, 0.1464986E-04f, -0.3235110E-06f, 0.5492141E-08f}},
    {5.0f, 18, 0.99935241f,
     {0.1115509E+02f, -0.1384861E+02f, 0.4289811E+01f, -0.6514303E-00f, 0.5844993E-01f,
      -0.3447736E-02f, 0.1435066E-03f, -0.4433680E-05f, 0.1056040E-06f, -0.1997173E-08f}},
    {6.0f, 18, 0.9990188f,
     {0.2495593E+02f, -0.3531124E+02f, 0.1383722E+02f, -0.2799028E+01f, 0.3437217E-00f,
      -0.2818024E-01f, 0.1645842E-02f, -0.7179160E-04f, 0.2424510E-05f, -0.6520875E-07f}},
    {7.0f, 20, 0.99998546f,
     {0.5767616E+02f, -0.8931343E+02f, 0.4167596E+02f, -0.1053599E+02f, 0.1662374E+01f,
      -0.1780527E-00f, 0.1372983E-01f, -0.7963169E-03f, 0.3593372E-04f, -0.1295941E-05f,
      0.3817796E-07f}},
};

/* get_pswf (pswf.cpp:36-47): first entry with |C - C_i| <= 0.01 in float. */
static const or_pswf *find_pswf(float C) {
    for (int i = 0; i < 5; i++) {
        float d = C - OR_PSWF_DB[i].C;
        if (d < 0) d = -d;
        if (!((double)d > 0.01)) return &OR_PSWF_DB[i];
    }
    return NULL;
}

/* ---------------------------------------------------------------------------
 * Filters (filters.cpp:10-24).  The reference evaluates in double wherever the
 * pi macro or a double literal enters, then returns float.
 * ------------------------------------------------------------------------- */
static float filt_f(int id, float x) {
    float ax = x > 0 ? x : -x;
    switch (id) {
    case OR_SHLO: { double v = sin(OR_PI * x) / OR_PI; return (float)(v > 0 ? v : -v); }
    case OR_HANN: return (float)(((double)ax * 0.5) * (1. + cos(2 * OR_PI * x)));
    case OR_HAMM: return (float)((double)ax * (0.54 + 0.46 * cos(2 * OR_PI * x)));
    case OR_RAMP: return ax;
    }
    return 0.0f;
}

/* Same filters, argument and value in double (FP64 oracle tables). */
static double filt_d(int id, double x) {
    double ax = fabs(x);
    switch (id) {
    case OR_SHLO: return fabs(sin(OR_PI * x) / OR_PI);
    case OR_HANN: return ax * 0.5 * (1. + cos(2 * OR_PI * x));
    case OR_HAMM: return ax * (0.54 + 0.46 * cos(2 * OR_PI * x));
    case OR_RAMP: return ax;
    }
    return 0.0;
}

/* ---------------------------------------------------------------------------
 * Sizes: recon_init, grid.cpp:87-120.
 * ------------------------------------------------------------------------- */
int or_sizes_get(const or_params *p, or_sizes *s) {
    if (!p || !s || p->n_ang <= 0 || p->n_det <= 0 || p->ltbl <= 0) return -1;
    if (p->geom < 0 || p->geom > 2 || (p->geom == 0 && !p->angles)) return -1;
    if (p->filter_id < 0 || p->filter_id > 3) return -1;
    const or_pswf *ps = find_pswf(p->C);
    if (!ps) return -1;
    if (!(p->MaxPixSiz > 0) || !(p->sampl > 0) || !(p->R > 0)) return -1;

    long pdim = 1, itmp = p->n_det - 1;             /* grid.cpp:88-94 */
    while (itmp) { pdim <<= 1; itmp >>= 1; }

    float D0 = p->R * (float)p->n_det;               /* grid.cpp:96 */
    float D1 = p->sampl * D0;                        /* grid.cpp:97 */
    float t = D1 / p->MaxPixSiz;                     /* grid.cpp:102, float */
    t = t - 1.0f;
    if (!(t >= 0.0f) || t > 1e9f) return -1;        /* the reference loops forever on t < 0 */
    itmp = (long)t;
    long M = 1;                                      /* grid.cpp:101-107 */
    while (itmp) { M <<= 1; itmp >>= 1; }

    float q = (float)(M / 2) / p->sampl;             /* grid.cpp:112: M/2/sampl */
    long M02 = (long)floor((double)q - 0.5);
    if (M02 < 0) return -1;
    long M0 = 2 * M02 + 1;                           /* grid.cpp:113 */

    float sampl = (float)M / (float)M0;              /* grid.cpp:117 */
    D1 = sampl * D0;                                 /* grid.cpp:118 */
    float twoC = 2.0f * ps->C;
    float L = (float)((double)(twoC * sampl) / OR_PI); /* grid.cpp:119 */
    float scale = D1 / (float)pdim;                  /* grid.cpp:120 */

    s->pdim = pdim; s->M = M; s->M0 = M0; s->M02 = M02;
    s->sampl = sampl; s->L = L; s->scale = scale;
    s->L2 = (float)((double)L / 2.);                 /* grid.cpp:223 */
    s->tblspcg = (float)(2 * p->ltbl) / L;           /* grid.cpp:224 */
    s->flag = (p->X0 != 0.0f || p->Y0 != 0.0f);      /* grid.cpp:84-85 */
    return 0;
}

/* ---------------------------------------------------------------------------
 * trig_su (grid.cpp:515-563).
 * ------------------------------------------------------------------------- */
static void trig_f(const or_params *p, float *S, float *C) {
    int n = p->n_ang;
    if (p->geom == 0) {
        float degtorad = (float)(OR_PI / 180);       /* grid.cpp:529 */
        for (int j = 0; j < n; j++) {
            float theta = degtorad * p->angles[j];   /* grid.cpp:534 */
            S[j] = sinf(theta);                      /* C++ float overloads */
            C[j] = cosf(theta);
        }
    } else {
        float dtheta = (float)(p->geom * OR_PI / n); /* grid.cpp:544 */
        float dcos = cosf(dtheta);
        float dsin = sinf(dtheta);
        S[0] = 0.0f; C[0] = 1.0f;
        for (int j = 1; j < n; j++) {                /* grid.cpp:549-553, float recurrence */
            float a = dcos * S[j - 1];
            float b = dsin * C[j - 1];
            float c = dcos * C[j - 1];
            float d = dsin * S[j - 1];
            S[j] = a + b;
            C[j] = c - d;
        }
    }
}

/* legendre (grid.cpp:566-606): sum_k coefs[k] P_2k(x), float recurrence. */
static float legendre_f(int n, const float *coefs, float x) {
    float y = coefs[0], penult = 1.0f, last = x, newer;
    int even = 1, k = 1;
    for (int j = 2; j <= n; j++) {
        float a = x * (float)(2 * j - 1);
        a = a * last;
        float b = (float)(j - 1) * penult;
        newer = (a - b) / (float)j;
        if (even) { y += newer * coefs[k]; even = 0; k++; }
        else even = 1;
        penult = last;
        last = newer;
    }
    return y;
}

static double legendre_d(int n, const float *coefs, double x) {
    double y = coefs[0], penult = 1.0, last = x, newer;
    int even = 1, k = 1;
    for (int j = 2; j <= n; j++) {
        newer = (x * (2 * j - 1) * last - (j - 1) * penult) / j;
        if (even) { y += newer * (double)coefs[k]; even = 0; k++; }
        else even = 1;
        penult = last;
        last = newer;
    }
    return y;
}

/* Cnvlvnt nearest-neighbour index, grid.cpp:22: (int)(X+0.5) with X float and
 * 0.5 double. */
static inline long cidx(float X) { return (long)((double)X + 0.5); }

/* Cnvlvnt of the INTERP variant (grid.cpp:19) read with its argument
 * parenthesised -- as written the macro does not compile: wtbl[(int)X] +
 * (X - (int)X) * dwtbl[(int)X], all float, with dwtbl[i] = wtbl[i] - wtbl[i-1]
 * (grid.cpp:488); dwtbl[0], never set by the reference, is 0 here. */
static inline float cnv_interp(const float *wtbl, const float *dwtbl, float X) {
    const int i = (int)X;
    return wtbl[i] + (X - (float)i) * dwtbl[i];
}

/* The table value of the convolvent at X (grid.cpp:18-24): nearest neighbour,
 * or linear interpolation when dwtbl is given. */
static inline float cnv(const float *wtbl, const float *dwtbl, float X) {
    return dwtbl ? cnv_interp(wtbl, dwtbl, X) : wtbl[cidx(X)];
}

/* pswf_su (grid.cpp:463-512).  sampl is the adjusted value (grid.cpp:117).
 * dwtbl (ltbl+1 floats) is filled and used when not NULL (INTERP). */
static void pswf_f(const or_pswf *ps, long ltbl, long linv, float sampl,
                   float *wtbl, float *winv, float *dwtbl) {
    float polyz = legendre_f(ps->nt, ps->coefs, 0.0f);
    wtbl[0] = 1.0f;
    if (dwtbl) dwtbl[0] = 0.0f;
    for (long i = 1; i <= ltbl; i++) {
        wtbl[i] = legendre_f(ps->nt, ps->coefs, (float)i / (float)ltbl) / polyz;
        if (dwtbl) dwtbl[i] = wtbl[i] - wtbl[i - 1];                   /* grid.cpp:488 */
    }
    float fac = (float)((double)(float)ltbl / ((double)linv + 0.5)); /* grid.cpp:492 */
    float norm = (float)(sqrt(OR_PI / 2 / (double)ps->C / (double)ps->lmbda) / (double)sampl);
    winv[linv] = norm / cnv(wtbl, dwtbl, 0.0f);
    for (long i = 1; i <= linv; i++) {
        norm = -norm;
        float v = norm / cnv(wtbl, dwtbl, (float)i * fac);
        winv[linv + i] = v;
        winv[linv - i] = v;
    }
}

/* FP64 tables at the same indices. */
static void pswf_d(const or_pswf *ps, long ltbl, long linv, float sampl_f,
                   double *wtbl, double *winv) {
    double polyz = legendre_d(ps->nt, ps->coefs, 0.0);
    wtbl[0] = 1.0;
    for (long i = 1; i <= ltbl; i++)
        wtbl[i] = legendre_d(ps->nt, ps->coefs, (double)i / (double)ltbl) / polyz;
    float fac = (float)((double)(float)ltbl / ((double)linv + 0.5));
    double norm = sqrt(OR_PI / 2 / (double)ps->C / (double)ps->lmbda) / (double)sampl_f;
    winv[linv] = norm / wtbl[0];
    for (long i = 1; i <= linv; i++) {
        norm = -norm;
        double v = norm / wtbl[cidx((float)i * fac)];
        winv[linv + i] = v;
        winv[linv - i] = v;
    }
}

/* filphase_su (grid.cpp:437-461). */
static void filphase_f(const or_params *p, long pd, float *A) {
    long pd2 = pd >> 1;
    float rtmp1 = (float)(2 * OR_PI * (double)p->center / (double)pd);
    float norm = (float)(OR_PI / (double)pd / (double)p->n_ang);
    for (long j = 0; j < pd2; j++) {
        float x = (float)j * rtmp1;
        float rtmp2 = filt_f(p->filter_id, (float)j / (float)pd) * norm;
        A[2 * j] = rtmp2 * cosf(x);                  /* grid.cpp:457-458, float overloads */
        A[2 * j + 1] = -rtmp2 * sinf(x);
    }
}

static void filphase_d(const or_params *p, long pd, double *A) {
    long pd2 = pd >> 1;
    double rtmp1 = 2 * OR_PI * (double)p->center / (double)pd;
    double norm = OR_PI / (double)pd / (double)p->n_ang;
    for (long j = 0; j < pd2; j++) {
        double x = (double)j * rtmp1;
        double rtmp2 = filt_d(p->filter_id, (double)j / (double)pd) * norm;
        A[2 * j] = rtmp2 * cos(x);
        A[2 * j + 1] = -rtmp2 * sin(x);
    }
}

int or_tables(const or_params *p, float *sine, float *cose, float *filphase,
              float *wtbl, float *winv) {
    or_sizes s;
    if (or_sizes_get(p, &s)) return -1;
    const or_pswf *ps = find_pswf(p->C);
    float *S = malloc(sizeof(float) * p->n_ang), *C = malloc(sizeof(float) * p->n_ang);
    float *W = malloc(sizeof(float) * (p->ltbl + 1)), *WI = malloc(sizeof(float) * s.M0);
    float *F = malloc(sizeof(float) * s.pdim);
    float *DW = p->interp ? malloc(sizeof(float) * (p->ltbl + 1)) : NULL;
    if (!S || !C || !W || !WI || !F || (p->interp && !DW)) {
        free(S); free(C); free(W); free(WI); free(F); free(DW);
        return -1;
    }
    trig_f(p, S, C);
    filphase_f(p, s.pdim, F);
    pswf_f(ps, p->ltbl, s.M02, s.sampl, W, WI, DW);
    free(DW);
    if (sine) memcpy(sine, S, sizeof(float) * p->n_ang);
    if (cose) memcpy(cose, C, sizeof(float) * p->n_ang);
    if (filphase) memcpy(filphase, F, sizeof(float) * s.pdim);
    if (wtbl) memcpy(wtbl, W, sizeof(float) * (p->ltbl + 1));
    if (winv) memcpy(winv, WI, sizeof(float) * s.M0);
    free(S); free(C); free(W); free(WI); free(F);
    return 0;
}

/* ---------------------------------------------------------------------------
 * FFTs: iterative radix-2 Cooley-Tukey with bit reversal, twiddles from a
 * double table.  sign=+1 computes sum x_k e^{+2 pi i jk/n}, sign=-1 e^{-...};
 * no normalisation.
 * ------------------------------------------------------------------------- */
static void bitrev_cf(cf *x, long n) {
    for (long i = 1, j = 0; i < n; i++) {
        long bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) { cf t = x[i]; x[i] = x[j]; x[j] = t; }
    }
}
static void bitrev_cd(cd *x, long n) {
    for (long i = 1, j = 0; i < n; i++) {
        long bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) { cd t = x[i]; x[i] = x[j]; x[j] = t; }
    }
}

/* tw: n/2 entries (cos, sin)(2 pi k / n) in double. */
static double *twiddles(long n) {
    long h = n > 1 ? n / 2 : 1;
    double *t = malloc(sizeof(double) * 2 * h);
    if (!t) return NULL;
    for (long k = 0; k < h; k++) {
        double a = 2.0 * M_PI * (double)k / (double)n;
        t[2 * k] = cos(a);
        t[2 * k + 1] = sin(a);
    }
    return t;
}
static float *twiddles_f(const double *t, long n) {
    long h = n > 1 ? n / 2 : 1;
    float *f = malloc(sizeof(float) * 2 * h);
    if (!f) return NULL;
    for (long k = 0; k < 2 * h; k++) f[k] = (float)t[k];
    return f;
}

static void fft_cf(cf *x, long n, int sign, const float *tw) {
    bitrev_cf(x, n);
    for (long len = 2; len <= n; len <<= 1) {
        long half = len >> 1, step = n / len;
        for (long i = 0; i < n; i += len) {
            for (long k = 0; k < half; k++) {
                float wr = tw[2 * k * step], wi = sign > 0 ? tw[2 * k * step + 1] : -tw[2 * k * step + 1];
                cf a = x[i + k], b = x[i + k + half];
                float tr = wr * b.r - wi * b.i;
                float ti = wr * b.i + wi * b.r;
                x[i + k + half].r = a.r - tr;
                x[i + k + half].i = a.i - ti;
                x[i + k].r = a.r + tr;
                x[i + k].i = a.i + ti;
            }
        }
    }
}

static void fft_cd(cd *x, long n, int sign, const double *tw) {
    bitrev_cd(x, n);
    for (long len = 2; len <= n; len <<= 1) {
        long half = len >> 1, step = n / len;
        for (long i = 0; i < n; i += len) {
            for (long k = 0; k < half; k++) {
                double wr = tw[2 * k * step], wi = sign > 0 ? tw[2 * k * step + 1] : -tw[2 * k * step + 1];
                cd a = x[i + k], b = x[i + k + half];
                double tr = wr * b.r - wi * b.i;
                double ti = wr * b.i + wi * b.r;
                x[i + k + half].r = a.r - tr;
                x[i + k + half].i = a.i - ti;
                x[i + k].r = a.r + tr;
                x[i + k].i = a.i + ti;
            }
        }
    }
}

#define COLBLK 16
/* 2D transform of an n x n row-major array: rows, then columns in blocks. */
static int fft2_cf(cf *h, long n, int sign, const float *tw) {
    for (long r = 0; r < n; r++) fft_cf(h + r * n, n, sign, tw);
    cf *tmp = malloc(sizeof(cf) * COLBLK * n);
    if (!tmp) return -1;
    for (long c0 = 0; c0 < n; c0 += COLBLK) {
        long nb = n - c0 < COLBLK ? n - c0 : COLBLK;
        for (long r = 0; r < n; r++)
            for (long b = 0; b < nb; b++) tmp[b * n + r] = h[r * n + c0 + b];
        for (long b = 0; b < nb; b++) fft_cf(tmp + b * n, n, sign, tw);
        for (long r = 0; r < n; r++)
            for (long b = 0; b < nb; b++) h[r * n + c0 + b] = tmp[b * n + r];
    }
    free(tmp);
    return 0;
}
static int fft2_cd(cd *h, long n, int sign, const double *tw) {
    for (long r = 0; r < n; r++) fft_cd(h + r * n, n, sign, tw);
    cd *tmp = malloc(sizeof(cd) * COLBLK * n);
    if (!tmp) return -1;
    for (long c0 = 0; c0 < n; c0 += COLBLK) {
        long nb = n - c0 < COLBLK ? n - c0 : COLBLK;
        for (long r = 0; r < n; r++)
            for (long b = 0; b < nb; b++) tmp[b * n + r] = h[r * n + c0 + b];
        for (long b = 0; b < nb; b++) fft_cd(tmp + b * n, n, sign, tw);
        for (long r = 0; r < n; r++)
            for (long b = 0; b < nb; b++) h[r * n + c0 + b] = tmp[b * n + r];
    }
    free(tmp);
    return 0;
}

/* Phase-3 row/column map (grid.cpp:386-409): output index j of M0 reads
 * H index M-M02+j for j < M02, then j-M02. */
static inline long corner(long j, long M, long M02) { return j < M02 ? M - M02 + j : j - M02; }

/* ---------------------------------------------------------------------------
 * do_recon, FP32 restatement (grid.cpp:152-415).
 * ------------------------------------------------------------------------- */
int or_recon_f32(const or_params *p, const float *g1, const float *g2, float *s1, float *s2) {
    or_sizes s;
    if (or_sizes_get(p, &s)) return -1;
    const or_pswf *ps = find_pswf(p->C);
    const long pdim = s.pdim, M = s.M, M0 = s.M0, M02 = s.M02, pdim2 = pdim >> 1, M2 = M >> 1;
    const int n_ang = p->n_ang, n_det = p->n_det;
    int rc = -1;
    float *SINE = malloc(sizeof(float) * n_ang), *COSE = malloc(sizeof(float) * n_ang);
    float *filphase = malloc(sizeof(float) * pdim);
    float *wtbl = malloc(sizeof(float) * (p->ltbl + 1)), *winv = malloc(sizeof(float) * M0);
    float *dwtbl = p->interp ? malloc(sizeof(float) * (p->ltbl + 1)) : NULL;
    cf *cproj = malloc(sizeof(cf) * pdim);
    cf *H = calloc((size_t)M * (size_t)M, sizeof(cf));            /* grid.cpp:165-170 */
    double *twr = twiddles(pdim), *twm = twiddles(M);
    float *twrf = twr ? twiddles_f(twr, pdim) : NULL, *twmf = twm ? twiddles_f(twm, M) : NULL;
    if (!SINE || !COSE || !filphase || !wtbl || !winv || !cproj || !H || !twrf || !twmf) goto out;
    if (p->interp && !dwtbl) goto out;
    trig_f(p, SINE, COSE);
    filphase_f(p, pdim, filphase);
    pswf_f(ps, p->ltbl, M02, s.sampl, wtbl, winv, dwtbl);

    const float L2 = s.L2, tblspcg = s.tblspcg, scale = s.scale, fM2 = (float)M2;
    for (int n = 0; n < n_ang; n++) {                                /* grid.cpp:234 */
        float offset = 0.0f;
        if (s.flag) {
            float a = p->X0 * COSE[n];
            float b = p->Y0 * SINE[n];
            offset = (float)((double)(a + b) * OR_PI);                /* grid.cpp:237 */
        }
        long j = 0;
        for (; j < n_det; j++) { cproj[j].r = g1[(long)n * n_det + j]; cproj[j].i = g2[(long)n * n_det + j]; }
        for (; j < pdim; j++) cproj[j].r = cproj[j].i = 0.0f;
        fft_cf(cproj, pdim, +1, twrf);                                /* grid.cpp:254 */
        for (j = 1; j < pdim2; j++) {                                 /* grid.cpp:257 */
            float ctr, cti;
            if (!s.flag) { ctr = filphase[2 * j]; cti = filphase[2 * j + 1]; }
            else {
                float arg = (float)j * offset;
                float phr = cosf(arg), phi = -sinf(arg);           /* grid.cpp:267-268 */
                float fr = filphase[2 * j], fi = filphase[2 * j + 1];
                ctr = fr * phr - fi * phi;
                cti = fr * phi + fi * phr;
            }
            cf P = cproj[j], Q = cproj[pdim - j];
            float d1r = ctr * P.r - cti * P.i, d1i = ctr * P.i + cti * P.r;   /* grid.cpp:272 */
            cti = -cti;
            float d2r = ctr * Q.r - cti * Q.i, d2i = ctr * Q.i + cti * Q.r;   /* grid.cpp:274 */
            float rtmp = scale * (float)j;
            float U = rtmp * COSE[n] + fM2;                            /* grid.cpp:276-277 */
            float V = rtmp * SINE[n] + fM2;
            long iul = (long)ceil((double)(U - L2)), iuh = (long)floor((double)(U + L2));
            long ivl = (long)ceil((double)(V - L2)), ivh = (long)floor((double)(V + L2));
            if (iul < 1) iul = 1;
            if (iuh >= M) iuh = M - 1;
            if (ivl < 1) ivl = 1;
            if (ivh >= M) ivh = M - 1;
            float work[64];
            long k = 0;
            for (long iv = ivl; iv <= ivh && k < 64; iv++, k++) {
                float d = V - (float)iv;
                work[k] = cnv(wtbl, dwtbl, (d > 0 ? d : -d) * tblspcg);  /* grid.cpp:290 */
            }
            for (long iu = iul; iu <= iuh; iu++) {
                float d = U - (float)iu;
                float rt = cnv(wtbl, dwtbl, (d > 0 ? d : -d) * tblspcg); /* grid.cpp:293 */
                cf *row = H + iu * M, *mrow = H + (M - iu) * M;
                k = 0;
                for (long iv = ivl; iv <= ivh && k < 64; iv++, k++) {
                    float cv = rt * work[k];
                    row[iv].r += cv * d1r;                             /* grid.cpp:297-300 */
                    row[iv].i += cv * d1i;
                    mrow[M - iv].r += cv * d2r;
                    mrow[M - iv].i += cv * d2i;
                }
            }
        }
    }
    if (fft2_cf(H, M, -1, twmf)) goto out;                             /* grid.cpp:326-328 */
    for (long j = 0; j < M0; j++) {                                    /* grid.cpp:383-409 */
        long iu = corner(j, M, M02);
        float cu = winv[j];
        for (long k = 0; k < M0; k++) {
            long iv = corner(k, M, M02);
            float corrn = cu * winv[k];
            s1[j * M0 + k] = corrn * H[iu * M + iv].r;
            s2[j * M0 + k] = corrn * H[iu * M + iv].i;
        }
    }
    rc = 0;
out:
    free(SINE); free(COSE); free(filphase); free(wtbl); free(winv); free(dwtbl); free(cproj); free(H);
    free(twr); free(twm); free(twrf); free(twmf);
    return rc;
}

/* ---------------------------------------------------------------------------
 * FP64 oracle.  exact == 0: same float SINE/COSE and float U, V for every box
 * and index decision as the FP32 path; everything else double.
 * exact == 1: exact angles and double U, V (decisions in double).
 * ------------------------------------------------------------------------- */
static int recon_d(const or_params *p, const float *g1, const float *g2, double *s1,
                   double *s2, int exact) {
    or_sizes s;
    if (or_sizes_get(p, &s)) return -1;
    const or_pswf *ps = find_pswf(p->C);
    const long pdim = s.pdim, M = s.M, M0 = s.M0, M02 = s.M02, pdim2 = pdim >> 1, M2 = M >> 1;
    const int n_ang = p->n_ang, n_det = p->n_det;
    int rc = -1;
    float *SINE = malloc(sizeof(float) * n_ang), *COSE = malloc(sizeof(float) * n_ang);
    double *Sd = malloc(sizeof(double) * n_ang), *Cd = malloc(sizeof(double) * n_ang);
    double *filphase = malloc(sizeof(double) * pdim);
    double *wtbl = malloc(sizeof(double) * (p->ltbl + 1)), *winv = malloc(sizeof(double) * M0);
    cd *cproj = malloc(sizeof(cd) * pdim);
    cd *H = calloc((size_t)M * (size_t)M, sizeof(cd));
    double *twr = twiddles(pdim), *twm = twiddles(M);
    if (!SINE || !COSE || !Sd || !Cd || !filphase || !wtbl || !winv || !cproj || !H || !twr || !twm) goto out;
    trig_f(p, SINE, COSE);
    for (int n = 0; n < n_ang; n++) {
        if (exact) {
            double th = p->geom == 0 ? (double)p->angles[n] * M_PI / 180.0
                                     : (double)n * p->geom * M_PI / (double)n_ang;
            Sd[n] = sin(th); Cd[n] = cos(th);
        } else { Sd[n] = SINE[n]; Cd[n] = COSE[n]; }
    }
    filphase_d(p, pdim, filphase);
    pswf_d(ps, p->ltbl, M02, s.sampl, wtbl, winv);

    /* exact-geometry constants in double */
    const double sampl_d = (double)M / (double)M0;
    const double L_d = 2.0 * (double)ps->C * sampl_d / M_PI;
    const double scale_d = sampl_d * (double)p->R * (double)n_det / (double)pdim;
    const double L2_d = L_d / 2.0, tbl_d = 2.0 * (double)p->ltbl / L_d;

    for (int n = 0; n < n_ang; n++) {
        double offset = 0.0;
        if (s.flag) offset = ((double)p->X0 * Cd[n] + (double)p->Y0 * Sd[n]) * OR_PI;
        long j = 0;
        for (; j < n_det; j++) { cproj[j].r = g1[(long)n * n_det + j]; cproj[j].i = g2[(long)n * n_det + j]; }
        for (; j < pdim; j++) cproj[j].r = cproj[j].i = 0.0;
        fft_cd(cproj, pdim, +1, twr);
        for (j = 1; j < pdim2; j++) {
            double ctr = filphase[2 * j], cti = filphase[2 * j + 1];
            if (s.flag) {
                double phr = cos((double)j * offset), phi = -sin((double)j * offset);
                double fr = ctr, fi = cti;
                ctr = fr * phr - fi * phi;
                cti = fr * phi + fi * phr;
            }
            cd P = cproj[j], Q = cproj[pdim - j];
            double d1r = ctr * P.r - cti * P.i, d1i = ctr * P.i + cti * P.r;
            cti = -cti;
            double d2r = ctr * Q.r - cti * Q.i, d2i = ctr * Q.i + cti * Q.r;
            long iul, iuh, ivl, ivh;
            double Ud, Vd, L2u, tblu;
            float U = 0, V = 0;
            if (!exact) {
                float rtmp = s.scale * (float)j;
                U = rtmp * COSE[n] + (float)M2;
                V = rtmp * SINE[n] + (float)M2;
                iul = (long)ceil((double)(U - s.L2)); iuh = (long)floor((double)(U + s.L2));
                ivl = (long)ceil((double)(V - s.L2)); ivh = (long)floor((double)(V + s.L2));
                Ud = U; Vd = V; L2u = s.L2; tblu = s.tblspcg;
            } else {
                Ud = scale_d * (double)j * Cd[n] + (double)M2;
                Vd = scale_d * (double)j * Sd[n] + (double)M2;
                iul = (long)ceil(Ud - L2_d); iuh = (long)floor(Ud + L2_d);
                ivl = (long)ceil(Vd - L2_d); ivh = (long)floor(Vd + L2_d);
                L2u = L2_d; tblu = tbl_d;
            }
            (void)L2u;
            if (iul < 1) iul = 1;
            if (iuh >= M) iuh = M - 1;
            if (ivl < 1) ivl = 1;
            if (ivh >= M) ivh = M - 1;
            double work[64];
            long k = 0;
            for (long iv = ivl; iv <= ivh && k < 64; iv++, k++) {
                long idx;
                if (!exact) { float d = V - (float)iv; idx = cidx((d > 0 ? d : -d) * s.tblspcg); }
                else idx = (long)(fabs(Vd - (double)iv) * tblu + 0.5);
                if (idx > p->ltbl) idx = p->ltbl;
                work[k] = wtbl[idx];
            }
            for (long iu = iul; iu <= iuh; iu++) {
                long idx;
                if (!exact) { float d = U - (float)iu; idx = cidx((d > 0 ? d : -d) * s.tblspcg); }
                else idx = (long)(fabs(Ud - (double)iu) * tblu + 0.5);
                if (idx > p->ltbl) idx = p->ltbl;
                double rt = wtbl[idx];
                cd *row = H + iu * M, *mrow = H + (M - iu) * M;
                k = 0;
                for (long iv = ivl; iv <= ivh && k < 64; iv++, k++) {
                    double cv = rt * work[k];
                    row[iv].r += cv * d1r;
                    row[iv].i += cv * d1i;
                    mrow[M - iv].r += cv * d2r;
                    mrow[M - iv].i += cv * d2i;
                }
            }
        }
    }
    if (fft2_cd(H, M, -1, twm)) goto out;
    for (long j = 0; j < M0; j++) {
        long iu = corner(j, M, M02);
        for (long k = 0; k < M0; k++) {
            long iv = corner(k, M, M02);
            double corrn = winv[j] * winv[k];
            s1[j * M0 + k] = corrn * H[iu * M + iv].r;
            s2[j * M0 + k] = corrn * H[iu * M + iv].i;
        }
    }
    rc = 0;
out:
    free(SINE); free(COSE); free(Sd); free(Cd); free(filphase); free(wtbl); free(winv);
    free(cproj); free(H); free(twr); free(twm);
    return rc;
}

int or_recon_f64(const or_params *p, const float *g1, const float *g2, double *s1, double *s2) {
    if (p->interp) return -1;
    return recon_d(p, g1, g2, s1, s2, 0);
}
int or_recon_exact64(const or_params *p, const float *g1, const float *g2, double *s1, double *s2) {
    if (p->interp) return -1;
    return recon_d(p, g1, g2, s1, s2, 1);
}

/* ---------------------------------------------------------------------------
 * Mixed precision: double pipeline, each stage optionally rounded to float
 * after every operation.
 * ------------------------------------------------------------------------- */
static inline double rnd(double x, int r32) { return r32 ? (double)(float)x : x; }

static void fft_mx(cd *x, long n, int sign, const double *tw, int r32) {
    bitrev_cd(x, n);
    for (long len = 2; len <= n; len <<= 1) {
        long half = len >> 1, step = n / len;
        for (long i = 0; i < n; i += len) {
            for (long k = 0; k < half; k++) {
                double wr = rnd(tw[2 * k * step], r32);
                double wi = rnd(sign > 0 ? tw[2 * k * step + 1] : -tw[2 * k * step + 1], r32);
                cd a = x[i + k], b = x[i + k + half];
                double tr = rnd(rnd(wr * b.r, r32) - rnd(wi * b.i, r32), r32);
                double ti = rnd(rnd(wr * b.i, r32) + rnd(wi * b.r, r32), r32);
                x[i + k + half].r = rnd(a.r - tr, r32);
                x[i + k + half].i = rnd(a.i - ti, r32);
                x[i + k].r = rnd(a.r + tr, r32);
                x[i + k].i = rnd(a.i + ti, r32);
            }
        }
    }
}

int or_recon_mixed(const or_params *p, const float *g1, const float *g2, double *s1, double *s2,
                   int mask) {
    if (p->interp) return -1;
    or_sizes s;
    if (or_sizes_get(p, &s)) return -1;
    const or_pswf *ps = find_pswf(p->C);
    const long pdim = s.pdim, M = s.M, M0 = s.M0, M02 = s.M02, pdim2 = pdim >> 1, M2 = M >> 1;
    const int n_ang = p->n_ang, n_det = p->n_det;
    const int rr = !(mask & OR_MX_ROWFFT), ra = !(mask & OR_MX_ACCUM), rf = !(mask & OR_MX_FFT2);
    const int rd = !(mask & OR_MX_DEAPOD), td = !!(mask & OR_MX_TABLES);
    int rc = -1;
    float *SINE = malloc(sizeof(float) * n_ang), *COSE = malloc(sizeof(float) * n_ang);
    float *fpf = malloc(sizeof(float) * pdim), *wtf = malloc(sizeof(float) * (p->ltbl + 1));
    float *wif = malloc(sizeof(float) * M0);
    double *fp = malloc(sizeof(double) * pdim), *wt = malloc(sizeof(double) * (p->ltbl + 1));
    double *wi = malloc(sizeof(double) * M0);
    cd *cproj = malloc(sizeof(cd) * pdim);
    cd *H = calloc((size_t)M * (size_t)M, sizeof(cd));
    double *twr = twiddles(pdim), *twm = twiddles(M);
    if (!SINE || !COSE || !fpf || !wtf || !wif || !fp || !wt || !wi || !cproj || !H || !twr || !twm) goto out;
    trig_f(p, SINE, COSE);
    if (td) { filphase_d(p, pdim, fp); pswf_d(ps, p->ltbl, M02, s.sampl, wt, wi); }
    else {
        filphase_f(p, pdim, fpf);
        pswf_f(ps, p->ltbl, M02, s.sampl, wtf, wif, NULL);
        for (long j = 0; j < pdim; j++) fp[j] = fpf[j];
        for (long j = 0; j <= p->ltbl; j++) wt[j] = wtf[j];
        for (long j = 0; j < M0; j++) wi[j] = wif[j];
    }
    for (int n = 0; n < n_ang; n++) {
        long j = 0;
        for (; j < n_det; j++) { cproj[j].r = g1[(long)n * n_det + j]; cproj[j].i = g2[(long)n * n_det + j]; }
        for (; j < pdim; j++) cproj[j].r = cproj[j].i = 0.0;
        fft_mx(cproj, pdim, +1, twr, rr);
        float offset = 0.0f;
        if (s.flag) offset = (float)((double)(p->X0 * COSE[n] + p->Y0 * SINE[n]) * OR_PI);
        for (j = 1; j < pdim2; j++) {
            double ctr = fp[2 * j], cti = fp[2 * j + 1];
            if (s.flag) {
                double phr, phi;
                if (rr) { float arg = (float)j * offset; phr = cosf(arg); phi = -sinf(arg); }
                else { phr = cos((double)j * offset); phi = -sin((double)j * offset); }
                double fr = ctr, fi = cti;
                ctr = rnd(rnd(fr * phr, rr) - rnd(fi * phi, rr), rr);
                cti = rnd(rnd(fr * phi, rr) + rnd(fi * phr, rr), rr);
            }
            cd P = cproj[j], Q = cproj[pdim - j];
            double d1r = rnd(rnd(ctr * P.r, rr) - rnd(cti * P.i, rr), rr);
            double d1i = rnd(rnd(ctr * P.i, rr) + rnd(cti * P.r, rr), rr);
            cti = -cti;
            double d2r = rnd(rnd(ctr * Q.r, rr) - rnd(cti * Q.i, rr), rr);
            double d2i = rnd(rnd(ctr * Q.i, rr) + rnd(cti * Q.r, rr), rr);
            float rtmp = s.scale * (float)j;
            float U = rtmp * COSE[n] + (float)M2;
            float V = rtmp * SINE[n] + (float)M2;
            long iul = (long)ceil((double)(U - s.L2)), iuh = (long)floor((double)(U + s.L2));
            long ivl = (long)ceil((double)(V - s.L2)), ivh = (long)floor((double)(V + s.L2));
            if (iul < 1) iul = 1;
            if (iuh >= M) iuh = M - 1;
            if (ivl < 1) ivl = 1;
            if (ivh >= M) ivh = M - 1;
            double work[64];
            long k = 0;
            for (long iv = ivl; iv <= ivh && k < 64; iv++, k++) {
                float d = V - (float)iv;
                work[k] = wt[cidx((d > 0 ? d : -d) * s.tblspcg)];
            }
            for (long iu = iul; iu <= iuh; iu++) {
                float d = U - (float)iu;
                double rt = wt[cidx((d > 0 ? d : -d) * s.tblspcg)];
                cd *row = H + iu * M, *mrow = H + (M - iu) * M;
                k = 0;
                for (long iv = ivl; iv <= ivh && k < 64; iv++, k++) {
                    double cv = rnd(rt * work[k], ra);
                    row[iv].r = rnd(row[iv].r + rnd(cv * d1r, ra), ra);
                    row[iv].i = rnd(row[iv].i + rnd(cv * d1i, ra), ra);
                    mrow[M - iv].r = rnd(mrow[M - iv].r + rnd(cv * d2r, ra), ra);
                    mrow[M - iv].i = rnd(mrow[M - iv].i + rnd(cv * d2i, ra), ra);
                }
            }
        }
    }
    {
        for (long r = 0; r < M; r++) fft_mx(H + r * M, M, -1, twm, rf);
        cd *tmp = malloc(sizeof(cd) * COLBLK * M);
        if (!tmp) goto out;
        for (long c0 = 0; c0 < M; c0 += COLBLK) {
            long nb = M - c0 < COLBLK ? M - c0 : COLBLK;
            for (long r = 0; r < M; r++)
                for (long b = 0; b < nb; b++) tmp[b * M + r] = H[r * M + c0 + b];
            for (long b = 0; b < nb; b++) fft_mx(tmp + b * M, M, -1, twm, rf);
            for (long r = 0; r < M; r++)
                for (long b = 0; b < nb; b++) H[r * M + c0 + b] = tmp[b * M + r];
        }
        free(tmp);
    }
    for (long j = 0; j < M0; j++) {
        long iu = corner(j, M, M02);
        for (long k = 0; k < M0; k++) {
            long iv = corner(k, M, M02);
            double corrn = rnd(wi[j] * wi[k], rd);
            s1[j * M0 + k] = rnd(corrn * H[iu * M + iv].r, rd);
            s2[j * M0 + k] = rnd(corrn * H[iu * M + iv].i, rd);
        }
    }
    rc = 0;
out:
    free(SINE); free(COSE); free(fpf); free(wtf); free(wif); free(fp); free(wt); free(wi);
    free(cproj); free(H); free(twr); free(twm);
    return rc;
}

/* ---------------------------------------------------------------------------
 * Caller conventions and the threaded batch driver.
 * ------------------------------------------------------------------------- */
static float or_normalise(int mode, unsigned short raw, float dark, float w1, float w2) {
    if (mode == 1) {  /* tomo_recon.cpp:1877 */
        const float proj = (float)raw;
        return 1.0f - 1.0f * (proj - dark) / (0.5f * w1 + 0.5f * w2 - dark);
    }
    /* find_shift.cpp:1685-1697: doubles, result stored as float */
    double dv = (double)raw - (double)dark;
    if (dv <= 0) return 0.0f;
    dv = (0.5 * ((double)w1 + (double)w2) - (double)dark) / dv;
    return dv > 0.0 ? (float)log(dv) : 0.0f;
}

int or_preprocess(int mode, int n_ang, int n_det, const unsigned short *raw, const float *dark,
                  const float *white1, const float *white2, float xoffset, float *out) {
    if (mode != 1 && mode != 2) return -1;
    if (!white2) white2 = white1;
    float *norm = (float *)malloc(sizeof(float) * (size_t)n_det);
    if (!norm) return -1;
    for (int j = 0; j < n_ang; j++) {
        for (int k = 0; k < n_det; k++)
            norm[k] = or_normalise(mode, raw[(size_t)j * n_det + k], dark[k], white1[k], white2[k]);
        /* tomo_recon.cpp:3708-3780: kk = k - xoffset, two-point linear
           interpolation, an index beyond either end reads the edge cell */
        for (int k = 0; k < n_det; k++) {
            const float kk = (float)k - xoffset;
            const int nkk = (int)floorf(kk);
            const int i0 = nkk < 0 ? 0 : (nkk >= n_det ? n_det - 1 : nkk);
            const int i1 = nkk + 1 < 0 ? 0 : (nkk + 1 >= n_det ? n_det - 1 : nkk + 1);
            float v = 0.0f;
            v += norm[i0] * (float)(nkk + 1 - kk);
            v += norm[i1] * (kk - (float)nkk);
            out[(size_t)j * n_det + k] = v;
        }
    }
    free(norm);
    return 0;
}

float or_entropy(const float *data, long n, float vmin, float vmax, int n_bins) {
    const float inv = (vmax - vmin) / (float)n_bins;
    long *hist = (long *)calloc((size_t)n_bins + 1, sizeof(long));
    if (!hist) return -1.0f;
    for (long i = 0; i < n; i++) {
        const float x = (data[i] - vmin) / inv;
        if (!(x > -1.0f && x < (float)(n_bins + 1))) continue;  /* (int)x outside 0..n_bins */
        hist[(int)x]++;
    }
    float ent = 0.0f;
    for (int b = 0; b <= n_bins; b++) {
        if (hist[b] <= 0) continue;
        const float pi = 1.0f * (float)hist[b] / (float)n;
        ent = (float)((double)ent + (double)(-pi * logf(pi)) / log(2.0));  /* float += double */
    }
    free(hist);
    return ent;
}

void or_ring_correct(float *data, int n_ang, int n_det, float ring_coeff, int iterations) {
    float *mv = (float *)malloc(sizeof(float) * (size_t)n_ang);
    float *ml = (float *)malloc(sizeof(float) * (size_t)n_det);
    float *lp = (float *)malloc(sizeof(float) * (size_t)n_det);
    if (!mv || !ml || !lp) { free(mv); free(ml); free(lp); return; }
    for (int m = 0; m < iterations; m++) {
        float mt = 0.0f;
        for (int i = 0; i < n_ang; i++) {
            mv[i] = 0.0f;
            for (int j = 0; j < n_det; j++) mv[i] += data[(size_t)i * n_det + j];
            mv[i] /= n_det;
            mt += mv[i];
        }
        mt /= n_ang;
        for (int i = 0; i < n_ang; i++)
            for (int j = 0; j < n_det; j++)
                if (mv[i] != 0.0f) data[(size_t)i * n_det + j] = data[(size_t)i * n_det + j] * mt / mv[i];
        for (int j = 0; j < n_det; j++) ml[j] = 0.0f;
        for (int i = 0; i < n_ang; i++)
            for (int j = 0; j < n_det; j++) ml[j] += data[(size_t)i * n_det + j];
        for (int j = 0; j < n_det; j++) ml[j] /= n_ang;
        for (int j = 1; j < n_det - 1; j++) lp[j] = (float)((ml[j - 1] + ml[j] + ml[j + 1]) / 3.0);
        lp[0] = ml[0];
        lp[n_det - 1] = ml[n_det - 1];
        for (int i = 0; i < n_ang; i++)
            for (int j = 0; j < n_det; j++) {
                const float t = ml[j] - lp[j];
                float *d = &data[(size_t)i * n_det + j];
                if ((*d - (t * ring_coeff)) > 0.0f) *d -= (t * ring_coeff);
                else *d = 0.0f;
            }
    }
    free(mv);
    free(ml);
    free(lp);
}

void or_pad_half(int n_ang, int n_det, const float *src, float *dst) {
    int off = n_det / 2, w = 2 * n_det;                   /* tomo_recon.cpp:3692,3698 */
    for (int a = 0; a < n_ang; a++) {
        const float *r = src + (long)a * n_det;
        float *o = dst + (long)a * w;
        for (int k = 0; k < w; k++) {
            int q = k - off;
            if (q < 0) q = 0;
            if (q > n_det - 1) q = n_det - 1;
            o[k] = r[q];                                  /* tomo_recon.cpp:3784-3807 */
        }
    }
}

typedef struct {
    const or_params *p;
    int pad_half, n_pairs, n_threads, tid;
    const float *sino;
    long crop_off, crop_n;
    float *out;
    int rc;
} batch_arg;

static void *batch_worker(void *va) {
    batch_arg *a = (batch_arg *)va;
    const or_params *p = a->p;
    or_params q = *p;
    int w = p->n_det;
    if (a->pad_half) { q.n_det = 2 * p->n_det; q.center = p->center + (float)(p->n_det / 2); w = q.n_det; }
    or_sizes s;
    a->rc = -1;
    if (or_sizes_get(&q, &s)) return NULL;
    if (a->crop_off < 0 || a->crop_off + a->crop_n > s.M0) return NULL;
    long plane = (long)p->n_ang * p->n_det;
    float *g1 = malloc(sizeof(float) * (long)p->n_ang * w), *g2 = malloc(sizeof(float) * (long)p->n_ang * w);
    float *s1 = malloc(sizeof(float) * s.M0 * s.M0), *s2 = malloc(sizeof(float) * s.M0 * s.M0);
    if (!g1 || !g2 || !s1 || !s2) goto done;
    for (int pr = a->tid; pr < a->n_pairs; pr += a->n_threads) {
        const float *r1 = a->sino + (2L * pr) * plane, *r2 = r1 + plane;
        if (a->pad_half) { or_pad_half(p->n_ang, p->n_det, r1, g1); or_pad_half(p->n_ang, p->n_det, r2, g2); }
        else { memcpy(g1, r1, sizeof(float) * plane); memcpy(g2, r2, sizeof(float) * plane); }
        if (or_recon_f32(&q, g1, g2, s1, s2)) goto done;
        float *o1 = a->out + (2L * pr) * a->crop_n * a->crop_n, *o2 = o1 + a->crop_n * a->crop_n;
        for (long j = 0; j < a->crop_n; j++) {
            memcpy(o1 + j * a->crop_n, s1 + (a->crop_off + j) * s.M0 + a->crop_off, sizeof(float) * a->crop_n);
            memcpy(o2 + j * a->crop_n, s2 + (a->crop_off + j) * s.M0 + a->crop_off, sizeof(float) * a->crop_n);
        }
    }
    a->rc = 0;
done:
    free(g1); free(g2); free(s1); free(s2);
    return NULL;
}

int or_recon_batch_f32(const or_params *p, int pad_half, const float *sino, int n_pairs,
                       long crop_off, long crop_n, float *out, int n_threads) {
    if (n_threads < 1) n_threads = 1;
    if (n_threads > 256) n_threads = 256;
    pthread_t th[256];
    batch_arg args[256];
    for (int t = 0; t < n_threads; t++) {
        args[t] = (batch_arg){p, pad_half, n_pairs, n_threads, t, sino, crop_off, crop_n, out, -1};
        if (pthread_create(&th[t], NULL, batch_worker, &args[t])) { n_threads = t; break; }
    }
    int rc = 0;
    for (int t = 0; t < n_threads; t++) { pthread_join(th[t], NULL); if (args[t].rc) rc = -1; }
    return n_threads > 0 ? rc : -1;
}

/* Test hooks: the oracle's own 1D transforms on interleaved complex data. */
int or_fft_f32(float *x, long n, int sign) {
    if (n < 1 || (n & (n - 1))) return -1;
    double *t = twiddles(n);
    float *tf = t ? twiddles_f(t, n) : NULL;
    if (!tf) { free(t); return -1; }
    fft_cf((cf *)x, n, sign, tf);
    free(t); free(tf);
    return 0;
}
int or_fft_f64(double *x, long n, int sign) {
    if (n < 1 || (n & (n - 1))) return -1;
    double *t = twiddles(n);
    if (!t) return -1;
    fft_cd((cd *)x, n, sign, t);
    free(t);
    return 0;
}
