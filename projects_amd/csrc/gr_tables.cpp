// gr_tables.cpp -- sizes and lookup tables of one gridrec geometry, built on the
// host exactly as recon_init builds them (gridrec/GridrecCPP/Gridrec5.2/grid.cpp:54-147).
//
// The reference is C++ compiled for IEEE single precision: float operands give
// float results, its `pi` macro (grid.h:20) is a double literal that promotes an
// expression to double, and sin/cos of a float argument are the float overloads.
// Each expression below keeps those types so the tables come out bit-identical;
// this file is built with -ffp-contract=off so no multiply-add is fused.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "gr_gather.hpp"
#include "gr_internal.h"

namespace {

const double kRefPi = 3.14159265359;  // grid.h:20

// PSWF database, data of gridrec/GridrecCPP/Gridrec5.2/pswf.cpp:5-33.
const gr_pswf kPswfDb[5] = {
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

// Filters of filters.cpp:10-24, with the reference's double promotion.
float filter_value(int id, float x) {
    const float ax = x > 0 ? x : -x;
    switch (id) {
    case 0: {
        const double v = sin(kRefPi * x) / kRefPi;
        return (float)(v > 0 ? v : -v);
    }
    case 1: return (float)(((double)ax * 0.5) * (1. + cos(2 * kRefPi * x)));
    case 2: return (float)((double)ax * (0.54 + 0.46 * cos(2 * kRefPi * x)));
    default: return ax;
    }
}

// legendre (grid.cpp:566-606) in float.
float legendre_sum(int n, const float *coefs, float x) {
    float y = coefs[0], penult = 1.0f, last = x;
    int even = 1, k = 1;
    for (int j = 2; j <= n; j++) {
        float t = x * (float)(2 * j - 1);
        t = t * last;
        const float u = (float)(j - 1) * penult;
        const float newer = (t - u) / (float)j;
        if (even) {
            y += newer * coefs[k];
            k++;
        }
        even = !even;
        penult = last;
        last = newer;
    }
    return y;
}

// Cnvlvnt nearest-neighbour index (grid.cpp:22): (int)(X + 0.5), X float, 0.5 double.
inline long nn_index(float X) { return (long)((double)X + 0.5); }

}  // namespace

int gr_find_pswf(float C, gr_pswf *out) {
    for (int i = 0; i < 5; i++) {
        float d = C - kPswfDb[i].C;  // pswf.cpp:38, float abs compared with 0.01
        if (d < 0) d = -d;
        if (!((double)d > 0.01)) {
            if (out) *out = kPswfDb[i];
            return 0;
        }
    }
    return GR_ENOPSWF;
}

int gr_filter_id_of(const char *name) {
    // filters.cpp:26-34: shlo (default), shepp, hann, hamm, hamming, ramp, ramlak
    static const struct { const char *name; int id; } tbl[] = {
        {"shlo", 0}, {"shepp", 0}, {"hann", 1}, {"hamm", 2}, {"hamming", 2}, {"ramp", 3}, {"ramlak", 3}};
    if (name)
        for (const auto &e : tbl)
            if (!strcmp(name, e.name)) return e.id;
    return -1;
}

int gr_geom_from_params(const gr_params *p, GrGeom *g, const char **why) {
    *why = "";
    if (!p || !g) { *why = "null argument"; return GR_EINVAL; }
    if (p->n_ang <= 0 || p->n_det <= 0) { *why = "n_ang and n_det must be positive"; return GR_EINVAL; }
    if (p->geom < 0 || p->geom > 2) { *why = "geom must be 0, 1 or 2 (grid.cpp:557-560)"; return GR_EINVAL; }
    if (p->geom == 0 && !p->angles) { *why = "geom 0 needs an angle list"; return GR_EINVAL; }
    if (p->ltbl < 1 || p->ltbl > 16384) { *why = "ltbl must be in [1, 16384]"; return GR_EINVAL; }
    if (!(p->sampl > 0) || !(p->MaxPixSiz > 0) || !(p->R > 0)) { *why = "sampl, MaxPixSiz, R must be > 0"; return GR_EINVAL; }
    if (p->pad != GR_PAD_NONE && p->pad != GR_PAD_HALF_EDGE) { *why = "unknown pad mode"; return GR_EINVAL; }
    if (p->accum != GR_ACCUM_COMPENSATED && p->accum != GR_ACCUM_PLAIN) { *why = "unknown accum mode"; return GR_EINVAL; }
    if (p->tables != GR_TABLES_REFERENCE && p->tables != GR_TABLES_FP64) { *why = "unknown tables mode"; return GR_EINVAL; }
    if (p->pswf) {
        g->pswf = *p->pswf;  // grid.cpp:80-81 uses the caller's record
        if (g->pswf.nt < 0 || g->pswf.nt > 28) { *why = "pswf->nt out of range"; return GR_EINVAL; }
    } else if (gr_find_pswf(p->C, &g->pswf)) {
        *why = "prolate parameter C not in the database (pswf.cpp:39-44)";
        return GR_ENOPSWF;
    }
    if (p->filter_fn && p->tables == GR_TABLES_FP64) { *why = "a filter function needs GR_TABLES_REFERENCE"; return GR_EINVAL; }
    if (p->interp != 0 && p->interp != 1) { *why = "interp must be 0 or 1"; return GR_EINVAL; }
    if (p->interp && p->tables == GR_TABLES_FP64) { *why = "interp needs GR_TABLES_REFERENCE"; return GR_EINVAL; }
    int fid = gr_filter_id_of(p->filter);
    g->filter_id = fid < 0 ? 0 : fid;  // unknown name: default shlo (filters.cpp:43-47)

    g->n_det_in = p->n_det;
    g->pad = p->pad;
    if (p->pad == GR_PAD_HALF_EDGE) {
        if (p->n_det > (1 << 29)) { *why = "n_det too large"; return GR_EINVAL; }
        g->n_det_eff = 2 * p->n_det;
        g->center_eff = p->center + (float)(p->n_det / 2);  // tomo_recon.cpp:3698, centre x_dim/2
    } else {
        g->n_det_eff = p->n_det;
        g->center_eff = p->center;
    }
    const int n_det = g->n_det_eff;

    long pdim = 1, itmp = n_det - 1;  // grid.cpp:88-94
    while (itmp) { pdim <<= 1; itmp >>= 1; }
    const float D0 = p->R * (float)n_det;  // grid.cpp:96-97
    float D1 = p->sampl * D0;
    float t = D1 / p->MaxPixSiz;  // grid.cpp:102
    t = t - 1.0f;
    if (!(t >= 0.0f) || t > 1e9f) { *why = "sampl*R*n_det/MaxPixSiz must be >= 1"; return GR_EINVAL; }
    itmp = (long)t;
    long M = 1;  // grid.cpp:101-107
    while (itmp) { M <<= 1; itmp >>= 1; }
    const float q = (float)(M / 2) / p->sampl;  // grid.cpp:112
    const long M02 = (long)floor((double)q - 0.5);
    if (M02 < 0 || M02 > M - 1) { *why = "sampl gives an image larger than the grid"; return GR_EINVAL; }
    const long M0 = 2 * M02 + 1;
    const float sampl = (float)M / (float)M0;  // grid.cpp:117
    D1 = sampl * D0;
    const float twoC = 2.0f * g->pswf.C;
    const float L = (float)((double)(twoC * sampl) / kRefPi);  // grid.cpp:119
    const float scale = D1 / (float)pdim;                      // grid.cpp:120
    if (pdim < 4 || M < 8) { *why = "geometry too small (pdim >= 4 and M >= 8 required)"; return GR_EINVAL; }
    if (pdim > 16384 || M > 16384) { *why = "pdim and M are limited to 16384"; return GR_EINVAL; }

    g->pdim = pdim;
    g->half = pdim / 2;
    g->M = M;
    g->M0 = M0;
    g->M02 = M02;
    g->sampl = sampl;
    g->L = L;
    g->scale = scale;
    g->L2 = (float)((double)L / 2.);               // grid.cpp:223
    g->tblspcg = (float)(2 * p->ltbl) / L;         // grid.cpp:224
    g->flag = (p->X0 != 0.0f || p->Y0 != 0.0f);    // grid.cpp:84-85

    g->out_off = p->crop_n > 0 ? p->crop_off : 0;
    g->out_n = p->crop_n > 0 ? p->crop_n : M0;
    if (g->out_off < 0 || g->out_off + g->out_n > M0) { *why = "crop window outside the M0 x M0 image"; return GR_EINVAL; }
    return GR_OK;
}

void gr_host_tables(const gr_params *p, const GrGeom &g, float *sine, float *cose, float *filphase,
                    float *wtbl, float *winv, float *roi_offset, float *dwtbl) {
    const int n_ang = p->n_ang;
    // trig_su (grid.cpp:515-563)
    if (p->geom == 0) {
        const float degtorad = (float)(kRefPi / 180);
        for (int j = 0; j < n_ang; j++) {
            const float theta = degtorad * p->angles[j];
            sine[j] = sinf(theta);
            cose[j] = cosf(theta);
        }
    } else {
        const float dtheta = (float)(p->geom * kRefPi / n_ang);
        const float dcos = cosf(dtheta), dsin = sinf(dtheta);
        sine[0] = 0.0f;
        cose[0] = 1.0f;
        for (int j = 1; j < n_ang; j++) {
            const float a = dcos * sine[j - 1], b = dsin * cose[j - 1];
            const float c = dcos * cose[j - 1], d = dsin * sine[j - 1];
            sine[j] = a + b;
            cose[j] = c - d;
        }
    }
    // ROI phase offset per angle (grid.cpp:237)
    for (int n = 0; n < n_ang; n++) {
        if (g.flag) {
            const float a = p->X0 * cose[n], b = p->Y0 * sine[n];
            roi_offset[n] = (float)((double)(a + b) * kRefPi);
        } else {
            roi_offset[n] = 0.0f;
        }
    }
    // filphase_su (grid.cpp:437-461), for the effective centre
    const long pd = g.pdim;
    const float rtmp1 = (float)(2 * kRefPi * (double)g.center_eff / (double)pd);
    const float norm = (float)(kRefPi / (double)pd / (double)n_ang);
    for (long j = 0; j < g.half; j++) {
        const float x = (float)j * rtmp1;
        const float fx = (float)j / (float)pd;
        const float rtmp2 = (p->filter_fn ? p->filter_fn(fx) : filter_value(g.filter_id, fx)) * norm;
        filphase[2 * j] = rtmp2 * cosf(x);
        filphase[2 * j + 1] = -rtmp2 * sinf(x);
    }
    // pswf_su (grid.cpp:463-512)
    const gr_pswf &ps = g.pswf;
    const long ltbl = p->ltbl, linv = g.M02;
    const float polyz = legendre_sum(ps.nt, ps.coefs, 0.0f);
    wtbl[0] = 1.0f;
    for (long i = 1; i <= ltbl; i++) wtbl[i] = legendre_sum(ps.nt, ps.coefs, (float)i / (float)ltbl) / polyz;
    // INTERP (grid.cpp:487-489): backward differences, dwtbl[0] = 0
    std::vector<float> dw(p->interp ? ltbl + 1 : 0);
    if (p->interp) {
        dw[0] = 0.0f;
        for (long i = 1; i <= ltbl; i++) dw[i] = wtbl[i] - wtbl[i - 1];
        if (dwtbl) std::copy(dw.begin(), dw.end(), dwtbl);
    }
    auto cnv = [&](float X) {  // Cnvlvnt (grid.cpp:18-24)
        if (!p->interp) return wtbl[nn_index(X)];
        return gr_cnv_interp(X, wtbl, dw.data(), (int)ltbl);
    };
    const float fac = (float)((double)(float)ltbl / ((double)linv + 0.5));
    float wn = (float)(sqrt(kRefPi / 2 / (double)ps.C / (double)ps.lmbda) / (double)g.sampl);
    winv[linv] = wn / cnv(0.0f);
    for (long i = 1; i <= linv; i++) {
        wn = -wn;
        const float v = wn / cnv((float)i * fac);
        winv[linv + i] = v;
        winv[linv - i] = v;
    }
}
