// k_tables.hip -- FP64 evaluation of the filter/phase, PSWF and deapodization
// tables on the GPU (GR_TABLES_FP64), rounded once to float.
//
// Same formulas and the same table indices as recon_init
// (gridrec/GridrecCPP/Gridrec5.2/grid.cpp:437-512, filters.cpp:10-24), but
// every intermediate in double: the reference's float Legendre recurrence and
// float filter products are the largest source of its FP32 error.  The winv
// lookup index keeps the reference's float rule (int)(i*fac + 0.5) with fac
// computed on the host, so both table modes make the same discrete choices.
// This is the only FP64 arithmetic of the GPU path.
#include "gr_internal.h"

#pragma clang fp contract(off)

namespace {

const double kRefPi = 3.14159265359;  // grid.h:20

struct TabArgs {
    float2 *filphase;
    float *wtbl, *winv;
    int half, filter_id, n_ang, nt;
    long ltbl, M02, pdim;
    double center, C, lmbda;
    float fac;       // (float)ltbl/(M02+0.5), grid.cpp:492
    float sampl;     // adjusted sampl
    float coefs[15];
};

__device__ double legendre_d(int n, const float *coefs, double x) {
    double y = coefs[0], penult = 1.0, last = x;
    int even = 1, k = 1;
    for (int j = 2; j <= n; j++) {
        const double newer = (x * (2 * j - 1) * last - (j - 1) * penult) / j;
        if (even) {
            y += newer * (double)coefs[k];
            k++;
        }
        even = !even;
        penult = last;
        last = newer;
    }
    return y;
}

__device__ double filter_d(int id, double x) {
    const double ax = fabs(x);
    switch (id) {
    case 0: return fabs(sin(kRefPi * x) / kRefPi);
    case 1: return ax * 0.5 * (1. + cos(2 * kRefPi * x));
    case 2: return ax * (0.54 + 0.46 * cos(2 * kRefPi * x));
    default: return ax;
    }
}

__device__ __forceinline__ long nn_index_f(float X) {
    const long i = (long)X;
    const float f = X - (float)i;
    return i + (f >= 0.5f ? 1 : 0);
}

__global__ void k_tables_fp64(TabArgs A) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const double polyz = legendre_d(A.nt, A.coefs, 0.0);
    if (i <= A.ltbl)
        A.wtbl[i] = i == 0 ? 1.0f : (float)(legendre_d(A.nt, A.coefs, (double)i / (double)A.ltbl) / polyz);
    if (i < A.half) {
        const double x = (double)i * (2 * kRefPi * A.center / (double)A.pdim);
        const double f = filter_d(A.filter_id, (double)i / (double)A.pdim) * (kRefPi / (double)A.pdim / A.n_ang);
        A.filphase[i] = make_float2((float)(f * cos(x)), (float)(-f * sin(x)));
    }
    const long M0 = 2 * A.M02 + 1;
    if (i < M0) {
        const long d = i > A.M02 ? i - A.M02 : A.M02 - i;
        long idx = nn_index_f((float)d * A.fac);
        idx = idx > A.ltbl ? A.ltbl : idx;
        const double w = idx == 0 ? 1.0 : legendre_d(A.nt, A.coefs, (double)idx / (double)A.ltbl) / polyz;
        const double norm = sqrt(kRefPi / 2 / A.C / A.lmbda) / (double)A.sampl;
        A.winv[i] = (float)(((d & 1) ? -norm : norm) / w);
    }
}

}  // namespace

hipError_t gr_launch_tables_fp64(const GrGeom &g, const gr_params *p, float2 *filphase, float *wtbl,
                                 float *winv, hipStream_t s) {
    TabArgs A;
    A.filphase = filphase;
    A.wtbl = wtbl;
    A.winv = winv;
    A.half = (int)g.half;
    A.filter_id = g.filter_id;
    A.n_ang = p->n_ang;
    A.nt = g.pswf.nt;
    A.ltbl = p->ltbl;
    A.M02 = g.M02;
    A.pdim = g.pdim;
    A.center = g.center_eff;
    A.C = g.pswf.C;
    A.lmbda = g.pswf.lmbda;
    A.fac = (float)((double)(float)p->ltbl / ((double)g.M02 + 0.5));
    A.sampl = g.sampl;
    for (int k = 0; k < 15; k++) A.coefs[k] = g.pswf.coefs[k];
    long n = g.half;
    if (p->ltbl + 1 > n) n = p->ltbl + 1;
    if (g.M0 > n) n = g.M0;
    hipLaunchKernelGGL(k_tables_fp64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, A);
    return hipGetLastError();
}
