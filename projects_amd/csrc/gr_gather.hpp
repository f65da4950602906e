// gr_gather.hpp -- per-tap logic of the gather form of gridrec's gridding.
//
// The reference scatters every filtered sample (angle n, frequency j) into the
// grid cells of an L x L box around (U, V) and its mirror
// (gridrec/GridrecCPP/Gridrec5.2/grid.cpp:276-302):
//
//   U = (scale*j)*COSE[n] + M2,  V = (scale*j)*SINE[n] + M2        (float)
//   box: ceil(U-L2) <= iu <= floor(U+L2), ceil(V-L2) <= iv <= floor(V+L2),
//        clipped to [1, M-1]
//   w  = wtbl[(int)(|U-iu|*tblspcg + 0.5)] * wtbl[(int)(|V-iv|*tblspcg + 0.5)]
//   H[iu][iv]     += w * Cdata1
//   H[M-iu][M-iv] += w * Cdata2
//
// The gather evaluates, for one cell (a, b), exactly the same float
// predicates and weights: the direct term uses (iu, iv) = (a, b) with Cdata1,
// the mirrored term uses (iu, iv) = (M-a, M-b) with Cdata2 and the weight of
// |U-(M-a)|, |V-(M-b)| -- the reference's weight for that deposit.  For an
// integer a, ceil(x) <= a <=> x <= a and a <= floor(y) <=> a <= y, so the box
// test is two float compares per axis; cells with a or b equal to 0 are never
// written (the clip to [1, M-1]).
//
// All of it is __host__ __device__ so a CPU build can check it against the
// reference's scatter; the file must be compiled without FP contraction.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#pragma clang fp contract(off)

struct GrGatherGeom {
    float scale, L2, tblspcg, fM2;
    int M, half, ltbl;
};

// (int)((double)X + 0.5) for X >= 0 without double arithmetic: X - trunc(X) is
// exact in float, so rounding half up is a compare of that fraction with 0.5.
__host__ __device__ __forceinline__ int gr_nn_index(float X) {
    const int i = (int)X;
    const float f = X - (float)i;
    return i + (f >= 0.5f ? 1 : 0);
}

// Sample coordinates (grid.cpp:276-277): U = (scale*j)*COSE + M2, V likewise.
__host__ __device__ __forceinline__ void gr_sample_uv(int j, float C, float S, const GrGatherGeom &g,
                                                      float &U, float &V) {
    const float rt = g.scale * (float)j;
    U = rt * C + g.fM2;
    V = rt * S + g.fM2;
}

// If sample j of the angle with table values (C, S) deposits into cell
// (qa, qb), return true and its weight.  qa, qb are exact small integers.
// gr_tap's table indices: the weight is wtbl[iu] * wtbl[iv].
__host__ __device__ __forceinline__ bool gr_tap_index(int j, float C, float S, float qa, float qb,
                                                      const GrGatherGeom &g, int &iu, int &iv) {
    float U, V;
    gr_sample_uv(j, C, S, g, U, V);
    if (!(U - g.L2 <= qa && qa <= U + g.L2 && V - g.L2 <= qb && qb <= V + g.L2)) return false;
    iu = gr_nn_index(fabsf(U - qa) * g.tblspcg);
    iv = gr_nn_index(fabsf(V - qb) * g.tblspcg);
    iu = iu > g.ltbl ? g.ltbl : iu;
    iv = iv > g.ltbl ? g.ltbl : iv;
    return true;
}

__host__ __device__ __forceinline__ bool gr_tap(int j, float C, float S, float qa, float qb,
                                                const GrGatherGeom &g, const float *wtbl,
                                                float &w) {
    int iu, iv;
    if (!gr_tap_index(j, C, S, qa, qb, g, iu, iv)) return false;
    w = wtbl[iu] * wtbl[iv];
    return true;
}


// Convolvent of the reference's compile-time INTERP variant (grid.cpp:18-24)
// with its argument parenthesised (as written the macro does not compile:
// Cnvlvnt(abs(V-iv)*tblspcg) would subscript wtbl with a float):
// wtbl[i] + (X - i) * dwtbl[i], i = (int)X for X >= 0, dwtbl[i] = wtbl[i] -
// wtbl[i-1] (grid.cpp:488) and dwtbl[0] = 0 (the reference never sets it).
__host__ __device__ __forceinline__ float gr_cnv_interp(float X, const float *wtbl, const float *dwtbl, int ltbl) {
    int i = (int)X;
    i = i > ltbl ? ltbl : i;
    return wtbl[i] + (X - (float)i) * dwtbl[i];
}

// gr_tap with the INTERP convolvent.
__host__ __device__ __forceinline__ bool gr_tap_interp(int j, float C, float S, float qa, float qb,
                                                       const GrGatherGeom &g, const float *wtbl,
                                                       const float *dwtbl, float &w) {
    float U, V;
    gr_sample_uv(j, C, S, g, U, V);
    if (!(U - g.L2 <= qa && qa <= U + g.L2 && V - g.L2 <= qb && qb <= V + g.L2)) return false;
    const float cu = gr_cnv_interp(fabsf(U - qa) * g.tblspcg, wtbl, dwtbl, g.ltbl);
    const float cv = gr_cnv_interp(fabsf(V - qb) * g.tblspcg, wtbl, dwtbl, g.ltbl);
    w = cu * cv;
    return true;
}

// Whether the 8 x 8 tile (tx, ty) can hold a nonzero cell: its centre within
// rmax (the largest sample radius plus the reach of a tap) of the grid centre.
// The gridding kernels write only such tiles and the row transform reads only
// them (the others are zero); both evaluate this same expression.
__host__ __device__ __forceinline__ bool gr_tile_active(int tx, int ty, float M2, float rmax) {
    const float cu = (float)(tx * 8) + 3.5f - M2, cv = (float)(ty * 8) + 3.5f - M2;
    return sqrtf(cu * cu + cv * cv) <= rmax;
}

// The tiles of tile row ty (ntx tiles wide) that gr_tile_active admits:
// [lo, hi], empty when lo > hi.  The predicate is monotone in |cu| (cu is
// exact; the rounded square, sum and square root are monotone), so these
// tiles form one interval around the column nearest the centre; a first
// guess from the real-valued circle is moved to the interval's exact ends by
// the predicate itself.  The row transform reads only these tiles.
__host__ __device__ __forceinline__ void gr_active_tile_range(int ty, int ntx, float M2, float rmax, int &lo,
                                                              int &hi) {
    const float c = (M2 - 3.5f) / 8.0f;  // |cu| is smallest at tile column c
    const int t0 = min(max((int)floorf(c), 0), ntx - 1), t1 = min(t0 + 1, ntx - 1);
    const int tc = gr_tile_active(t0, ty, M2, rmax) ? t0 : (gr_tile_active(t1, ty, M2, rmax) ? t1 : -1);
    if (tc < 0) {
        lo = 1;
        hi = 0;
        return;
    }
    const float cv = (float)(ty * 8) + 3.5f - M2;
    const float w = sqrtf(fmaxf(rmax * rmax - cv * cv, 0.0f)) / 8.0f;
    lo = max(0, min(tc, (int)floorf(c - w)));
    hi = min(ntx - 1, max(tc, (int)ceilf(c + w)));
    while (lo < tc && !gr_tile_active(lo, ty, M2, rmax)) lo++;
    while (lo > 0 && gr_tile_active(lo - 1, ty, M2, rmax)) lo--;
    while (hi > tc && !gr_tile_active(hi, ty, M2, rmax)) hi--;
    while (hi < ntx - 1 && gr_tile_active(hi + 1, ty, M2, rmax)) hi++;
}

// Angle window of an 8 x 8 tile with origin (tx*8, ty*8): the ranges
// [r0, r1) and [r2, r3) of sorted angle keys (akey, ascending in [0, pi)) whose
// sample line passes within dmax of the tile centre.  A line of direction
// theta through the grid centre is at distance R |sin(phi - theta)| from a
// point at polar (R, phi), so the window is |phi - theta| <= asin(dmax/R)
// modulo pi; central tiles take every angle.
__host__ __device__ __forceinline__ int gr_lower_bound(const float *a, int n, float x) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

__host__ __device__ __forceinline__ int gr_upper_bound(const float *a, int n, float x) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] <= x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

__host__ __device__ __forceinline__ void gr_angle_window(int tx, int ty, float M2, float dmax,
                                                         const float *akey, int n_ang, int &r0,
                                                         int &r1, int &r2, int &r3) {
    const float cu = (float)(tx * 8) + 3.5f - M2, cv = (float)(ty * 8) + 3.5f - M2;
    const float Rt = sqrtf(cu * cu + cv * cv);
    r0 = 0;
    r1 = n_ang;
    r2 = 0;
    r3 = 0;
    if (Rt <= dmax + 1.0f) return;
    const float kPi = 3.14159265f;
    float phi = atan2f(cv, cu);
    if (phi < 0.0f) phi += kPi;
    if (phi >= kPi) phi -= kPi;
    const float alpha = asinf(fminf(1.0f, dmax / Rt)) + 1e-4f;
    const float lo = phi - alpha, hi = phi + alpha;
    if (lo < 0.0f) {
        r0 = gr_lower_bound(akey, n_ang, lo + kPi);
        r3 = gr_upper_bound(akey, n_ang, hi);
    } else if (hi >= kPi) {
        r0 = gr_lower_bound(akey, n_ang, lo);
        r3 = gr_upper_bound(akey, n_ang, hi - kPi);
    } else {
        r0 = gr_lower_bound(akey, n_ang, lo);
        r1 = gr_upper_bound(akey, n_ang, hi);
    }
}

// Candidate samples of one angle for a cell at offset (du, dv) from the grid
// centre: every j whose (U, V) can lie within L2 of the cell in both axes,
// i.e. s = scale*j in {s : |s C - du| <= L2m, |s S - dv| <= L2m} with a margin
// L2m > L2 covering float rounding of U, V.  ic = 1/C and is = 1/S, or 0 when
// |C| (|S|) <= 1e-6: that axis then bounds nothing (the samples stay within
// 2048*1e-6 of the centre line) and the cell must lie within L2m of it.  The
// mirrored cell (M-a, M-b) sits at (-du, -dv), whose interval is the negation
// of this one.  Returns the direct range [jd0, jd1] and the mirrored range
// [jm0, jm1] (empty when j0 > j1), both clipped to 1 .. half-1.
__host__ __device__ __forceinline__ void gr_candidates_r(float du, float dv, float ic, float is,
                                                         float L2m, const GrGatherGeom &g,
                                                         int &jd0, int &jd1, int &jm0, int &jm1) {
    float lo = -1e9f, hi = 1e9f;
    bool empty = false;
    if (ic != 0.0f) {
        const float x0 = (du - L2m) * ic, x1 = (du + L2m) * ic;
        lo = fmaxf(lo, fminf(x0, x1));
        hi = fminf(hi, fmaxf(x0, x1));
    } else if (fabsf(du) > L2m) {
        empty = true;
    }
    if (is != 0.0f) {
        const float y0 = (dv - L2m) * is, y1 = (dv + L2m) * is;
        lo = fmaxf(lo, fminf(y0, y1));
        hi = fminf(hi, fmaxf(y0, y1));
    } else if (fabsf(dv) > L2m) {
        empty = true;
    }
    const float inv = 1.0f / g.scale;
    const float jmax = (float)(g.half - 1);
    jd0 = jm0 = 1;
    jd1 = jm1 = 0;
    if (empty || lo > hi || jmax < 1.0f) return;
    lo = fmaxf(lo, -1e9f);
    hi = fminf(hi, 1e9f);
    const float d0 = fmaxf(1.0f, ceilf(lo * inv - 1e-3f));
    const float d1 = fminf(jmax, floorf(hi * inv + 1e-3f));
    const float m0 = fmaxf(1.0f, ceilf(-hi * inv - 1e-3f));
    const float m1 = fminf(jmax, floorf(-lo * inv + 1e-3f));
    if (d0 <= d1) { jd0 = (int)d0; jd1 = (int)d1; }
    if (m0 <= m1) { jm0 = (int)m0; jm1 = (int)m1; }
}
