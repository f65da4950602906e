// gather_emul.cpp -- TEST ONLY.  Runs the gather logic of the gridding kernel
// (projects_amd/csrc/gr_gather.hpp, the same __host__ __device__ functions
// k_grid.hip calls) on the CPU, and the reference's scatter
// (gridrec/GridrecCPP/Gridrec5.2/grid.cpp:276-302) on the same filtered
// samples, both accumulating in double, so a test can check that the two
// visit exactly the same (sample, cell) pairs with the same weights.
//
// Built by __graft_entry__.build() into tests/gather_emul/libgather_emul.so.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "../../projects_amd/csrc/gr_gather.hpp"

extern "C" {

// spec1/spec2: [n_ang][half] complex (interleaved floats), entry 0 unused.
// H_*: [M][M] complex doubles, H[a][b] (a along U/COSE, b along V/SINE).
int gather_emul(int M, int half, int n_ang, float scale, float L2, float tblspcg, int ltbl,
                const float *wtbl, const float *cose, const float *sine, const float *spec1,
                const float *spec2, double *H_gather, double *H_scatter, long *taps_gather,
                long *taps_scatter) {
    const long MM = (long)M * M;
    memset(H_gather, 0, sizeof(double) * 2 * MM);
    memset(H_scatter, 0, sizeof(double) * 2 * MM);
    *taps_gather = *taps_scatter = 0;

    // --- reference scatter (grid.cpp:276-302), float decisions, double sums
    const long M2 = M >> 1;
    for (int n = 0; n < n_ang; n++) {
        for (long j = 1; j < half; j++) {
            const float *c1 = spec1 + 2 * ((long)n * half + j), *c2 = spec2 + 2 * ((long)n * half + j);
            const float rtmp = scale * (float)j;
            const float U = rtmp * cose[n] + (float)M2;
            const float V = rtmp * sine[n] + (float)M2;
            long iul = (long)ceil((double)(U - L2)), iuh = (long)floor((double)(U + L2));
            long ivl = (long)ceil((double)(V - L2)), ivh = (long)floor((double)(V + L2));
            if (iul < 1) iul = 1;
            if (iuh >= M) iuh = M - 1;
            if (ivl < 1) ivl = 1;
            if (ivh >= M) ivh = M - 1;
            for (long iu = iul; iu <= iuh; iu++) {
                float du = U - (float)iu;
                const float wu = wtbl[(long)((double)((du > 0 ? du : -du) * tblspcg) + 0.5)];
                for (long iv = ivl; iv <= ivh; iv++) {
                    float dv = V - (float)iv;
                    const float wv = wtbl[(long)((double)((dv > 0 ? dv : -dv) * tblspcg) + 0.5)];
                    const float w = wu * wv;
                    double *h = H_scatter + 2 * (iu * M + iv);
                    h[0] += (double)w * c1[0];
                    h[1] += (double)w * c1[1];
                    double *g = H_scatter + 2 * ((M - iu) * M + (M - iv));
                    g[0] += (double)w * c2[0];
                    g[1] += (double)w * c2[1];
                    *taps_scatter += 2;
                }
            }
        }
    }

    // --- the kernel's gather: angle keys and window exactly as the plan builds them
    std::vector<double> key(n_ang);
    for (int n = 0; n < n_ang; n++) {
        double k = atan2((double)sine[n], (double)cose[n]);
        if (k < 0) k += M_PI;
        if (k >= M_PI) k -= M_PI;
        key[n] = k;
    }
    std::vector<int> ord(n_ang);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return key[x] < key[y]; });
    std::vector<float> akey(n_ang);
    for (int i = 0; i < n_ang; i++) akey[i] = (float)key[ord[i]];
    for (int i = 1; i < n_ang; i++) akey[i] = std::max(akey[i], akey[i - 1]);

    // per-angle table in sorted order, as the plan uploads it
    std::vector<float> C4(4 * (size_t)n_ang);
    for (int i = 0; i < n_ang; i++) {
        const float c = cose[ord[i]], sn = sine[ord[i]];
        C4[4 * i] = c;
        C4[4 * i + 1] = sn;
        C4[4 * i + 2] = fabsf(c) > 1e-6f ? 1.0f / c : 0.0f;
        C4[4 * i + 3] = fabsf(sn) > 1e-6f ? 1.0f / sn : 0.0f;
    }

    GrGatherGeom g;
    g.scale = scale;
    g.L2 = L2;
    g.tblspcg = tblspcg;
    g.fM2 = (float)M2;
    g.M = M;
    g.half = half;
    g.ltbl = ltbl;
    // the launch constants of k_grid.hip
    const float L2m = L2 + 0.01f;
    const float dmax = 3.5f * 1.41421356f + L2m * 1.41421356f + 0.05f;
    const float rmax = g.scale * (float)(g.half - 1) * 1.0001f + dmax + 0.5f;
    const float reach = 3.5f + L2m;
    const int nt = (M + 7) / 8;
    for (int ty = 0; ty < nt; ty++) {
        for (int tx = 0; tx < nt; tx++) {
            const float cu = (float)(tx * 8) + 3.5f - g.fM2, cv = (float)(ty * 8) + 3.5f - g.fM2;
            const bool reachable = sqrtf(cu * cu + cv * cv) <= rmax;
            int r0 = 0, r1 = 0, r2 = 0, r3 = 0;
            if (reachable) gr_angle_window(tx, ty, g.fM2, dmax, akey.data(), n_ang, r0, r1, r2, r3);
            for (int lane = 0; lane < 64; lane++) {
                const int a = tx * 8 + (lane & 7), b = ty * 8 + (lane >> 3);
                if (a >= M || b >= M) continue;
                const bool valid = a >= 1 && b >= 1;
                const float pa = (float)a, pb = (float)b, ma = (float)(M - a), mb = (float)(M - b);
                const float du = pa - g.fM2, dv = pb - g.fM2;
                double sr = 0, si = 0;
                for (int part = 0; part < 2; part++) {
                    const int k0 = part ? r2 : r0, k1 = part ? r3 : r1;
                    for (int k = k0; k < k1; k++) {
                        const float *cs = &C4[4 * (size_t)k];
                        if (fabsf(cu * cs[1] - cv * cs[0]) > reach * (fabsf(cs[0]) + fabsf(cs[1])) + 0.05f) continue;
                        int td0, td1, tm0, tm1;
                        gr_candidates_r(cu, cv, cs[2], cs[3], reach, g, td0, td1, tm0, tm1);
                        if (td0 > td1 && tm0 > tm1) continue;
                        int jd0, jd1, jm0, jm1;
                        gr_candidates_r(du, dv, cs[2], cs[3], L2m, g, jd0, jd1, jm0, jm1);
                        if (!valid) continue;
                        const int n = ord[k];
                        for (int mir = 0; mir < 2; mir++) {
                            const int j0 = mir ? std::max(jm0, tm0) : std::max(jd0, td0);
                            const int j1 = mir ? std::min(jm1, tm1) : std::min(jd1, td1);
                            for (int j = j0; j <= j1; j++) {
                                float w;
                                if (gr_tap(j, cs[0], cs[1], mir ? ma : pa, mir ? mb : pb, g, wtbl, w)) {
                                    const float *v = (mir ? spec2 : spec1) + 2 * ((long)n * half + j);
                                    sr += (double)w * v[0];
                                    si += (double)w * v[1];
                                    *taps_gather += 1;
                                }
                            }
                        }
                    }
                }
                H_gather[2 * ((long)a * M + b)] = sr;
                H_gather[2 * ((long)a * M + b) + 1] = si;
            }
        }
    }
    return 0;
}

// For every tile row of an M x M grid: does gr_active_tile_range give exactly
// the tiles gr_tile_active admits?  Returns the number of rows where it does
// not (0 expected); *n_active receives the admitted tiles in all.
int tile_range_check(int M, float rmax, long *n_active) {
    const int ntx = M / 8;
    const float M2 = (float)(M >> 1);
    int bad = 0;
    long count = 0;
    for (int ty = 0; ty < ntx; ty++) {
        int lo, hi;
        gr_active_tile_range(ty, ntx, M2, rmax, lo, hi);
        for (int tx = 0; tx < ntx; tx++) {
            const bool act = gr_tile_active(tx, ty, M2, rmax);
            count += act;
            if (act != (tx >= lo && tx <= hi)) {
                bad++;
                break;
            }
        }
    }
    *n_active = count;
    return bad;
}

}  // extern "C"
