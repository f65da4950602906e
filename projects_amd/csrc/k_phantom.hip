// k_phantom.hip -- synthetic sinograms on the device: seeded random-ellipse
// phantoms plus an optional band-limited random texture, analytic Radon
// transform, Poisson photon noise and -log.
//
// The same generator as projects_amd/phantom.py (which documents the counter
// layout): Philox4x32-10 counters give identical ellipse parameters and noise
// draws on CPU and GPU; the arithmetic here is float, so sinograms agree with
// the numpy version to float rounding (and noisy ones statistically).  The
// reference reads its sinograms from files; this exists to produce the
// benchmark workloads at full size in HBM.
#include "gr_internal.h"

namespace {

struct Philox {
    unsigned x[4];
};

__device__ __forceinline__ Philox philox4x32_10(unsigned c0, unsigned c1, unsigned c2, unsigned c3,
                                                unsigned k0, unsigned k1) {
    for (int r = 0; r < 10; r++) {
        const unsigned hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
        const unsigned hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
        const unsigned n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    Philox o;
    o.x[0] = c0;
    o.x[1] = c1;
    o.x[2] = c2;
    o.x[3] = c3;
    return o;
}

__device__ __forceinline__ float u01(unsigned x) { return ((float)(x >> 8) + 0.5f) * (1.0f / 16777216.0f); }

__global__ void k_ellipses(GrEllipse *E, int n_slices, long slice0, int n_ell, unsigned k0, unsigned k1,
                           int density) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    if (e >= n_ell || s >= n_slices) return;
    const unsigned sl = (unsigned)(slice0 + s);
    const Philox r = philox4x32_10((unsigned)e, sl, 0u, 0u, k0, k1);
    const Philox q = philox4x32_10((unsigned)e, sl, 1u, 0u, k0, k1);
    GrEllipse o;
    o.a = 0.05f + 0.4f * u01(r.x[0]);
    o.b = 0.05f + 0.4f * u01(r.x[1]);
    o.rho = density ? exp10f(-4.0f + 4.0f * u01(r.x[2])) : -0.3f + 1.3f * u01(r.x[2]);
    o.phi = 3.14159265f * u01(r.x[3]);
    const float rad = sqrtf(u01(q.x[0]));
    float sa, ca;
    sincospif(2.0f * u01(q.x[1]), &sa, &ca);
    o.x0 = rad * ca;
    o.y0 = rad * sa;
    E[(long)s * n_ell + e] = o;
}

// log of the Poisson probability k log(lam) - lam - log(k!).  For k >= 9 the
// Stirling series of log Gamma(z), z = k+1, turns it into
//   k log1p((lam-z)/z) + (z-lam) - log(2 pi z)/2 - 1/(12z) + 1/(360z^3),
// whose terms are O(sqrt(lam)) instead of O(lam log lam), so float keeps it to
// ~1e-5 absolute where the direct form loses everything at lam ~ 1e4.
__device__ __forceinline__ float log_poisson_pmf(float kk, float lam) {
    if (kk < 9.0f) return kk * logf(lam) - lam - lgammaf(kk + 1.0f);
    const float z = kk + 1.0f, iz = 1.0f / z;
    return kk * log1pf((lam - z) * iz) + (z - lam) - 0.5f * logf(6.28318531f * z) - iz / 12.0f +
           iz * iz * iz / 360.0f;
}

// Poisson draw: inversion below 10, PTRS (Hoermann 1993) above.
__device__ float poisson_draw(float lam, unsigned k, unsigned n, unsigned sl, unsigned k0, unsigned k1) {
    if (lam < 10.0f) {
        const Philox w = philox4x32_10(k, n, sl, 2u, k0, k1);
        const float u = u01(w.x[0]);
        float p = expf(-lam), F = p, kk = 0.0f;
        for (int it = 0; it < 1000 && u > F; it++) {
            kk += 1.0f;
            p *= lam / kk;
            F += p;
        }
        return kk;
    }
    const float slam = sqrtf(lam);
    const float b = 0.931f + 2.53f * slam;
    const float a = -0.059f + 0.02483f * b;
    const float inv_alpha = 1.1239f + 1.1328f / (b - 3.4f);
    const float vr = 0.9277f - 3.6224f / (b - 2.0f);
    for (int r = 0; r < 64; r++) {
        const Philox w = philox4x32_10(k, n, sl, 2u + (unsigned)(r >> 1), k0, k1);
        const float U = u01(w.x[(r & 1) * 2]) - 0.5f, V = u01(w.x[(r & 1) * 2 + 1]);
        const float us = 0.5f - fabsf(U);
        const float kk = floorf((2.0f * a / us + b) * U + lam + 0.43f);
        if (us >= 0.07f && V <= vr) return kk;
        if (kk < 0.0f || (us < 0.013f && V > us)) continue;
        if (logf(V) + logf(inv_alpha) - logf(a / (us * us) + b) <= log_poisson_pmf(kk, lam)) return kk;
    }
    return rintf(lam);
}

// Texture term k of a slice: amplitude, frequency (cycles per unit length)
// and phase of amp * cos(2 pi f.r + phase) inside the unit disk.
__global__ void k_texture(GrWave *T, int n_slices, long slice0, int n_tex, float amp, float fmax, unsigned k0,
                          unsigned k1) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    if (e >= n_tex || s >= n_slices) return;
    const unsigned sl = (unsigned)(slice0 + s);
    const Philox r = philox4x32_10((unsigned)e, sl, 3u, 0u, k0, k1);
    const float fr = fmax * sqrtf(u01(r.x[0]));  // uniform over the disk |f| <= fmax
    float sd, cd;
    sincospif(2.0f * u01(r.x[1]), &sd, &cd);
    GrWave w;
    w.fx = fr * cd;
    w.fy = fr * sd;
    w.phase = 6.28318531f * u01(r.x[2]);
    const Philox q = philox4x32_10((unsigned)e, sl, 4u, 0u, k0, k1);
    const float z = sqrtf(-2.0f * logf(u01(q.x[0]))) * cospif(2.0f * u01(q.x[1]));
    w.amp = amp * z / sqrtf((float)n_tex);
    T[(long)s * n_tex + e] = w;
}

struct SinoArgs {
    float *sino;
    const GrEllipse *E;
    const GrWave *T;
    const float *centers;
    long slice_stride, slice0;
    int n_ang, n_det, n_ell, n_tex;
    float photons, gauss_rel;
    unsigned k0, k1;
};

__global__ void __launch_bounds__(256) k_sinogram(SinoArgs A) {
    __shared__ GrEllipse se[64];
    __shared__ GrWave sw[256];
    const int s = blockIdx.z, n = blockIdx.y;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (threadIdx.x < A.n_ell) se[threadIdx.x] = A.E[(long)s * A.n_ell + threadIdx.x];
    if (threadIdx.x < A.n_tex) sw[threadIdx.x] = A.T[(long)s * A.n_tex + threadIdx.x];
    __syncthreads();
    if (k >= A.n_det) return;
    float st, ct;
    sincospif((float)n / (float)A.n_ang, &st, &ct);
    const float t = ((float)k - A.centers[s]) / (0.5f * (float)A.n_det);
    float p = 0.0f;
    for (int e = 0; e < A.n_ell; e++) {
        const GrEllipse el = se[e];
        float sp, cp;
        sincosf(el.phi, &sp, &cp);
        const float c = ct * cp + st * sp;  // cos(theta - phi)
        const float sn = st * cp - ct * sp; // sin(theta - phi)
        const float s2 = el.a * el.a * c * c + el.b * el.b * sn * sn;
        const float tp = t - (el.x0 * ct + el.y0 * st);
        const float d = s2 - tp * tp;
        if (d > 0.0f) p += 2.0f * el.rho * el.a * el.b * sqrtf(d) / s2;
    }
    // texture: the chord of the unit disk at offset t has half-length h; along
    // it cos(alpha + beta s) integrates to 2 cos(alpha) sin(beta h) / beta
    if (A.n_tex > 0 && fabsf(t) < 1.0f) {
        const float h = sqrtf(1.0f - t * t);
        for (int e = 0; e < A.n_tex; e++) {
            const GrWave w = sw[e];
            const float alpha = 6.28318531f * t * (w.fx * ct + w.fy * st) + w.phase;
            const float beta = 6.28318531f * (w.fy * ct - w.fx * st);
            const float bh = beta * h;
            const float sinc = fabsf(bh) > 1e-4f ? sinf(bh) / bh : 1.0f - bh * bh / 6.0f;
            p += w.amp * 2.0f * h * cosf(alpha) * sinc;
        }
    }
    if (A.photons > 0.0f) {
        const float I = poisson_draw(A.photons * expf(-p), (unsigned)k, (unsigned)n,
                                     (unsigned)(A.slice0 + s), A.k0, A.k1);
        p = -logf(fmaxf(I, 1.0f) / A.photons);
    }
    if (A.gauss_rel > 0.0f) {  // p * (1 + gauss_rel * z), z by Box-Muller on counter word 1<<16
        const Philox w = philox4x32_10((unsigned)k, (unsigned)n, (unsigned)(A.slice0 + s), 1u << 16, A.k0, A.k1);
        const float z = sqrtf(-2.0f * logf(u01(w.x[0]))) * cospif(2.0f * u01(w.x[1]));
        p = p * (1.0f + A.gauss_rel * z);
    }
    A.sino[(long)s * A.slice_stride + (long)n * A.n_det + k] = p;
}

}  // namespace

hipError_t gr_launch_phantom(float *sino, long slice_stride, int n_slices, long slice0,
                             unsigned long long seed, int n_ang, int n_det, const float *centers_dev,
                             int n_ell, float photons, int density, float gauss_rel, int n_tex, float tex_amp,
                             float tex_fmax, hipStream_t s) {
    if (n_ell < 1 || n_ell > 64 || n_tex < 0 || n_tex > 256 || n_slices < 1 || n_ang < 1 || n_det < 1)
        return hipErrorInvalidValue;
    GrEllipse *E = nullptr;
    GrWave *T = nullptr;
    hipError_t e = hipMallocAsync((void **)&E, sizeof(GrEllipse) * (size_t)n_slices * n_ell, s);
    if (e != hipSuccess) return e;
    e = hipMallocAsync((void **)&T, sizeof(GrWave) * (size_t)n_slices * (n_tex ? n_tex : 1), s);
    if (e != hipSuccess) {
        (void)hipFreeAsync(E, s);
        return e;
    }
    const unsigned k0 = (unsigned)(seed & 0xffffffffu), k1 = (unsigned)(seed >> 32);
    hipLaunchKernelGGL(k_ellipses, dim3((n_ell + 63) / 64, n_slices), dim3(64), 0, s, E, n_slices, slice0,
                       n_ell, k0, k1, density);
    if (n_tex > 0)
        hipLaunchKernelGGL(k_texture, dim3((n_tex + 63) / 64, n_slices), dim3(64), 0, s, T, n_slices, slice0, n_tex,
                           tex_amp, tex_fmax, k0, k1);
    SinoArgs A{sino, E, T, centers_dev, slice_stride, slice0, n_ang, n_det, n_ell, n_tex, photons, gauss_rel, k0, k1};
    for (int s0 = 0; s0 < n_slices; s0 += 65535) {
        const int ns = n_slices - s0 < 65535 ? n_slices - s0 : 65535;
        SinoArgs B = A;
        B.sino = sino + (long)s0 * slice_stride;
        B.E = E + (long)s0 * n_ell;
        B.T = T + (long)s0 * n_tex;
        B.centers = centers_dev + s0;
        B.slice0 = slice0 + s0;
        hipLaunchKernelGGL(k_sinogram, dim3((n_det + 255) / 256, n_ang, ns), dim3(256), 0, s, B);
    }
    e = hipGetLastError();
    hipError_t f = hipFreeAsync(E, s);
    hipError_t f2 = hipFreeAsync(T, s);
    return e != hipSuccess ? e : (f != hipSuccess ? f : f2);
}
