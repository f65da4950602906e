// gr_noalloc.cpp -- gr_recon_device_noalloc: gr_recon_device without any
// allocation, for callers that cannot allocate device memory (code under HIP
// graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The kernels' scratch -- the filtered spectra, the gridded planes and the
// row-transformed planes of a chunk of slice pairs, normally the plan's own
// workspace -- lives in one file-scope device array of kScratchBytes.  Each
// GPU holds its own copy for the life of the process; its address is looked up
// once per device (hipGetSymbolAddress, with that device current) and only
// that address is handed to kernels: in host code the array's name is not a
// device address.
//
// Every call on a GPU shares that GPU's array and leaves its contents behind:
// two calls that overlap on one GPU (different streams, or threads) race, so
// callers serialise them.  Nothing in the scratch has to start at zero: each
// stage writes every element the next one reads, as with the plan's own
// (never cleared) workspace.  A plan whose gather operator is not built yet
// keeps evaluating the geometry per launch here (building it would allocate),
// and no per-kernel timing is recorded (events are resources too); the
// images are bitwise those of gr_recon_device either way.
#include <mutex>

#include "gr_plan.h"

namespace {

constexpr size_t kScratchBytes = size_t(128) << 20;  // 128 MiB per GPU
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void *g_scratch_ptr[kMaxDevices] = {};

#define NA_HIP(call)                                                                    \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) return gr_set_error(GR_EHIP, hipGetErrorString(e_));      \
    } while (0)

// The device address of this GPU's scratch (the current device must be dev).
int scratch_of(int dev, void **ptr) {
    if (dev < 0 || dev >= kMaxDevices) return gr_set_error(GR_EINVAL, "device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        NA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return gr_set_error(GR_EHIP, "no device address for the static scratch");
        g_scratch_ptr[dev] = p;
    }
    *ptr = g_scratch_ptr[dev];
    return GR_OK;
}

size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

extern "C" {

size_t gr_noalloc_scratch_bytes(void) { return kScratchBytes; }

int gr_recon_device_noalloc(gr_plan *pl, const float *sino, long sino_slice_stride, long sino_row_stride, float *out,
                            long out_slice_stride, long out_row_stride, int n_slices, void *stream) {
    if (!pl || !sino || !out) return gr_set_error(GR_EINVAL, "null argument");
    if (n_slices < 0) return gr_set_error(GR_EINVAL, "negative slice count");
    if (n_slices == 0) return GR_OK;
    const GrGeom &g = pl->g;
    if (sino_row_stride < g.n_det_in || sino_slice_stride < sino_row_stride * (long)pl->prm.n_ang)
        return gr_set_error(GR_EINVAL, "sinogram strides smaller than the rows they hold");
    if (out_row_stride < g.out_n || out_slice_stride < out_row_stride * g.out_n)
        return gr_set_error(GR_EINVAL, "image strides smaller than the window they hold");
    // pairs per launch: as many as the plan takes and the scratch holds
    int mp = pl->max_pairs;
    while (mp > 0 && gr_plan_workspace_bytes(pl, mp) > kScratchBytes) mp--;
    if (mp < 1) {
        char msg[160];
        snprintf(msg, sizeof msg, "one slice pair needs %zu bytes of scratch, more than the %zu of the static scratch",
                 gr_plan_workspace_bytes(pl, 1), kScratchBytes);
        return gr_set_error(GR_EINVAL, msg);
    }
    NA_HIP(hipSetDevice(pl->prm.device));
    void *base = nullptr;
    const int rc = scratch_of(pl->prm.device, &base);
    if (rc) return rc;
    // the layout of gr_plan_bind_workspace for mp pairs
    const size_t spec_pp = (size_t)pl->prm.n_ang * g.half, ht_pp = (size_t)g.M * g.M;
    const size_t spec_pairs = (size_t)(mp + kSpecGroup - 1) / kSpecGroup * kSpecGroup;
    char *p = (char *)base;
    float2 *spec1 = (float2 *)p;
    p += up256(sizeof(float2) * spec_pp * spec_pairs);
    float2 *spec2 = (float2 *)p;
    p += up256(sizeof(float2) * spec_pp * spec_pairs);
    float2 *ht = (float2 *)p;
    p += up256(sizeof(float2) * ht_pp * mp);
    float2 *z = (float2 *)p;
    hipStream_t s = (hipStream_t)stream;
    const GrDevTables t = pl->tables();
    const int n_pairs = (n_slices + 1) / 2;
    for (int p0 = 0; p0 < n_pairs; p0 += mp) {
        const int np = n_pairs - p0 < mp ? n_pairs - p0 : mp;
        NA_HIP(gr_launch_rowfft(g, t, pl->prm.n_ang, sino, sino_slice_stride, sino_row_stride, nullptr, n_slices, p0,
                                np, spec1, spec2, s));
        NA_HIP(gr_launch_grid(g, t, pl->prm.n_ang, pl->prm.ltbl, spec1, spec2, ht, z, np, pl->prm.accum, s));
        NA_HIP(gr_launch_ifft_rows(g, t, ht, z, np, s));
        NA_HIP(gr_launch_ifft_cols(g, t, z, out, out_slice_stride, out_row_stride, n_slices, p0, np,
                                   pl->cols_per_block, s));
    }
    return GR_OK;
}

}  // extern "C"
