// gr_capi.cpp -- the C-ABI of include/gridrec_hip.h.
//
// A plan is what recon_init sets up (gridrec/GridrecCPP/Gridrec5.2/grid.cpp:54-147):
// sizes, the sine/cosine, filter-phase, PSWF and deapodization tables -- here
// resident in HBM -- plus this library's own device tables (sorted angle keys
// for the gather, the tile order, FFT twiddles) and a workspace for
// `max_pairs` slice pairs.  gr_recon_device runs K1..K4 over a batch of pairs
// in chunks of max_pairs; nothing is allocated per call.  Plans are independent
// objects: unlike the reference's globals, any number may coexist.
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "gr_fft.hpp"
#include "gr_plan.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

int hip_fail(hipError_t e, const char *what) {
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return GR_EHIP;
}

#define GR_HIP(call)                                        \
    do {                                                    \
        hipError_t e_ = (call);                             \
        if (e_ != hipSuccess) return hip_fail(e_, #call);   \
    } while (0)

}  // namespace


namespace {

template <typename T>
int dev_upload(gr_plan *pl, T **dst, const T *src, size_t n) {
    GR_HIP(hipMalloc((void **)dst, sizeof(T) * (n ? n : 1)));
    pl->dev_bytes += sizeof(T) * (n ? n : 1);
    if (n) GR_HIP(hipMemcpy(*dst, src, sizeof(T) * n, hipMemcpyHostToDevice));
    return GR_OK;
}

template <typename T>
int dev_alloc(gr_plan *pl, T **dst, size_t n) {
    hipError_t e = hipMalloc((void **)dst, sizeof(T) * (n ? n : 1));
    if (e != hipSuccess) return fail(GR_ENOMEM, std::string("device allocation failed: ") + hipGetErrorString(e));
    pl->dev_bytes += sizeof(T) * (n ? n : 1);
    return GR_OK;
}

// Tables that depend only on a transform size -- twiddles and the gridding
// tile order -- are built once per (device, kind, size) and shared by every
// plan that needs them, reference-counted: a batch of hundreds of datasets
// holds one copy per size instead of one per dataset.
struct SharedTable {
    void *ptr;
    int refs;
    size_t count;  // elements
};
std::mutex g_shared_mu;
std::map<std::tuple<int, int, long>, SharedTable> g_shared;
enum { kTwPlus = 0, kTwMinus = 1, kTiles = 2 };

template <typename T, typename Build>
int shared_table(int device, int kind, long n, Build build, T **out, int *count) {
    std::lock_guard<std::mutex> lock(g_shared_mu);
    auto key = std::make_tuple(device, kind, n);
    auto it = g_shared.find(key);
    if (it == g_shared.end()) {
        std::vector<T> h = build();
        void *d = nullptr;
        hipError_t e = hipMalloc(&d, sizeof(T) * (h.empty() ? 1 : h.size()));
        if (e == hipSuccess && !h.empty()) e = hipMemcpy(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            if (d) (void)hipFree(d);
            return fail(GR_ENOMEM, std::string("shared table: ") + hipGetErrorString(e));
        }
        it = g_shared.emplace(key, SharedTable{d, 0, h.size()}).first;
    }
    if (count) *count = (int)it->second.count;
    it->second.refs++;
    *out = (T *)it->second.ptr;
    return GR_OK;
}

void shared_release(const void *ptr) {
    if (!ptr) return;
    std::lock_guard<std::mutex> lock(g_shared_mu);
    for (auto it = g_shared.begin(); it != g_shared.end(); ++it) {
        if (it->second.ptr != ptr) continue;
        if (--it->second.refs == 0) {
            (void)hipFree(it->second.ptr);
            g_shared.erase(it);
        }
        return;
    }
}

int plan_stream(gr_plan *pl, hipStream_t *s) {
    if (!pl->stream && hipStreamCreateWithFlags(&pl->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(GR_EHIP, "stream creation failed");
    *s = pl->stream;
    return GR_OK;
}

// Tiles with more rows than this are gridded one pair per wave: a few dozen
// near the centre of the grid, whose thousands of rows would otherwise make
// the launch wait for their waves.
constexpr unsigned kGopHeavyRows = 64;

// Build K2's gather operator on stream s: count pass, exclusive sum of the
// per-tile entry counts on the host, fill pass.  Blocks the host until it is
// built.  With `required` false a geometry whose operator would not fit (its
// sample index range, its entry count, a third of the free device memory, or
// max_bytes) is declined quietly and the plan keeps evaluating the geometry
// per launch.
int build_gop(gr_plan *pl, hipStream_t s, bool required, size_t max_bytes) {
    const auto t0 = std::chrono::steady_clock::now();
    const GrGeom &g = pl->g;
    const int n_ang = pl->prm.n_ang;
    size_t sop_bytes = 0;
    auto free_sop = [&]() {
        for (void *p : {(void *)pl->d_sop_tile, (void *)pl->d_sop_ch, (void *)pl->d_sop_keys, (void *)pl->d_sop_ent,
                        (void *)pl->d_hmeta, (void *)pl->d_hseg, (void *)pl->d_hsplit, (void *)pl->d_hitems})
            if (p) (void)hipFree(p);
        pl->d_sop_tile = nullptr;
        pl->d_sop_ch = nullptr;
        pl->d_sop_keys = pl->d_sop_ent = nullptr;
        pl->d_hmeta = nullptr;
        pl->d_hseg = nullptr;
        pl->d_hsplit = nullptr;
        pl->d_hitems = nullptr;
        pl->n_hitems = pl->n_split = pl->split_groups = 0;
        pl->dev_bytes -= sop_bytes;
        sop_bytes = 0;
        pl->sop_chunks = 0;
    };
    auto decline = [&](const char *why) {
        pl->gop_state = -1;
        (void)hipStreamSynchronize(s);
        for (unsigned **pp : {&pl->d_gop_n, &pl->d_gop_off}) {
            if (*pp) {
                (void)hipFree(*pp);
                *pp = nullptr;
                pl->dev_bytes -= sizeof(unsigned) * pl->ntiles;
            }
        }
        free_sop();
        pl->gop_aligned_tiles = pl->gop_heavy = 0;
        return required ? fail(GR_ENOMEM, std::string("gather operator: ") + why) : GR_OK;
    };
    auto dev_reserve = [&](void **p, size_t bytes) {
        if (hipMalloc(p, bytes ? bytes : 1) != hipSuccess) {
            *p = nullptr;
            (void)hipGetLastError();
            return false;
        }
        sop_bytes += bytes ? bytes : 1;
        pl->dev_bytes += bytes ? bytes : 1;
        return true;
    };
    pl->gop_state = -1;  // until it is built: a failed build is not tried again
    if (!gr_gop_fits(g, n_ang)) return decline("sample index range too large");
    const int nt = pl->ntiles;
    GR_HIP(hipMalloc((void **)&pl->d_gop_n, sizeof(unsigned) * nt));
    GR_HIP(hipMalloc((void **)&pl->d_gop_off, sizeof(unsigned) * nt));
    pl->dev_bytes += 2 * sizeof(unsigned) * nt;
    GrDevTables t = pl->tables();
    // count pass: rows of both layouts per tile (the aligned counts go to the
    // offsets array, overwritten below) and merge groups per tile
    std::vector<unsigned> cnt(nt), cnt_al(nt), off(nt), groups(nt);
    {
        unsigned *d_groups = nullptr;
        if (!dev_reserve((void **)&d_groups, sizeof(unsigned) * nt)) return decline("device allocation failed");
        hipError_t e = gr_launch_gop_count(g, t, n_ang, pl->prm.ltbl, pl->d_gop_n, pl->d_gop_off, d_groups, s);
        if (e == hipSuccess) e = hipMemcpyAsync(cnt.data(), pl->d_gop_n, sizeof(unsigned) * nt, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(cnt_al.data(), pl->d_gop_off, sizeof(unsigned) * nt, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipMemcpyAsync(groups.data(), d_groups, sizeof(unsigned) * nt, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        (void)hipFree(d_groups);
        sop_bytes -= sizeof(unsigned) * nt;
        pl->dev_bytes -= sizeof(unsigned) * nt;
        if (e != hipSuccess) return hip_fail(e, "gather operator count pass");
    }
    // per tile the aligned layout (wave-uniform merges) where it costs at
    // most a quarter more rows than the ragged one
    unsigned long long total = 0;
    for (int i = 0; i < nt; i++) {
        if (cnt[i] >= kGopAlignedBit || cnt_al[i] >= kGopAlignedBit) return decline("tile row count overflow");
        if (cnt_al[i] && 4ull * cnt_al[i] <= 5ull * cnt[i]) {
            cnt[i] = cnt_al[i] | kGopAlignedBit;
            pl->gop_aligned_tiles++;
        }
        // the leading slots of every sector (nearest to the centre first) up to
        // the last tile of more than kGopHeavyRows rows take the
        // one-pair-per-wave path
        if ((cnt[i] & ~kGopAlignedBit) > kGopHeavyRows) pl->gop_heavy = std::max(pl->gop_heavy, i % pl->sector_len + 1);
    }
    // with compensated sums the heavy tiles are gridded in segments of
    // kGopSegGroups merge groups (the merges are the cut points), all P pairs
    // of a group per wave, their partials in the row transform's output
    // buffer, which must hold them for every pair; otherwise one pair per wave
    std::vector<uint4> hmeta;
    std::vector<unsigned> hsplit;
    unsigned long long seg_total = 0, part_total = 0;
    if (pl->prm.accum == GR_ACCUM_COMPENSATED) {
        hmeta.assign(nt, make_uint4(0u, 0u, 0u, 0u));
        for (int i = 0; i < nt; i++) {
            if (i % pl->sector_len >= pl->gop_heavy || groups[i] == 0) continue;
            const unsigned S = (groups[i] + kGopSegGroups - 1) / kGopSegGroups;
            hmeta[i] = make_uint4(S, (unsigned)seg_total, (unsigned)part_total, (unsigned)hsplit.size());
            seg_total += 64ull * (S + 1);
            part_total += groups[i];
            hsplit.push_back((unsigned)i);
        }
        if (hsplit.empty() || part_total * 64 > (unsigned long long)g.M * g.out_n || seg_total > 0xffffffffull) {
            hmeta.clear();
            hsplit.clear();
        }
    }
    // the heavy work items, slot by slot across the sectors (the centre-most
    // tiles first): a cut tile's segments side by side, or per pair
    std::vector<uint2> hitems;
    for (int slot = 0; slot < pl->gop_heavy; slot++)
        for (int sx = 0; sx < kGridSectors; sx++) {
            const int i = sx * pl->sector_len + slot;
            const unsigned S = hmeta.empty() ? (unsigned)kSpecGroup : hmeta[i].x;
            for (unsigned k = 0; k < S; k++) hitems.push_back(make_uint2((unsigned)i, k));
        }
    // the staged form for the tiles past the heavy slots, where the plan
    // admits it: count pass, exclusive sums of extra chunks and rows, fill
    // pass (below, once the budget is checked); those tiles then take no
    // entries of the streamed form
    size_t sop_need = 0;
    unsigned long long sop_rows = 0, sop_ch = nt;
    std::vector<uint4> sop(nt);
    if (gr_sop_fits(g, t, pl->prm.ltbl)) {
        if (!dev_reserve((void **)&pl->d_sop_tile, sizeof(uint4) * nt)) return decline("device allocation failed");
        GR_HIP(gr_launch_sop_count(g, t, n_ang, pl->prm.ltbl, pl->gop_heavy, pl->d_sop_tile, s));
        GR_HIP(hipMemcpyAsync(sop.data(), pl->d_sop_tile, sizeof(uint4) * nt, hipMemcpyDeviceToHost, s));
        GR_HIP(hipStreamSynchronize(s));
        for (int i = 0; i < nt; i++) {
            const uint4 c = sop[i];
            sop[i] = make_uint4((unsigned)(sop_ch - nt), c.y, 0u, (unsigned)sop_rows);
            sop_ch += c.y ? c.y - 1 : 0;
            sop_rows += c.w;
            if (i % pl->sector_len >= pl->gop_heavy) cnt[i] = 0;
        }
        sop_rows += 1;  // a spare row: the clamped entry loads of a chunk without rows stay inside
        if (sop_ch * kSopKeysPerChunk > 0xffffffffull || sop_rows > 0xffffffffull / 64)
            return decline("staged operator too large");
        sop_need = (sizeof(uint2) + sizeof(unsigned) * kSopKeysPerChunk) * sop_ch + sizeof(unsigned) * 64ull * sop_rows;
    }
    for (int i = 0; i < nt; i++) {
        off[i] = (unsigned)total;
        total += 64ull * (cnt[i] & ~kGopAlignedBit);
        if (total > 0xffffffffull) return decline("more than 2^32 entries");
    }
    const size_t bytes = sizeof(uint2) * (size_t)std::max<unsigned long long>(total, 1);
    size_t free_b = 0, total_b = 0;
    if (!required && (hipMemGetInfo(&free_b, &total_b) != hipSuccess || bytes + sop_need > free_b / 3 ||
                      bytes + sop_need > max_bytes))
        return decline("larger than a third of the free device memory or the caller's budget");
    if (hipMalloc((void **)&pl->d_gop, bytes) != hipSuccess) {
        pl->d_gop = nullptr;
        (void)hipGetLastError();
        return decline("device allocation failed");
    }
    pl->dev_bytes += bytes;
    if (pl->d_sop_tile) {
        if (!dev_reserve((void **)&pl->d_sop_ch, sizeof(uint2) * sop_ch) ||
            !dev_reserve((void **)&pl->d_sop_keys, sizeof(unsigned) * kSopKeysPerChunk * sop_ch) ||
            !dev_reserve((void **)&pl->d_sop_ent, sizeof(unsigned) * 64ull * sop_rows)) {
            (void)hipFree(pl->d_gop);
            pl->d_gop = nullptr;
            pl->dev_bytes -= bytes;
            return decline("device allocation failed");
        }
        pl->sop_chunks = (int)std::min<unsigned long long>(sop_ch, INT32_MAX);
        GR_HIP(hipMemcpyAsync(pl->d_sop_tile, sop.data(), sizeof(uint4) * nt, hipMemcpyHostToDevice, s));
        // the spare entry row: slot 0
        static const unsigned zeros[64] = {};
        GR_HIP(hipMemcpyAsync(pl->d_sop_ent + 64ull * (sop_rows - 1), zeros, sizeof zeros, hipMemcpyHostToDevice, s));
        GR_HIP(gr_launch_sop_fill(g, t, n_ang, pl->prm.ltbl, pl->gop_heavy, pl->d_sop_tile, pl->d_sop_ch,
                                  pl->d_sop_keys, pl->d_sop_ent, s));
    }
    if (!hitems.empty()) {
        if (!dev_reserve((void **)&pl->d_hitems, sizeof(uint2) * hitems.size())) {
            (void)hipFree(pl->d_gop);
            pl->d_gop = nullptr;
            pl->dev_bytes -= bytes;
            return decline("device allocation failed");
        }
        GR_HIP(hipMemcpyAsync(pl->d_hitems, hitems.data(), sizeof(uint2) * hitems.size(), hipMemcpyHostToDevice, s));
        pl->n_hitems = (int)hitems.size();
    }
    if (!hsplit.empty()) {
        if (!dev_reserve((void **)&pl->d_hmeta, sizeof(uint4) * nt) ||
            !dev_reserve((void **)&pl->d_hseg, sizeof(uint2) * seg_total) ||
            !dev_reserve((void **)&pl->d_hsplit, sizeof(unsigned) * hsplit.size())) {
            (void)hipFree(pl->d_gop);
            pl->d_gop = nullptr;
            pl->dev_bytes -= bytes;
            return decline("device allocation failed");
        }
        GR_HIP(hipMemcpyAsync(pl->d_hmeta, hmeta.data(), sizeof(uint4) * nt, hipMemcpyHostToDevice, s));
        GR_HIP(hipMemcpyAsync(pl->d_hsplit, hsplit.data(), sizeof(unsigned) * hsplit.size(), hipMemcpyHostToDevice,
                              s));
        pl->n_split = (int)hsplit.size();
        pl->split_groups = (int)part_total;
    }
    pl->gop_bytes = bytes + 2 * sizeof(unsigned) * nt + sop_bytes;
    GR_HIP(hipMemcpyAsync(pl->d_gop_n, cnt.data(), sizeof(unsigned) * nt, hipMemcpyHostToDevice, s));
    GR_HIP(hipMemcpyAsync(pl->d_gop_off, off.data(), sizeof(unsigned) * nt, hipMemcpyHostToDevice, s));
    GR_HIP(gr_launch_gop_fill(g, t, n_ang, pl->prm.ltbl, pl->d_gop_n, pl->d_gop_off, pl->d_gop, pl->d_hmeta,
                              pl->d_hseg, kGopSegGroups, s));
    GR_HIP(hipStreamSynchronize(s));
    pl->gop_state = 1;
    pl->gop_build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return GR_OK;
}

// The gridding tile list: kGridSectors sectors of the polar angle of the
// tile centre (seen from the grid centre), each ordered by the squared
// distance of the tile centre from the grid centre (exact integers in units of
// 1/4 cell) and then row-major, padded with kGridNoTile to the longest
// sector's length.  Entries are tx | ty << 16.
std::vector<unsigned> sector_tiles(int nt, int ts, long M) {
    const long c2 = 2 * (M >> 1), h2 = ts - 1;
    std::vector<std::vector<unsigned long long>> sec(kGridSectors);
    for (int ty = 0; ty < nt; ty++)
        for (int tx = 0; tx < nt; tx++) {
            const long x = 2L * tx * ts + h2 - c2, y = 2L * ty * ts + h2 - c2;
            double phi = atan2((double)y, (double)x);
            if (phi < 0) phi += 2 * M_PI;
            const int sx = std::min(kGridSectors - 1, (int)(phi / (2 * M_PI / kGridSectors)));
            const size_t i = (size_t)ty * nt + tx;
            sec[sx].push_back(((unsigned long long)(x * x + y * y) << 24) | i);
        }
    size_t len = 0;
    for (auto &v : sec) {
        std::sort(v.begin(), v.end());
        len = std::max(len, v.size());
    }
    std::vector<unsigned> out(kGridSectors * len, kGridNoTile);
    for (int sx = 0; sx < kGridSectors; sx++)
        for (size_t k = 0; k < sec[sx].size(); k++) {
            const unsigned idx = (unsigned)(sec[sx][k] & 0xffffff);
            out[sx * len + k] = (idx % (unsigned)nt) | ((idx / (unsigned)nt) << 16);
        }
    return out;
}

// The host-row entry points' staging slots, streams and events; every handle
// is reset, so a partly built set can be freed and built again.
void free_staging(gr_plan *pl) {
    for (int b = 0; b < 2; b++) {
        if (pl->h_in[b]) (void)hipHostFree(pl->h_in[b]);
        if (pl->h_out[b]) (void)hipHostFree(pl->h_out[b]);
        if (pl->d_in[b]) (void)hipFree(pl->d_in[b]);
        if (pl->d_out[b]) (void)hipFree(pl->d_out[b]);
        for (hipEvent_t e : {pl->ev_up[b], pl->ev_done[b], pl->ev_down[b]})
            if (e) (void)hipEventDestroy(e);
        pl->h_in[b] = pl->h_out[b] = pl->d_in[b] = pl->d_out[b] = nullptr;
        pl->ev_up[b] = pl->ev_done[b] = pl->ev_down[b] = nullptr;
    }
    if (pl->up_stream) (void)hipStreamDestroy(pl->up_stream);
    if (pl->down_stream) (void)hipStreamDestroy(pl->down_stream);
    pl->up_stream = pl->down_stream = nullptr;
    pl->stage_pairs = 0;
}

void free_plan(gr_plan *pl) {
    if (!pl) return;
    free_staging(pl);
    void *ptrs[] = {pl->d_cose, pl->d_sine, pl->d_roi, pl->d_wtbl, pl->d_winv, pl->d_filphase, pl->d_akey,
                    pl->d_aord, pl->d_ang, pl->d_gop, pl->d_gop_n, pl->d_gop_off, pl->d_dwtbl, pl->d_rowtiles,
                    pl->d_sop_tile, pl->d_sop_ch, pl->d_sop_keys, pl->d_sop_ent, pl->d_hmeta, pl->d_hseg,
                    pl->d_hsplit, pl->d_hitems};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    shared_release(pl->d_tw_row);
    shared_release(pl->d_tw_grid);
    shared_release(pl->d_tiles);
    if (pl->owns_workspace) {
        void *ws[] = {pl->d_spec1, pl->d_spec2, pl->d_ht, pl->d_z};
        for (void *p : ws)
            if (p) (void)hipFree(p);
    }
    for (auto &q : pl->pending) {
        (void)hipEventDestroy(q.e0);
        (void)hipEventDestroy(q.e1);
    }
    for (hipEvent_t e : pl->event_pool) (void)hipEventDestroy(e);
    if (pl->stream) (void)hipStreamDestroy(pl->stream);
    delete pl;
}

std::vector<float2> twiddles(long n, int sign) {
    std::vector<float2> t(n);
    for (long m = 0; m < n; m++) {
        const double a = 2.0 * M_PI * (double)m / (double)n;
        t[m] = make_float2((float)cos(a), (float)(sign * sin(a)));
    }
    return t;
}

// The twiddles of an n-point transform in the pass order gr_lds_fft reads them
// (gr_fft.hpp): the values of twiddles(n, sign), regrouped per pass, first
// for passes up to radix 16, then for passes up to radix 4.
std::vector<float2> pass_twiddles(long n, int sign) {
    const std::vector<float2> w = twiddles(n, sign);
    std::vector<float2> out;
    out.reserve(gr_twp_len(n, 16) + gr_twp_len(n, 4));
    for (int max_radix : {16, 4}) {
        long Ns = 1;
        while (max_radix >= 16 && Ns * 16 <= n) {
            for (int r = 1; Ns > 1 && r < 16; r++)
                for (long k = 0; k < Ns; k++) out.push_back(w[r * k * (n / (Ns * 16))]);
            Ns *= 16;
        }
        while (Ns * 4 <= n) {
            for (int r = 1; Ns > 1 && r < 4; r++)
                for (long k = 0; k < Ns; k++) out.push_back(w[r * k * (n / (Ns * 4))]);
            Ns *= 4;
        }
        for (long k = 0; Ns < n && k < Ns; k++) out.push_back(w[k]);
    }
    return out;
}

hipEvent_t pool_event(gr_plan *pl) {
    if (!pl->event_pool.empty()) {
        hipEvent_t e = pl->event_pool.back();
        pl->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

// Per-launch timing: an event pair around one kernel on its own stream; no
// synchronisation here, the pairs are resolved by gr_get_timing.
struct KernelTimer {
    gr_plan *pl;
    hipStream_t s;
    int kind;
    hipEvent_t e0 = nullptr;
    KernelTimer(gr_plan *p, hipStream_t st, int k) : pl(p), s(st), kind(k) {
        if (pl->timing) {
            e0 = pool_event(pl);
            (void)hipEventRecord(e0, s);
        }
    }
    void done() {
        if (!pl->timing) return;
        hipEvent_t e1 = pool_event(pl);
        (void)hipEventRecord(e1, s);
        pl->pending.push_back({kind, e0, e1});
    }
};

void resolve_timing(gr_plan *pl) {
    for (auto &q : pl->pending) {
        (void)hipEventSynchronize(q.e1);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, q.e0, q.e1);
        switch (q.kind) {
        case 0: pl->t.rowfft_ms += ms; pl->t.rowfft_launches++; break;
        case 1: pl->t.grid_ms += ms; pl->t.grid_launches++; break;
        case 2: pl->t.ifft_rows_ms += ms; pl->t.ifft_rows_launches++; break;
        default: pl->t.ifft_cols_ms += ms; pl->t.ifft_cols_launches++; break;
        }
        pl->event_pool.push_back(q.e0);
        pl->event_pool.push_back(q.e1);
    }
    pl->pending.clear();
}

}  // namespace

extern "C" {

const char *gr_last_error(void) { return g_err.c_str(); }
const char *gr_version(void) { return "gridrec-mi355x 0.1 (gfx950)"; }

int gr_get_pswf(float C, gr_pswf *out) {
    if (!out) return fail(GR_EINVAL, "null output");
    if (gr_find_pswf(C, out)) {
        char buf[96];
        snprintf(buf, sizeof buf, "Prolate parameter, C = %f not in data base", C);
        return fail(GR_ENOPSWF, buf);
    }
    return GR_OK;
}

int gr_filter_lookup(const char *name, char *name_io, int *filter_id) {
    int id = gr_filter_id_of(name);
    if (id < 0) {
        id = 0;  // filters.cpp:43-47: default shlo, name rewritten
        if (name_io) strcpy(name_io, "shlo");
    }
    if (filter_id) *filter_id = id;
    return GR_OK;
}

}  // extern "C"

int gr_set_error(int code, const char *msg) { return fail(code, msg); }

size_t gr_plan_workspace_bytes(const gr_plan *pl, int max_pairs) {
    const GrGeom &g = pl->g;
    const size_t spec_pp = (size_t)pl->prm.n_ang * g.half, ht_pp = (size_t)g.M * g.M, z_pp = (size_t)g.M * gr_z_cols(g.out_n);
    const size_t spec_pairs = (size_t)(max_pairs + kSpecGroup - 1) / kSpecGroup * kSpecGroup;
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    return 2 * up(sizeof(float2) * spec_pp * spec_pairs) + up(sizeof(float2) * ht_pp * max_pairs) +
           up(sizeof(float2) * z_pp * max_pairs);
}

int gr_plan_bind_workspace(gr_plan *pl, void *base, size_t bytes, int max_pairs) {
    if (!base || max_pairs < 1 || bytes < gr_plan_workspace_bytes(pl, max_pairs))
        return fail(GR_EINVAL, "workspace smaller than the plan needs for its pairs per launch");
    const GrGeom &g = pl->g;
    const size_t spec_pp = (size_t)pl->prm.n_ang * g.half, ht_pp = (size_t)g.M * g.M;
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t spec_pairs = (size_t)(max_pairs + kSpecGroup - 1) / kSpecGroup * kSpecGroup;
    char *p = (char *)base;
    pl->d_spec1 = (float2 *)p;
    p += up(sizeof(float2) * spec_pp * spec_pairs);
    pl->d_spec2 = (float2 *)p;
    p += up(sizeof(float2) * spec_pp * spec_pairs);
    pl->d_ht = (float2 *)p;
    p += up(sizeof(float2) * ht_pp * max_pairs);
    pl->d_z = (float2 *)p;
    pl->max_pairs = max_pairs;
    pl->owns_workspace = false;
    return GR_OK;
}

int gr_plan_build_operator(gr_plan *pl, hipStream_t s, bool required, size_t max_bytes) {
    if (pl->gop_state != 0) return GR_OK;
    return build_gop(pl, s, required, max_bytes);
}

extern "C" int gr_plan_create(const gr_params *params, gr_plan **plan_out) {
    return gr_plan_create_impl(params, plan_out, true);
}

int gr_plan_create_impl(const gr_params *params, gr_plan **plan_out, bool alloc_workspace) {
    if (!plan_out || !params) return fail(GR_EINVAL, "null argument");
    *plan_out = nullptr;
    const char *why = "";
    gr_plan *pl = new gr_plan();
    pl->prm = *params;
    int rc = gr_geom_from_params(params, &pl->g, &why);
    if (rc) {
        delete pl;
        return fail(rc, why);
    }
    const GrGeom &g = pl->g;
    const int n_ang = params->n_ang;
    if (params->geom == 0) {
        pl->angles.assign(params->angles, params->angles + n_ang);
        pl->prm.angles = pl->angles.data();
    } else {
        pl->prm.angles = nullptr;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        delete pl;
        return fail(GR_ENODEV, "no HIP device");
    }
    if (params->device < 0 || params->device >= ndev) {
        delete pl;
        return fail(GR_EINVAL, "device ordinal out of range");
    }
#define CK(x)                  \
    do {                       \
        rc = (x);              \
        if (rc) {              \
            free_plan(pl);     \
            return rc;         \
        }                      \
    } while (0)
    CK(hipSetDevice(params->device) == hipSuccess ? 0 : fail(GR_EHIP, "hipSetDevice failed"));

    // reference tables (host, bit-identical to recon_init)
    std::vector<float> sine(n_ang), cose(n_ang), roi(n_ang), filphase(2 * g.half), wtbl(params->ltbl + 1),
        winv(g.M0);
    std::vector<float> dwtbl(params->interp ? params->ltbl + 1 : 0);
    gr_host_tables(&pl->prm, g, sine.data(), cose.data(), filphase.data(), wtbl.data(), winv.data(), roi.data(),
                   params->interp ? dwtbl.data() : nullptr);
    CK(dev_upload(pl, &pl->d_sine, sine.data(), n_ang));
    CK(dev_upload(pl, &pl->d_cose, cose.data(), n_ang));
    CK(dev_upload(pl, &pl->d_roi, roi.data(), n_ang));
    CK(dev_upload(pl, &pl->d_filphase, (const float2 *)filphase.data(), g.half));
    CK(dev_upload(pl, &pl->d_wtbl, wtbl.data(), wtbl.size()));
    CK(dev_upload(pl, &pl->d_winv, winv.data(), winv.size()));
    if (params->interp) CK(dev_upload(pl, &pl->d_dwtbl, dwtbl.data(), dwtbl.size()));
    if (params->tables == GR_TABLES_FP64) {
        hipStream_t ts;
        CK(plan_stream(pl, &ts));
        hipError_t e = gr_launch_tables_fp64(g, &pl->prm, pl->d_filphase, pl->d_wtbl, pl->d_winv, ts);
        if (e == hipSuccess) e = hipStreamSynchronize(ts);
        if (e != hipSuccess) {
            free_plan(pl);
            return hip_fail(e, "FP64 tables");
        }
    }
    // twiddles (shared per size)
    CK(shared_table(params->device, kTwPlus, g.pdim, [&] { return pass_twiddles(g.pdim, +1); }, &pl->d_tw_row, nullptr));
    CK(shared_table(params->device, kTwMinus, g.M, [&] { return pass_twiddles(g.M, -1); }, &pl->d_tw_grid, nullptr));
    pl->dev_bytes += sizeof(float2) * (gr_twp_len(g.pdim, 16) + gr_twp_len(g.pdim, 4) + gr_twp_len(g.M, 16) +
                                       gr_twp_len(g.M, 4));
    // angle keys: direction of each sample line, atan2(SINE, COSE) mod pi, sorted
    {
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
        // keep the float keys sorted after rounding
        for (int i = 1; i < n_ang; i++) akey[i] = std::max(akey[i], akey[i - 1]);
        CK(dev_upload(pl, &pl->d_akey, akey.data(), n_ang));
        CK(dev_upload(pl, &pl->d_aord, ord.data(), n_ang));
        std::vector<float4> ang(n_ang);
        for (int i = 0; i < n_ang; i++) {
            const float c = cose[ord[i]], sn = sine[ord[i]];
            ang[i] = make_float4(c, sn, fabsf(c) > 1e-6f ? 1.0f / c : 0.0f, fabsf(sn) > 1e-6f ? 1.0f / sn : 0.0f);
        }
        CK(dev_upload(pl, &pl->d_ang, ang.data(), n_ang));
    }
    // tiles of the gridding kernel over the M x M grid (shared per size):
    // eight sectors of the polar angle of the tile centre, each listed nearest
    // to the centre first and padded to a common length with kNoTile; the
    // kernels give sector x to the workgroups that the hardware places on
    // XCD x, so each XCD's L2 holds the samples of its own sector
    {
        const int ts = gr_grid_tile_size();
        const int nt = (int)((g.M + ts - 1) / ts);
        if ((long)nt * nt >= (1L << 24)) CK(fail(GR_EINVAL, "grid too large for the tile list"));
        int slots = 0;
        CK(shared_table(params->device, kTiles, g.M, [&] { return sector_tiles(nt, ts, g.M); }, &pl->d_tiles,
                        &slots));
        pl->sector_len = slots / kGridSectors;
        pl->ntiles = slots;
        pl->dev_bytes += sizeof(unsigned) * (size_t)pl->ntiles;
        // the row transform's active-tile interval of every tile row
        CK(dev_alloc(pl, &pl->d_rowtiles, (size_t)nt));
        hipStream_t rs;
        CK(plan_stream(pl, &rs));
        hipError_t e = gr_launch_row_tiles(g, pl->d_rowtiles, rs);
        if (e == hipSuccess) e = hipStreamSynchronize(rs);
        if (e != hipSuccess) {
            free_plan(pl);
            return hip_fail(e, "row tile intervals");
        }
    }
    // workspace
    const size_t spec_pp = (size_t)n_ang * g.half, ht_pp = (size_t)g.M * g.M, z_pp = (size_t)g.M * gr_z_cols(g.out_n);
    const size_t per_pair = sizeof(float2) * (2 * spec_pp + ht_pp + z_pp);
    int mp = params->max_pairs;
    if (mp <= 0) {
        // 32 pairs per launch -- four groups of the widest gridding kernel
        // (which reads each sample's geometry once for all 8 pairs of a
        // group), interleaved so that they read the gather operator's entries
        // through L2 once (at 1500 x 2048 on MI355X: 0.638 ms per 8 pairs
        // against 0.645 with two groups) -- when a quarter of the free device
        // memory (at most 8 GiB) holds them, fewer otherwise
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = (size_t)8 << 30;
        const size_t budget = std::min<size_t>(free_b / 4, (size_t)8 << 30);
        mp = (int)std::max<size_t>(1, std::min<size_t>(32, budget / per_pair));
    }
    pl->max_pairs = mp;
    if (alloc_workspace) {
        const size_t spec_pairs = (size_t)(mp + kSpecGroup - 1) / kSpecGroup * kSpecGroup;
        CK(dev_alloc(pl, &pl->d_spec1, spec_pp * spec_pairs));
        CK(dev_alloc(pl, &pl->d_spec2, spec_pp * spec_pairs));
        CK(dev_alloc(pl, &pl->d_ht, ht_pp * mp));
        CK(dev_alloc(pl, &pl->d_z, z_pp * mp));
    } else {
        pl->owns_workspace = false;
    }
    if (params->grid_op < GR_GRIDOP_AUTO || params->grid_op > GR_GRIDOP_ON) CK(fail(GR_EINVAL, "unknown grid_op"));
    if (params->grid_op == GR_GRIDOP_ON) {
        hipStream_t bs;
        CK(plan_stream(pl, &bs));
        CK(build_gop(pl, bs, true, SIZE_MAX));
    }
#undef CK
    *plan_out = pl;
    return GR_OK;
}

extern "C" {

int gr_plan_info(const gr_plan *pl, gr_info *info) {
    if (!pl || !info) return fail(GR_EINVAL, "null argument");
    const GrGeom &g = pl->g;
    info->pdim = g.pdim;
    info->M = g.M;
    info->M0 = g.M0;
    info->M02 = g.M02;
    info->sampl = g.sampl;
    info->L = g.L;
    info->scale = g.scale;
    info->n_det_eff = g.n_det_eff;
    info->center_eff = g.center_eff;
    info->out_off = g.out_off;
    info->out_n = g.out_n;
    info->max_pairs = pl->max_pairs;
    info->device_bytes = pl->dev_bytes;
    info->grid_op_bytes = pl->gop_bytes;
    info->grid_op_build_ms = pl->gop_build_ms;
    return GR_OK;
}

int gr_plan_tables(const gr_plan *pl, float *sine, float *cose, float *filphase, float *wtbl, float *winv) {
    if (!pl) return fail(GR_EINVAL, "null plan");
    GR_HIP(hipSetDevice(pl->prm.device));
    const int n_ang = pl->prm.n_ang;
    if (sine) GR_HIP(hipMemcpy(sine, pl->d_sine, sizeof(float) * n_ang, hipMemcpyDeviceToHost));
    if (cose) GR_HIP(hipMemcpy(cose, pl->d_cose, sizeof(float) * n_ang, hipMemcpyDeviceToHost));
    if (filphase) GR_HIP(hipMemcpy(filphase, pl->d_filphase, sizeof(float2) * pl->g.half, hipMemcpyDeviceToHost));
    if (wtbl) GR_HIP(hipMemcpy(wtbl, pl->d_wtbl, sizeof(float) * (pl->prm.ltbl + 1), hipMemcpyDeviceToHost));
    if (winv) GR_HIP(hipMemcpy(winv, pl->d_winv, sizeof(float) * pl->g.M0, hipMemcpyDeviceToHost));
    return GR_OK;
}

}  // extern "C"

namespace {

// Pairs after which a plan in GR_GRIDOP_AUTO builds its gather operator: two
// full launches of the widest gridding kernel, about where the build (two
// geometry passes) has paid for itself.
constexpr long kGopAutoPairs = 16;

// The chunk loop of gr_recon_device / gr_recon_device_raw: K1..K4 per chunk
// of max_pairs pairs; K1 reads float rows (pre null) or detector counts.
int recon_chunks(gr_plan *pl, const float *sino, long sino_slice_stride, long sino_row_stride, const GrPre *pre,
                 float *out, long out_slice_stride, long out_row_stride, int n_slices, void *stream) {
    if (!pl->d_spec1) return fail(GR_EINVAL, "plan has no workspace");
    const GrGeom &g = pl->g;
    if (out_row_stride < g.out_n || out_slice_stride < out_row_stride * g.out_n)
        return fail(GR_EINVAL, "image strides smaller than the window they hold");
    GR_HIP(hipSetDevice(pl->prm.device));
    hipStream_t s = (hipStream_t)stream;  // NULL: the null stream
    const int n_pairs = (n_slices + 1) / 2;
    if (pl->gop_state == 0 && pl->prm.grid_op == GR_GRIDOP_AUTO && pl->pairs_done + n_pairs >= kGopAutoPairs) {
        // not while a graph is being captured: the build synchronises
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        GR_HIP(hipStreamIsCapturing(s, &cs));
        if (cs == hipStreamCaptureStatusNone) {
            const int rc = build_gop(pl, s, false, SIZE_MAX);
            if (rc) return rc;
        }
    }
    pl->pairs_done += n_pairs;
    const GrDevTables t = pl->tables();
    for (int p0 = 0; p0 < n_pairs; p0 += pl->max_pairs) {
        const int np = std::min(pl->max_pairs, n_pairs - p0);
        {
            KernelTimer kt(pl, s, 0);
            GR_HIP(gr_launch_rowfft(g, t, pl->prm.n_ang, sino, sino_slice_stride, sino_row_stride, pre, n_slices, p0, np,
                                    pl->d_spec1, pl->d_spec2, s));
            kt.done();
        }
        {
            KernelTimer kt(pl, s, 1);
            GR_HIP(gr_launch_grid(g, t, pl->prm.n_ang, pl->prm.ltbl, pl->d_spec1, pl->d_spec2, pl->d_ht, pl->d_z,
                                  np, pl->prm.accum, s));
            kt.done();
        }
        {
            KernelTimer kt(pl, s, 2);
            GR_HIP(gr_launch_ifft_rows(g, t, pl->d_ht, pl->d_z, np, s));
            kt.done();
        }
        {
            KernelTimer kt(pl, s, 3);
            GR_HIP(gr_launch_ifft_cols(g, t, pl->d_z, out, out_slice_stride, out_row_stride, n_slices, p0, np,
                                       pl->cols_per_block, s));
            kt.done();
        }
        if (pl->timing) pl->t.pairs += np;
    }
    return GR_OK;
}

// Run fn(lo, hi) over [0, n) split across up to 16 host threads.
template <typename F>
void parallel_rows(long n, F &&fn) {
    const long nt = std::max(1L, std::min<long>(16, std::min<long>((long)std::thread::hardware_concurrency(), n / 64)));
    if (nt <= 1) {
        fn(0L, n);
        return;
    }
    std::vector<std::thread> th;
    for (long t = 0; t < nt; t++) th.emplace_back([&, t] { fn(n * t / nt, n * (t + 1) / nt); });
    for (auto &x : th) x.join();
}

// Host rows in, host rows out, for npairs pairs, in chunks of stage_pairs
// pairs through two staging slots: while the GPU uploads, reconstructs and
// downloads chunk c (upload stream, compute stream `cs`, download stream,
// ordered by events; uploads and downloads use the two DMA directions at
// once), the host unpacks chunk c-1's images and packs chunk c+1's rows into
// the other slot.  Synchronous on return.
int exec_pairs(gr_plan *pl, const float *const *G1, const float *const *G2, float *const *S1, float *const *S2,
               int npairs, hipStream_t cs) {
    const GrGeom &g = pl->g;
    const int n_ang = pl->prm.n_ang, nd = g.n_det_in;
    const long on = g.out_n;
    const size_t in_pp = (size_t)2 * n_ang * nd, out_pp = (size_t)2 * on * on;  // floats per pair
    if (!pl->stage_pairs) {
        // chunks of at most 8 pairs: the host's packing and PCIe, not the
        // kernels, bound this path, and more chunks overlap more of them
        // (the pipeline's fill and drain are one chunk each)
        const int sp = std::min(pl->max_pairs, 8);
        auto setup = [&]() -> int {
            for (int b = 0; b < 2; b++) {
                GR_HIP(hipHostMalloc((void **)&pl->h_in[b], sizeof(float) * in_pp * sp));
                GR_HIP(hipHostMalloc((void **)&pl->h_out[b], sizeof(float) * out_pp * sp));
                int rc = dev_alloc(pl, &pl->d_in[b], in_pp * sp);
                if (rc) return rc;
                rc = dev_alloc(pl, &pl->d_out[b], out_pp * sp);
                if (rc) return rc;
                GR_HIP(hipEventCreateWithFlags(&pl->ev_up[b], hipEventDisableTiming));
                GR_HIP(hipEventCreateWithFlags(&pl->ev_done[b], hipEventDisableTiming));
                GR_HIP(hipEventCreateWithFlags(&pl->ev_down[b], hipEventDisableTiming));
            }
            GR_HIP(hipStreamCreateWithFlags(&pl->up_stream, hipStreamNonBlocking));
            GR_HIP(hipStreamCreateWithFlags(&pl->down_stream, hipStreamNonBlocking));
            return GR_OK;
        };
        const int rc = setup();
        if (rc) {  // a failure part-way frees what was built; the next call starts afresh
            free_staging(pl);
            return rc;
        }
        pl->stage_pairs = sp;
    }
    const int sp = pl->stage_pairs, nch = (npairs + sp - 1) / sp;
    hipStream_t up = pl->up_stream, down = pl->down_stream;
    auto pairs_of = [&](int c) { return std::min(sp, npairs - c * sp); };
    auto pack = [&](int c, int b) {  // slice 2q <- G1 rows of pair p0+q, slice 2q+1 <- G2 rows
        const long p0 = (long)c * sp, rows = 2L * pairs_of(c) * n_ang;
        float *dst = pl->h_in[b];
        parallel_rows(rows, [&](long lo, long hi) {
            for (long r = lo; r < hi; r++) {
                const long sl = r / n_ang, n = r - sl * n_ang, pair = p0 + sl / 2;
                const float *src = (sl & 1 ? G2 : G1)[pair * n_ang + n];
                memcpy(dst + r * nd, src, sizeof(float) * nd);
            }
        });
    };
    auto unpack = [&](int c, int b) {
        const long p0 = (long)c * sp, rows = 2L * pairs_of(c) * on;
        const float *src = pl->h_out[b];
        parallel_rows(rows, [&](long lo, long hi) {
            for (long r = lo; r < hi; r++) {
                const long sl = r / on, j = r - sl * on, pair = p0 + sl / 2;
                memcpy((sl & 1 ? S2 : S1)[pair * on + j], src + r * on, sizeof(float) * on);
            }
        });
    };
    // slot b's chunk c: upload (after chunk c-2's gridrec has read d_in[b]),
    // gridrec, download; chunk c-2's download, which d_out[b] also waits for,
    // the host has already waited for
    auto launch = [&](int c, int b) -> int {
        const int np = pairs_of(c);
        if (c >= 2) GR_HIP(hipStreamWaitEvent(up, pl->ev_done[b], 0));
        GR_HIP(hipMemcpyAsync(pl->d_in[b], pl->h_in[b], sizeof(float) * in_pp * np, hipMemcpyHostToDevice, up));
        GR_HIP(hipEventRecord(pl->ev_up[b], up));
        GR_HIP(hipStreamWaitEvent(cs, pl->ev_up[b], 0));
        const int rc = recon_chunks(pl, pl->d_in[b], (long)n_ang * nd, nd, nullptr, pl->d_out[b], on * on, on,
                                    2 * np, cs);
        if (rc) return rc;
        GR_HIP(hipEventRecord(pl->ev_done[b], cs));
        GR_HIP(hipStreamWaitEvent(down, pl->ev_done[b], 0));
        GR_HIP(hipMemcpyAsync(pl->h_out[b], pl->d_out[b], sizeof(float) * out_pp * np, hipMemcpyDeviceToHost, down));
        GR_HIP(hipEventRecord(pl->ev_down[b], down));
        return GR_OK;
    };
    for (int c = 0; c < nch; c++) {
        const int b = c & 1;
        if (c >= 2) {  // slot b last held chunk c-2: its download (and so its upload) is done
            GR_HIP(hipEventSynchronize(pl->ev_down[b]));
            unpack(c - 2, b);
        }
        pack(c, b);
        const int rc = launch(c, b);
        if (rc) return rc;
    }
    for (int c = std::max(0, nch - 2); c < nch; c++) {
        GR_HIP(hipEventSynchronize(pl->ev_down[c & 1]));
        unpack(c, c & 1);
    }
    return GR_OK;
}

}  // namespace

extern "C" {

int gr_recon_device(gr_plan *pl, const float *sino, long sino_slice_stride, long sino_row_stride, float *out,
                    long out_slice_stride, long out_row_stride, int n_slices, void *stream) {
    if (!pl || !sino || !out) return fail(GR_EINVAL, "null argument");
    if (n_slices < 0) return fail(GR_EINVAL, "negative slice count");
    if (n_slices == 0) return GR_OK;
    if (sino_row_stride < pl->g.n_det_in || sino_slice_stride < sino_row_stride * (long)pl->prm.n_ang)
        return fail(GR_EINVAL, "sinogram strides smaller than the rows they hold");
    return recon_chunks(pl, sino, sino_slice_stride, sino_row_stride, nullptr, out, out_slice_stride, out_row_stride,
                        n_slices, stream);
}

int gr_ring_correct(float *sino, long slice_stride, long row_stride, int n_slices, int n_ang, int n_det,
                    float ring_coeff, int iterations, void *stream) {
    if (!sino) return fail(GR_EINVAL, "null argument");
    if (n_slices < 1 || n_ang < 1 || n_det < 1 || iterations < 0 || iterations > 1000 || n_ang > 65535)
        return fail(GR_EINVAL, "invalid ring-correction sizes");
    if (row_stride < n_det || slice_stride < row_stride * (long)n_ang)
        return fail(GR_EINVAL, "sinogram strides smaller than the rows they hold");
    if (!(fabsf(ring_coeff) < 1e30f)) return fail(GR_EINVAL, "ring coefficient not finite");
    if (iterations == 0) return GR_OK;
    hipStream_t s = (hipStream_t)stream;
    float *rm = nullptr, *cm = nullptr;
    hipError_t e = hipMallocAsync((void **)&rm, sizeof(float) * (size_t)n_slices * n_ang, s);
    if (e == hipSuccess) e = hipMallocAsync((void **)&cm, sizeof(float) * (size_t)n_slices * n_det, s);
    if (e == hipSuccess)
        e = gr_launch_ring(sino, slice_stride, row_stride, n_slices, n_ang, n_det, ring_coeff, iterations, rm, cm, s);
    if (rm) (void)hipFreeAsync(rm, s);
    if (cm) (void)hipFreeAsync(cm, s);
    if (e != hipSuccess) return hip_fail(e, "ring correction");
    return GR_OK;
}

int gr_center_entropy(gr_plan *pl, const float *sino, long sino_slice_stride, long sino_row_stride,
                      const float *shifts, int n_shifts, float hist_min, float hist_max, int n_bins, float *entropy,
                      void *stream) {
    if (!pl || !sino || !shifts || !entropy) return fail(GR_EINVAL, "null argument");
    if (n_shifts < 1 || n_bins < 1 || n_bins > 16383 || !(hist_max > hist_min))
        return fail(GR_EINVAL, "need n_shifts >= 1, 1 <= n_bins <= 16383 and hist_max > hist_min");
    for (int i = 0; i < n_shifts; i++)
        if (!(fabsf(shifts[i]) < 1e6f)) return fail(GR_EINVAL, "shift out of range");
    const GrGeom &g = pl->g;
    if (sino_row_stride < g.n_det_in || sino_slice_stride < sino_row_stride * (long)pl->prm.n_ang)
        return fail(GR_EINVAL, "sinogram strides smaller than the rows they hold");
    GR_HIP(hipSetDevice(pl->prm.device));
    hipStream_t s = (hipStream_t)stream;
    const long on = g.out_n, nb1 = n_bins + 1;
    float *d_shift = nullptr, *d_img = nullptr;
    unsigned *d_hist = nullptr;
    hipError_t e = hipMallocAsync((void **)&d_shift, sizeof(float) * n_shifts, s);
    if (e == hipSuccess) e = hipMallocAsync((void **)&d_img, sizeof(float) * 2 * on * on * n_shifts, s);
    if (e == hipSuccess) e = hipMallocAsync((void **)&d_hist, sizeof(unsigned) * nb1 * n_shifts, s);
    if (e == hipSuccess) e = hipMemcpyAsync(d_shift, shifts, sizeof(float) * n_shifts, hipMemcpyHostToDevice, s);
    int rc = GR_OK;
    if (e == hipSuccess) {
        GrPre p{3, nullptr, 0, 0, nullptr, nullptr, nullptr, 0, 0.0f, d_shift};
        rc = recon_chunks(pl, sino, sino_slice_stride, sino_row_stride, &p, d_img, on * on, on, 2 * n_shifts, s);
    }
    std::vector<unsigned> h((size_t)nb1 * n_shifts);
    if (e == hipSuccess && rc == GR_OK)
        e = gr_launch_hist(d_img, 2 * on * on, on * on, on, (int)on, n_shifts, hist_min, hist_max, n_bins, d_hist, s);
    if (e == hipSuccess && rc == GR_OK)
        e = hipMemcpyAsync(h.data(), d_hist, sizeof(unsigned) * h.size(), hipMemcpyDeviceToHost, s);
    if (d_shift) (void)hipFreeAsync(d_shift, s);
    if (d_img) (void)hipFreeAsync(d_img, s);
    if (d_hist) (void)hipFreeAsync(d_hist, s);
    const hipError_t w = hipStreamSynchronize(s);
    if (rc) return rc;
    if (e != hipSuccess) return hip_fail(e, "centre sweep");
    if (w != hipSuccess) return hip_fail(w, "centre sweep sync");
    // entropy in bits with the reference's arithmetic (find_shift.cpp:2101-2108):
    // p in float, -p log p in float, divided by log 2 in double, summed into a float
    const float npix = (float)(2 * on * on);
    for (int c = 0; c < n_shifts; c++) {
        float ent = 0.0f;
        for (long b = 0; b < nb1; b++) {
            const unsigned k = h[(size_t)c * nb1 + b];
            if (!k) continue;
            const float pi = 1.0f * (float)k / npix;
            ent = (float)((double)ent + (double)(-pi * logf(pi)) / log(2.0));
        }
        entropy[c] = ent;
    }
    return GR_OK;
}

int gr_recon_device_raw(gr_plan *pl, const gr_preproc *pre, float *out, long out_slice_stride, long out_row_stride,
                        int n_slices, void *stream) {
    if (!pl || !pre || !out || !pre->raw || !pre->dark || !pre->white1) return fail(GR_EINVAL, "null argument");
    if (pre->mode != GR_PRE_FLATDARK && pre->mode != GR_PRE_FLATDARK_LOG)
        return fail(GR_EINVAL, "unknown preprocessing mode");
    if (n_slices < 0) return fail(GR_EINVAL, "negative slice count");
    if (n_slices == 0) return GR_OK;
    if (!(fabsf(pre->xoffset) < 1e6f)) return fail(GR_EINVAL, "xoffset out of range");
    if (pre->raw_row_stride < pl->g.n_det_in || pre->raw_slice_stride < pre->raw_row_stride * (long)pl->prm.n_ang)
        return fail(GR_EINVAL, "count strides smaller than the rows they hold");
    if (pre->field_stride < pl->g.n_det_in) return fail(GR_EINVAL, "field stride smaller than a row");
    GrPre p;
    p.mode = pre->mode;
    p.raw = pre->raw;
    p.raw_slice_stride = pre->raw_slice_stride;
    p.raw_row_stride = pre->raw_row_stride;
    p.dark = pre->dark;
    p.white1 = pre->white1;
    p.white2 = pre->white2 ? pre->white2 : pre->white1;
    p.field_stride = pre->field_stride;
    p.xoffset = pre->xoffset;
    p.pair_xoffset = nullptr;
    return recon_chunks(pl, nullptr, 0, 0, &p, out, out_slice_stride, out_row_stride, n_slices, stream);
}

int gr_recon_host(gr_plan *pl, const float *const *G1, const float *const *G2, float *const *S1,
                  float *const *S2) {
    if (!pl || !G1 || !G2 || !S1 || !S2) return fail(GR_EINVAL, "null argument");
    GR_HIP(hipSetDevice(pl->prm.device));
    hipStream_t st;
    const int rc = plan_stream(pl, &st);
    if (rc) return rc;
    return exec_pairs(pl, G1, G2, S1, S2, 1, st);
}

int gr_exec_pairs(gr_plan *pl, const float *const *G1, const float *const *G2, float *const *S1,
                  float *const *S2, int npairs, void *stream) {
    if (!pl || !G1 || !G2 || !S1 || !S2) return fail(GR_EINVAL, "null argument");
    if (npairs < 0) return fail(GR_EINVAL, "negative pair count");
    if (npairs == 0) return GR_OK;
    GR_HIP(hipSetDevice(pl->prm.device));
    return exec_pairs(pl, G1, G2, S1, S2, npairs, (hipStream_t)stream);
}

int gr_plan_build_grid_operator(gr_plan *pl, void *stream) {
    if (!pl) return fail(GR_EINVAL, "null plan");
    GR_HIP(hipSetDevice(pl->prm.device));
    if (pl->gop_state != 0) return GR_OK;
    return build_gop(pl, (hipStream_t)stream, true, SIZE_MAX);
}

int gr_plan_set_cols_per_block(gr_plan *pl, int cols) {
    if (!pl) return fail(GR_EINVAL, "null plan");
    if (cols < 0 || (cols > 0 && !gr_cols_width_ok(pl->g.M, cols))) return fail(GR_EINVAL, "unsupported column width");
    pl->cols_per_block = cols;
    return GR_OK;
}

int gr_plan_destroy(gr_plan *pl) {
    if (!pl) return GR_OK;
    if (pl->stream) (void)hipStreamSynchronize(pl->stream);
    free_plan(pl);
    return GR_OK;
}

int gr_set_timing(gr_plan *pl, int enable) {
    if (!pl) return fail(GR_EINVAL, "null plan");
    pl->timing = enable ? 1 : 0;
    return GR_OK;
}

int gr_get_timing(gr_plan *pl, gr_timing *out, int reset) {
    if (!pl || !out) return fail(GR_EINVAL, "null argument");
    GR_HIP(hipSetDevice(pl->prm.device));
    resolve_timing(pl);
    *out = pl->t;
    if (reset) pl->t = gr_timing{};
    return GR_OK;
}

int gr_phantom_sinograms(float *sino, long slice_stride, int n_slices, long slice0, unsigned long long seed,
                         int n_ang, int n_det, const float *centers, int n_ell, float photons, int density,
                         float gauss_rel, int n_tex, float tex_amp, float tex_fmax, void *stream) {
    if (!sino || !centers || n_slices < 1 || n_ang < 1 || n_det < 1 || n_ell < 1 || n_ell > 64 || n_tex < 0 ||
        n_tex > 256)
        return fail(GR_EINVAL, "invalid phantom arguments (1 <= n_ell <= 64, 0 <= n_tex <= 256)");
    if (slice_stride < (long)n_ang * n_det) return fail(GR_EINVAL, "slice stride too small");
    hipStream_t s = (hipStream_t)stream;
    float *dc = nullptr;
    GR_HIP(hipMallocAsync((void **)&dc, sizeof(float) * n_slices, s));
    GR_HIP(hipMemcpyAsync(dc, centers, sizeof(float) * n_slices, hipMemcpyHostToDevice, s));
    hipError_t e = gr_launch_phantom(sino, slice_stride, n_slices, slice0, seed, n_ang, n_det, dc, n_ell, photons,
                                     density, gauss_rel, n_tex, tex_amp, tex_fmax, s);
    hipError_t f = hipFreeAsync(dc, s);
    hipError_t w = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "phantom kernels");
    if (f != hipSuccess) return hip_fail(f, "hipFreeAsync");
    if (w != hipSuccess) return hip_fail(w, "phantom sync");
    return GR_OK;
}

}  // extern "C"
