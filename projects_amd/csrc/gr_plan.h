// gr_plan.h -- the plan object behind the C-ABI, shared by gr_capi.cpp (plans)
// and gr_queue.cpp (batched queues of plans).  Internal.
#pragma once

#include <vector>

#include "gr_internal.h"

struct gr_plan {
    gr_params prm;
    GrGeom g;
    std::vector<float> angles;  // copy of the caller's list (geom 0)
    hipStream_t stream = nullptr;
    // device tables
    float *d_cose = nullptr, *d_sine = nullptr, *d_roi = nullptr, *d_wtbl = nullptr, *d_winv = nullptr;
    float *d_dwtbl = nullptr;  // INTERP only
    float2 *d_filphase = nullptr, *d_tw_row = nullptr, *d_tw_grid = nullptr;
    float *d_akey = nullptr;
    int *d_aord = nullptr;
    float4 *d_ang = nullptr;
    unsigned *d_tiles = nullptr;
    int2 *d_rowtiles = nullptr;  // active-tile interval per tile row (K3)
    int ntiles = 0;      // slots: kGridSectors * sector_len
    int sector_len = 0;
    // K2's precomputed gather operator (GR_GRIDOP_*): entries, taps per tile,
    // first entry per tile; gop_state 0 = not built yet, 1 = built,
    // -1 = declined (too large), not tried again
    uint2 *d_gop = nullptr;
    unsigned *d_gop_n = nullptr, *d_gop_off = nullptr;
    size_t gop_bytes = 0;
    double gop_build_ms = 0.0;
    int gop_state = 0;
    int gop_aligned_tiles = 0;  // tiles in the aligned layout
    int gop_heavy = 0;          // leading slots of each sector gridded one pair per wave
    // its staged form for the other tiles (null when the plan has none)
    uint4 *d_sop_tile = nullptr;
    uint2 *d_sop_ch = nullptr;
    unsigned *d_sop_keys = nullptr, *d_sop_ent = nullptr;
    int sop_chunks = 0;
    // heavy work items and tiles gridded in segments (GrDevTables::hitems)
    uint2 *d_hitems = nullptr;
    uint4 *d_hmeta = nullptr;
    uint2 *d_hseg = nullptr;
    unsigned *d_hsplit = nullptr;
    int n_hitems = 0, n_split = 0, split_groups = 0;
    long pairs_done = 0;  // pairs reconstructed so far (the automatic build's trigger)
    // workspace for max_pairs pairs: owned, or a region bound by a queue
    int max_pairs = 1;
    bool owns_workspace = true;
    float2 *d_spec1 = nullptr, *d_spec2 = nullptr, *d_ht = nullptr, *d_z = nullptr;
    size_t dev_bytes = 0;
    // launch tuning (0 = automatic)
    int cols_per_block = 0;
    // host-row entry points (gr_exec_pairs, gr_recon_host): two staging slots
    // of stage_pairs pairs each, pinned on the host and on the device, an
    // upload and a download stream beside the compute stream, and per slot the
    // events that order upload -> gridrec -> download
    int stage_pairs = 0;
    float *h_in[2] = {nullptr, nullptr}, *h_out[2] = {nullptr, nullptr};
    float *d_in[2] = {nullptr, nullptr}, *d_out[2] = {nullptr, nullptr};
    hipStream_t up_stream = nullptr, down_stream = nullptr;  // host-to-device, device-to-host
    hipEvent_t ev_up[2] = {nullptr, nullptr}, ev_done[2] = {nullptr, nullptr}, ev_down[2] = {nullptr, nullptr};
    // timing: event pairs recorded around launches, resolved in gr_get_timing
    int timing = 0;
    gr_timing t{};
    struct Pending {
        int kind;  // 0 rowfft, 1 grid, 2 ifft rows, 3 ifft cols
        hipEvent_t e0, e1;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> event_pool;

    GrDevTables tables() const {
        GrDevTables t;
        t.cose = d_cose;
        t.sine = d_sine;
        t.roi_offset = d_roi;
        t.filphase = d_filphase;
        t.wtbl = d_wtbl;
        t.dwtbl = d_dwtbl;
        t.winv = d_winv;
        t.tw_row = d_tw_row;
        t.tw_grid = d_tw_grid;
        t.akey = d_akey;
        t.aord = d_aord;
        t.ang = d_ang;
        t.tiles = d_tiles;
        t.ntiles = ntiles;
        t.sector_len = sector_len;
        t.row_tiles = d_rowtiles;
        if (gop_state == 1) {
            t.gop = d_gop;
            t.gop_n = d_gop_n;
            t.gop_off = d_gop_off;
            t.gop_heavy = gop_heavy;
            t.hitems = d_hitems;
            t.n_hitems = n_hitems;
            t.hmeta = d_hmeta;
            t.hseg = d_hseg;
            t.hsplit = d_hsplit;
            t.n_split = n_split;
            t.split_groups = split_groups;
            t.sop_tile = d_sop_tile;
            t.sop_ch = d_sop_ch;
            t.sop_keys = d_sop_keys;
            t.sop_ent = d_sop_ent;
        }
        return t;
    }
};

// Create a plan; with alloc_workspace false the plan gets no workspace and
// must be bound to one (gr_plan_bind_workspace) before it runs.
int gr_plan_create_impl(const gr_params *params, gr_plan **plan, bool alloc_workspace);
// Workspace bytes for max_pairs slice pairs per launch (the spectra are held
// in whole groups of kSpecGroup pairs).
size_t gr_plan_workspace_bytes(const gr_plan *plan, int max_pairs);
// Point the plan's workspace at base (`bytes` long, 256-byte aligned), owned by
// the caller; GR_EINVAL when bytes < gr_plan_workspace_bytes(plan, max_pairs).
int gr_plan_bind_workspace(gr_plan *plan, void *base, size_t bytes, int max_pairs);
int gr_set_error(int code, const char *msg);
// Build the plan's gather operator now, on stream s (once: a plan that has
// one, or declined one, keeps that state).  With required false it declines
// quietly beyond max_bytes or a third of the free device memory.
int gr_plan_build_operator(gr_plan *plan, hipStream_t s, bool required, size_t max_bytes);
