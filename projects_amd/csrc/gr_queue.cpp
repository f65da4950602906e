// gr_queue.cpp -- batched gridrec over many datasets of uneven geometry.
//
// The reference reconstructs one slice pair per do_recon call (its drivers
// loop over pairs and over datasets: Multitomo_unit.txt:823-856,
// tomo_recon.cpp:3700-3863).  A queue takes a whole batch of datasets -- each
// with its own detector width, angle count, centre and slice count, all
// resident in HBM -- and turns it into one hipGraph:
//
//  * one plan per dataset (its own tables), with workspaces shared per stream
//    instead of one per plan;
//  * an autotuner: for every size class (pdim, M, output window) it times, on
//    the class's largest dataset, the slice pairs per launch, the column-FFT
//    width, the largest FFT radix and the gridding workgroup size (coordinate
//    descent) and keeps the fastest;
//  * datasets are spread over several streams, heaviest first onto the least
//    loaded stream, and all their launches are captured into one graph whose
//    branches fork from and join back into the caller's stream, so small
//    datasets run side by side and the host issues one launch per batch.
//
// A queue's results are bitwise those of per-dataset gr_recon_device calls
// where the tuner keeps the default FFT radix (16): the same kernels, and the
// pairs per launch, column width and gridding workgroup size change no sum
// order.  A class tuned to radix 4 computes its FFTs in another order and
// agrees with the per-dataset images to FP32 rounding.
#include <algorithm>
#include <map>
#include <numeric>
#include <string>
#include <tuple>
#include <vector>

#include "gr_plan.h"

struct gr_queue {
    int device = 0;
    std::vector<gr_dataset> ds;
    std::vector<gr_plan *> plans;
    std::vector<hipStream_t> streams;
    std::vector<void *> workspaces;
    std::vector<int> stream_of;
    std::vector<double> cost_ms;  // per dataset, measured while tuning
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<hipEvent_t> events;
    gr_queue_info info{};
};

namespace {

#define QHIP(call)                                                                  \
    do {                                                                            \
        hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess) return gr_set_error(GR_EHIP, hipGetErrorString(e_)); \
    } while (0)

void free_queue(gr_queue *q) {
    if (!q) return;
    if (q->exec) (void)hipGraphExecDestroy(q->exec);
    if (q->graph) (void)hipGraphDestroy(q->graph);
    for (gr_plan *p : q->plans)
        if (p) gr_plan_destroy(p);
    for (void *w : q->workspaces)
        if (w) (void)hipFree(w);
    for (hipEvent_t e : q->events) (void)hipEventDestroy(e);
    for (hipStream_t s : q->streams) (void)hipStreamDestroy(s);
    delete q;
}

int run_dataset(gr_queue *q, int d, hipStream_t s) {
    const gr_dataset &x = q->ds[d];
    return gr_recon_device(q->plans[d], x.sino, x.sino_slice_stride, x.sino_row_stride, x.out, x.out_slice_stride,
                           x.out_row_stride, x.n_slices, s);
}

// Time one dataset with the plan's current settings (ms, after a warm-up run).
int time_dataset(gr_queue *q, int d, hipStream_t s, hipEvent_t e0, hipEvent_t e1, float *ms) {
    int rc = run_dataset(q, d, s);
    if (rc) return rc;
    QHIP(hipEventRecord(e0, s));
    rc = run_dataset(q, d, s);
    if (rc) return rc;
    QHIP(hipEventRecord(e1, s));
    QHIP(hipEventSynchronize(e1));
    QHIP(hipEventElapsedTime(ms, e0, e1));
    return GR_OK;
}

}  // namespace

extern "C" {

int gr_queue_create(const gr_dataset *datasets, int n_datasets, int device, int n_streams, int autotune,
                    gr_queue **out) {
    if (!out || !datasets || n_datasets < 1) return gr_set_error(GR_EINVAL, "invalid queue arguments");
    *out = nullptr;
    n_streams = std::max(1, std::min(n_streams, 8));
    gr_queue *q = new gr_queue();
    q->device = device;
    q->ds.assign(datasets, datasets + n_datasets);
    int rc;
#define QHIP_Q(call)                                            \
    do {                                                        \
        hipError_t e_ = (call);                                 \
        if (e_ != hipSuccess) {                                 \
            (void)hipFree(tune_ws);                             \
            free_queue(q);                                      \
            return gr_set_error(GR_EHIP, hipGetErrorString(e_)); \
        }                                                       \
    } while (0)
#define QCK(x)              \
    do {                    \
        rc = (x);           \
        if (rc) {           \
            free_queue(q);  \
            return rc;      \
        }                   \
    } while (0)
    if (hipSetDevice(device) != hipSuccess) {
        delete q;
        return gr_set_error(GR_ENODEV, "hipSetDevice failed");
    }
    // plans without workspaces
    q->plans.assign(n_datasets, nullptr);
    std::vector<char> queue_op(n_datasets, 0);  // GR_GRIDOP_AUTO: the queue decides below
    for (int d = 0; d < n_datasets; d++) {
        gr_params p = datasets[d].params;
        p.device = device;
        // the plan's own automatic build must not run: it synchronises, which a
        // captured graph cannot, and would change kernels between the tuner's
        // trials; the queue builds AUTO datasets' operators itself
        if (p.grid_op == GR_GRIDOP_AUTO) {
            p.grid_op = GR_GRIDOP_OFF;
            queue_op[d] = 1;
        }
        if (datasets[d].n_slices < 1 || !datasets[d].sino || !datasets[d].out) {
            free_queue(q);
            return gr_set_error(GR_EINVAL, "dataset without slices or buffers");
        }
        QCK(gr_plan_create_impl(&p, &q->plans[d], false));
    }
    for (int i = 0; i < n_streams; i++) {
        hipStream_t s;
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
            free_queue(q);
            return gr_set_error(GR_EHIP, "stream creation failed");
        }
        q->streams.push_back(s);
    }
    for (int i = 0; i < 2 + n_streams; i++) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) {
            free_queue(q);
            return gr_set_error(GR_EHIP, "event creation failed");
        }
        q->events.push_back(e);
    }
    hipStream_t s0 = q->streams[0];

    // size classes (pdim, M, output window); each is tuned on its largest dataset
    std::map<std::tuple<long, long, long>, std::vector<int>> classes;
    for (int d = 0; d < n_datasets; d++) {
        const GrGeom &g = q->plans[d]->g;
        classes[std::make_tuple(g.pdim, g.M, g.out_n)].push_back(d);
    }
    auto tuned_dataset = [&](const std::vector<int> &members) {
        int d0 = members[0];
        for (int d : members)
            if (q->ds[d].n_slices > q->ds[d0].n_slices) d0 = d;
        return d0;
    };

    // gather operators: a queue is built once and replayed, so a dataset's
    // operator pays off from the first replay on.  The dataset each class is
    // tuned on gets its operator first (the tuner times the kernels the class
    // will run), within three quarters of the device memory left after a
    // generous estimate of the workspaces; the others get theirs after
    // tuning, when the workspaces' sizes are known, within seven eighths of
    // what those leave -- in order of slices (the operator's benefit per byte
    // grows with the pairs it serves), then of size.
    std::vector<int> op_order(n_datasets);
    std::iota(op_order.begin(), op_order.end(), 0);
    std::stable_sort(op_order.begin(), op_order.end(), [&](int a, int b) {
        if (q->ds[a].n_slices != q->ds[b].n_slices) return q->ds[a].n_slices > q->ds[b].n_slices;
        const GrGeom &ga = q->plans[a]->g, &gb = q->plans[b]->g;
        return (long)q->plans[a]->prm.n_ang * ga.half > (long)q->plans[b]->prm.n_ang * gb.half;
    });
    auto build_operators = [&](const std::vector<char> &want, size_t budget) {
        for (int d : op_order) {
            if (!want[d] || !queue_op[d] || q->plans[d]->gop_state != 0 || budget == 0) continue;
            const int brc = gr_plan_build_operator(q->plans[d], s0, false, budget);
            if (brc) return brc;
            const size_t b = q->plans[d]->gop_state == 1 ? q->plans[d]->gop_bytes : 0;
            budget = b < budget ? budget - b : 0;
        }
        return (int)GR_OK;
    };
    {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 0;
        size_t ws_est = 0;
        for (int d = 0; d < n_datasets; d++)
            ws_est = std::max(ws_est, gr_plan_workspace_bytes(q->plans[d], std::min((q->ds[d].n_slices + 1) / 2, 8)));
        ws_est *= (size_t)(n_streams + 1);
        std::vector<char> rep(n_datasets, 0);
        for (auto &kv : classes) rep[tuned_dataset(kv.second)] = 1;
        QCK(build_operators(rep, free_b > ws_est ? (free_b - ws_est) / 4 * 3 : 0));
    }

    // the autotuner
    const int kPairChoices[] = {1, 2, 4, 8};
    size_t tune_bytes = 0;
    for (int d = 0; d < n_datasets; d++) {
        const int pairs = (q->ds[d].n_slices + 1) / 2;
        tune_bytes = std::max(tune_bytes, gr_plan_workspace_bytes(q->plans[d], std::min(pairs, 8)));
    }
    void *tune_ws = nullptr;
    if (hipMalloc(&tune_ws, tune_bytes) != hipSuccess) {
        free_queue(q);
        return gr_set_error(GR_ENOMEM, "autotune workspace allocation failed");
    }
    q->cost_ms.assign(n_datasets, 0.0);
    int tuned = 0;
    for (auto &kv : classes) {
        const int d0 = tuned_dataset(kv.second);
        gr_plan *p = q->plans[d0];
        const int pairs = (q->ds[d0].n_slices + 1) / 2;
        int best_mp = std::min(pairs, 8), best_cw = 0, best_radix = 16, best_wpb = 4;
        float best_ms = 0.0f;
        // one timing of the class's dataset with a candidate setting
        auto trial = [&](int mp, int cw, int radix, int wpb, float *ms) {
            // mp <= min(pairs, 8): tune_bytes covers it
            int brc = gr_plan_bind_workspace(p, tune_ws, tune_bytes, mp);
            if (brc) return brc;
            p->cols_per_block = cw;
            p->g.fft_radix = radix;
            p->g.grid_wpb = wpb;
            tuned++;
            return time_dataset(q, d0, s0, q->events[0], q->events[1], ms);
        };
        if (autotune) {
            // coordinate descent: pairs per launch, then the column-FFT width,
            // the largest FFT radix and the gridding workgroup size in turn,
            // each with the others at their best so far
            best_ms = 1e30f;
            const long M = p->g.M;
            std::vector<int> cws = {0};
            for (int c = 1; c <= 16 && c * M <= 16384; c *= 2) cws.push_back(c);
            const int radixes[] = {16, 4}, wpbs[] = {4, 2, 1};
            for (int stage = 0; stage < 4 && !rc; stage++) {
                std::vector<int> cand;
                if (stage == 0)
                    for (int mp : kPairChoices) {
                        if (mp <= pairs || mp == 1) cand.push_back(mp);
                    }
                if (stage == 1) cand = cws;
                if (stage == 2) cand.assign(std::begin(radixes), std::end(radixes));
                if (stage == 3) cand.assign(std::begin(wpbs), std::end(wpbs));
                for (int v : cand) {
                    int mp = best_mp, cw = best_cw, radix = best_radix, wpb = best_wpb;
                    (stage == 0 ? mp : stage == 1 ? cw : stage == 2 ? radix : wpb) = v;
                    float ms = 0.0f;
                    rc = trial(mp, cw, radix, wpb, &ms);
                    if (rc) break;
                    if (ms < best_ms) {
                        best_ms = ms;
                        best_mp = mp;
                        best_cw = cw;
                        best_radix = radix;
                        best_wpb = wpb;
                    }
                }
            }
        } else {
            rc = trial(best_mp, best_cw, best_radix, best_wpb, &best_ms);
        }
        if (rc) {
            (void)hipFree(tune_ws);
            free_queue(q);
            return rc;
        }
        for (int d : kv.second) {
            // a dataset with fewer pairs than the tuned count launches them all
            // at once either way; its workspace holds only its own pairs
            q->plans[d]->max_pairs = std::min(best_mp, std::max(1, (q->ds[d].n_slices + 1) / 2));
            q->plans[d]->cols_per_block = best_cw;
            q->plans[d]->g.fft_radix = best_radix;
            q->plans[d]->g.grid_wpb = best_wpb;
        }
    }
    // one pass over every dataset with its tuned settings before capture: the
    // per-dataset costs for the stream balance, and every kernel instance the
    // graph will hold has been launched (and its LDS limit raised) once.  The
    // tuned pairs per launch of a class can exceed what the tuning buffer was
    // sized for on another member (more angles, fewer slices than the class's
    // tuned dataset), so the buffer is grown to the largest tuned workspace.
    {
        size_t pass_bytes = 0;
        for (int d = 0; d < n_datasets; d++)
            pass_bytes = std::max(pass_bytes, gr_plan_workspace_bytes(q->plans[d], q->plans[d]->max_pairs));
        // the other datasets' operators: the stream workspaces allocated below
        // hold at most n_streams times the largest tuned workspace, and the
        // tuning buffer is freed before them
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 0;
        const size_t avail = free_b + tune_bytes, need = (size_t)n_streams * pass_bytes;
        rc = build_operators(queue_op, avail > need ? (avail - need) / 8 * 7 : 0);
        if (rc) {
            (void)hipFree(tune_ws);
            free_queue(q);
            return rc;
        }
        for (int d = 0; d < n_datasets; d++) {
            if (q->plans[d]->gop_state != 1) continue;
            q->info.datasets_with_operator++;
            q->info.operator_bytes += q->plans[d]->gop_bytes;
        }
        if (pass_bytes > tune_bytes) {
            (void)hipFree(tune_ws);
            tune_ws = nullptr;
            if (hipMalloc(&tune_ws, pass_bytes) != hipSuccess) {
                free_queue(q);
                return gr_set_error(GR_ENOMEM, "queue workspace allocation failed");
            }
            tune_bytes = pass_bytes;
        }
    }
    for (int d = 0; d < n_datasets; d++) {
        rc = gr_plan_bind_workspace(q->plans[d], tune_ws, tune_bytes, q->plans[d]->max_pairs);
        if (rc) {
            (void)hipFree(tune_ws);
            free_queue(q);
            return rc;
        }
        float ms = 0.0f;
        QHIP_Q(hipEventRecord(q->events[0], s0));
        rc = run_dataset(q, d, s0);
        if (rc) {
            (void)hipFree(tune_ws);
            free_queue(q);
            return rc;
        }
        QHIP_Q(hipEventRecord(q->events[1], s0));
        QHIP_Q(hipEventSynchronize(q->events[1]));
        QHIP_Q(hipEventElapsedTime(&ms, q->events[0], q->events[1]));
        q->cost_ms[d] = ms;
    }
    (void)hipFree(tune_ws);

    // heaviest datasets first, each onto the least loaded stream
    std::vector<int> order(n_datasets);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return q->cost_ms[a] > q->cost_ms[b]; });
    std::vector<double> load(n_streams, 0.0);
    std::vector<size_t> ws_bytes(n_streams, 0);
    q->stream_of.assign(n_datasets, 0);
    for (int d : order) {
        const int s = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        q->stream_of[d] = s;
        load[s] += q->cost_ms[d];
        ws_bytes[s] = std::max(ws_bytes[s], gr_plan_workspace_bytes(q->plans[d], q->plans[d]->max_pairs));
    }
    q->workspaces.assign(n_streams, nullptr);
    for (int s = 0; s < n_streams; s++) {
        if (!ws_bytes[s]) continue;
        if (hipMalloc(&q->workspaces[s], ws_bytes[s]) != hipSuccess) {
            free_queue(q);
            return gr_set_error(GR_ENOMEM, "queue workspace allocation failed");
        }
    }
    for (int d = 0; d < n_datasets; d++)
        QCK(gr_plan_bind_workspace(q->plans[d], q->workspaces[q->stream_of[d]], ws_bytes[q->stream_of[d]],
                                   q->plans[d]->max_pairs));

    // one graph: fork from stream 0, every dataset on its stream, join
    if (hipStreamBeginCapture(s0, hipStreamCaptureModeRelaxed) != hipSuccess) {
        free_queue(q);
        return gr_set_error(GR_EHIP, "graph capture failed to start");
    }
    bool ok = hipEventRecord(q->events[0], s0) == hipSuccess;
    for (int s = 1; s < n_streams && ok; s++) ok = hipStreamWaitEvent(q->streams[s], q->events[0], 0) == hipSuccess;
    for (int d : order) {
        if (!ok) break;
        ok = run_dataset(q, d, q->streams[q->stream_of[d]]) == GR_OK;
    }
    for (int s = 1; s < n_streams && ok; s++) {
        ok = hipEventRecord(q->events[1 + s], q->streams[s]) == hipSuccess &&
             hipStreamWaitEvent(s0, q->events[1 + s], 0) == hipSuccess;
    }
    hipError_t e = hipStreamEndCapture(s0, &q->graph);
    if (!ok || e != hipSuccess) {
        free_queue(q);
        return gr_set_error(GR_EHIP, "graph capture failed");
    }
    if (hipGraphInstantiate(&q->exec, q->graph, nullptr, nullptr, 0) != hipSuccess) {
        free_queue(q);
        return gr_set_error(GR_EHIP, "graph instantiation failed");
    }
    size_t nodes = 0;
    (void)hipGraphGetNodes(q->graph, nullptr, &nodes);
    q->info.n_datasets = n_datasets;
    q->info.n_streams = n_streams;
    q->info.size_classes = (int)classes.size();
    q->info.tuned_candidates = tuned;
    q->info.graph_nodes = (long)nodes;
    size_t wsb = 0;
    for (size_t b : ws_bytes) wsb += b;
    q->info.workspace_bytes = wsb;
    long slices = 0;
    for (auto &x : q->ds) slices += x.n_slices;
    q->info.slices = slices;
#undef QCK
#undef QHIP_Q
    *out = q;
    return GR_OK;
}

int gr_queue_run(gr_queue *q, void *stream) {
    if (!q) return gr_set_error(GR_EINVAL, "null queue");
    QHIP(hipSetDevice(q->device));
    QHIP(hipGraphLaunch(q->exec, (hipStream_t)stream));
    return GR_OK;
}

int gr_queue_run_sequential(gr_queue *q, void *stream) {
    if (!q) return gr_set_error(GR_EINVAL, "null queue");
    QHIP(hipSetDevice(q->device));
    hipStream_t s = (hipStream_t)stream;
    // every dataset on one stream, one after the other, with the workspace of
    // the stream the graph gives it (one stream serialises them)
    for (int d = 0; d < (int)q->ds.size(); d++) {
        int rc = run_dataset(q, d, s);
        if (rc) return rc;
    }
    return GR_OK;
}

int gr_queue_get_info(const gr_queue *q, gr_queue_info *info) {
    if (!q || !info) return gr_set_error(GR_EINVAL, "null argument");
    *info = q->info;
    return GR_OK;
}

int gr_queue_dataset_settings(const gr_queue *q, int d, int *max_pairs, int *cols_per_block, int *fft_radix,
                              int *grid_wpb, int *stream) {
    if (!q || d < 0 || d >= (int)q->ds.size()) return gr_set_error(GR_EINVAL, "invalid dataset index");
    if (max_pairs) *max_pairs = q->plans[d]->max_pairs;
    if (cols_per_block) *cols_per_block = q->plans[d]->cols_per_block;
    if (fft_radix) *fft_radix = q->plans[d]->g.fft_radix;
    if (grid_wpb) *grid_wpb = q->plans[d]->g.grid_wpb;
    if (stream) *stream = q->stream_of[d];
    return GR_OK;
}

int gr_queue_destroy(gr_queue *q) {
    if (!q) return GR_OK;
    (void)hipSetDevice(q->device);
    for (hipStream_t s : q->streams) (void)hipStreamSynchronize(s);
    free_queue(q);
    return GR_OK;
}

}  // extern "C"
