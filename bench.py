"""Benchmark of the MI355X gridrec path: plain FP32 gridrec of a batch of slices.

Workload (one "step" = one pass of gridrec over the whole batch):
  2048 slices, each 1500 angles x 2048 detectors float32 (25.2 GB resident in
  HBM), seeded 20-ellipse phantoms projected analytically, Poisson noise at
  1e4 photons per cell and -log; rotation centre 1024 + U[-3,3] for the
  dataset; each row half-padded to 4096 (edge replication, generated on chip),
  PSWF C = 6, Shepp-Logan filter; output the 2048 x 2048 window of each
  4095 x 4095 image (slices processed in pairs, two per complex transform).

Prints one JSON line: slices/s over the timed steps, per-kernel roofline of the
dominant kernel from HIP events recorded around every launch in the timed
region, accuracy of one pair against the reference's gridrec (its FP32
restatement, and the reference build when present) and the FP64 oracle, and
the oracle's CPU throughput on a bounded sample on this host.

With --workload all (the default) the line also carries "batched": the uneven
queue (4096 slices in datasets of 1-16 slices with widths 512..2560, 180-1800
angles, centres n/2 + U[-5,5]) replayed as one HIP graph, against one plain
plan per dataset on the same data, with bitwise equality of the images checked.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "slices/s @2048² (1500 ang) plain vs batched vs learned; rel-L2 vs FP64, PSNR"


def progress(msg: str):
    """A progress line on stderr (stdout carries only the result line)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--slices", type=int, default=2048)
    ap.add_argument("--n-ang", type=int, default=1500)
    ap.add_argument("--n-det", type=int, default=2048)
    ap.add_argument("--seed", type=int, default=20260)
    ap.add_argument("--accum", default="compensated", choices=["compensated", "plain"])
    ap.add_argument("--grid-op", default="auto", choices=["auto", "off", "on"],
                    help="gridding from the precomputed gather operator (plain workload)")
    ap.add_argument("--cpu-pairs", type=int, default=0, help="pairs in the CPU baseline sample (0 = one per thread)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--h2h-slices", type=int, default=256,
                    help="slices timed host-to-host through gr_exec_pairs (0: skip)")
    ap.add_argument("--workload", default="all", choices=["all", "plain", "batched", "precision", "learned"],
                    help="all: the plain headline line with the batched queue, the precision study and the "
                         "learned sparse-angle path after it")
    ap.add_argument("--learned-pairs", type=int, default=20000, help="training pairs (512^2, 90 / 900 angles)")
    ap.add_argument("--learned-seconds", type=float, default=150.0, help="training time cap")
    ap.add_argument("--learned-steps", type=int, default=8000,
                    help="training steps (0 = time budget only); a fixed count makes the trained network reproducible")
    ap.add_argument("--learned-slices", type=int, default=256, help="inference slices (2048^2, 150 angles)")
    ap.add_argument("--learned-truth-slices", type=int, default=32,
                    help="inference slices also compared with the analytic phantom")
    ap.add_argument("--learned-threshold", type=float, default=0.5,
                    help="fallback when ||residual|| / ||image|| exceeds this")
    ap.add_argument("--precision-slices", type=int, default=256)
    ap.add_argument("--precision-fp64-pairs", type=int, default=0,
                    help="pairs checked against the FP64 oracle (0 = all)")
    ap.add_argument("--queue-slices", type=int, default=4096)
    ap.add_argument("--queue-streams", type=int, default=4)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 4 plain device copies of 1 GiB (reference for the HBM byte counters)")
    return ap.parse_args()


def copy_calibration():
    """A plain copy of known size: 1 GiB read + 1 GiB written per launch, the
    scale for FETCH_SIZE / WRITE_SIZE counter readings."""
    import torch

    src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(4):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_per_launch": 2 * src.numel() * 4, "launches": 4}), flush=True)


def host_cpu():
    """This host's CPU model (from /proc/cpuinfo), logical CPUs and those usable here."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "logical_cpus": os.cpu_count(), "usable_cpus": len(os.sched_getaffinity(0))}


def ref_worker(job):
    """Child process: the reference's own do_recon (oracle/_ref, one global
    context per process) on pairs of a saved batch; returns (slices, seconds)."""
    path, pairs, n_ang, n_det, center = job
    import oracle as O

    data = np.load(path, mmap_mode="r")
    g = O.Geometry(n_ang=n_ang, n_det=n_det, center=center).padded()
    t0 = time.perf_counter()
    for p in pairs:
        O.ref_recon(g, O.pad_half(data[2 * p]), O.pad_half(data[2 * p + 1]))
    return 2 * len(pairs), time.perf_counter() - t0


def cpu_baseline(a, O, sino, ns, nang, ndet, center, out_n):
    """The reference's gridrec timed on this host: single core (one process,
    2 pairs) and one process per usable core (up to 16, one pair each, run
    together; the rate is the sum of the processes' own rates), with the
    caller's half padding on the host as in tomo_recon.  Beside it the FP32
    restatement (oracle/liboracle.so) threaded over all usable cores."""
    import multiprocessing as mp
    import tempfile

    host = host_cpu()
    procs = max(1, min(16, host["usable_cpus"]))
    out = {"host": host}
    if O.ref_available():
        npairs = min(max(2, procs), ns // 2)
        with tempfile.TemporaryDirectory() as td:
            path = os.path.join(td, "sino.npy")
            np.save(path, sino[0:2 * npairs].cpu().numpy())
            ctx = mp.get_context("spawn")
            with ctx.Pool(1) as pool:
                n1, t1 = pool.map(ref_worker, [(path, [0, 1], nang, ndet, center)])[0]
            with ctx.Pool(procs) as pool:
                res = pool.map(ref_worker, [(path, [w % npairs], nang, ndet, center) for w in range(procs)])
        rate = sum(n / t for n, t in res)
        out.update({"value": round(rate, 3), "unit": "slices/s", "cores": procs, "kind": "reference",
                    "sample": f"{procs} processes x 1 pair ({nang} angles x {ndet} detectors, half-padded "
                              f"to {2 * ndet} on the host) of GridrecCPP 5.2 do_recon compiled from its sources, "
                              f"run together; {max(t for _, t in res):.1f} s",
                    "single_core": {"value": round(n1 / t1, 4), "unit": "slices/s",
                                    "sample": f"{n1} slices ({n1 // 2} pairs) in one process, {t1:.1f} s"}})
    threads = procs
    pairs = min(a.cpu_pairs or 4 * threads, ns // 2)   # ~5-10 s of CPU work at this size
    s = sino[0:2 * pairs].cpu().numpy()
    g = O.Geometry(n_ang=nang, n_det=ndet, center=center)
    t0 = time.perf_counter()
    O.recon_batch_f32(g, s, True, ndet // 2, out_n, threads=threads)
    cdt = time.perf_counter() - t0
    port = {"value": round(2 * pairs / cdt, 3), "unit": "slices/s", "cores": threads, "kind": "port",
            "sample": f"{2 * pairs} slices ({pairs} pairs) of the same workload, FP32 restatement of "
                      f"GridrecCPP, {threads} threads, {cdt:.1f} s"}
    if "value" in out:
        out["port_all_cores"] = port
    else:
        out.update(port)
    return out


def kernel_bytes(kind, n_ang, n_det, half, M, out_n, pairs):
    """Algorithmic HBM bytes of one launch covering `pairs` slice pairs:
    every input read once, every output written once."""
    spec = 2 * n_ang * half * 8          # Cdata1 + Cdata2, complex64
    if kind == "rowfft":
        b = 2 * n_ang * n_det * 4 + spec  # two unpadded float rows in, spectra out
    elif kind == "grid":
        b = spec + M * M * 8              # spectra in, complex grid out
    elif kind == "ifft_rows":
        b = M * M * 8 + M * out_n * 8     # grid in, pruned rows out
    else:
        b = M * out_n * 8 + 2 * out_n * out_n * 4  # pruned rows in, two float images out
    return b * pairs


QUEUE_WIDTHS = (512, 1024, 1536, 2048, 2560)


def queue_datasets(gr, torch, seed, total):
    """The uneven batch: datasets of 1-16 slices until `total` slices, each
    with detector width uniform over QUEUE_WIDTHS, angle count uniform in
    [180, 1800], centre n/2 + U[-5, 5], 20-ellipse phantoms with Poisson(1e4)
    noise; half-padded on chip, output the n-wide window (round(M0/2) pixels)."""
    rng = np.random.default_rng(seed)
    sets, done, d = [], 0, 0
    while done < total:
        ns = int(min(rng.integers(1, 17), total - done))
        n = int(rng.choice(QUEUE_WIDTHS))
        na = int(rng.integers(180, 1801))
        c = float(np.float32(n / 2 + rng.uniform(-5.0, 5.0)))
        sino = torch.empty((ns, na, n), dtype=torch.float32, device="cuda")
        gr.phantom_sinograms(sino, seed + 7919 * d, na, n, [c] * ns, 20, photons=1e4)
        kw = dict(n_ang=na, n_det=n, center=c, pad="half")
        M0 = (1 << (2 * n - 1).bit_length()) - 1  # M0 = pdim - 1 at sampl 1, pdim = 2^ceil(log2(2n))
        cn = int(round(n * M0 / (2 * n)))
        kw["crop"] = ((M0 - cn + 1) // 2, cn)
        out = torch.empty((ns, cn, cn), dtype=torch.float32, device="cuda")
        sets.append((kw, sino, out))
        done += ns
        d += 1
    torch.cuda.synchronize()
    return sets


def bit_checksums(torch, sets):
    """Per dataset: (sum of the images' bit patterns as int64, float64 sum)."""
    return [(int(o.view(torch.int32).sum(dtype=torch.int64)), float(o.sum(dtype=torch.float64))) for _, _, o in sets]


def bench_batched(a, gr, O, torch):
    """The batched uneven queue against per-dataset plain plans on the same
    datasets: execution time of both (HIP events), queue build time, bitwise
    equality of the images (checksums), one pair against the oracle."""
    progress("batched: generating the uneven datasets")
    t0 = time.time()
    sets = queue_datasets(gr, torch, a.seed + 1, a.queue_slices)
    gen_s = time.time() - t0
    n_slices = sum(x[1].shape[0] for x in sets)
    widths = {w: sum(x[1].shape[0] for x in sets if x[0]["n_det"] == w) for w in QUEUE_WIDTHS}

    # per-dataset plain launches: one plan per dataset (built outside the
    # timed region), recon_device on one stream
    stream = torch.cuda.Stream()  # a stream of its own: events and launches on one queue
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    # second baseline: the same plans after building their gather operators
    # (outside the timed region), as a caller that reuses a plan would
    plain_ms, plain_op_ms, setup_s, op_build_s = 0.0, 0.0, 0.0, 0.0
    for kw, sino, out in sets:
        ts = time.perf_counter()
        p = gr.Plan(**kw)
        torch.cuda.synchronize()
        setup_s += time.perf_counter() - ts
        e0.record(stream)
        p.recon(sino, out, stream=stream)
        e1.record(stream)
        e1.synchronize()
        plain_ms += e0.elapsed_time(e1)
        ts = time.perf_counter()
        p.build_operator(stream)
        op_build_s += time.perf_counter() - ts
        e0.record(stream)
        p.recon(sino, out, stream=stream)
        e1.record(stream)
        e1.synchronize()
        plain_op_ms += e0.elapsed_time(e1)
        p.close()
    ref_sums = bit_checksums(torch, sets)
    progress(f"batched: per-dataset plain plans done ({plain_ms:.0f} ms of kernels)")
    for _, _, o in sets:
        o.zero_()
    torch.cuda.synchronize()

    ts = time.perf_counter()
    q = gr.Queue(sets, n_streams=a.queue_streams, autotune=True)
    torch.cuda.synchronize()
    build_s = time.perf_counter() - ts
    for _ in range(a.warmup):
        q.run(stream)
    torch.cuda.synchronize()
    steps = max(1, a.steps)
    e0.record(stream)
    for _ in range(steps):
        q.run(stream)
    e1.record(stream)
    e1.synchronize()
    queue_ms = e0.elapsed_time(e1) / steps
    progress(f"batched: queue built in {build_s:.1f} s, {queue_ms:.0f} ms per batch")
    # bitwise where the tuner kept the plain path's FFT radix; elsewhere the FFT
    # sums in another order (accuracy: the oracle comparison below)
    qsums = bit_checksums(torch, sets)
    keep = [d for d in range(len(sets)) if q.settings(d)["fft_radix"] == 16]
    same = all(qsums[d] == ref_sums[d] for d in keep)

    # one pair of a small dataset against the oracle restatement and FP64
    acc = None
    if not a.no_accuracy:
        for kw, sino, out in sets:
            if sino.shape[0] >= 2 and kw["n_det"] <= 1024:
                g = O.Geometry(n_ang=kw["n_ang"], n_det=kw["n_det"], center=kw["center"]).padded()
                s = sino[0:2].cpu().numpy()
                o = out[0:2].cpu().numpy()
                p1, p2 = O.pad_half(s[0]), O.pad_half(s[1])
                off, cn = kw["crop"]
                c = slice(off, off + cn)
                r1, r2 = O.recon_f32(g, p1, p2)
                d1, d2 = O.recon_f64(g, p1, p2)
                acc = {"dataset": {"n_det": kw["n_det"], "n_ang": kw["n_ang"], "center": kw["center"]},
                       "rel_l2_vs_reference_restatement": float(f"{max(O.rel_l2(o[0], r1[c, c]), O.rel_l2(o[1], r2[c, c])):.3e}"),
                       "rel_l2_vs_fp64": float(f"{max(O.rel_l2(o[0], d1[c, c]), O.rel_l2(o[1], d2[c, c])):.3e}")}
                break
    # whole-pipeline algorithmic bytes of the batch (the plain line's formula
    # per slice pair: 8 n_ang n_det + 16 n_ang pdim + 32 M^2 + 8 n_out^2, with
    # pdim = M = the power of two >= the half-padded width)
    batch_bytes = 0
    for kw, sino, out in sets:
        pd = 1 << (2 * kw["n_det"] - 1).bit_length()
        pairs = (sino.shape[0] + 1) // 2
        batch_bytes += pairs * (8 * kw["n_ang"] * kw["n_det"] + 16 * kw["n_ang"] * pd + 32 * pd * pd +
                                8 * out.shape[1] * out.shape[1])
    info = q.info
    settings = {}
    for d, (kw, _, _) in enumerate(sets):
        st = q.settings(d)
        settings.setdefault(f"{kw['n_det']}", (st["max_pairs"], st["cols_per_block"], st["fft_radix"], st["grid_wpb"]))
    q.close()
    return {
        "workload": f"batched uneven queue: {n_slices} slices in {len(sets)} datasets of 1-16 slices, "
                    f"width in {list(QUEUE_WIDTHS)}, 180-1800 angles, centre n/2+U[-5,5], Poisson(1e4)",
        "value": round(n_slices / (queue_ms * 1e-3), 2), "unit": "slices/s",
        "per_dataset_plain_slices_s": round(n_slices / (plain_ms * 1e-3), 2),
        "speedup_vs_per_dataset_plain": round(plain_ms / queue_ms, 3),
        "per_dataset_plain_with_operator_slices_s": round(n_slices / (plain_op_ms * 1e-3), 2),
        "speedup_vs_per_dataset_plain_with_operator": round(plain_op_ms / queue_ms, 3),
        "per_dataset_operator_build_s": round(op_build_s, 2),
        "queue_datasets_with_operator": info["datasets_with_operator"],
        "queue_operator_GB": round(info["operator_bytes"] / 1e9, 2),
        "pipeline_alg_GBs": round(batch_bytes / (queue_ms * 1e-3) / 1e9, 1),
        "per_dataset_plain_pipeline_alg_GBs": round(batch_bytes / (plain_ms * 1e-3) / 1e9, 1),
        "queue_ms_per_batch": round(queue_ms, 2), "per_dataset_plain_ms": round(plain_ms, 2),
        "per_dataset_plan_setup_s": round(setup_s, 2), "queue_build_s": round(build_s, 2),
        "bitwise_equal_to_per_dataset_plain": bool(same),
        "datasets_compared_bitwise": len(keep),
        "slices_by_width": widths, "streams": info["n_streams"], "size_classes": info["size_classes"],
        "tuned_candidates": info["tuned_candidates"], "graph_nodes": info["graph_nodes"],
        "tuned_pairs_cols_radix_wpb_by_width": settings,
        "workspace_GB": round(info["workspace_bytes"] / 1e9, 2), "gen_s": round(gen_s, 2), "accuracy": acc,
    }


def bench_precision(a, gr, O, torch):
    """Precision study: 1800 angles x 2560 detectors, ellipse densities
    log-uniform over 4 decades, 1% multiplicative Gaussian noise, rows
    half-padded to 5120 (pdim = M = 8192, M0 = 8191), output the 4096^2
    window that covers the detector's field of view.  Four GPU variants -- compensated or plain tap accumulation, float
    tables built as the reference builds them or evaluated in FP64 -- each
    against the FP64 oracle gridrec per slice; plus, on pair 0, the error each
    FP32 stage contributes alone (the oracle's mixed-precision pipeline)."""
    from concurrent.futures import ThreadPoolExecutor

    ns, nang, ndet = a.precision_slices, 1800, 2560
    M0 = (1 << (2 * ndet - 1).bit_length()) - 1
    cn = int(round(ndet * M0 / (2 * ndet)))
    crop = ((M0 - cn + 1) // 2, cn)
    rng = np.random.default_rng(a.seed + 2)
    center = float(np.float32(ndet / 2 + rng.uniform(-3.0, 3.0)))
    sino = torch.empty((ns, nang, ndet), dtype=torch.float32, device="cuda")
    for s0 in range(0, ns, 64):
        k = min(64, ns - s0)
        gr.phantom_sinograms(sino[s0:s0 + k], a.seed + 2, nang, ndet, [center] * k, 20, photons=0.0, density=1,
                             gauss_rel=0.01, slice0=s0)
    torch.cuda.synchronize()
    stream = torch.cuda.Stream()
    variants = [("compensated", "reference"), ("plain", "reference"), ("compensated", "fp64"), ("plain", "fp64")]
    progress("precision: GPU variants")
    outs, rates = {}, {}
    for acc, tab in variants:
        p = gr.Plan(n_ang=nang, n_det=ndet, center=center, accum=acc, tables=tab, pad="half", crop=crop)
        out = torch.empty((ns, p.out_n, p.out_n), dtype=torch.float32, device="cuda")
        p.recon(sino, out, stream=stream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        p.recon(sino, out, stream=stream)
        torch.cuda.synchronize()
        rates[f"{acc}/{tab}"] = round(ns / (time.perf_counter() - t0), 1)
        outs[f"{acc}/{tab}"] = out  # stays in HBM
        p.close()
    s_host = sino.cpu().numpy()
    del sino
    torch.cuda.empty_cache()
    g = O.Geometry(n_ang=nang, n_det=ndet, center=center).padded()
    c = slice(crop[0], crop[0] + cn)
    yy, xx = torch.meshgrid(torch.arange(cn, device="cuda"), torch.arange(cn, device="cuda"), indexing="ij")
    disk = torch.hypot((yy - cn // 2).double(), (xx - cn // 2).double()) < 0.9 * cn / 2
    n_pairs = ns // 2 if a.precision_fp64_pairs <= 0 else min(ns // 2, a.precision_fp64_pairs)

    def errors(k, d):
        """(whole rel-L2, disk rel-L2, whole max-abs, disk max-abs) of slices 2k, 2k+1, per variant (GPU)."""
        res = {}
        dd = [torch.from_numpy(np.ascontiguousarray(x[c, c])).cuda() for x in d]
        for name, o in outs.items():
            r = []
            for j in (0, 1):
                e = o[2 * k + j].double() - dd[j]
                r.append((float(e.norm() / dd[j].norm()), float(e[disk].norm() / dd[j][disk].norm()),
                          float(e.abs().max()), float(e[disk].abs().max())))
            res[name] = r
        return res

    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    t0 = time.perf_counter()
    per_pair = []
    with ThreadPoolExecutor(threads) as ex:
        for k0 in range(0, n_pairs, threads):  # FP64 oracle on the host cores, a wave of pairs at a time
            ks = list(range(k0, min(n_pairs, k0 + threads)))
            ds = list(ex.map(lambda k: O.recon_f64(g, O.pad_half(s_host[2 * k]), O.pad_half(s_host[2 * k + 1])), ks))
            progress(f"precision: FP64 oracle pairs {ks[0]}-{ks[-1]} of {n_pairs}")
            per_pair += [errors(k, d) for k, d in zip(ks, ds)]
    oracle_s = time.perf_counter() - t0
    summary = {}
    for name in outs:
        arr = np.array([x for pp in per_pair for x in pp[name]])  # [slices, 4]
        summary[name] = {
            "slices_per_s": rates[name],
            "rel_l2_whole": {"median": float(f"{np.median(arr[:, 0]):.3e}"), "max": float(f"{arr[:, 0].max():.3e}")},
            "rel_l2_disk": {"median": float(f"{np.median(arr[:, 1]):.3e}"), "max": float(f"{arr[:, 1].max():.3e}")},
            "max_abs_whole": float(f"{arr[:, 2].max():.3e}"), "max_abs_disk": float(f"{arr[:, 3].max():.3e}"),
            "per_slice_rel_l2_disk": [float(f"{v:.3e}") for v in arr[:, 1]],
        }
    # error each stage contributes when it alone runs in FP32 (pair 0, disk)
    h0, h1 = O.pad_half(s_host[0]), O.pad_half(s_host[1])
    d0 = [x[c, c] for x in O.recon_f64(g, h0, h1)]
    stages = {"row FFT + filter": O.MX_ROWFFT, "gridding accumulation": O.MX_ACCUM, "2D inverse FFT": O.MX_FFT2,
              "filter/PSWF/deapodization tables": O.MX_TABLES, "deapodization product": O.MX_DEAPOD}
    budget = {}
    dk = disk.cpu().numpy()
    with ThreadPoolExecutor(len(stages)) as ex:
        mixed = list(ex.map(lambda bit: [x[c, c] for x in O.recon_mixed(g, h0, h1, O.MX_ALL & ~bit)],
                            stages.values()))
    for nm, (m1, m2) in zip(stages, mixed):
        budget[nm] = float(f"{max(O.rel_l2(m1[dk], d0[0][dk]), O.rel_l2(m2[dk], d0[1][dk])):.3e}")
    del outs
    torch.cuda.empty_cache()
    return {
        "workload": f"precision study: {ns} slices x {nang} angles x {ndet} det, ellipse densities log-uniform "
                    f"over 4 decades, 1% Gaussian noise, half-padded to {2 * ndet} (M {M0 + 1}), {cn}^2 window "
                    f"of {M0}^2 (the field of view)",
        "fp64_on_gpu": "filter coefficients (filphase), PSWF table wtbl, deapodization winv: evaluated in FP64 "
                       "and rounded to float in the fp64-tables variants; everything else FP32",
        "slices_vs_fp64_oracle": 2 * n_pairs, "oracle_s": round(oracle_s, 1), "oracle_threads": threads,
        "variants": summary,
        "fp32_stage_alone_rel_l2_disk_pair0": budget,
    }


def bench_learned(a, gr, torch):
    """Learned sparse-angle path.  Training: 512^2 seeded phantoms (20 ellipses
    + band-limited texture, noise-free), gridrec from 90 angles as input and
    from 900 angles as target, Adam (batch 32, learning rate on a cosine from
    1e-3 to 5e-5) for a fixed number of steps on the GPU (a time cap bounds
    the run).  Inference: 2048^2 slices from 150-angle sinograms, gridrec + the
    network (with the fallback), judged by PSNR / rel-L2 against 1500-angle
    plain gridrec of the same phantoms; both paths timed with HIP events."""
    from projects_amd import learned as Lrn

    res = {}
    progress("learned: generating training pairs")
    t0 = time.perf_counter()
    inp, tgt = Lrn.make_pairs(a.learned_pairs, 512, 90, 900, seed=a.seed + 5)
    progress(f"learned: {a.learned_pairs} pairs in {time.perf_counter() - t0:.1f} s; training")
    res["train_pairs"] = a.learned_pairs
    res["pair_gen_s"] = round(time.perf_counter() - t0, 2)
    scale = float(tgt[:256].std())
    net = Lrn.Corrector(seed=a.seed, scale=scale, threshold=a.learned_threshold)
    hist, steps, train_s = Lrn.train(net, inp, tgt, steps=a.learned_steps, seconds=a.learned_seconds, batch=32,
                                     lr=1e-3, progress=progress)
    progress(f"learned: {steps} steps in {train_s:.1f} s; inference")
    res.update(train_steps=steps, train_s=round(train_s, 1), epochs=round(steps * 32 / a.learned_pairs, 2),
               train_ms_per_step=round(train_s / steps * 1e3, 2),
               loss_first=float(f"{hist[0][1]:.4e}"), loss_last=float(f"{hist[-1][1]:.4e}"),
               loss_curve=[(k, float(f"{v:.4e}")) for k, v in hist[:: max(1, len(hist) // 12)]])
    del inp, tgt
    torch.cuda.empty_cache()
    # held-out pairs at the training size
    vi, vt = Lrn.make_pairs(64, 512, 90, 900, seed=a.seed + 5, slice0=10_000_000)
    vy, _ = net.correct(vi)
    torch.cuda.synchronize()
    res["heldout_512"] = {"psnr_plain90_vs_900": round(np.mean([Lrn.psnr(vi[i], vt[i]) for i in range(64)]), 2),
                          "psnr_learned90_vs_900": round(np.mean([Lrn.psnr(vy[i], vt[i]) for i in range(64)]), 2)}
    del vi, vt, vy

    # inference at 2048^2: 150 angles (+ network) against 1500-angle plain gridrec
    ns, n = a.learned_slices, 2048
    stream = torch.cuda.Stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    outs = {}
    for na in (150, 1500):
        sino = torch.empty((ns, na, n), dtype=torch.float32, device="cuda")
        for s0 in range(0, ns, 64):
            k = min(64, ns - s0)
            gr.phantom_sinograms(sino[s0:s0 + k], a.seed + 6, na, n, [n / 2.0] * k, 20, slice0=s0, n_tex=64,
                                 tex_amp=0.15, tex_fmax=24.0)
        plan = gr.Plan(n_ang=na, n_det=n, center=n / 2.0, pad="half", crop=(n // 2, n))
        out = torch.empty((ns, n, n), dtype=torch.float32, device="cuda")
        y = torch.empty_like(out) if na == 150 else None
        for rep in range(2):  # warm-up, then timed
            e0.record(stream)
            plan.recon(sino, out, stream=stream)
            e1.record(stream)
            e1.synchronize()
            t_recon = e0.elapsed_time(e1)
            t_cnn = 0.0
            if na == 150:
                with torch.cuda.stream(stream):
                    e0.record(stream)
                    for s0 in range(0, ns, 16):  # 16 slices of 2048^2 per network launch
                        net.correct(out[s0:s0 + 16], out=y[s0:s0 + 16], stream=stream)
                    e1.record(stream)
                e1.synchronize()
                t_cnn = e0.elapsed_time(e1)
        outs[na] = (out, y, t_recon, t_cnn)
        plan.close()
        del sino
    plain, _, t1500, _ = outs[1500]
    sparse, corr, t150, tcnn = outs[150]
    _, used = net.correct(sparse[0:16].contiguous())
    # against the phantom itself: the analytic image on the window's pixel
    # grid ([-1, 1]^2, density = reconstruction * n/2), inside the 0.95 disk
    progress("learned: PSNR against the analytic phantoms")
    from projects_amd import phantom as Ph
    cc = (np.arange(n) - (n - 1) / 2.0) / (n / 2.0)
    disk = np.add.outer(cc * cc, cc * cc) < 0.95 ** 2
    disk_t = torch.from_numpy(disk).cuda()
    truth_ps = {"plain1500": [], "plain150": [], "learned150": []}
    for i in range(min(a.learned_truth_slices, ns)):
        t = Ph.phantom_image(a.seed + 6, i, 20, n, n_tex=64, tex_amp=0.15, tex_fmax=24.0, device="cuda")[disk_t]
        rngt = float(t.max() - t.min())
        for key, img in (("plain1500", plain[i]), ("plain150", sparse[i]), ("learned150", corr[i])):
            r = img.double()[disk_t] * (n / 2.0)
            truth_ps[key].append(10.0 * math.log10(rngt * rngt / float(torch.mean((r - t) ** 2))))
    ps_sparse = [Lrn.psnr(sparse[i], plain[i]) for i in range(ns)]
    ps_learned = [Lrn.psnr(corr[i], plain[i]) for i in range(ns)]
    rl = lambda u, v: float(torch.linalg.norm((u - v).double()) / torch.linalg.norm(v.double()))
    cnn_flops = 2.0 * (9 * 32 + 3 * 9 * 32 * 32 + 9 * 32) * n * n * ns
    res["inference_2048"] = {
        "slices": ns,
        "plain1500_slices_s": round(ns / (t1500 * 1e-3), 1),
        "learned150_slices_s": round(ns / ((t150 + tcnn) * 1e-3), 1),
        "speedup_vs_plain1500": round(t1500 / (t150 + tcnn), 3),
        "gridrec150_ms": round(t150, 2), "cnn_ms": round(tcnn, 2),
        "cnn_TFLOPs": round(cnn_flops / (tcnn * 1e-3) / 1e12, 1),
        "psnr_plain150_vs_plain1500": round(float(np.mean(ps_sparse)), 2),
        "psnr_learned150_vs_plain1500": round(float(np.mean(ps_learned)), 2),
        "rel_l2_plain150_vs_plain1500": float(f"{np.mean([rl(sparse[i], plain[i]) for i in range(ns)]):.4e}"),
        "rel_l2_learned150_vs_plain1500": float(f"{np.mean([rl(corr[i], plain[i]) for i in range(ns)]):.4e}"),
        "psnr_vs_phantom": {k: round(float(np.mean(v)), 2) for k, v in truth_ps.items()},
        "psnr_vs_phantom_spread": {k: {"std": round(float(np.std(v)), 2), "min": round(float(np.min(v)), 2),
                                       "max": round(float(np.max(v)), 2)} for k, v in truth_ps.items()},
        "learned150_minus_plain1500_dB_vs_phantom_per_slice": {
            "min": round(float(np.min(np.subtract(truth_ps["learned150"], truth_ps["plain1500"]))), 2),
            "max": round(float(np.max(np.subtract(truth_ps["learned150"], truth_ps["plain1500"]))), 2)},
        "psnr_vs_phantom_slices": len(truth_ps["plain1500"]),
        "learned150_minus_plain1500_dB_vs_phantom": round(float(np.mean(truth_ps["learned150"]) -
                                                            np.mean(truth_ps["plain1500"])), 2),
        "fallback_slices_of_first16": int((~used).sum()),
        "threshold": a.learned_threshold,
    }
    del outs, plain, sparse, corr
    torch.cuda.empty_cache()
    return res


def main():
    a = parse()
    if a.copy_calibration:
        copy_calibration()
        return
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              "(slice pairs are independent: several GPUs would each take a share of the batch)", file=sys.stderr)
        sys.exit(2)
    import torch

    import oracle as O
    from projects_amd import gridrec as gr

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if a.workload == "learned":
        b = bench_learned(a, gr, torch)
        inf = b["inference_2048"]
        line = {"metric": METRIC, "value": inf["learned150_slices_s"], "unit": "slices/s", "n_gpus": 1, "steps": 1,
                "warmup": 1, "ms_per_step": round(inf["gridrec150_ms"] + inf["cnn_ms"], 2), "higher_is_better": True,
                "scaling": "weak", "vs_baseline": None, "dtype": "bf16 network / f32 gridrec",
                "data": "synthetic: seeded ellipse + texture phantoms, analytic Radon, noise-free",
                "config": {"workload": "learned sparse-angle path: 2048^2 from 150 angles + 5-layer residual CNN"},
                "learned": b}
        print(json.dumps(line), flush=True)
        return
    if a.workload == "precision":
        b = bench_precision(a, gr, O, torch)
        comp = b["variants"]["compensated/reference"]
        line = {"metric": METRIC, "value": comp["slices_per_s"], "unit": "slices/s", "n_gpus": 1, "steps": 1,
                "warmup": 1, "ms_per_step": round(a.precision_slices / comp["slices_per_s"] * 1e3, 2),
                "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
                "data": "synthetic: seeded 20-ellipse phantoms (log-uniform densities), analytic Radon, 1% Gaussian",
                "config": {"workload": b.pop("workload")}, "precision": b}
        print(json.dumps(line), flush=True)
        return
    if a.workload == "batched":
        b = bench_batched(a, gr, O, torch)
        line = {"metric": METRIC, "value": b.pop("value"), "unit": "slices/s", "n_gpus": 1, "steps": a.steps,
                "warmup": a.warmup, "ms_per_step": b["queue_ms_per_batch"], "higher_is_better": True,
                "scaling": "weak", "vs_baseline": None, "dtype": "f32",
                "data": "synthetic: seeded 20-ellipse phantoms, analytic Radon, Poisson(1e4)+log",
                "config": {"workload": b.pop("workload")}, "batched": b}
        print(json.dumps(line), flush=True)
        return
    dev = torch.device("cuda:0")
    ns, nang, ndet = a.slices, a.n_ang, a.n_det
    rng = np.random.default_rng(a.seed)
    center = float(np.float32(ndet / 2 + rng.uniform(-3.0, 3.0)))

    # synthetic sinograms, generated in HBM
    sino = torch.empty((ns, nang, ndet), dtype=torch.float32, device=dev)
    t0 = time.time()
    chunk = 256
    for s0 in range(0, ns, chunk):
        k = min(chunk, ns - s0)
        gr.phantom_sinograms(sino[s0:s0 + k], a.seed, nang, ndet, [center] * k, 20, photons=1e4, slice0=s0)
    torch.cuda.synchronize()
    gen_s = time.time() - t0

    plan = gr.Plan(n_ang=nang, n_det=ndet, center=center, pad="half", crop=(ndet // 2, ndet), accum=a.accum,
                   grid_op=a.grid_op)
    inf = plan.info
    out_n = plan.out_n
    out = torch.empty((ns, out_n, out_n), dtype=torch.float32, device=dev)
    stream = torch.cuda.Stream()  # a stream of its own: events and launches on one queue

    progress(f"plain: {ns} slices generated; warm-up")
    for _ in range(a.warmup):
        plan.recon(sino, out)
    torch.cuda.synchronize()
    progress("plain: timed steps")
    plan.set_timing(True)
    plan.timing(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        plan.recon(sino, out, stream=stream)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    tim = plan.timing(reset=True)
    plan.set_timing(False)
    inf = plan.info  # after the warm-up, which builds the gather operator in automatic mode
    grid_op_used = "on" if inf["grid_op_bytes"] > 0 else "off"
    slices_s = ns * a.steps / dt

    # roofline of the dominant kernel (events recorded in the timed region)
    half, M = inf["pdim"] // 2, inf["M"]
    pairs_total = tim["pairs"]
    kinds = {"rowfft": ("rowfft_ms", "rowfft_launches"), "grid": ("grid_ms", "grid_launches"),
             "ifft_rows": ("ifft_rows_ms", "ifft_rows_launches"), "ifft_cols": ("ifft_cols_ms", "ifft_cols_launches")}
    per_kernel = {}
    for kname, (ms_f, n_f) in kinds.items():
        launches = max(1, tim[n_f])
        avg_ms = tim[ms_f] / launches
        pairs_per_launch = pairs_total / launches
        b = kernel_bytes(kname, nang, ndet, half, M, out_n, pairs_per_launch)
        per_kernel[kname] = dict(total_ms=round(tim[ms_f], 3), launches=tim[n_f], avg_ms=round(avg_ms, 4),
                                 alg_bytes=int(b), gbs=round(b / (avg_ms * 1e-3) / 1e9, 1))
        if kname == "grid" and grid_op_used == "on":
            # the precomputed gather operator, streamed once per launch on top
            # of the algorithmic bytes
            per_kernel[kname]["operator_bytes"] = int(inf["grid_op_bytes"])
    dom = max(per_kernel, key=lambda k: per_kernel[k]["total_ms"])
    d = per_kernel[dom]
    roofline = {"bound": "hbm", "kernel": dom, "achieved": d["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(d["gbs"] / HBM_PEAK_GBS, 5), "traffic": None}
    # measured HBM bytes per launch of this kernel at this workload, from the
    # FETCH_SIZE / WRITE_SIZE counter runs kept beside the profile summaries
    tf = os.path.join(ROOT, "profiles", "gridrec_plain_1500x2048_hbm_traffic.json")
    if os.path.exists(tf) and (nang, ndet, a.accum) == (1500, 2048, "compensated"):
        with open(tf) as f:
            k = json.load(f)["kernels"].get(dom)
        if k and k["pairs_per_launch"] == inf["max_pairs"] and k.get("grid_op", "off") == grid_op_used:
            roofline["traffic"] = k["total_bytes"]
            roofline["traffic_read_write"] = [k["read_bytes"], k["write_bytes"]]
            roofline["traffic_source"] = "profiles/gridrec_plain_1500x2048_hbm_traffic.md"
    # whole-pipeline algorithmic bytes per pair (8 n_ang n_det + 16 n_ang pdim + 32 M^2 + 8 n_out^2)
    b_pair = 8 * nang * ndet + 16 * nang * inf["pdim"] + 32 * M * M + 8 * out_n * out_n
    pipeline_gbs = b_pair * (slices_s / 2) / 1e9

    result = {
        "metric": METRIC, "value": round(slices_s, 2), "unit": "slices/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(dt / a.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: seeded 20-ellipse phantoms, analytic Radon, Poisson(1e4)+log",
        "config": {"workload": "plain gridrec, 2048 slices x 1500 angles x 2048 det, half-pad to 4096, "
                               "PSWF C=6, shlo, 2048^2 window of 4095^2",
                   "slices": ns, "n_ang": nang, "n_det": ndet, "center": center, "pdim": inf["pdim"], "M": M,
                   "M0": inf["M0"], "out_n": out_n, "accum": a.accum, "tables": "reference",
                   "pairs_per_launch": inf["max_pairs"], "grid_op": grid_op_used,
                   "grid_op_MB": round(inf["grid_op_bytes"] / 1e6, 1),
                   "grid_op_build_ms": round(inf["grid_op_build_ms"], 1)},
        "roofline": roofline,
        "kernels": per_kernel,
        "pipeline_alg_GBs": round(pipeline_gbs, 1),
        "gen_s": round(gen_s, 2),
    }

    # accuracy on pair 0 (the timed output)
    progress(f"plain: {slices_s:.0f} slices/s; accuracy and CPU baseline")
    if not a.no_accuracy:
        s = sino[0:2].cpu().numpy()
        o = out[0:2].cpu().numpy()
        g = O.Geometry(n_ang=nang, n_det=ndet, center=center).padded()
        p1, p2 = O.pad_half(s[0]), O.pad_half(s[1])
        c = slice(ndet // 2, ndet // 2 + out_n)
        r1, r2 = O.recon_f32(g, p1, p2)
        d1, d2 = O.recon_f64(g, p1, p2)
        acc = {"rel_l2_vs_reference_restatement": max(O.rel_l2(o[0], r1[c, c]), O.rel_l2(o[1], r2[c, c])),
               "rel_l2_vs_fp64": max(O.rel_l2(o[0], d1[c, c]), O.rel_l2(o[1], d2[c, c])),
               "max_abs_vs_fp64": float(max(np.abs(o[0] - d1[c, c]).max(), np.abs(o[1] - d2[c, c]).max())),
               "reference_restatement_rel_l2_vs_fp64": max(O.rel_l2(r1[c, c], d1[c, c]), O.rel_l2(r2[c, c], d2[c, c]))}
        if O.ref_available():
            q1, q2 = O.ref_recon(g, p1, p2)
            acc["rel_l2_vs_reference_build"] = max(O.rel_l2(o[0], q1[c, c]), O.rel_l2(o[1], q2[c, c]))
        result["accuracy"] = {k: float(f"{v:.3e}") for k, v in acc.items()}

    # PCIe-inclusive rate: the reference's own boundary hands over host rows;
    # gr_exec_pairs uploads, reconstructs and downloads in a pipeline
    if a.h2h_slices > 0:
        progress("plain: host-to-host (gr_exec_pairs)")
        k = min(a.h2h_slices, ns)
        h_sino = sino[0:k].cpu().numpy()
        h_out = np.empty((k, out_n, out_n), np.float32)
        plan.exec_pairs(h_sino[0:16], h_out[0:16])  # staging allocated
        t0 = time.perf_counter()
        plan.exec_pairs(h_sino, h_out)
        hdt = time.perf_counter() - t0
        same = bool(np.array_equal(h_out[0:2], out[0:2].cpu().numpy()))
        result["host_to_host"] = {"slices": k, "slices_s": round(k / hdt, 1), "bitwise_equal_to_device_path": same,
                                  "what": "host sinogram rows in, host image rows out: PCIe both ways, pinned "
                                          "staging, row packing on 16 host threads (gr_exec_pairs)"}
        del h_sino, h_out

    # CPU baseline on this host, bounded samples of the same workload: the
    # reference's own gridrec (oracle/_ref, GridrecCPP compiled from its
    # sources) on one core and one process per core, and the FP32 restatement
    # on all cores
    if not a.no_cpu:
        progress("CPU baseline")
        result["cpu_baseline"] = cpu_baseline(a, O, sino, ns, nang, ndet, center, out_n)
    plan.close()
    if a.workload == "all":
        del sino, out
        torch.cuda.empty_cache()
        result["batched"] = bench_batched(a, gr, O, torch)
        torch.cuda.empty_cache()
        if a.precision_fp64_pairs <= 0:
            a.precision_fp64_pairs = 32  # a 64-slice sample against FP64 in the combined run
        pr = bench_precision(a, gr, O, torch)
        for v in pr["variants"].values():
            v.pop("per_slice_rel_l2_disk")  # keep the headline line short
        result["precision"] = pr
        torch.cuda.empty_cache()
        result["learned"] = bench_learned(a, gr, torch)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
