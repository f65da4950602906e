"""GPU tests of the batched uneven queue (gr_queue_*), through the C-ABI.

The queue's images must be bitwise those of a plain per-dataset plan
(gr_recon_device with default settings) wherever the tuner kept the default
FFT radix: slice pairs per launch, the column-FFT width, the gridding
workgroup size and the gather operator the queue builds for its datasets
change no sum order.  Where it chose radix 4 the FFTs add in
another order and the images agree to FP32 rounding: rel-L2 <= 1e-5, the
parity gate against the reference (over whole windows, whose corners the
deapodization amplifies, either order is ~6e-6 from FP64; the two orders
have been seen 1.6e-6 apart at M = 4096).  The
graph is replayed several times with eager work of the program between
replays, and every replay is checked against the eager result.
"""
import numpy as np
import pytest

import oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if torch.cuda.is_available():
    from projects_amd import gridrec as gr


def uneven_batch(seed, n_sets, widths=(64, 96, 128, 192), slices=(1, 2, 3, 5, 8), angles=(24, 120)):
    """Datasets of random width, angle count, centre and slice count, with
    their sinograms generated on the device and their output buffers."""
    rng = np.random.default_rng(seed)
    sets = []
    for d in range(n_sets):
        n = int(rng.choice(widths))
        na = int(rng.integers(*angles))
        ns = int(rng.choice(slices))
        c = n / 2 + float(rng.uniform(-3, 3))
        sino = torch.empty((ns, na, n), dtype=torch.float32, device="cuda")
        gr.phantom_sinograms(sino, seed=seed * 1000 + d, n_ang=na, n_det=n, centers=[c] * ns, n_ell=8,
                             photons=1e4)
        kw = dict(n_ang=na, n_det=n, center=c, filter=str(rng.choice(["shlo", "hann", "ramp"])), pad="half")
        M0 = gr.Plan(**kw).info["M0"]
        cn = int(round(n * M0 / (2 * n)))
        kw["crop"] = ((M0 - cn + 1) // 2, cn)
        out = torch.full((ns, cn, cn), float("nan"), dtype=torch.float32, device="cuda")
        sets.append((kw, sino, out))
    return sets


def plain_results(sets):
    res = []
    for kw, sino, _ in sets:
        p = gr.Plan(**kw)
        res.append(p.recon(sino))
        p.close()
    torch.cuda.synchronize()
    return [r.cpu() for r in res]


class Differs(str):
    def __bool__(self):
        return False


def same_as_plain(q, d, out, ref):
    """Bitwise where the tuned radix is the plain default, else FP32 rounding.
    A mismatch comes back as a false string naming the settings and the error."""
    o = out.cpu()
    st = q.settings(d)
    rel = float((o - ref).norm() / ref.norm())
    if st["fft_radix"] == 16:
        ok = torch.equal(o, ref)
    else:
        ok = bool(torch.isfinite(o).all()) and rel <= 1e-5
    return True if ok else Differs(f"dataset {d} settings {st} rel-L2 {rel:.3g}")


def test_queue_matches_plain_plans_and_replays():
    sets = uneven_batch(11, 10)
    ref = plain_results(sets)
    q = gr.Queue(sets, n_streams=3, autotune=True)
    torch.cuda.synchronize()
    info = q.info
    print("queue info", info, [q.settings(d) for d in range(10)], flush=True)
    assert info["n_datasets"] == 10 and info["slices"] == sum(s[1].shape[0] for s in sets)
    # small datasets: every one fits the operator budget
    assert info["datasets_with_operator"] == 10 and info["operator_bytes"] > 0
    for d, (_, _, out) in enumerate(sets):  # create ran the batch once
        assert same_as_plain(q, d, out, ref[d]), d
    first = [out.cpu() for _, _, out in sets]
    for rep in range(3):
        for _, _, out in sets:
            out.fill_(float("nan"))
        scratch = torch.randn(1 << 20, device="cuda").sum()  # eager work between replays
        q.run()
        torch.cuda.synchronize()
        assert torch.isfinite(scratch)
        for d, (_, _, out) in enumerate(sets):
            assert torch.equal(out.cpu(), first[d]), (rep, d)  # every replay bitwise the first run
    for _, _, out in sets:
        out.fill_(float("nan"))
    q.run_sequential()
    torch.cuda.synchronize()
    for d, (_, _, out) in enumerate(sets):
        assert torch.equal(out.cpu(), first[d]), ("sequential", d)
    for d in range(10):
        s = q.settings(d)
        assert s["max_pairs"] in (1, 2, 4, 8) and 0 <= s["stream"] < 3
        assert s["fft_radix"] in (4, 16) and s["grid_wpb"] in (1, 2, 4)
    q.close()


def test_queue_without_autotune_and_single_stream():
    """One stream, default settings, and datasets that refuse the gather
    operator (grid_op off): the geometry is evaluated per launch."""
    sets = [(dict(kw, grid_op="off"), sino, out) for kw, sino, out in uneven_batch(12, 5, slices=(2, 4))]
    ref = plain_results(sets)
    q = gr.Queue(sets, n_streams=1, autotune=False)
    assert q.info["datasets_with_operator"] == 0
    for _, _, out in sets:
        out.fill_(float("nan"))
    q.run()
    torch.cuda.synchronize()
    for d, (_, _, out) in enumerate(sets):
        assert q.settings(d)["fft_radix"] == 16
        assert torch.equal(out.cpu(), ref[d]), d
    q.close()


def test_queue_images_match_oracle():
    """A queue dataset against the C oracle (FP32 restatement) on the padded rows."""
    sets = uneven_batch(13, 4, slices=(2,))
    q = gr.Queue(sets, n_streams=2)
    q.run()
    torch.cuda.synchronize()
    for kw, sino, out in sets:
        g = O.Geometry(n_ang=kw["n_ang"], n_det=kw["n_det"], center=kw["center"], filter=kw["filter"]).padded()
        s = sino.cpu().numpy()
        a, b = O.pad_half(s[0]), O.pad_half(s[1])
        r1, r2 = O.recon_f32(g, a, b)
        off, cn = kw["crop"]
        o = out.cpu().numpy()
        for img, r in ((o[0], r1), (o[1], r2)):
            e = O.rel_l2(img, r[off:off + cn, off:off + cn])
            assert e <= 1e-5, (kw, e)
    q.close()


def test_queue_rejects_bad_datasets():
    sets = uneven_batch(14, 2, slices=(2,))
    kw, sino, out = sets[1]
    bad = dict(kw, n_ang=0)
    with pytest.raises(gr.GridrecError):
        gr.Queue([sets[0], (bad, sino, out)])


def test_queue_with_angle_lists():
    """Datasets with explicit (non-uniform) angle lists (geom 0) through the
    queue: equal to their plain plans, and one against the oracle."""
    rng = np.random.default_rng(15)
    sets = []
    for d in range(4):
        n, na, ns = int(rng.choice([64, 96])), int(rng.integers(30, 80)), int(rng.choice([2, 3, 4]))
        ang = np.sort(rng.uniform(0, 180, na)).astype(np.float32)
        sino = torch.from_numpy(rng.standard_normal((ns, na, n)).astype(np.float32)).cuda()
        kw = dict(n_ang=na, n_det=n, center=n / 2 + float(rng.uniform(-2, 2)), geom=0, angles=ang)
        M0 = gr.Plan(**kw).info["M0"]
        out = torch.full((ns, M0, M0), float("nan"), device="cuda")
        sets.append((kw, sino, out))
    ref = plain_results(sets)
    q = gr.Queue(sets, n_streams=2)
    torch.cuda.synchronize()
    for d, (_, _, out) in enumerate(sets):
        assert same_as_plain(q, d, out, ref[d]), d
    kw, sino, out = sets[0]
    g = O.Geometry(n_ang=kw["n_ang"], n_det=kw["n_det"], center=kw["center"], geom=0, angles=kw["angles"])
    s = sino.cpu().numpy()
    r1, r2 = O.recon_f32(g, s[0], s[1])
    o = out.cpu().numpy()
    M0 = o.shape[1]
    yy, xx = np.mgrid[:M0, :M0]
    disk = np.hypot(yy - M0 // 2, xx - M0 // 2) < 0.9 * M0 / 2
    assert O.rel_l2(o[0][disk], r1[disk]) < 1e-5 and O.rel_l2(o[1][disk], r2[disk]) < 1e-5
    q.close()


@pytest.mark.parametrize("autotune", [False, True])
def test_queue_one_size_class_uneven_angles(autotune):
    """One size class whose dataset with the most slices has the fewest angles:
    the class is tuned on that dataset (up to 8 pairs per launch), and the
    others -- one pair with 20 times the angles -- must still get a workspace
    that holds their spectra and grids; their images equal their plain plans."""
    specs = [(16, 30), (2, 600), (1, 450), (3, 200)]  # (slices, angles), width 128
    sets = []
    for d, (ns, na) in enumerate(specs):
        c = 64.0 + 0.7 * d
        sino = torch.empty((ns, na, 128), dtype=torch.float32, device="cuda")
        gr.phantom_sinograms(sino, seed=300 + d, n_ang=na, n_det=128, centers=[c] * ns, n_ell=6, photons=1e4)
        kw = dict(n_ang=na, n_det=128, center=c, pad="half", crop=(64, 128))
        out = torch.full((ns, 128, 128), float("nan"), dtype=torch.float32, device="cuda")
        sets.append((kw, sino, out))
    ref = plain_results(sets)
    q = gr.Queue(sets, n_streams=2, autotune=autotune)
    torch.cuda.synchronize()
    assert q.info["size_classes"] == 1
    for d, (ns, _) in enumerate(specs):
        assert q.settings(d)["max_pairs"] <= max(1, (ns + 1) // 2), (d, q.settings(d))
    for rep in range(2):
        if rep:
            for _, _, out in sets:
                out.fill_(float("nan"))
            q.run()
            torch.cuda.synchronize()
        for d, (_, _, out) in enumerate(sets):
            assert same_as_plain(q, d, out, ref[d]), (rep, d)
    q.close()


def test_queue_at_production_widths():
    """The uneven queue at the sizes of the production batch: detector widths
    512 to 2560 (all five, two of them not powers of two), 180 to 1800 angles,
    1 to 4 slices per dataset, centres off the middle.  Every dataset's images
    are those of its plain plan (bitwise at the default radix), and three of
    them match the FP32 oracle on the window to 1e-5."""
    specs = [(512, 180, 3), (1024, 1800, 1), (1536, 700, 2), (2048, 1500, 4), (2560, 400, 2),
             (512, 1333, 4), (2048, 250, 1), (1024, 960, 2)]
    rng = np.random.default_rng(31)
    sets = []
    for d, (n, na, ns) in enumerate(specs):
        c = float(np.float32(n / 2 + rng.uniform(-5, 5)))
        sino = torch.empty((ns, na, n), dtype=torch.float32, device="cuda")
        gr.phantom_sinograms(sino, seed=500 + d, n_ang=na, n_det=n, centers=[c] * ns, n_ell=20, photons=1e4)
        kw = dict(n_ang=na, n_det=n, center=c, pad="half")
        M0 = (1 << (2 * n - 1).bit_length()) - 1
        cn = int(round(n * M0 / (2 * n)))
        kw["crop"] = ((M0 - cn + 1) // 2, cn)
        out = torch.full((ns, cn, cn), float("nan"), dtype=torch.float32, device="cuda")
        sets.append((kw, sino, out))
    ref = plain_results(sets)
    q = gr.Queue(sets, n_streams=4, autotune=True)
    q.run()
    torch.cuda.synchronize()
    print("queue info", q.info, flush=True)
    # size classes (pdim, M, window): 1536 and 2048 share one -- both pdim =
    # M = 4096, and the 1536 window round(1536 * 4095 / 3072) = 2048
    assert q.info["size_classes"] == 4
    for d, (_, _, out) in enumerate(sets):
        assert same_as_plain(q, d, out, ref[d]), d
    for d in (0, 7, 2):
        kw, sino, out = sets[d]
        g = O.Geometry(n_ang=kw["n_ang"], n_det=kw["n_det"], center=kw["center"]).padded()
        s = sino.cpu().numpy()
        r1, r2 = O.recon_f32(g, O.pad_half(s[0]), O.pad_half(s[1]))
        off, cn = kw["crop"]
        o = out.cpu().numpy()
        for img, r in ((o[0], r1), (o[1], r2)):
            e = O.rel_l2(img, r[off:off + cn, off:off + cn])
            print(f"dataset {d} ({kw['n_det']} x {kw['n_ang']}): rel-L2 vs oracle {e:.2e}", flush=True)
            assert e <= 1e-5, (d, e)
    q.close()
