"""GPU parity tests of the HIP gridrec path, through the C-ABI.

Bars: tables bit-exact against the reference's; images to FP32 rounding against
the reference's own gridrec (golden vectors and oracle/_ref where built) and
the C oracle -- rel-L2 <= 1e-5 on the reconstruction window of the caller
convention and on the disk of full M0 x M0 images.  Outside that disk the
deapodization factor winv^2 grows steeply (1/PSWF^2 towards the corners) and
amplifies the FFT rounding of any two FP32 implementations -- the reference's
own error against FP64 is largest there -- so whole images are held to 1e-4.
"""
import numpy as np
import pytest

import oracle as O
from projects_amd import phantom

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if torch.cuda.is_available():
    from projects_amd import gridrec as gr

DISK_TOL, FULL_TOL = 1e-5, 1e-4


def disk_mask(M0, frac=0.9):
    yy, xx = np.mgrid[:M0, :M0]
    return np.hypot(yy - M0 // 2, xx - M0 // 2) < frac * M0 / 2


def image_errors(out, ref):
    """(whole-image, disk) rel-L2 of one M0 x M0 image."""
    d = disk_mask(ref.shape[0])
    return O.rel_l2(out, ref), O.rel_l2(out[d], ref[d])


def geometry_of(case, arr, i):
    c = {k: v for k, v in case.items() if k not in ("M0", "pdim", "M", "sampl_out", "L", "scale")}
    return O.Geometry(angles=arr.get(f"case{i:02d}_angles"), **c)


def plan_for(g: O.Geometry, **kw):
    return gr.Plan(n_ang=g.n_ang, n_det=g.n_det, center=g.center, geom=g.geom, angles=g.angles, filter=g.filter,
                   C_=g.C, sampl=g.sampl, MaxPixSiz=g.MaxPixSiz, R=g.R, X0=g.X0, Y0=g.Y0, ltbl=g.ltbl, **kw)


def gpu_pair(plan, a, b):
    out = plan.recon(torch.from_numpy(np.stack([a, b]).astype(np.float32)).cuda())
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_library_loads_and_reports():
    assert gr.lib().gr_version().decode().startswith("gridrec-mi355x")
    p = gr.Plan(n_ang=16, n_det=64, center=32.0)
    assert p.info["M0"] == 63 and p.info["pdim"] == 64
    p.close()


def test_plan_tables_bit_exact(golden):
    meta, arr = golden
    for i, case in enumerate(meta["cases"]):
        g = geometry_of(case, arr, i)
        p = plan_for(g)
        t = p.tables()
        for tn in ("sine", "cose", "filphase", "wtbl", "winv"):
            assert np.array_equal(t[tn].view(np.uint32), arr[f"case{i:02d}_{tn}"].view(np.uint32)), (i, tn)
        p.close()


def test_golden_small_cases(golden):
    """Every golden case of the reference's gridrec: against its output (disk
    1e-5, whole image 1e-4), and -- whole M0 x M0 images -- against the FP64
    oracle, where the GPU must be at least as accurate as the reference
    itself.  That is the claim behind the whole-image allowance: two FP32
    implementations, each about e from the exact result, differ by about
    sqrt(2) e, and on whole images e is ~1e-5 for the reference too (the
    deapodization factor winv^2 amplifies FFT rounding in the corners).
    Over the 58 images: the RMS of the GPU's errors is at most the
    reference's, its largest error at most the reference's largest, and no
    image's error exceeds 3x the reference's on the same image (each error is
    dominated by a few corner pixels, so per-image ratios of two independent
    roundings scatter; measured: RMS 6.6e-6 against 8.3e-6, the GPU more
    accurate on 74% of the images, largest ratio 2.4)."""
    meta, arr = golden
    worse, errs = [], []
    for i, case in enumerate(meta["cases"]):
        g = geometry_of(case, arr, i)
        p = plan_for(g)
        a, b = arr[f"case{i:02d}_in"]
        out = gpu_pair(p, a, b)
        r1, r2 = arr[f"case{i:02d}_out"]
        (f1, d1), (f2, d2) = image_errors(out[0], r1), image_errors(out[1], r2)
        x1, x2 = O.recon_f64(g, a, b)
        gpu64 = (O.rel_l2(out[0], x1), O.rel_l2(out[1], x2))
        ref64 = (O.rel_l2(r1, x1), O.rel_l2(r2, x2))
        print(f"case {i}: rel-L2 vs reference: disk {d1:.2e} {d2:.2e}, whole {f1:.2e} {f2:.2e}; "
              f"vs FP64: GPU {gpu64[0]:.2e} {gpu64[1]:.2e}, reference {ref64[0]:.2e} {ref64[1]:.2e}", flush=True)
        assert d1 < DISK_TOL and d2 < DISK_TOL and f1 < FULL_TOL and f2 < FULL_TOL, (i, d1, d2, f1, f2)
        for k in range(2):
            errs.append((gpu64[k], ref64[k]))
            if gpu64[k] > 3.0 * ref64[k]:
                worse.append((i, k, gpu64[k], ref64[k]))
        p.close()
    g, r = np.array(errs).T
    print(f"whole images vs FP64: RMS GPU {np.sqrt(np.mean(g ** 2)):.2e}, reference {np.sqrt(np.mean(r ** 2)):.2e}; "
          f"max GPU {g.max():.2e}, reference {r.max():.2e}; GPU better on {np.mean(g <= r):.0%}", flush=True)
    assert np.sqrt(np.mean(g ** 2)) <= np.sqrt(np.mean(r ** 2))
    assert g.max() <= r.max()
    assert not worse, worse


def test_golden_cfg1_onchip_padding(golden):
    """256 detectors, 180 angles, Shepp-Logan: half padding generated while
    loading, 256 x 256 window of the 511 x 511 image, vs the reference."""
    meta, arr = golden
    m = meta["cfg1"]
    p = gr.Plan(n_ang=m["n_ang"], n_det=m["n_det"], center=m["center"], filter=m["filter"], pad="half",
                crop=(m["crop_off"], m["crop_n"]))
    assert p.info["M0"] == m["M0"]
    out = gpu_pair(p, *arr["cfg1_in"])
    r1, r2 = arr["cfg1_out"]
    e1, e2 = O.rel_l2(out[0], r1), O.rel_l2(out[1], r2)
    print(f"cfg1: rel-L2 vs reference {e1:.2e} {e2:.2e}", flush=True)
    assert e1 < 1e-5 and e2 < 1e-5
    p.close()


CASES = [
    dict(n_ang=60, n_det=100, center=50.3, filter="hann"),
    dict(n_ang=45, n_det=128, center=60.0, geom=2, filter="hamm"),
    dict(n_ang=30, n_det=96, center=47.0, X0=0.1, Y0=-0.07, filter="ramp"),
    dict(n_ang=50, n_det=200, center=100.5, sampl=1.2),
    dict(n_ang=40, n_det=64, center=31.0, C=4.2, ltbl=300),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "-".join(f"{k}{v}" for k, v in c.items()))
@pytest.mark.parametrize("accum", ["compensated", "plain"])
def test_random_cases_vs_oracle(case, accum):
    g = O.Geometry(**case)
    rng = np.random.default_rng(len(case) * 7 + g.n_ang)
    a, b = rng.standard_normal((2, g.n_ang, g.n_det)).astype(np.float32)
    p = plan_for(g, accum=accum)
    out = gpu_pair(p, a, b)
    r1, r2 = O.recon_f32(g, a, b)
    (f1, d1), (f2, d2) = image_errors(out[0], r1), image_errors(out[1], r2)
    print(f"{case} {accum}: disk {d1:.2e} {d2:.2e} whole {f1:.2e} {f2:.2e}", flush=True)
    assert max(d1, d2) < DISK_TOL and max(f1, f2) < FULL_TOL
    p.close()


def test_angle_list_geom0():
    rng = np.random.default_rng(11)
    ang = np.sort(rng.uniform(0, 360, 57)).astype(np.float32)
    g = O.Geometry(n_ang=57, n_det=150, center=74.2, geom=0, angles=ang, filter="shlo")
    a, b = rng.standard_normal((2, 57, 150)).astype(np.float32)
    p = plan_for(g)
    out = gpu_pair(p, a, b)
    r1, r2 = O.recon_f32(g, a, b)
    (f1, d1), (f2, d2) = image_errors(out[0], r1), image_errors(out[1], r2)
    assert max(d1, d2) < DISK_TOL and max(f1, f2) < FULL_TOL, (d1, d2, f1, f2)
    p.close()


def test_batch_odd_count_strided_and_pairing():
    """7 slices in a strided buffer (row stride 70 > n_det 64), small max_pairs
    so several chunks run; the odd last slice pairs with zeros; outputs land in
    a strided image buffer.  Each pair equals its own reconstruction."""
    n_ang, n_det, ns = 40, 64, 7
    g = O.Geometry(n_ang=n_ang, n_det=n_det, center=32.0)
    rng = np.random.default_rng(5)
    sino = rng.standard_normal((ns, n_ang, n_det)).astype(np.float32)
    buf = torch.zeros((ns, n_ang, 70), device="cuda")
    buf[:, :, :n_det] = torch.from_numpy(sino).cuda()
    p = plan_for(g, max_pairs=2)
    M0 = p.info["M0"]
    out = torch.full((ns, M0 + 3, M0 + 5), 7.0, device="cuda")
    p.recon(buf, out=out)
    torch.cuda.synchronize()
    o = out.cpu().numpy()
    for k in range(0, ns, 2):
        b = sino[k + 1] if k + 1 < ns else np.zeros_like(sino[k])
        r1, r2 = O.recon_f32(g, sino[k], b)
        f1, d1 = image_errors(o[k, :M0, :M0], r1)
        assert d1 < DISK_TOL and f1 < FULL_TOL, (k, d1, f1)
        if k + 1 < ns:
            f2, d2 = image_errors(o[k + 1, :M0, :M0], r2)
            assert d2 < DISK_TOL and f2 < FULL_TOL, (k + 1, d2, f2)
    assert (o[:, M0:, :] == 7.0).all() and (o[:, :, M0:] == 7.0).all()  # padding untouched
    p.close()


@pytest.mark.parametrize("grid_op", ["off", "on"])
@pytest.mark.parametrize("accum", ["compensated", "plain"])
def test_pair_grouping_is_bitwise_neutral(accum, grid_op):
    """Pairs per launch 1..16 group the gridding into launches of 8, 4, 2, 1
    pairs and partial groups of 8 (5-7 pairs) or 4 (3 pairs); every grouping
    gives bitwise the images of one pair per launch, with the geometry
    evaluated per launch or streamed from the precomputed operator."""
    g = O.Geometry(n_ang=72, n_det=96, center=47.25)
    rng = np.random.default_rng(21)
    ns = 31  # 16 pairs, the last one odd
    sino = torch.from_numpy(rng.standard_normal((ns, 72, 96)).astype(np.float32)).cuda()
    ref = plan_for(g, max_pairs=1, accum=accum, grid_op=grid_op).recon(sino)
    torch.cuda.synchronize()
    for mp in (2, 3, 4, 5, 6, 7, 8, 13, 16):
        out = plan_for(g, max_pairs=mp, accum=accum, grid_op=grid_op).recon(sino)
        torch.cuda.synchronize()
        assert torch.equal(out, ref), mp


def test_many_interleaved_groups_are_bitwise_neutral():
    """Launches of 24 and 64 pairs grid in 3 and 8 interleaved groups of 8
    pairs (the staged operator's workgroups of a slot side by side on one
    XCD), with a one-pair remainder launch; bitwise the images of one pair
    per launch."""
    g = O.Geometry(n_ang=72, n_det=96, center=47.25)
    rng = np.random.default_rng(65)
    ns = 130  # 65 pairs
    sino = torch.from_numpy(rng.standard_normal((ns, 72, 96)).astype(np.float32)).cuda()
    ref = plan_for(g, max_pairs=1, grid_op="on").recon(sino)
    torch.cuda.synchronize()
    for mp in (24, 64):
        out = plan_for(g, max_pairs=mp, grid_op="on").recon(sino)
        torch.cuda.synchronize()
        assert torch.equal(out, ref), mp


@pytest.mark.parametrize("grid_op", ["off", "on"])
def test_pair_grouping_at_4096_is_bitwise_neutral(grid_op):
    """At the 2048-detector geometry (pdim = M = 4096: the persistent row and
    column transforms over items of 4 rows / 4 columns of every pair of the
    launch) and, with the operator, gridding in interleaved groups of 8 pairs
    (16 pairs: two groups, 3 pairs: one partial group), every grouping gives
    bitwise the images of one pair per launch."""
    rng = np.random.default_rng(4096)
    n_ang, n_det, ns = 40, 2048, 33  # 17 pairs, the last one odd
    sino = torch.from_numpy(rng.standard_normal((ns, n_ang, n_det)).astype(np.float32)).cuda()
    kw = dict(n_ang=n_ang, n_det=n_det, center=1023.6, pad="half", crop=(1024, n_det), grid_op=grid_op)
    p = gr.Plan(max_pairs=1, **kw)
    ref = p.recon(sino)
    torch.cuda.synchronize()
    p.close()
    assert torch.isfinite(ref).all() and ref.abs().max() > 0
    for mp in (3, 4, 5, 8, 16, 40):
        p = gr.Plan(max_pairs=mp, **kw)
        out = p.recon(sino)
        torch.cuda.synchronize()
        assert torch.equal(out, ref), mp
        p.close()


def test_persistent_transforms_with_ragged_window_match_oracle():
    """A 2047-detector half-padded geometry (pdim = M = 4096) with a 2047-wide
    window: the persistent row and column transforms run with a padded last
    column block of Z (2047 is not a multiple of 4).  The images match the
    FP32 oracle's padded-and-cropped reconstruction to FP32 rounding, and the
    persistent column transform is bitwise the one-column-per-workgroup
    form."""
    rng = np.random.default_rng(2047)
    n_ang, n_det, center = 24, 2047, 1023.4
    sino = rng.standard_normal((2, n_ang, n_det)).astype(np.float32)
    p = gr.Plan(n_ang=n_ang, n_det=n_det, center=center, pad="half", crop=(n_det // 2, n_det))
    assert p.info["M"] == 4096 and p.out_n == n_det
    out = p.recon(torch.from_numpy(sino).cuda())
    torch.cuda.synchronize()
    p.set_cols_per_block(1)
    out1 = p.recon(torch.from_numpy(sino).cuda())
    torch.cuda.synchronize()
    assert torch.equal(out, out1)
    ref = O.recon_batch_f32(O.Geometry(n_ang=n_ang, n_det=n_det, center=center), sino, True, n_det // 2, n_det,
                            threads=8)
    o = out.cpu().numpy()
    for k in range(2):
        err = float(np.linalg.norm(o[k].astype(np.float64) - ref[k]) / np.linalg.norm(ref[k].astype(np.float64)))
        print(f"slice {k}: rel-L2 vs oracle {err:.2e}", flush=True)
        assert err < 1e-5, (k, err)
    p.close()


def test_column_widths_are_bitwise_neutral():
    """The column transform's workgroup width (columns side by side, or one
    after another) changes only the launch shape: every width a plan accepts
    gives bitwise the images of one column per workgroup, from M = 128 to
    M = 8192; widths neither form takes are refused."""
    cases = [dict(n_ang=40, n_det=96, center=47.25), dict(n_ang=60, n_det=300, center=150.5),
             dict(n_ang=50, n_det=700, center=349.0), dict(n_ang=40, n_det=1500, center=751.0),
             dict(n_ang=45, n_det=2048, center=1024.3, pad="half"),
             dict(n_ang=30, n_det=2560, center=1280.0, pad="half")]
    rng = np.random.default_rng(33)
    seen = set()
    for case in cases:
        sino = torch.from_numpy(rng.standard_normal((3, case["n_ang"], case["n_det"])).astype(np.float32)).cuda()
        p = gr.Plan(**case)
        M = p.info["M"]
        seen.add(M)
        p.set_cols_per_block(1)
        ref = p.recon(sino)
        torch.cuda.synchronize()
        assert torch.isfinite(ref).all()
        for cols in (0, 2, 4, 8, 16):
            try:
                p.set_cols_per_block(cols)
            except gr.GridrecError:
                print(f"M={M}: width {cols} refused", flush=True)
                assert cols * M > 16384, (M, cols)
                continue
            out = p.recon(sino)
            torch.cuda.synchronize()
            assert torch.equal(out, ref), (M, cols)
        p.close()
    assert {4096, 8192} <= seen, seen


def test_recon_host_equals_device_and_is_deterministic():
    g = O.Geometry(n_ang=90, n_det=128, center=63.5)
    rng = np.random.default_rng(9)
    a, b = rng.standard_normal((2, 90, 128)).astype(np.float32)
    p = plan_for(g)
    d1 = gpu_pair(p, a, b)
    d2 = gpu_pair(p, a, b)
    h1, h2 = p.recon_host(a, b)
    assert np.array_equal(d1, d2)                       # no atomics: bitwise repeatable
    assert np.array_equal(d1[0], h1) and np.array_equal(d1[1], h2)
    p.close()


def test_fp64_tables_mode():
    """GR_TABLES_FP64 computes filphase/wtbl/winv in double on the GPU: the
    result then tracks the oracle whose tables are double."""
    g = O.Geometry(n_ang=120, n_det=256, center=128.0)
    sino = np.stack([phantom.sinogram(3, s, 120, 256, 128.0, 10) for s in range(2)])
    pf = plan_for(g, tables="fp64")
    out = gpu_pair(pf, sino[0], sino[1])
    m1, _ = O.recon_mixed(g, sino[0], sino[1], O.MX_TABLES)   # FP32 pipeline, FP64 tables
    d1, _ = O.recon_f64(g, sino[0], sino[1])
    r1, _ = O.recon_f32(g, sino[0], sino[1])
    M0 = pf.info["M0"]
    yy, xx = np.mgrid[:M0, :M0]
    disk = np.hypot(yy - M0 // 2, xx - M0 // 2) < 0.45 * M0
    e_mixed = O.rel_l2(out[0][disk], m1[disk])
    e64, e_ref64 = O.rel_l2(out[0][disk], d1[disk]), O.rel_l2(r1[disk], d1[disk])
    print(f"fp64 tables: vs mixed oracle {e_mixed:.2e}, vs FP64 {e64:.2e} (reference tables: {e_ref64:.2e})")
    assert e_mixed < 1e-5 and e64 <= e_ref64
    t = pf.tables()
    ref = O.tables(g)
    assert np.abs(t["wtbl"] - ref["wtbl"]).max() < 1e-3   # same table, rounding of the float recurrence
    pf.close()


def test_phantom_kernel_matches_numpy():
    n_ang, n_det, ns = 90, 128, 3
    centers = [64.0, 63.3, 65.1]
    out = torch.empty((ns, n_ang, n_det), device="cuda")
    gr.phantom_sinograms(out, 1234, n_ang, n_det, centers, 12, slice0=5)
    got = out.cpu().numpy()
    for s in range(ns):
        ref = phantom.sinogram(1234, 5 + s, n_ang, n_det, centers[s], 12)
        # float vs double evaluation; sqrt(s^2 - t^2) is steep at ellipse edges
        assert O.rel_l2(got[s], ref) < 1e-4, s
        assert np.abs(got[s] - ref).max() < 5e-3 * np.abs(ref).max(), s


def test_phantom_kernel_texture_matches_numpy():
    n_ang, n_det = 60, 256
    out = torch.empty((2, n_ang, n_det), device="cuda")
    gr.phantom_sinograms(out, 77, n_ang, n_det, [128.0, 127.4], 6, n_tex=48, tex_amp=0.2, tex_fmax=12.0)
    got = out.cpu().numpy()
    for s, c in enumerate([128.0, 127.4]):
        ref = phantom.sinogram(77, s, n_ang, n_det, c, 6, n_tex=48, tex_amp=0.2, tex_fmax=12.0)
        assert O.rel_l2(got[s], ref) < 1e-4, s


def test_phantom_kernel_noise_statistics():
    n_ang, n_det = 64, 512
    out = torch.empty((1, n_ang, n_det), device="cuda")
    gr.phantom_sinograms(out, 99, n_ang, n_det, [256.0], 20, photons=1e4)
    noisy = out.cpu().numpy()[0]
    clean = phantom.sinogram(99, 0, n_ang, n_det, 256.0, 20)
    lam = 1e4 * np.exp(-clean)
    z = (np.exp(-noisy) * 1e4 - lam) / np.sqrt(lam)       # standardised photon counts
    m = lam > 50
    assert abs(z[m].mean()) < 0.05 and abs(z[m].std() - 1) < 0.05


def test_cfg2_pair_full_size_parity():
    """One slice pair of the benchmark workload at full size: 1500 angles x 2048
    detectors, Poisson(1e4) noise, centre 1024 + 1.37, half-padded on chip to
    4096, 2048 x 2048 window of the 4095 x 4095 image."""
    n, nang, ctr = 2048, 1500, 1024.0 + 1.37
    sino = torch.empty((2, nang, n), device="cuda")
    gr.phantom_sinograms(sino, 2026, nang, n, [ctr, ctr], 20, photons=1e4)
    p = gr.Plan(n_ang=nang, n_det=n, center=ctr, pad="half", crop=(n // 2, n))
    assert p.info["M"] == 4096 and p.info["M0"] == 4095
    out = p.recon(sino)
    torch.cuda.synchronize()
    o = out.cpu().numpy()
    s = sino.cpu().numpy()
    g = O.Geometry(n_ang=nang, n_det=n, center=ctr).padded()
    p1, p2 = O.pad_half(s[0]), O.pad_half(s[1])
    r1, r2 = O.recon_f32(g, p1, p2)
    c = slice(n // 2, n // 2 + n)
    e1, e2 = O.rel_l2(o[0], r1[c, c]), O.rel_l2(o[1], r2[c, c])
    print(f"cfg2 pair: rel-L2 vs oracle {e1:.2e} {e2:.2e}", flush=True)
    assert e1 < 1e-5 and e2 < 1e-5
    if O.ref_available():
        q1, q2 = O.ref_recon(g, p1, p2)
        f1, f2 = O.rel_l2(o[0], q1[c, c]), O.rel_l2(o[1], q2[c, c])
        print(f"cfg2 pair: rel-L2 vs reference build {f1:.2e} {f2:.2e}", flush=True)
        assert f1 < 1e-5 and f2 < 1e-5
    p.close()


@pytest.mark.parametrize("accum", ["compensated", "plain"])
def test_interp_convolvent_matches_oracle(accum):
    """interp=1: the deapodization table is bitwise the oracle's INTERP table
    (the host computes both with the reference's float expressions) and the
    images match the oracle's INTERP restatement; they differ from the
    nearest-neighbour images by the convolvent change only."""
    g = O.Geometry(n_ang=120, n_det=256, center=127.25, interp=True)
    sino = np.stack([phantom.sinogram(8, s, 120, 256, 127.25, 10) for s in range(2)])
    p = plan_for(g, accum=accum, interp=True)
    t = p.tables()
    ot = O.tables(g)
    assert np.array_equal(t["winv"], ot["winv"]) and np.array_equal(t["wtbl"], ot["wtbl"])
    out = gpu_pair(p, sino[0], sino[1])
    r1, r2 = O.recon_f32(g, sino[0], sino[1])
    for o, r in ((out[0], r1), (out[1], r2)):
        f, d = image_errors(o, r)
        assert d < DISK_TOL and f < FULL_TOL, (d, f)
    n1, _ = O.recon_f32(O.Geometry(n_ang=120, n_det=256, center=127.25), sino[0], sino[1])
    assert O.rel_l2(out[0], n1) > 10 * O.rel_l2(out[0], r1)
    p.close()


def test_exec_pairs_host_rows_match_device_path():
    """gr_exec_pairs (host rows in, host rows out, two staging slots): 9 slices
    through a plan of 2 pairs per launch -- three chunks, both slots reused,
    an odd last slice -- from rows with a stride wider than the detector,
    bitwise the device path's images."""
    g = O.Geometry(n_ang=90, n_det=128, center=63.75)
    rng = np.random.default_rng(31)
    wide = rng.standard_normal((9, 90, 140)).astype(np.float32)
    sino = wide[:, :, :128]
    p = plan_for(g, max_pairs=2)
    host = p.exec_pairs(sino)
    dev = p.recon(torch.from_numpy(np.ascontiguousarray(sino)).cuda()).cpu().numpy()
    assert np.array_equal(host, dev)
    again = p.exec_pairs(sino, out=np.full_like(host, np.nan))  # staging reused
    assert np.array_equal(again, dev)
    p.close()
