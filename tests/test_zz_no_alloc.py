"""gr_recon_device_noalloc, through the C-ABI: the entry point that allocates
nothing (its kernels' scratch is the library's static per-GPU array) gives
bitwise the images of gr_recon_device on the same plan and input -- on two
calls in a row, the second with a smaller input, on a second geometry that
lays the shared scratch out differently, and on a geometry where the scratch
holds fewer pairs per launch than the plan takes -- and refuses a geometry
whose single pair does not fit."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if torch.cuda.is_available():
    from projects_amd import gridrec as gr


def sinos(seed, ns, n_ang, n_det, center):
    s = torch.empty((ns, n_ang, n_det), dtype=torch.float32, device="cuda")
    gr.phantom_sinograms(s, seed, n_ang, n_det, [center] * ns, 12, photons=1e4)
    return s


def check(plan, sino):
    ref = plan.recon(sino)
    out = torch.full_like(ref, float("nan"))
    plan.recon_noalloc(sino, out)
    torch.cuda.synchronize()
    assert torch.isfinite(ref).all()
    assert torch.equal(out, ref)


def test_noalloc_matches_device_entry_point():
    assert gr.lib().gr_noalloc_scratch_bytes() == 128 << 20
    p = gr.Plan(n_ang=180, n_det=256, center=128.4, pad="half", crop=(128, 256))
    check(p, sinos(1, 6, 180, 256, 128.4))   # first call: 3 pairs
    check(p, sinos(2, 3, 180, 256, 128.4))   # second call, smaller input: 2 pairs, the last slice alone
    p.close()
    q = gr.Plan(n_ang=90, n_det=128, center=63.0, filter="hann")  # another layout of the same scratch
    check(q, sinos(3, 2, 90, 128, 63.0))
    q.close()


def test_noalloc_fewer_pairs_per_launch_than_the_plan():
    """512 detectors half-padded, 720 angles: 8 pairs per launch need more than
    the scratch, so the entry point launches fewer at a time."""
    p = gr.Plan(n_ang=720, n_det=512, center=255.0, pad="half", crop=(256, 512), max_pairs=8)
    check(p, sinos(4, 16, 720, 512, 255.0))
    p.close()


def test_noalloc_refuses_a_pair_larger_than_the_scratch():
    p = gr.Plan(n_ang=1500, n_det=2048, center=1024.0, pad="half", crop=(1024, 2048), max_pairs=1)
    sino = torch.zeros((2, 1500, 2048), dtype=torch.float32, device="cuda")
    out = torch.zeros((2, 2048, 2048), dtype=torch.float32, device="cuda")
    with pytest.raises(gr.GridrecError) as e:
        p.recon_noalloc(sino, out)
    assert e.value.code == gr.GR_EINVAL
    p.close()
