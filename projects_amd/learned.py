"""Sparse-angle streak corrector: a small residual CNN on gridrec images.

Host side of the kernels declared in include/gridrec_cnn.h (k_cnn.hip).  The
reference has no learned component; this module adds one on top of the
gridrec path: a sparse-angle gridrec image x goes through five 3x3
convolutions (1 -> 32 -> 32 -> 32 -> 32 -> 1 channels, ReLU after the first
four) and the result is added back, y = x + scale * residual.  Trained with
Adam on pairs that gridrec itself produces (sparse-angle input, dense-angle
target).  At inference a slice whose correction is large compared with the
image itself (||residual|| / ||x/scale|| above a threshold) falls back to the
plain gridrec image.

Device memory comes from PyTorch tensors (plumbing); every forward, backward
and Adam step runs in the library's HIP kernels.
"""
from __future__ import annotations

import ctypes as C
import json
from typing import Optional

import numpy as np

from . import gridrec as gr

CH = 32
PARAM_SHAPES = [("W1", (CH, 9)), ("b1", (CH,)), ("W2", (CH, CH, 9)), ("b2", (CH,)), ("W3", (CH, CH, 9)),
                ("b3", (CH,)), ("W4", (CH, CH, 9)), ("b4", (CH,)), ("W5", (CH, 9)), ("b5", (1,))]


def _ptr(t) -> C.c_void_p:
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(None)


def init_params(seed: int) -> np.ndarray:
    """He-normal initial weights (std sqrt(2 / fan_in)), zero biases, the last
    layer scaled by 0.1 (the untrained network is close to the identity)."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shape in PARAM_SHAPES:
        if name.startswith("b"):
            parts.append(np.zeros(int(np.prod(shape)), np.float32))
            continue
        fan_in = 9 * (1 if name == "W1" else CH)
        w = rng.standard_normal(shape) * np.sqrt(2.0 / fan_in) * (0.1 if name == "W5" else 1.0)
        parts.append(w.astype(np.float32).ravel())
    return np.concatenate(parts)


def unflatten(flat: np.ndarray) -> dict:
    out, o = {}, 0
    for name, shape in PARAM_SHAPES:
        k = int(np.prod(shape))
        out[name] = np.asarray(flat[o:o + k]).reshape(shape)
        o += k
    return out


class Corrector:
    """The network's parameters, Adam state and scratch on one GPU."""

    def __init__(self, params: Optional[np.ndarray] = None, seed: int = 0, scale: float = 1.0, device: int = 0,
                 threshold: float = 1.0):
        import torch

        self.torch = torch
        self.dev = torch.device("cuda", device)
        L = gr.lib()
        self.n_params = L.gr_cnn_param_count()
        p = init_params(seed) if params is None else np.asarray(params, np.float32)
        assert p.shape == (self.n_params,)
        self.params = torch.from_numpy(p.copy()).to(self.dev)
        self.packed = torch.empty(L.gr_cnn_packed_count(), dtype=torch.int16, device=self.dev)
        self.grads = torch.zeros_like(self.params)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.loss = torch.zeros(1, dtype=torch.float32, device=self.dev)
        self.step = 0
        self.scale = float(scale)       # images are divided by scale before the network
        self.threshold = float(threshold)
        self._bufs = {}
        self.pack()

    # -- buffers ----------------------------------------------------------------
    def _buffers(self, n, H, W, train, acts=True):
        key = (n, H, W, train, acts)
        if key not in self._bufs:
            torch = self.torch
            for k in [k for k in self._bufs if k[3] == train]:  # keep one shape per mode
                del self._bufs[k]
            na = 0 if not acts else (4 if train else 2)
            acts_l = [torch.empty((n, H, W, CH), dtype=torch.int16, device=self.dev) for _ in range(na)]
            wb = gr.lib().gr_cnn_work_bytes(n, H, W) if na else gr.lib().gr_cnn_infer_work_bytes(n, H, W)
            work = torch.empty(wb, dtype=torch.uint8, device=self.dev)
            yn = torch.empty((n, H, W), dtype=torch.float32, device=self.dev) if train else None
            stats = torch.empty((n, 2), dtype=torch.float32, device=self.dev)
            self._bufs[key] = (acts_l, work, yn, stats)
        return self._bufs[key]

    def pack(self, stream=None):
        gr._check(gr.lib().gr_cnn_pack(_ptr(self.params), _ptr(self.packed), C.c_void_p(gr._stream_handle(stream))),
                  "gr_cnn_pack")

    # -- passes -----------------------------------------------------------------
    def forward(self, x, out=None, train=False, stream=None):
        """x: cuda float32 [n, H, W] gridrec images.  Returns (y, stats) with
        y = the corrected images and stats [n, 2] = (sum residual^2, sum
        (x/scale)^2) per image."""
        torch = self.torch
        assert x.is_cuda and x.dtype == torch.float32 and x.dim() == 3 and x.is_contiguous()
        n, H, W = x.shape
        acts, work, yn, stats = self._buffers(n, H, W, train)
        a = acts if train else [acts[0], acts[1], acts[0], acts[1]]
        if out is None:
            out = torch.empty_like(x)
        rc = gr.lib().gr_cnn_forward(_ptr(self.params), _ptr(self.packed), _ptr(x), n, H, W, self.scale,
                                     *[_ptr(t) for t in a], _ptr(out), _ptr(yn), _ptr(stats), _ptr(work),
                                     C.c_void_p(gr._stream_handle(stream)))
        gr._check(rc, "gr_cnn_forward")
        return out, stats

    def infer(self, x, out=None, stream=None):
        """Inference in one fused kernel, bitwise forward()'s y, with no
        activation buffers (for memory-tight callers; forward() is faster on
        MI355X at 2048^2): returns (y, stats)."""
        torch = self.torch
        assert x.is_cuda and x.dtype == torch.float32 and x.dim() == 3 and x.is_contiguous()
        n, H, W = x.shape
        _, work, _, stats = self._buffers(n, H, W, False, acts=False)
        if out is None:
            out = torch.empty_like(x)
        rc = gr.lib().gr_cnn_infer(_ptr(self.params), _ptr(self.packed), _ptr(x), n, H, W, self.scale, _ptr(out),
                                   _ptr(stats), _ptr(work), C.c_void_p(gr._stream_handle(stream)))
        gr._check(rc, "gr_cnn_infer")
        return out, stats

    def backward(self, x, target, stream=None):
        """Gradients of mean((y - target)^2 / scale^2) after forward(x, train=True)."""
        n, H, W = x.shape
        acts, work, yn, _ = self._buffers(n, H, W, True)
        rc = gr.lib().gr_cnn_backward(_ptr(self.params), _ptr(self.packed), _ptr(x), _ptr(target), n, H, W,
                                      self.scale, *[_ptr(t) for t in acts], _ptr(yn), _ptr(self.grads),
                                      _ptr(self.loss), _ptr(work), C.c_void_p(gr._stream_handle(stream)))
        gr._check(rc, "gr_cnn_backward")

    def adam(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, stream=None):
        self.step += 1
        rc = gr.lib().gr_cnn_adam(_ptr(self.params), _ptr(self.grads), _ptr(self.m), _ptr(self.v), self.n_params,
                                  self.step, lr, beta1, beta2, eps, C.c_void_p(gr._stream_handle(stream)))
        gr._check(rc, "gr_cnn_adam")
        self.pack(stream)

    def train_step(self, x, target, lr=1e-3, stream=None):
        """One Adam step on a batch; returns the loss tensor (device, before the step)."""
        self.forward(x, train=True, stream=stream)
        self.backward(x, target, stream=stream)
        self.adam(lr=lr, stream=stream)
        return self.loss

    def correct(self, x, out=None, stream=None):
        """Inference with the fallback: images whose relative correction norm
        ||residual|| / ||x/scale|| exceeds the threshold keep the plain image.
        Returns (y, used) with used[i] True where the network's output was kept."""
        torch = self.torch
        y, stats = self.forward(x, out=out, stream=stream)
        s = torch.cuda.current_stream() if stream is None else stream
        with torch.cuda.stream(s):  # device-side selection, no host synchronisation
            rel = torch.sqrt(stats[:, 0] / torch.clamp(stats[:, 1], min=1e-30))
            used = rel <= self.threshold
            torch.where(used.view(-1, 1, 1), y, x, out=y)
        return y, used

    # -- persistence --------------------------------------------------------------
    def state(self) -> dict:
        return {"params": self.params.cpu().numpy(), "m": self.m.cpu().numpy(), "v": self.v.cpu().numpy(),
                "step": self.step, "scale": self.scale, "threshold": self.threshold}

    def save(self, path: str, meta: Optional[dict] = None):
        st = self.state()
        np.savez(path, params=st["params"], m=st["m"], v=st["v"],
                 meta=np.frombuffer(json.dumps({"step": st["step"], "scale": st["scale"],
                                                "threshold": st["threshold"], **(meta or {})}).encode(), np.uint8))

    @classmethod
    def load(cls, path: str, device: int = 0) -> "Corrector":
        z = np.load(path, allow_pickle=False)
        meta = json.loads(bytes(z["meta"]).decode())
        c = cls(params=z["params"], scale=meta["scale"], device=device, threshold=meta["threshold"])
        c.m.copy_(c.torch.from_numpy(z["m"]))
        c.v.copy_(c.torch.from_numpy(z["v"]))
        c.step = int(meta["step"])
        c.meta = meta
        return c


# -- data and training ------------------------------------------------------------

def gridrec_images(sino, n_det: int, center: float, stream=None):
    """Plain gridrec of [n, n_ang, n_det] sinograms under the half-padding caller
    convention: [n, n_det, n_det] images (the detector's field of view)."""
    import torch

    n, n_ang, _ = sino.shape
    M0 = (1 << (2 * n_det - 1).bit_length()) - 1
    cn = int(round(n_det * M0 / (2 * n_det)))
    plan = gr.Plan(n_ang=n_ang, n_det=n_det, center=center, pad="half", crop=((M0 - cn + 1) // 2, cn))
    out = torch.empty((n, cn, cn), dtype=torch.float32, device=sino.device)
    plan.recon(sino, out, stream=stream)
    torch.cuda.synchronize()
    plan.close()
    return out


def make_pairs(n: int, size: int = 512, sparse: int = 90, dense: int = 900, seed: int = 1, n_ell: int = 20,
               n_tex: int = 64, tex_amp: float = 0.15, tex_fmax: float = 24.0, slice0: int = 0, chunk: int = 512):
    """n training pairs of size x size gridrec images of the same seeded
    phantoms (ellipses + texture, noise-free): sparse-angle input and
    dense-angle target, both on the GPU."""
    import torch

    inp = torch.empty((n, size, size), dtype=torch.float32, device="cuda")
    tgt = torch.empty_like(inp)
    c = size / 2.0
    for s0 in range(0, n, chunk):
        k = min(chunk, n - s0)
        for n_ang, dst in ((sparse, inp), (dense, tgt)):
            sino = torch.empty((k, n_ang, size), dtype=torch.float32, device="cuda")
            gr.phantom_sinograms(sino, seed, n_ang, size, [c] * k, n_ell, slice0=slice0 + s0, n_tex=n_tex,
                                 tex_amp=tex_amp, tex_fmax=tex_fmax)
            dst[s0:s0 + k] = gridrec_images(sino, size, c)
            del sino
    torch.cuda.synchronize()
    return inp, tgt


def train(net: Corrector, inp, tgt, steps: int = 0, seconds: float = 0.0, batch: int = 32, lr: float = 1e-3,
          seed: int = 0, log_every: int = 50, progress=None, lr_final: float = 0.05):
    """Adam steps on random batches of (inp, tgt); stops after `steps` steps or
    `seconds` of wall time, whichever is set (both: the first reached).  The
    learning rate follows a cosine from lr down to lr * lr_final over the
    budget -- the step budget when `steps` is set (the time then only caps
    the run, so the schedule, and with it the trained network, does not
    depend on timing), else the time budget -- so the weights the run ends
    with are a settled iterate, not one of the noisy ones a constant rate
    leaves (lr_final = 1: constant).
    Returns the loss history [(step, loss)] sampled every log_every steps."""
    import math
    import time

    import torch

    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    n = inp.shape[0]
    hist = []
    t0 = time.perf_counter()
    k = 0
    while True:
        idx = torch.randint(0, n, (batch,), device="cuda", generator=g)
        xb, tb = inp.index_select(0, idx).contiguous(), tgt.index_select(0, idx).contiguous()
        f = k / steps if steps else ((time.perf_counter() - t0) / seconds if seconds else 0.0)
        rate = lr * (lr_final + (1.0 - lr_final) * 0.5 * (1.0 + math.cos(math.pi * min(f, 1.0))))
        loss = net.train_step(xb, tb, lr=rate)
        k += 1
        if k % log_every == 0 or k == 1:
            hist.append((k, float(loss.cpu())))
            if progress and k % (log_every * 10) == 0:
                progress(f"step {k}, loss {hist[-1][1]:.4e}")
        if (steps and k >= steps) or (seconds and time.perf_counter() - t0 >= seconds):
            break
    torch.cuda.synchronize()
    return hist, k, time.perf_counter() - t0


def psnr(img, ref) -> float:
    """PSNR (dB) of img against ref, peak = the dynamic range of ref."""
    import torch

    mse = torch.mean((img.double() - ref.double()) ** 2)
    rng = ref.max().double() - ref.min().double()
    return float(10.0 * torch.log10(rng * rng / mse))
