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
    def _buffers(self, n, H, W, train):
        key = (n, H, W, train)
        if key not in self._bufs:
            torch = self.torch
            for k in [k for k in self._bufs if k[3] == train]:  # keep one shape per mode
                del self._bufs[k]
            na = 4 if train else 2
            acts = [torch.empty((n, H, W, CH), dtype=torch.int16, device=self.dev) for _ in range(na)]
            work = torch.empty(gr.lib().gr_cnn_work_bytes(n, H, W), dtype=torch.uint8, device=self.dev)
            yn = torch.empty((n, H, W), dtype=torch.float32, device=self.dev) if train else None
            stats = torch.empty((n, 2), dtype=torch.float32, device=self.dev)
            self._bufs[key] = (acts, work, yn, stats)
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
        rel = torch.sqrt(stats[:, 0] / torch.clamp(stats[:, 1], min=1e-30))
        used = rel <= self.threshold
        if not bool(used.all()):
            y[~used] = x[~used]
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
