"""Seeded synthetic sinograms: random ellipse phantoms, analytic Radon transform,
Poisson photon noise.  CPU (numpy) implementation; ``projects_amd.gridrec``
exposes the GPU kernel that produces the same sinograms at scale.

The reference has no data generator (its drivers read HDF5/netCDF files); this
module provides the synthetic workloads the benchmarks are quoted on.

Random numbers come from Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel
random numbers: as easy as 1, 2, 3", SC'11), a counter-based generator: every
value is a pure function of (key, counter), so the CPU and GPU generators agree
on every ellipse parameter and every noise draw without sharing state.

Counter layout (c0, c1, c2, c3), key = (seed low 32 bits, seed high 32 bits):

* ellipse e of slice s:  (e, s, 0, 0) -> a, b, density, rotation;
                         (e, s, 1, 0) -> centre radius, centre angle
* texture wave e of slice s: (e, s, 3, 0) -> |f|, direction, phase;
                             (e, s, 4, 0) -> Gaussian amplitude (Box-Muller)
* noise of detector k, angle n, slice s, round r: (k, n, s, 2 + r)

Geometry: the unit disk maps to n_det/2 detector pixels around ``center``;
angle n is n*pi/n_ang (the reference's geom 1).  The Radon transform of an
ellipse with semi-axes (a, b), rotation phi, centre (x0, y0), density rho is
p(t, theta) = 2 rho a b sqrt(s^2 - t'^2) / s^2 for |t'| < s, with
s^2 = a^2 cos^2(theta-phi) + b^2 sin^2(theta-phi) and
t' = t - (x0 cos theta + y0 sin theta).

Texture (a band-limited random field inside the unit disk): a sum of n_tex
plane waves amp_e cos(2 pi f_e.r + phase_e), frequencies uniform over the disk
|f| <= fmax (cycles per unit length), amplitudes N(0, tex_amp^2 / n_tex).
Along the chord at offset t (half-length h = sqrt(1 - t^2)) a wave integrates
to 2 amp cos(alpha) sin(beta h) / beta with alpha = 2 pi t f.u + phase,
beta = 2 pi f.u_perp, u = (cos theta, sin theta).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

PHILOX_M0 = np.uint64(0xD2511F53)
PHILOX_M1 = np.uint64(0xCD9E8D57)
PHILOX_W0 = np.uint32(0x9E3779B9)
PHILOX_W1 = np.uint32(0xBB67AE85)
_MASK32 = np.uint64(0xFFFFFFFF)

DENSITY_UNIFORM = 0   # density U[-0.3, 1.0]
DENSITY_LOG4 = 1      # density log-uniform over 4 decades, [1e-4, 1]


def philox4x32(c0, c1, c2, c3, k0: int, k1: int):
    """Philox4x32-10 on arrays of counters; returns four uint32 arrays."""
    c0 = np.asarray(c0, np.uint32).astype(np.uint64)
    c1 = np.asarray(c1, np.uint32).astype(np.uint64)
    c2 = np.asarray(c2, np.uint32).astype(np.uint64)
    c3 = np.asarray(c3, np.uint32).astype(np.uint64)
    c0, c1, c2, c3 = np.broadcast_arrays(c0, c1, c2, c3)
    c0, c1, c2, c3 = c0.copy(), c1.copy(), c2.copy(), c3.copy()
    k0 = np.uint64(k0 & 0xFFFFFFFF)
    k1 = np.uint64(k1 & 0xFFFFFFFF)
    for _ in range(10):
        p0 = PHILOX_M0 * c0
        p1 = PHILOX_M1 * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & _MASK32
        hi1, lo1 = p1 >> np.uint64(32), p1 & _MASK32
        n0 = hi1 ^ c1 ^ k0
        n1 = lo1
        n2 = hi0 ^ c3 ^ k1
        n3 = lo0
        c0, c1, c2, c3 = n0, n1, n2, n3
        k0 = (k0 + np.uint64(PHILOX_W0)) & _MASK32
        k1 = (k1 + np.uint64(PHILOX_W1)) & _MASK32
    return (c0.astype(np.uint32), c1.astype(np.uint32), c2.astype(np.uint32), c3.astype(np.uint32))


def u01(x) -> np.ndarray:
    """uint32 -> float64 uniform in (0, 1): (x >> 8 + 0.5) * 2^-24 (exact in float32)."""
    return ((np.asarray(x, np.uint32) >> np.uint32(8)).astype(np.float64) + 0.5) * (2.0 ** -24)


@dataclass
class Ellipses:
    a: np.ndarray
    b: np.ndarray
    rho: np.ndarray
    phi: np.ndarray
    x0: np.ndarray
    y0: np.ndarray


def ellipses(seed: int, slice_idx: int, n_ell: int, density: int = DENSITY_UNIFORM) -> Ellipses:
    """Parameters of the n_ell ellipses of one slice (float32 values)."""
    k0, k1 = seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF
    e = np.arange(n_ell, dtype=np.uint32)
    r = philox4x32(e, slice_idx, 0, 0, k0, k1)
    q = philox4x32(e, slice_idx, 1, 0, k0, k1)
    u = [u01(x) for x in r]
    v = [u01(x) for x in q]
    f = np.float32
    a = (0.05 + 0.4 * u[0]).astype(f)
    b = (0.05 + 0.4 * u[1]).astype(f)
    if density == DENSITY_LOG4:
        rho = (10.0 ** (-4.0 + 4.0 * u[2])).astype(f)
    else:
        rho = (-0.3 + 1.3 * u[2]).astype(f)
    phi = (math.pi * u[3]).astype(f)
    rad = np.sqrt(v[0])
    ang = 2.0 * math.pi * v[1]
    x0 = (rad * np.cos(ang)).astype(f)
    y0 = (rad * np.sin(ang)).astype(f)
    return Ellipses(a, b, rho, phi, x0, y0)


def radon_ellipses(E: Ellipses, n_ang: int, n_det: int, center: float) -> np.ndarray:
    """Analytic line integrals (unit-disk length units) on the geom-1 grid, float64."""
    th = np.arange(n_ang, dtype=np.float64) * (math.pi / n_ang)
    t = (np.arange(n_det, dtype=np.float64) - float(center)) / (n_det / 2.0)
    ct, st = np.cos(th)[:, None], np.sin(th)[:, None]
    p = np.zeros((n_ang, n_det))
    for i in range(len(E.a)):
        a, b, rho, phi = float(E.a[i]), float(E.b[i]), float(E.rho[i]), float(E.phi[i])
        x0, y0 = float(E.x0[i]), float(E.y0[i])
        c = np.cos(th - phi)[:, None]
        s = np.sin(th - phi)[:, None]
        s2 = a * a * c * c + b * b * s * s
        tp = t[None, :] - (x0 * ct + y0 * st)
        d = s2 - tp * tp
        p += np.where(d > 0, 2.0 * rho * a * b * np.sqrt(np.maximum(d, 0.0)) / s2, 0.0)
    return p


@dataclass
class Waves:
    amp: np.ndarray
    fx: np.ndarray
    fy: np.ndarray
    phase: np.ndarray


def waves(seed: int, slice_idx: int, n_tex: int, amp: float, fmax: float) -> Waves:
    """Texture waves of one slice (float32 values)."""
    k0, k1 = seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF
    e = np.arange(n_tex, dtype=np.uint32)
    r = [u01(x) for x in philox4x32(e, slice_idx, 3, 0, k0, k1)]
    q = [u01(x) for x in philox4x32(e, slice_idx, 4, 0, k0, k1)]
    f = np.float32
    fr = fmax * np.sqrt(r[0])
    d = 2.0 * math.pi * r[1]
    z = np.sqrt(-2.0 * np.log(q[0])) * np.cos(2.0 * math.pi * q[1])
    return Waves((amp * z / math.sqrt(n_tex)).astype(f), (fr * np.cos(d)).astype(f), (fr * np.sin(d)).astype(f),
                 (2.0 * math.pi * r[2]).astype(f))


def radon_waves(T: Waves, n_ang: int, n_det: int, center: float) -> np.ndarray:
    """Analytic line integrals of the texture on the geom-1 grid, float64."""
    th = np.arange(n_ang, dtype=np.float64) * (math.pi / n_ang)
    t = (np.arange(n_det, dtype=np.float64) - float(center)) / (n_det / 2.0)
    ct, st = np.cos(th)[:, None], np.sin(th)[:, None]
    inside = np.abs(t) < 1.0
    h = np.sqrt(np.where(inside, 1.0 - t * t, 0.0))[None, :]
    p = np.zeros((n_ang, n_det))
    for i in range(len(T.amp)):
        fx, fy = float(T.fx[i]), float(T.fy[i])
        alpha = 2.0 * math.pi * t[None, :] * (fx * ct + fy * st) + float(T.phase[i])
        beta = 2.0 * math.pi * (fy * ct - fx * st)
        bh = beta * h
        sinc = np.where(np.abs(bh) > 1e-4, np.sin(bh) / np.where(bh == 0, 1.0, bh), 1.0 - bh * bh / 6.0)
        p += float(T.amp[i]) * 2.0 * h * np.cos(alpha) * sinc
    return np.where(inside[None, :], p, 0.0)


def _ptrs_constants(lam):
    slam = np.sqrt(lam)
    b = 0.931 + 2.53 * slam
    a = -0.059 + 0.02483 * b
    inv_alpha = 1.1239 + 1.1328 / (b - 3.4)
    vr = 0.9277 - 3.6224 / (b - 2.0)
    return a, b, inv_alpha, vr


def log_poisson_pmf(k: np.ndarray, lam: np.ndarray) -> np.ndarray:
    """k log(lam) - lam - log(k!), through the Stirling series of log Gamma(k+1)
    for k >= 9 (the form the GPU generator evaluates in float)."""
    from math import lgamma

    k = np.asarray(k, np.float64)
    lam = np.asarray(lam, np.float64)
    z = k + 1.0
    big = k >= 9
    out = np.empty(np.broadcast(k, lam).shape)
    zb, lb, kb = np.broadcast_to(z, out.shape)[big], np.broadcast_to(lam, out.shape)[big], np.broadcast_to(k, out.shape)[big]
    out[big] = (kb * np.log1p((lb - zb) / zb) + (zb - lb) - 0.5 * np.log(2 * math.pi * zb) - 1 / (12 * zb)
                + 1 / (360 * zb ** 3))
    ks, ls = np.broadcast_to(k, out.shape)[~big], np.broadcast_to(lam, out.shape)[~big]
    out[~big] = ks * np.log(ls) - ls - np.array([lgamma(x + 1.0) if x >= 0 else 0.0 for x in ks.ravel()])
    return out


def poisson(lam: np.ndarray, seed: int, slice_idx: int, max_rounds: int = 64) -> np.ndarray:
    """Poisson(lam) draws for an (n_ang, n_det) array of means.

    lam >= 10: transformed rejection with squeeze, PTRS (W. Hoermann, "The
    transformed rejection method for generating Poisson random variables",
    Insurance: Mathematics and Economics 12, 1993), two uniforms per round.
    lam < 10: inversion by sequential search on one uniform.
    Round r of detector k, angle n uses counter (k, n, slice, 2 + r//2);
    the even round takes words 0-1, the odd round words 2-3.
    """
    lam = np.asarray(lam, np.float64)
    n_ang, n_det = lam.shape
    k0, k1 = seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF
    kk, nn = np.meshgrid(np.arange(n_det, dtype=np.uint32), np.arange(n_ang, dtype=np.uint32))
    out = np.zeros(lam.shape, np.float64)

    small = lam < 10.0
    if small.any():
        w = philox4x32(kk[small], nn[small], slice_idx, 2, k0, k1)
        u = u01(w[0])
        ls = lam[small]
        k = np.zeros_like(ls)
        p = np.exp(-ls)
        F = p.copy()
        active = u > F
        it = 0
        while active.any() and it < 1000:
            k[active] += 1
            p[active] *= ls[active] / k[active]
            F[active] += p[active]
            active = active & (u > F)
            it += 1
        out[small] = k

    big = ~small
    if big.any():
        lb = lam[big]
        a, b, inv_alpha, vr = _ptrs_constants(lb)
        res = np.full(lb.shape, -1.0)
        kb, nb = kk[big], nn[big]
        for r in range(max_rounds):
            act = res < 0
            if not act.any():
                break
            w = philox4x32(kb[act], nb[act], slice_idx, 2 + r // 2, k0, k1)
            if r % 2 == 0:
                U, V = u01(w[0]) - 0.5, u01(w[1])
            else:
                U, V = u01(w[2]) - 0.5, u01(w[3])
            la, aa, ba, ia, va = lb[act], a[act], b[act], inv_alpha[act], vr[act]
            us = 0.5 - np.abs(U)
            k = np.floor((2.0 * aa / us + ba) * U + la + 0.43)
            acc = (us >= 0.07) & (V <= va)
            rej = (k < 0) | ((us < 0.013) & (V > us))
            test = (np.log(V) + np.log(ia) - np.log(aa / (us * us) + ba)) <= log_poisson_pmf(np.maximum(k, 0), la)
            ok = acc | (~rej & test)
            cur = res[act]
            cur[ok] = k[ok]
            res[act] = cur
        res[res < 0] = np.round(lb[res < 0])
        out[big] = res
    return out


def gaussian(seed: int, slice_idx: int, n_ang: int, n_det: int) -> np.ndarray:
    """Standard normal draws per cell: Box-Muller on words 0-1 of counter
    (k, n, slice, 1 << 16) (a counter word no Poisson round reaches)."""
    k0, k1 = seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF
    kk, nn = np.meshgrid(np.arange(n_det, dtype=np.uint32), np.arange(n_ang, dtype=np.uint32))
    w = philox4x32(kk, nn, slice_idx, 1 << 16, k0, k1)
    u1, u2 = u01(w[0]), u01(w[1])
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


def sinogram(seed: int, slice_idx: int, n_ang: int, n_det: int, center: float, n_ell: int,
             photons: float = 0.0, density: int = DENSITY_UNIFORM, gauss_rel: float = 0.0, n_tex: int = 0,
             tex_amp: float = 0.0, tex_fmax: float = 0.0) -> np.ndarray:
    """One float32 sinogram (n_ang x n_det).

    photons > 0: I ~ Poisson(photons * exp(-p)), value -log(max(I,1)/photons).
    gauss_rel > 0: multiplicative Gaussian noise, p * (1 + gauss_rel * z) with
    z ~ N(0,1) per cell (the precision study's "1% Gaussian noise").
    """
    E = ellipses(seed, slice_idx, n_ell, density)
    p = radon_ellipses(E, n_ang, n_det, center)
    if n_tex > 0:
        p = p + radon_waves(waves(seed, slice_idx, n_tex, tex_amp, tex_fmax), n_ang, n_det, center)
    if photons > 0:
        I = poisson(photons * np.exp(-p), seed, slice_idx)
        p = -np.log(np.maximum(I, 1.0) / photons)
    if gauss_rel > 0:
        p = p * (1.0 + gauss_rel * gaussian(seed, slice_idx, n_ang, n_det))
    return p.astype(np.float32)


def phantom_image(seed: int, slice_idx: int, n_ell: int, n: int, density: int = DENSITY_UNIFORM, n_tex: int = 0,
                  tex_amp: float = 0.0, tex_fmax: float = 0.0, device=None):
    """The phantom itself on an n x n grid covering [-1, 1]^2 (pixel centres),
    row index = y... returned as img[x_index][y_index] to match gridrec's
    S[j][k] with j along cos (x) and k along sin (y).  With a torch device the
    same float64 expressions are evaluated there and a float64 tensor is
    returned (large images for the benchmark's comparisons)."""
    E = ellipses(seed, slice_idx, n_ell, density)
    if device is not None:
        import torch

        xp, kw = torch, dict(dtype=torch.float64, device=device)
        c = (torch.arange(n, **kw) - (n - 1) / 2.0) / (n / 2.0)
        X, Y = torch.meshgrid(c, c, indexing="ij")
        img = torch.zeros((n, n), **kw)
        where = lambda m, v: torch.where(m, torch.tensor(v, **kw), torch.tensor(0.0, **kw))
    else:
        xp = np
        c = (np.arange(n) - (n - 1) / 2.0) / (n / 2.0)
        X, Y = np.meshgrid(c, c, indexing="ij")
        img = np.zeros((n, n))
        where = lambda m, v: np.where(m, v, 0.0)
    for i in range(len(E.a)):
        ca, sa = math.cos(float(E.phi[i])), math.sin(float(E.phi[i]))
        dx, dy = X - float(E.x0[i]), Y - float(E.y0[i])
        xr = dx * ca + dy * sa
        yr = -dx * sa + dy * ca
        img += where((xr / float(E.a[i])) ** 2 + (yr / float(E.b[i])) ** 2 <= 1.0, float(E.rho[i]))
    if n_tex > 0:
        T = waves(seed, slice_idx, n_tex, tex_amp, tex_fmax)
        tex = sum(float(T.amp[i]) * xp.cos(2.0 * math.pi * (float(T.fx[i]) * X + float(T.fy[i]) * Y) +
                                           float(T.phase[i])) for i in range(n_tex))
        img += xp.where(X * X + Y * Y <= 1.0, tex, xp.zeros_like(tex))
    return img
