"""CPU pins of the synthetic-data generator (projects_amd/phantom.py).

The analytic projections are checked against brute-force quadrature of the
phantom image along the same lines, so a wrong sign, offset or scale in either
closed form fails here.
"""
import math

import numpy as np
import pytest

from projects_amd import phantom


def line_integral(f, t, theta, m=20001, h=None):
    """Midpoint rule of f(x, y) along the line at offset t, over |s| <= h
    (default: the chord of the unit disk)."""
    h = math.sqrt(max(0.0, 1.0 - t * t)) if h is None else h
    s = (np.arange(m) + 0.5) / m * 2 * h - h
    x = t * math.cos(theta) - s * math.sin(theta)
    y = t * math.sin(theta) + s * math.cos(theta)
    return float(np.sum(f(x, y)) * (2 * h / m))


def test_texture_projection_is_the_line_integral():
    T = phantom.waves(5, 3, 16, 0.3, 9.0)

    def f(x, y):
        return sum(float(T.amp[i]) * np.cos(2 * math.pi * (float(T.fx[i]) * x + float(T.fy[i]) * y)
                                            + float(T.phase[i])) for i in range(16))

    n_ang, n_det, c = 7, 40, 19.6
    p = phantom.radon_waves(T, n_ang, n_det, c)
    for n in (0, 2, 5):
        for k in (1, 9, 19, 30, 38):
            t = (k - c) / (n_det / 2)
            assert abs(p[n, k] - line_integral(f, t, n * math.pi / n_ang)) < 1e-6, (n, k)


def test_ellipse_projection_is_the_line_integral():
    E = phantom.ellipses(11, 2, 4)

    def f(x, y):
        out = np.zeros_like(x)
        for i in range(4):
            ca, sa = math.cos(float(E.phi[i])), math.sin(float(E.phi[i]))
            dx, dy = x - float(E.x0[i]), y - float(E.y0[i])
            xr, yr = dx * ca + dy * sa, -dx * sa + dy * ca
            out += np.where((xr / float(E.a[i])) ** 2 + (yr / float(E.b[i])) ** 2 <= 1, float(E.rho[i]), 0.0)
        return out

    n_ang, n_det, c = 9, 48, 24.3
    p = phantom.radon_ellipses(E, n_ang, n_det, c)
    for n in (0, 4, 7):
        for k in (3, 17, 24, 31, 44):
            t = (k - c) / (n_det / 2)
            # ellipses reach beyond the unit disk: integrate over the whole line
            ref = line_integral(f, t, n * math.pi / n_ang, m=400001, h=2.0)
            assert abs(p[n, k] - ref) < 1e-4, (n, k, p[n, k], ref)


def test_texture_amplitude_statistics():
    T = phantom.waves(1, 0, 20000, 0.5, 10.0)
    assert abs(np.sqrt(np.sum(T.amp.astype(np.float64) ** 2)) - 0.5) < 0.02
    f = np.hypot(T.fx, T.fy)
    assert f.max() <= 10.0 and abs(np.mean(f <= 5.0) - 0.25) < 0.02  # uniform over the disk


def test_phantom_image_torch_matches_numpy():
    """The torch evaluation of the phantom image (used for large benchmark
    images) is the numpy one up to the last bits of cos."""
    torch = pytest.importorskip("torch")
    a = phantom.phantom_image(5, 3, 20, 128, n_tex=64, tex_amp=0.15, tex_fmax=24.0)
    b = phantom.phantom_image(5, 3, 20, 128, n_tex=64, tex_amp=0.15, tex_fmax=24.0, device="cpu")
    assert b.dtype == torch.float64 and np.abs(a - b.numpy()).max() < 1e-12
