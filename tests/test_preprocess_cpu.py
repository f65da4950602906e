"""Pins of the oracle's restatement of the callers' preprocessing
(oracle.preprocess: flat/dark normalisation, tomo_recon.cpp:1877 and
find_shift.cpp:1685-1697; sub-pixel shift, tomo_recon.cpp:3708-3780), by
closed forms: flat and dark fields map to 0 and 1, the log form inverts the
Beer-Lambert counts, a zero shift is the identity, an integer shift moves
cells with edge extension, and a linear row is shifted exactly.
"""
import numpy as np
import pytest

import oracle as O


def fields(n, rng):
    dark = rng.uniform(90, 110, n).astype(np.float32)
    w1 = rng.uniform(30000, 40000, n).astype(np.float32)
    w2 = (w1 + rng.uniform(-500, 500, n)).astype(np.float32)
    return dark, w1, w2


def test_flat_and_dark_map_to_zero_and_one():
    rng = np.random.default_rng(0)
    n = 64
    dark, w1, _ = fields(n, rng)
    w = np.round(w1).astype(np.float32)
    raw = np.stack([w.astype(np.uint16), np.round(dark).astype(np.uint16)])
    d = np.round(dark).astype(np.float32)
    out = O.preprocess(1, raw, d, w, w)
    assert np.array_equal(out[0], np.zeros(n, np.float32)) and np.array_equal(out[1], np.ones(n, np.float32))


def test_log_form_inverts_beer_lambert():
    rng = np.random.default_rng(1)
    n = 200
    dark, w1, w2 = fields(n, rng)
    mu = rng.uniform(0.01, 3.0, (5, n))
    raw = np.round(dark + (0.5 * (w1 + w2) - dark) * np.exp(-mu)).astype(np.uint16)
    v = O.preprocess(2, raw, dark, w1, w2).astype(np.float64)
    back = dark + (0.5 * (w1.astype(np.float64) + w2) - dark) * np.exp(-v)
    assert np.max(np.abs(back - raw) / raw) < 1e-6
    low = raw.copy()
    low[0, :5] = np.floor(dark[:5]).astype(np.uint16) - 1   # counts below dark
    assert np.array_equal(O.preprocess(2, low, dark, w1, w2)[0, :5], np.zeros(5, np.float32))


def test_zero_and_integer_shifts():
    rng = np.random.default_rng(2)
    n = 50
    dark, w1, w2 = fields(n, rng)
    raw = rng.integers(1000, 30000, (3, n)).astype(np.uint16)
    base = O.preprocess(1, raw, dark, w1, w2)
    for s in (-3, 2, 7):
        out = O.preprocess(1, raw, dark, w1, w2, xoffset=float(s))
        idx = np.clip(np.arange(n) - s, 0, n - 1)
        assert np.array_equal(out, base[:, idx]), s


@pytest.mark.parametrize("xoff", [0.25, -1.6, 3.5])
def test_linear_row_shifts_exactly(xoff):
    n = 40
    dark = np.zeros(n, np.float32)
    w = np.full(n, 40000.0, np.float32)
    k = np.arange(n)
    raw = (1000 + 500 * k)[None, :].astype(np.uint16)
    v = O.preprocess(1, raw, dark, w, w, xoffset=xoff)[0].astype(np.float64)
    lin = lambda x: 1.0 - (1000 + 500 * x) / 40000.0
    interior = (k - xoff >= 0) & (k - xoff <= n - 1)
    assert np.max(np.abs(v[interior] - lin(k[interior] - xoff))) < 1e-6
    edge = k - xoff < 0
    assert np.allclose(v[edge], lin(0.0), atol=1e-6)   # edge extension


def test_entropy_closed_forms():
    assert O.entropy(np.zeros(1000, np.float32), -1.0, 1.0, 10) == 0.0           # one bin
    half = np.r_[np.full(500, 0.05), np.full(500, 0.95)].astype(np.float32)
    assert abs(O.entropy(half, 0.0, 1.0, 10) - 1.0) < 1e-6                         # two equal bins
    u = (np.arange(4096, dtype=np.float32) + 0.5) / 4096
    assert abs(O.entropy(u, 0.0, 1.0, 64) - 6.0) < 1e-5                            # uniform over 64 bins
    # outside the range not counted but still in the total; the upper edge
    # counts in bin n_bins; values within one bin below the lower edge truncate
    # into bin 0, as (int) does: three singleton bins out of five values
    x = np.array([0.5, 2.0, -5.0, 1.0, -0.05], np.float32)
    assert abs(O.entropy(x, 0.0, 1.0, 10) - 3 * (-0.2 * np.log2(0.2))) < 1e-6


def test_ring_correction_closed_forms():
    """RingCorrectionSinogram restatement: a sinogram of identical linear rows
    has no ring and is left as is; one bright column on a flat background is
    spread over its neighbours as the 3-tap low pass prescribes; values pushed
    below zero are clamped."""
    n_ang, n_det = 30, 40
    ramp = np.tile(1.0 + 0.01 * np.arange(n_det, dtype=np.float32), (n_ang, 1))
    assert np.max(np.abs(O.ring_correct(ramp, 1.0) - ramp)) < 1e-5
    B, delta, k = 2.0, 0.3, 17
    s = np.full((n_ang, n_det), B, np.float32)
    s[:, k] += delta
    one = O.ring_correct(s, 1.0, iterations=1)
    expect = np.full(n_det, B, np.float64)
    expect[k - 1:k + 2] = B + delta / 3
    assert np.max(np.abs(one[0] - expect)) < 1e-5 and np.allclose(one, one[0][None, :])
    clamp = O.ring_correct(s, 100.0, iterations=1)
    assert clamp[0, k] == 0.0 and clamp[0, k - 1] > B and clamp[0, k + 1] > B
