"""CPU checks of the oracle's INTERP convolvent (or_params.interp).

The reference's compile-time INTERP variant (gridrec/GridrecCPP/Gridrec5.2/
grid.cpp:18-24, 487-489) does not compile as written -- the macro's argument
is not parenthesised, so Cnvlvnt(abs(V-iv)*tblspcg) subscripts wtbl with a
float -- so no build of it can pin the restatement.  What pins it here is
mathematics: read back through the deapodization table (winv[linv +- i] =
+-norm / Cnvlvnt(i * fac), grid.cpp:497-510), the convolvent must be the
second-order *backward* extrapolation wtbl[i] + f (wtbl[i] - wtbl[i-1]) of the
smooth PSWF table: against a cubic spline through the table its error must
match -f (1 + f) / 2 times the table's second difference in size and sign (a
forward difference gives the opposite sign, a dropped fraction term an error
of first order), and it must be far closer to the spline than the nearest-
neighbour convolvent.
"""
import numpy as np
import pytest
from scipy.interpolate import CubicSpline

import oracle as O


def convolvent_from_winv(g):
    t = O.tables(g)
    s = O.sizes(g)
    linv = s["M02"]
    wi = t["winv"].astype(np.float64)
    norm = wi[linv]
    i = np.arange(1, linv + 1)
    cnv = norm * np.where(i % 2 == 1, -1.0, 1.0) / wi[linv + i]
    fac = np.float32(np.float32(g.ltbl) / (np.float64(linv) + 0.5))
    X = (i.astype(np.float32) * fac).astype(np.float64)
    return X, cnv, t["wtbl"].astype(np.float64)


@pytest.mark.parametrize("C_, n_det", [(6.0, 256), (4.0, 512), (7.0, 300)])
def test_interp_convolvent_is_backward_extrapolation(C_, n_det):
    g = O.Geometry(n_ang=60, n_det=n_det, center=n_det / 2, C=C_, interp=True)
    X, cnv, w = convolvent_from_winv(g)
    spline = CubicSpline(np.arange(len(w)), w)
    k = np.floor(X).astype(int)
    f = X - k
    ok = (k >= 2) & (k <= len(w) - 3)
    d2 = w[k + 1] - 2 * w[k] + w[k - 1]
    # w(k+f) = w_k + f w' + f^2 w''/2 + ..., w_k - w_(k-1) = w' - w''/2 + ...:
    # backward extrapolation leaves cnv - w(X) = -f (1 + f) / 2 w''; a forward
    # difference would leave +f (1 - f) / 2 w''
    predicted = -f * (1 + f) / 2 * d2
    err = cnv - spline(X)
    tol = 0.05 * np.abs(predicted) + 3e-7
    assert np.all(np.abs(err[ok] - predicted[ok]) <= tol[ok])
    big = ok & (np.abs(predicted) > 1e-6)  # well above the float rounding of winv (~1e-7)
    assert big.sum() > 10 and np.all(np.sign(err[big]) == np.sign(predicted[big]))


def test_interp_is_second_order_nearest_first_order():
    g = O.Geometry(n_ang=60, n_det=256, center=128.0, interp=True)
    X, cnv_i, w = convolvent_from_winv(g)
    _, cnv_n, _ = convolvent_from_winv(O.Geometry(n_ang=60, n_det=256, center=128.0))
    s = CubicSpline(np.arange(len(w)), w)(X)
    ei, en = np.abs(cnv_i - s).max(), np.abs(cnv_n - s).max()
    assert ei < en / 20, (ei, en)
    # at 0 (the centre of the deapodization) both are wtbl[0] = 1: dwtbl[0] = 0
    t = O.tables(g)
    assert t["winv"][O.sizes(g)["M02"]] == O.tables(O.Geometry(n_ang=60, n_det=256, center=128.0))["winv"][
        O.sizes(g)["M02"]]


def test_interp_reconstruction_close_to_nearest():
    """Both convolvents approximate the same PSWF: images within 1e-2 of each
    other, and not identical."""
    rng = np.random.default_rng(4)
    a, b = rng.standard_normal((2, 90, 128)).astype(np.float32)
    gn = O.Geometry(n_ang=90, n_det=128, center=63.5)
    gi = O.Geometry(n_ang=90, n_det=128, center=63.5, interp=True)
    n1, _ = O.recon_f32(gn, a, b)
    i1, _ = O.recon_f32(gi, a, b)
    e = O.rel_l2(i1, n1)
    assert 1e-6 < e < 1e-2, e
    with pytest.raises(RuntimeError):
        O.recon_f64(gi, a, b)
