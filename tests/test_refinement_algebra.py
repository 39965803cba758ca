"""The correction problem used by FP64 refinement of FP32-mixed solves is the
original LP in shifted, rescaled coordinates: PDHG on it must reproduce PDHG on
the original problem step for step.  CPU, numpy, FP64.

Original step (projections of pdhg.cu:431-545):
    x+ = clamp(x - tau (c - A^T y), l, u);  xb = 2 x+ - x
    y+ = sigma (t - clamp(t, -u_c, -l_c)),  t = y / sigma - A xb
Correction around (x*, y*) with scales Dp, Dd: d = Dp (x - x*), w = Dd (y - y*),
objective Dd (c - A^T y*), bounds Dp ([l, u] - x*), constraint bounds
Dp ([l_c, u_c] - A x*), steps tau' = tau Dp / Dd, sigma' = sigma Dd / Dp, and the
dual projection moved to the kink k = -Dd y*, in the form the kernel uses:
    w+ = clamp(k, s + sigma' l_c', s + sigma' u_c'),  s = w - sigma' A db
"""

import numpy as np
import pytest


def kinked_dual(w, adb, sig, lo, hi, k):
    s = w - sig * adb
    return np.maximum(np.minimum(k, s + sig * hi), s + sig * lo)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_correction_pdhg_reproduces_original(seed):
    rng = np.random.default_rng(seed)
    m, n = 30, 60
    A = rng.normal(size=(m, n)) * (rng.random((m, n)) < 0.3)
    c = rng.normal(size=n)
    l = np.where(rng.random(n) < 0.5, 0.0, -np.inf)
    u = np.where(rng.random(n) < 0.5, 5.0, np.inf)
    lc = rng.normal(size=m) - 1
    uc = lc + np.where(rng.random(m) < 0.5, 0.0, 2.0)  # equality and ranged rows
    lc[rng.random(m) < 0.3] = -np.inf                   # and one-sided rows
    tau = sigma = 0.9 / np.linalg.norm(A, 2)
    x = np.clip(rng.normal(size=n), l, u)
    y = rng.normal(size=m)
    xs, ys = x.copy(), y.copy()
    Dp, Dd = 10.0 ** rng.uniform(0, 4), 10.0 ** rng.uniform(0, 4)
    ch = Dd * (c - A.T @ ys)
    lo_p, hi_p = Dp * (l - xs), Dp * (u - xs)
    lo_c, hi_c = Dp * (lc - A @ xs), Dp * (uc - A @ xs)
    kink = -Dd * ys
    taup, sigp = tau * Dp / Dd, sigma * Dd / Dp
    d, w = np.zeros(n), np.zeros(m)
    for _ in range(50):
        xn = np.clip(x - tau * (c - A.T @ y), l, u)
        xb = 2 * xn - x
        t = y / sigma - A @ xb
        yn = sigma * (t - np.clip(t, -uc, -lc))
        dn = np.clip(d - taup * (ch - A.T @ w), lo_p, hi_p)
        db = 2 * dn - d
        wn = kinked_dual(w, A @ db, sigp, lo_c, hi_c, kink)
        x, y, d, w = 2 * xn - x, 2 * yn - y, 2 * dn - d, 2 * wn - w
        assert np.max(np.abs(xs + d / Dp - x)) <= 1e-12 * (1 + np.max(np.abs(x)))
        assert np.max(np.abs(ys + w / Dd - y)) <= 1e-12 * (1 + np.max(np.abs(y)))


def test_kinked_projection_reduces_to_the_reference_form_at_zero():
    rng = np.random.default_rng(3)
    y, adb = rng.normal(size=100), rng.normal(size=100)
    lo = np.where(rng.random(100) < 0.3, -np.inf, rng.normal(size=100) - 1)
    hi = np.where(rng.random(100) < 0.3, np.inf, lo + rng.uniform(0, 2, 100))
    sig = 0.7
    t = y / sig - adb
    ref = sig * (t - np.clip(t, -hi, -lo))
    assert np.allclose(kinked_dual(y, adb, sig, lo, hi, 0.0), ref, atol=1e-14)
