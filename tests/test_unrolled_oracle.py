"""The unrolled-PDHG restatement (oracle/unrolled.py) against what the
mathematics fixes: central finite differences of the map (FP64, h = 1e-6) and
the envelope identities d obj/d c = x*, d obj/d b = y* once the ridge-regularised
LP has converged.  CPU only."""

import numpy as np
import pytest

from cuopt_amd import generators as G
from oracle import unrolled as U


def small(seed=0, m=30, n=75, lam=1e-3):
    lp = G.random_lp(m=m, n=n, seed=seed)
    A = lp.to_scipy()
    return lp, A, lam, U.step_sizes(A)


@pytest.mark.parametrize("seed", [0, 1])
def test_vjp_matches_central_differences(seed):
    lp, A, lam, eta = small(seed)
    rng = np.random.default_rng(seed)
    B, K = 3, 60
    c = lp.c[:, None] + 0.1 * rng.normal(size=(lp.n, B))
    b = lp.con_lb.copy()
    target = rng.uniform(0, 1, size=(lp.n, B))

    def loss(cc, bb):
        x, _ = U.forward(A, cc, bb, bb, lp.var_lb, lp.var_ub, lam, eta, K)
        return float(np.sum((x - target) ** 2))

    x, y, xs, ys = U.forward(A, c, b, b, lp.var_lb, lp.var_ub, lam, eta, K, keep=True)
    gc, glc, guc = U.backward(A, b, b, lp.var_lb, lp.var_ub, lam, eta, xs, ys, 2.0 * (x - target))
    gb = (glc + guc).sum(axis=1)  # b shared by the batch
    h = 1e-6
    for _ in range(3):
        vc = rng.normal(size=c.shape)
        fd = (loss(c + h * vc, b) - loss(c - h * vc, b)) / (2 * h)
        assert abs(fd - np.sum(gc * vc)) <= 1e-6 * (1 + abs(fd))
        vb = rng.normal(size=b.shape)
        fd = (loss(c, b + h * vb) - loss(c, b - h * vb)) / (2 * h)
        assert abs(fd - np.sum(gb * vb)) <= 1e-6 * (1 + abs(fd))


def test_inequality_rows_and_bounds():
    # ranged rows exercise the inside / lower / upper clamp states of the dual step
    lp, A, lam, eta = small(2)
    rng = np.random.default_rng(2)
    B, K = 2, 40
    c = lp.c[:, None] + 0.1 * rng.normal(size=(lp.n, B))
    lc = lp.con_lb - rng.uniform(0, 0.5, size=lp.m)
    uc = lp.con_ub + rng.uniform(0, 0.5, size=lp.m)
    x, y, xs, ys = U.forward(A, c, lc, uc, lp.var_lb, lp.var_ub, lam, eta, K, keep=True)
    states = np.concatenate([s.ravel() for s in ys[1:]])
    assert (states == 0).any() and (states > 0).any() and (states < 0).any()
    gx = rng.normal(size=x.shape)
    gc, glc, guc = U.backward(A, lc, uc, lp.var_lb, lp.var_ub, lam, eta, xs, ys, gx)

    def f(cc, l, u):
        return float(np.sum(gx * U.forward(A, cc, l, u, lp.var_lb, lp.var_ub, lam, eta, K)[0]))

    h = 1e-6
    for g, which in ((gc, 0), (glc, 1), (guc, 2)):
        v = rng.normal(size=g.shape)
        args_p = [c, lc[:, None].repeat(B, 1), uc[:, None].repeat(B, 1)]
        args_m = [a.copy() for a in args_p]
        args_p[which] = args_p[which] + h * v
        args_m[which] = args_m[which] - h * v
        fd = (f(*args_p) - f(*args_m)) / (2 * h)
        assert abs(fd - np.sum(g * v)) <= 1e-6 * (1 + abs(fd))


def test_envelope_gradients_at_convergence():
    # The ridge makes the solution unique and PDHG linearly convergent at a rate
    # set by lam; with lam = 1e-2 the unrolled derivative has converged after 20k steps.
    lp, A, lam, eta = small(3, m=20, n=50, lam=1e-2)
    c = lp.c[:, None]
    b = lp.con_lb.copy()
    K = 20000
    x, y, xs, ys = U.forward(A, c, b, b, lp.var_lb, lp.var_ub, lam, eta, K, keep=True)
    # obj = c.x + lam/2 ||x||^2: explicit d/dc = x, plus the VJP through x_K
    gc, glc, guc = U.backward(A, b, b, lp.var_lb, lp.var_ub, lam, eta, xs, ys, c + lam * x)
    assert np.abs(gc).max() <= 1e-9  # the implicit term vanishes: d obj / d c = x*
    assert np.allclose((glc + guc)[:, 0], y[:, 0], atol=1e-9 * (1 + np.abs(y).max()))  # d obj / d b = y*
    assert np.linalg.norm(A @ x[:, 0] - b) <= 1e-10
