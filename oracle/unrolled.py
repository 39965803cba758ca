"""CPU restatement of unrolled PDHG and its adjoint (test infrastructure only).

Checks the HIP forward / backward of ``cuopt_amd.autograd``; nothing in the
product imports this module.

The map is K steps of the PDHG iteration the reference runs
(pdhg.cu:431-545 projections, reflection x-bar = 2 x+ - x), without restarts or
Halpern averaging so that x_K is a piecewise-linear function of (c, l_c, u_c),
plus a ridge term lam/2 ||x||^2 whose proximal step divides by 1 + tau lam
(the reference has no QP prox on this path; solve.cu:1142-1152 hands QPs to
barrier):

    x+ = clamp((x - tau (c - A^T y)) / (1 + tau lam), l, u)
    xb = 2 x+ - x
    t  = y / sigma - A xb
    y+ = sigma (t - clamp(t, -u_c, -l_c))

The adjoint runs the same steps backwards with the masks of the forward
clamps; only the clamp states of x+ and y+ are needed:

    x+ strictly inside (l, u)           <=> d x+ / d v = 1
    y+ == 0 with l_c < u_c              <=> t inside (-u_c, -l_c)
    y+ > 0 (or y+ == 0 with l_c == u_c) <=> the -l_c side is active  (d y+/d l_c = sigma)
    y+ < 0                              <=> the -u_c side is active  (d y+/d u_c = sigma)

Pinned in tests/test_unrolled_oracle.py by central finite differences of this
map (FP64, h = 1e-6) and by the envelope identities at convergence
(d obj/d c = x*, d obj/d b = y*).
"""

from __future__ import annotations

import numpy as np


def step_sizes(A, iters: int = 200, seed: int = 0) -> float:
    """0.998 / sigma_max(A) by dense SVD (small problems) -- tau = sigma = eta."""
    s = np.linalg.norm(A.toarray() if hasattr(A, "toarray") else A, 2)
    return 0.998 / s


def forward(A, c, lc, uc, lb, ub, lam, eta, K, keep=False):
    """K steps from x = y = 0.  c: (n, B); lc, uc: (m, B) or (m,).  Returns x_K, y_K
    and, with keep, the lists of x_k, y_k for k = 0..K."""
    n, B = c.shape
    m = A.shape[0]
    lc = np.broadcast_to(lc.reshape(m, -1), (m, B))
    uc = np.broadcast_to(uc.reshape(m, -1), (m, B))
    lo, hi = lb[:, None], ub[:, None]
    tau = sigma = eta
    x = np.clip(np.zeros((n, B)), lo, hi)
    y = np.zeros((m, B))
    xs, ys = [x], [y]
    AT = A.T.tocsr() if hasattr(A, "tocsr") else A.T
    for _ in range(K):
        v = (x - tau * (c - AT @ y)) / (1.0 + tau * lam)
        xn = np.clip(v, lo, hi)
        xb = 2.0 * xn - x
        t = y / sigma - A @ xb
        yn = sigma * (t - np.clip(t, -uc, -lc))
        x, y = xn, yn
        if keep:
            xs.append(x)
            ys.append(y)
    return (x, y, xs, ys) if keep else (x, y)


def backward(A, lc, uc, lb, ub, lam, eta, xs, ys, gx, gy=None):
    """VJP of (x_K, y_K) -> (gx, gy) through the K steps recorded in xs, ys.
    Returns (g_c, g_lc, g_uc), shaped (n, B), (m, B), (m, B)."""
    n, B = xs[0].shape
    m = ys[0].shape[0]
    lcb = np.broadcast_to(lc.reshape(m, -1), (m, B))
    ucb = np.broadcast_to(uc.reshape(m, -1), (m, B))
    lo, hi = lb[:, None], ub[:, None]
    tau = sigma = eta
    K = len(xs) - 1
    ax = gx.copy()
    ay = np.zeros((m, B)) if gy is None else gy.copy()
    gc = np.zeros((n, B))
    glc = np.zeros((m, B))
    guc = np.zeros((m, B))
    AT = A.T.tocsr() if hasattr(A, "tocsr") else A.T
    for k in range(K - 1, -1, -1):
        xn, yn = xs[k + 1], ys[k + 1]
        # dual: y+ = sigma (t - clamp(t, -u_c, -l_c)), t = y / sigma - A xb
        eq = lcb == ucb
        inside = (yn == 0.0) & ~eq
        low = (yn > 0.0) | ((yn == 0.0) & eq)
        up = yn < 0.0
        gt = np.where(inside, 0.0, sigma * ay)
        glc += np.where(low, sigma * ay, 0.0)
        guc += np.where(up, sigma * ay, 0.0)
        gy_prev = gt / sigma
        gxb = -(AT @ gt)
        # reflection xb = 2 x+ - x
        gxn = ax + 2.0 * gxb
        gx_prev = -gxb
        # primal: x+ = clamp(v, l, u), v = (x - tau (c - A^T y)) / (1 + tau lam)
        inner = (xn > lo) & (xn < hi)
        gw = np.where(inner, gxn, 0.0) / (1.0 + tau * lam)
        gx_prev = gx_prev + gw
        gc -= tau * gw
        gy_prev = gy_prev + tau * (A @ gw)
        ax, ay = gx_prev, gy_prev
    return gc, glc, guc
