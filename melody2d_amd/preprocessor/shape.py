"""Meshfree shape functions with first derivatives, vectorised over evaluation points.

The step itself never recomputes shape functions: it consumes the values and
derivatives that the preprocessor wrote into STATIC_DATA.asc
(Code/IO.h:1688-1698 for Gauss points, Code/IO.h:1805-1823 for DOF_TO_DISP).
This module plays the part of Preprocessor/MLS2DWithDerivatives.m and
Preprocessor/RPIM2D.m: given node positions, per-node influence radii and a
set of evaluation points, it returns the support lists and phi, dphi/dx,
dphi/dy.

MLS: linear basis [1, x, y] (Number_Monomials = 3, AnalyseSimpleBody.m:143)
with a quartic-spline weight w(s) = 1 - 6s^2 + 8s^3 - 3s^4 (s = r/R <= 1).
RPIM: Gaussian radial basis exp(-(c*d/D)^2) augmented with the linear
polynomial basis (Gauss_RBF.m:8-17 uses c = 3; the plate workload uses c = 1).
"""
from __future__ import annotations

import numpy as np
from scipy.spatial import cKDTree


def supports(points, nodes, radius):
    """Support lists: node j supports point x when |x - X_j| < R_j.

    Mirrors Preprocessor/SupportDomain.m:17-18 (strict inequality on the
    squared distance).  Returns CSR (ptr, idx) with indices ascending per
    point, which is the order SupportDomain.m produces.
    """
    tree = cKDTree(nodes)
    rmax = float(np.max(radius))
    cand = tree.query_ball_point(points, rmax * (1.0 + 1e-12))
    ptr = np.zeros(len(points) + 1, dtype=np.int64)
    out = []
    for i, c in enumerate(cand):
        c = np.asarray(sorted(c), dtype=np.int64)
        if len(c):
            d2 = np.sum((nodes[c] - points[i]) ** 2, axis=1)
            c = c[d2 - radius[c] ** 2 < 0.0]
        out.append(c)
        ptr[i + 1] = ptr[i] + len(c)
    idx = np.concatenate(out) if out else np.zeros(0, dtype=np.int64)
    return ptr, idx


def _pad(ptr, idx):
    n = len(ptr) - 1
    cnt = np.diff(ptr)
    k = int(cnt.max()) if n else 0
    mask = np.arange(k)[None, :] < cnt[:, None]
    pad = np.zeros((n, k), dtype=np.int64)
    pad[mask] = idx
    return pad, mask


def _quartic(s):
    w = 1.0 - 6.0 * s**2 + 8.0 * s**3 - 3.0 * s**4
    dw = -12.0 * s + 24.0 * s**2 - 12.0 * s**3
    inside = s < 1.0
    return np.where(inside, w, 0.0), np.where(inside, dw, 0.0)


def mls(points, nodes, radius, ptr, idx):
    """MLS values and derivatives (linear basis, quartic spline weight).

    The basis is centred at the evaluation point and scaled by the mean radius
    so the 3x3 moment matrix stays well conditioned; derivatives follow the
    usual gamma = A^-1 p, gamma_x = A^-1 (p_x - A_x gamma) identities.
    Returns phi, phix, phiy in the CSR order of (ptr, idx).
    """
    pad, mask = _pad(ptr, idx)
    n, k = pad.shape
    xj = nodes[pad]                              # (n,k,2)
    dx = xj - points[:, None, :]                 # X_j - x
    r = np.sqrt(np.sum(dx * dx, axis=2))
    Rj = radius[pad]
    s = r / Rj
    w, dws = _quartic(s)
    w = np.where(mask, w, 0.0)
    dws = np.where(mask, dws, 0.0)
    # dw/dx = dw/ds * ds/dx ; ds/dx = (x - X_j)/(r R) = -dx/(r R)
    rsafe = np.where(r > 0, r, 1.0)
    wx = np.where(r > 0, dws * (-dx[..., 0]) / (rsafe * Rj), 0.0)
    wy = np.where(r > 0, dws * (-dx[..., 1]) / (rsafe * Rj), 0.0)
    h = np.mean(radius)
    pj = np.stack([np.ones_like(r), dx[..., 0] / h, dx[..., 1] / h], axis=2)   # (n,k,3)
    A = np.einsum('nk,nki,nkj->nij', w, pj, pj)
    Ax = np.einsum('nk,nki,nkj->nij', wx, pj, pj)
    Ay = np.einsum('nk,nki,nkj->nij', wy, pj, pj)
    # basis at x, in coordinates centred at x: p = [1,0,0]; relative to a
    # fixed centre c = x the derivative of p(x) is [0, 1/h, 0] and [0, 0, 1/h]
    p = np.zeros((n, 3)); p[:, 0] = 1.0
    px = np.zeros((n, 3)); px[:, 1] = 1.0 / h
    py = np.zeros((n, 3)); py[:, 2] = 1.0 / h
    Ainv = np.linalg.inv(A)
    g = np.einsum('nij,nj->ni', Ainv, p)
    gx = np.einsum('nij,nj->ni', Ainv, px - np.einsum('nij,nj->ni', Ax, g))
    gy = np.einsum('nij,nj->ni', Ainv, py - np.einsum('nij,nj->ni', Ay, g))
    # Holding the centre fixed while differentiating, p_j depends on x only
    # through the centre, which is constant; so B_j,x = w_j,x p_j.
    gp = np.einsum('ni,nki->nk', g, pj)
    phi = gp * w
    phix = np.einsum('ni,nki->nk', gx, pj) * w + gp * wx
    phiy = np.einsum('ni,nki->nk', gy, pj) * w + gp * wy
    return phi[mask], phix[mask], phiy[mask]


def rpim(points, nodes, ptr, idx, shape_c=1.0, dscale=None):
    """RPIM values and derivatives: Gaussian RBF exp(-(c d / D)^2) plus the
    linear polynomial basis (moment conditions), solved per point as in
    Preprocessor/RPIM2D.m.  D is the support scale (dscale per point, default
    the largest support distance)."""
    pad, mask = _pad(ptr, idx)
    n, k = pad.shape
    out_phi = np.zeros((n, k)); out_x = np.zeros((n, k)); out_y = np.zeros((n, k))
    cnt = np.diff(ptr)
    for kk in np.unique(cnt):
        sel = np.nonzero(cnt == kk)[0]
        if kk == 0:
            continue
        P = nodes[pad[sel, :kk]]                           # (m,kk,2)
        x = points[sel]
        if dscale is None:
            D = np.sqrt(np.max(np.sum((P - x[:, None, :]) ** 2, axis=2), axis=1))
        else:
            D = dscale[sel]
        a = (shape_c / D) ** 2                              # (m,)
        dij = P[:, :, None, :] - P[:, None, :, :]
        r2 = np.sum(dij * dij, axis=3)
        m = len(sel)
        G = np.zeros((m, kk + 3, kk + 3))
        G[:, :kk, :kk] = np.exp(-a[:, None, None] * r2)
        h = D[:, None]
        Pm = np.stack([np.ones((m, kk)), (P[..., 0] - x[:, 0:1]) / h, (P[..., 1] - x[:, 1:2]) / h], axis=2)
        G[:, :kk, kk:] = Pm
        G[:, kk:, :kk] = np.transpose(Pm, (0, 2, 1))
        dx = x[:, None, :] - P                               # x - X_j
        rx2 = np.sum(dx * dx, axis=2)
        rb = np.exp(-a[:, None] * rx2)
        rbx = rb * (-2.0 * a[:, None] * dx[..., 0])
        rby = rb * (-2.0 * a[:, None] * dx[..., 1])
        rhs = np.zeros((m, kk + 3, 3))
        rhs[:, :kk, 0] = rb; rhs[:, :kk, 1] = rbx; rhs[:, :kk, 2] = rby
        rhs[:, kk, 0] = 1.0
        rhs[:, kk + 1, 1] = 1.0 / D
        rhs[:, kk + 2, 2] = 1.0 / D
        sol = np.linalg.solve(G, rhs)                        # symmetric system
        out_phi[sel, :kk] = sol[:, :kk, 0]
        out_x[sel, :kk] = sol[:, :kk, 1]
        out_y[sel, :kk] = sol[:, :kk, 2]
    return out_phi[mask], out_x[mask], out_y[mask]


def rpim_fixed_torch(points, nodes, nn, shape_c=1.0, device="cuda", chunk=200000):
    """rpim() for supports of one fixed size (rows of nn), with the systems
    built and solved in batches by torch on `device` (the 1M-node plate has
    ~7M of them).  Same formulas as rpim(); the solver's rounding differs
    from LAPACK's, so the two agree to rounding, not bit for bit."""
    import torch
    n, k = nn.shape
    out = [np.zeros((n, k)) for _ in range(3)]
    Xn = torch.as_tensor(nodes, dtype=torch.float64, device=device)
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        idx = torch.as_tensor(nn[a:b], dtype=torch.int64, device=device)
        x = torch.as_tensor(points[a:b], dtype=torch.float64, device=device)
        P = Xn[idx]                                              # (m,k,2)
        m = b - a
        D = torch.sqrt(torch.max(torch.sum((P - x[:, None, :]) ** 2, dim=2), dim=1).values)
        al = (shape_c / D) ** 2
        dij = P[:, :, None, :] - P[:, None, :, :]
        r2 = torch.sum(dij * dij, dim=3)
        del dij
        G = torch.zeros((m, k + 3, k + 3), dtype=torch.float64, device=device)
        G[:, :k, :k] = torch.exp(-al[:, None, None] * r2)
        del r2
        h = D[:, None]
        Pm = torch.stack([torch.ones((m, k), dtype=torch.float64, device=device), (P[..., 0] - x[:, 0:1]) / h,
                          (P[..., 1] - x[:, 1:2]) / h], dim=2)
        G[:, :k, k:] = Pm
        G[:, k:, :k] = Pm.transpose(1, 2)
        dx = x[:, None, :] - P
        rx2 = torch.sum(dx * dx, dim=2)
        rb = torch.exp(-al[:, None] * rx2)
        rhs = torch.zeros((m, k + 3, 3), dtype=torch.float64, device=device)
        rhs[:, :k, 0] = rb
        rhs[:, :k, 1] = rb * (-2.0 * al[:, None] * dx[..., 0])
        rhs[:, :k, 2] = rb * (-2.0 * al[:, None] * dx[..., 1])
        rhs[:, k, 0] = 1.0
        rhs[:, k + 1, 1] = 1.0 / D
        rhs[:, k + 2, 2] = 1.0 / D
        sol = torch.linalg.solve(G, rhs)
        for q in range(3):
            out[q][a:b] = sol[:, :k, q].cpu().numpy()
        del G, rhs, sol
    return out[0].reshape(-1), out[1].reshape(-1), out[2].reshape(-1)
