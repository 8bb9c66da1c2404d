"""Oracle: Mohr-Coulomb potential per integration point (test
infrastructure only; the product computes it with k_potential{2,3}d).

Vectorised numpy restatement of potential_3D.m (E point-major 6 x n with
engineering shears, halved into the tensor; trigonometric eigenvalues; the
invariant I1 in the yield test and the multipliers, e1+e2+e3 in the
energies) and potential_2D.m (closed-form 2x2 eigenvalues plus the
out-of-plane zero, reordered).  Branches and their precedence as in the .m
files: elastic, smooth, left edge, right edge, apex.

Pinned by: S = dPsi/dE against the oracle return mapping (itself bitwise
equal to the reference's compiled kernels) by central differences, and the
elastic / apex closed forms (tests/test_oracle_potential.py).
"""
from __future__ import annotations

import numpy as np


def _branches(e1, e2, e3, tI, tr, cb, sp, G, K, lam, den_s):
    f_tr = 2 * G * ((1 + sp) * e1 - (1 - sp) * e3) + 2 * (lam * sp) * tI - cb
    g_sl = (e1 - e2) / (1 + sp)
    g_sr = (e2 - e3) / (1 - sp)
    g_la = (e1 + e2 - 2 * e3) / (3 - sp)
    g_ra = (2 * e1 - e2 - e3) / (3 + sp)
    den_l = 4 * lam * sp ** 2 + G * (1 + sp) ** 2 + 2 * G * (1 - sp) ** 2
    den_r = 4 * lam * sp ** 2 + 2 * G * (1 + sp) ** 2 + G * (1 - sp) ** 2
    den_a = 4 * K * sp ** 2
    with np.errstate(divide="ignore", invalid="ignore"):
        l_s = f_tr / den_s
        l_l = (G * ((1 + sp) * (e1 + e2) - 2 * (1 - sp) * e3) + 2 * lam * sp * tI - cb) / den_l
        l_r = (G * (2 * (1 + sp) * e1 - (1 - sp) * (e2 + e3)) + 2 * lam * sp * tI - cb) / den_r
        l_a = (2 * K * sp * tI - cb) / den_a
    el = 0.5 * lam * tr ** 2 + G * (e1 ** 2 + e2 ** 2 + e3 ** 2)
    psi = np.zeros_like(e1)
    t_el = f_tr <= 0
    psi[t_el] = el[t_el]
    t_s = (l_s <= np.minimum(g_sl, g_sr)) & ~t_el
    psi[t_s] = (el - 0.5 * den_s * l_s ** 2)[t_s]
    t_l = (g_sl < g_sr) & (l_l >= g_sl) & (l_l <= g_la) & ~(t_el | t_s)
    psi[t_l] = (0.5 * lam * tr ** 2 + G * (e3 ** 2 + 0.5 * (e1 + e2) ** 2) - 0.5 * den_l * l_l ** 2)[t_l]
    t_r = (g_sl > g_sr) & (l_r >= g_sr) & (l_r <= g_ra) & ~(t_el | t_s)
    psi[t_r] = (0.5 * lam * tr ** 2 + G * (e1 ** 2 + 0.5 * (e2 + e3) ** 2) - 0.5 * den_r * l_r ** 2)[t_r]
    t_a = ~(t_el | t_s | t_l | t_r)
    psi[t_a] = (0.5 * K * tr ** 2 - 0.5 * den_a * l_a ** 2)[t_a]
    br = np.select([t_el, t_s, t_l, t_r], [0, 1, 2, 3], 4)
    return psi, br


def potential_3d(E, c_bar, sin_phi, shear, bulk, lame):
    """E: 6 x n (engineering shears).  Returns (Psi, branch)."""
    E = np.asarray(E, dtype=np.float64)
    t = E * np.array([1, 1, 1, 0.5, 0.5, 0.5])[:, None]
    I1 = t[0] + t[1] + t[2]
    I2 = t[0] * t[1] + t[0] * t[2] + t[1] * t[2] - t[3] ** 2 - t[4] ** 2 - t[5] ** 2
    I3 = t[0] * t[1] * t[2] - t[2] * t[3] ** 2 - t[1] * t[5] ** 2 - t[0] * t[4] ** 2 + 2 * t[3] * t[4] * t[5]
    Q = np.maximum(0.0, (I1 ** 2 - 3 * I2) / 9.0)
    R = (-2 * I1 ** 3 + 9 * I1 * I2 - 27 * I3) / 54.0
    th0 = np.zeros_like(Q)
    pos = Q > 0
    th0[pos] = R[pos] / np.sqrt(Q[pos] ** 3)
    th = np.arccos(np.clip(th0, -1, 1)) / 3
    e1 = -2 * np.sqrt(Q) * np.cos(th + 2 * np.pi / 3) + I1 / 3
    e2 = -2 * np.sqrt(Q) * np.cos(th - 2 * np.pi / 3) + I1 / 3
    e3 = -2 * np.sqrt(Q) * np.cos(th) + I1 / 3
    sp, G, lam = sin_phi, shear, lame
    den_s = 4 * lam * sp ** 2 + 4 * G * (1 + sp ** 2)
    return _branches(e1, e2, e3, I1, e1 + e2 + e3, c_bar, sp, G, bulk, lam, den_s)


def potential_2d(E, c_bar, sin_phi, shear, bulk, lame):
    """E: 3 x n (eps11, eps22, gamma12).  Returns (Psi, branch)."""
    E = np.asarray(E, dtype=np.float64)
    I1 = E[0] + E[1]
    I2 = np.sqrt((E[0] - E[1]) ** 2 + E[2] ** 2)
    a1, a2, a3 = (I1 + I2) / 2, (I1 - I2) / 2, np.zeros_like(I1)
    e1, e2, e3 = a1.copy(), a2.copy(), a3.copy()
    t2 = (a1 >= a3) & (a3 > a2)
    e2[t2], e3[t2] = a3[t2], a2[t2]
    t3 = a3 > a1
    e1[t3], e2[t3], e3[t3] = a3[t3], a1[t3], a2[t3]
    tr = e1 + e2 + e3
    sp, G, lam = sin_phi, shear, lame
    den_s = 4 * lam * sp ** 2 + 2 * G * (1 + sp) ** 2 + 2 * G * (1 - sp) ** 2
    return _branches(e1, e2, e3, tr, tr, c_bar, sp, G, bulk, lam, den_s)
