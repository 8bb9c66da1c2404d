"""P2 finite-element setup data (numpy, run once per problem).

* quadrature: quadrature_volume_2D.m:24-33 (7 points), quadrature_volume_3D.m:33-42
  (11 points, negative centroid weight)
* reference basis: local_basis_volume_2D.m (P2), local_basis_volume_3D.m:51-107
* physical gradients DPhi and WEIGHT = |det J| w_q: elastic_stiffness_matrix_2D.m,
  elastic_stiffness_matrix_3D.m (the Jacobian is formed at every point from
  the P2 basis derivatives, as there)
* material fields per point: heterogenous_materials.m:35-87
* volume force f_V = sum_q HatP w_q f: vector_volume_{2D,3D}.m
"""
from __future__ import annotations

import numpy as np


def quadrature_P2(dim):
    if dim == 2:
        Xi = np.array([[0.1012865073235, 0.7974269853531, 0.1012865073235, 0.4701420641051, 0.4701420641051,
                        0.0597158717898, 1 / 3],
                       [0.1012865073235, 0.1012865073235, 0.7974269853531, 0.0597158717898, 0.4701420641051,
                        0.4701420641051, 1 / 3]])
        WF = np.array([0.1259391805448, 0.1259391805448, 0.1259391805448, 0.1323941527885, 0.1323941527885,
                       0.1323941527885, 0.225]) / 2
        return Xi, WF
    a, b = 0.0714285714285714, 0.785714285714286
    c, d = 0.399403576166799, 0.100596423833201
    Xi = np.array([[1 / 4, a, b, a, a, c, d, d, c, c, d],
                   [1 / 4, a, a, b, a, d, c, d, c, d, c],
                   [1 / 4, a, a, a, b, d, d, c, d, c, c]])
    WF = np.array([-0.013155555555555] + [0.007622222222222] * 4 + [0.024888888888888] * 6)
    return Xi, WF


def basis_P2(dim, Xi):
    """(HatP, [DHatP1, DHatP2(, DHatP3)]) each n_p x n_q."""
    if dim == 2:
        x1, x2 = Xi
        x0 = 1 - x1 - x2
        z = np.zeros_like(x1)
        HatP = np.array([x0 * (2 * x0 - 1), x1 * (2 * x1 - 1), x2 * (2 * x2 - 1), 4 * x1 * x2, 4 * x0 * x2, 4 * x0 * x1])
        D1 = np.array([-4 * x0 + 1, 4 * x1 - 1, z, 4 * x2, -4 * x2, 4 * (x0 - x1)])
        D2 = np.array([-4 * x0 + 1, z, 4 * x2 - 1, 4 * x1, 4 * (x0 - x2), -4 * x1])
        return HatP, [D1, D2]
    x1, x2, x3 = Xi
    x0 = 1 - x1 - x2 - x3
    z = np.zeros_like(x1)
    HatP = np.array([x0 * (2 * x0 - 1), x1 * (2 * x1 - 1), x2 * (2 * x2 - 1), x3 * (2 * x3 - 1),
                     4 * x0 * x1, 4 * x1 * x2, 4 * x0 * x2, 4 * x1 * x3, 4 * x2 * x3, 4 * x0 * x3])
    D1 = np.array([-4 * x0 + 1, 4 * x1 - 1, z, z, 4 * (x0 - x1), 4 * x2, -4 * x2, 4 * x3, z, -4 * x3])
    D2 = np.array([-4 * x0 + 1, z, 4 * x2 - 1, z, -4 * x1, 4 * x1, 4 * (x0 - x2), z, 4 * x3, -4 * x3])
    D3 = np.array([-4 * x0 + 1, z, z, 4 * x3 - 1, -4 * x1, z, -4 * x2, 4 * x1, 4 * x2, 4 * (x0 - x3)])
    return HatP, [D1, D2, D3]


def gradients(coord, elem, dim):
    """DPhi (list of dim arrays n_p x n_int) and WEIGHT (n_int) for P2 elements."""
    Xi, WF = quadrature_P2(dim)
    HatP, DH = basis_P2(dim, Xi)
    n_p, n_e = elem.shape
    n_q = WF.size
    # coordinates per element node: (dim, n_p, n_e)
    X = coord[:, elem]
    # Jacobian J[r][c] = sum_a X_c[a] dHat_r[a, q], per (q, e)
    J = np.einsum("rap,cae->rcpe", np.array(DH), X)  # (dim, dim, n_q, n_e)
    J = np.moveaxis(J, (2, 3), (1, 0))  # (n_e, n_q, dim, dim) with [r][c]
    if dim == 2:
        J11, J12, J21, J22 = J[..., 0, 0], J[..., 0, 1], J[..., 1, 0], J[..., 1, 1]
        DET = J11 * J22 - J12 * J21
        inv = np.empty_like(J)
        inv[..., 0, 0] = J22 / DET
        inv[..., 0, 1] = -J12 / DET
        inv[..., 1, 0] = -J21 / DET
        inv[..., 1, 1] = J11 / DET
    else:
        J11, J12, J13 = J[..., 0, 0], J[..., 0, 1], J[..., 0, 2]
        J21, J22, J23 = J[..., 1, 0], J[..., 1, 1], J[..., 1, 2]
        J31, J32, J33 = J[..., 2, 0], J[..., 2, 1], J[..., 2, 2]
        DET = J11 * (J22 * J33 - J32 * J23) - J12 * (J21 * J33 - J23 * J31) + J13 * (J21 * J32 - J22 * J31)
        inv = np.empty_like(J)
        inv[..., 0, 0] = (J22 * J33 - J23 * J32) / DET
        inv[..., 0, 1] = -(J12 * J33 - J13 * J32) / DET
        inv[..., 0, 2] = (J12 * J23 - J13 * J22) / DET
        inv[..., 1, 0] = -(J21 * J33 - J23 * J31) / DET
        inv[..., 1, 1] = (J11 * J33 - J13 * J31) / DET
        inv[..., 1, 2] = -(J11 * J23 - J13 * J21) / DET
        inv[..., 2, 0] = (J21 * J32 - J22 * J31) / DET
        inv[..., 2, 1] = -(J11 * J32 - J12 * J31) / DET
        inv[..., 2, 2] = (J11 * J22 - J12 * J21) / DET
    # DPhi_k[a, (e,q)] = sum_r inv[k][r] dHat_r[a, q]
    DHa = np.array(DH)  # (dim, n_p, n_q)
    DPhi = []
    for k in range(dim):
        v = np.einsum("eqr,rpq->peq", inv[..., k, :], DHa)  # (n_p, n_e, n_q)
        DPhi.append(np.ascontiguousarray(v.reshape(n_p, n_e * n_q)))
    WEIGHT = (np.abs(DET) * WF[None, :]).reshape(-1)
    return DPhi, WEIGHT, HatP, WF


def materials_per_point(mat_id, n_q, mats, saturation=None):
    """c0, phi, psi (rad), shear, bulk, lame, gamma per point
    (heterogenous_materials.m).  mats: list of dicts with keys c0, phi, psi
    (deg), young, poisson, gamma_sat, gamma_unsat."""
    mat_id = np.asarray(mat_id, dtype=np.int64)
    out = {k: [] for k in ("c0", "phi", "psi", "shear", "bulk", "lame", "gamma_sat", "gamma_unsat")}
    for m in mats:
        young, nu = m["young"], m["poisson"]
        shear = young / (2 * (1 + nu))
        bulk = young / (3 * (1 - 2 * nu))
        out["c0"].append(m["c0"])
        out["phi"].append(m["phi"] * np.pi / 180)
        out["psi"].append(m["psi"] * np.pi / 180)
        out["shear"].append(shear)
        out["bulk"].append(bulk)
        out["lame"].append(bulk - 2 * shear / 3)
        out["gamma_sat"].append(m["gamma_sat"])
        out["gamma_unsat"].append(m["gamma_unsat"])
    res = {k: np.repeat(np.asarray(v)[mat_id], n_q) for k, v in out.items()}
    sat = np.ones(mat_id.size * n_q, dtype=bool) if saturation is None else saturation
    res["gamma"] = np.where(sat, res["gamma_sat"], res["gamma_unsat"])
    return res


def volume_force(elem, n_n, HatP, WEIGHT, f_int):
    """f_V (dim x n_n) = sum over points of HatP * WEIGHT * f_int (dim x n_int)."""
    dim = f_int.shape[0]
    n_p, n_e = elem.shape
    n_q = HatP.shape[1]
    f = np.zeros((dim, n_n))
    Hp = np.tile(HatP, (1, n_e))  # n_p x n_int
    nodes = np.repeat(elem, n_q, axis=1)  # n_p x n_int
    for c in range(dim):
        v = Hp * (WEIGHT * f_int[c])[None, :]
        np.add.at(f[c], nodes.ravel(), v.ravel())
    return f
