"""Global-matrix oracle for the tangent assembly (test infrastructure only).

Restates, with scipy sparse products on small meshes, the reference's global
formulation that the element-level MEX replaces:

* ``global_B``      -- the strain-displacement matrix exactly as assembled in
  elastic_stiffness_matrix_3D.m:105-132 / elastic_stiffness_matrix_2D.m:90-104
  (rows ns*n_int, columns dim*n_n, node-major DOFs);
* ``ref_pattern``   -- CONSTITUTIVE.init_K_r_pattern (CONSTITUTIVE.m:407-454):
  symbolic pattern of B_Q'*D*B_Q for any block-diagonal D, then column-major
  ``find`` (ref_I, ref_J), V_elast_QQ and transpose_perm;
* ``scatter_map``   -- CONSTITUTIVE.build_scatter_map (CONSTITUTIVE.m:582-635);
* ``global_tangent``-- CONSTITUTIVE.build_K_tangent_QQ_vals_global
  (CONSTITUTIVE.m:503-513): full(K_t(ref_idx)) with K_t = B_Q' D_t B_Q;
* ``build_K_r``     -- CONSTITUTIVE.build_K_r_QQ_vals (CONSTITUTIVE.m:515-532).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def _ns(dim):
    return 6 if dim == 3 else 3


def global_B(elem, DPhi, dim, n_n):
    n_p, n_e = elem.shape
    n_int = DPhi[0].shape[1]
    n_q = n_int // n_e
    ns = _ns(dim)
    rows, cols, vals = [], [], []
    g = np.arange(n_int)
    nodes = np.repeat(elem, n_q, axis=1)  # n_p x n_int
    if dim == 3:
        # (strain row, displacement component, gradient index)
        entries = [(0, 0, 0), (3, 0, 1), (5, 0, 2), (1, 1, 1), (3, 1, 0), (4, 1, 2), (2, 2, 2), (4, 2, 1), (5, 2, 0)]
    else:
        entries = [(0, 0, 0), (2, 0, 1), (1, 1, 1), (2, 1, 0)]
    for a in range(n_p):
        for (r, c, k) in entries:
            rows.append(ns * g + r)
            cols.append(dim * nodes[a] + c)
            vals.append(DPhi[k][a])
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(ns * n_int, dim * n_n))


def block_D(DS, WEIGHT, ns):
    """Block diagonal D with blocks WEIGHT_p * reshape(DS(:,p), ns, ns)."""
    n_int = DS.shape[1]
    base = np.arange(n_int) * ns
    i = np.tile(np.arange(ns), ns)  # column-major k = i + ns*j
    j = np.repeat(np.arange(ns), ns)
    rows = (base[None, :] + i[:, None]).ravel(order="F")
    cols = (base[None, :] + j[:, None]).ravel(order="F")
    vals = (DS * WEIGHT[None, :]).ravel(order="F")
    return sp.csr_matrix((vals, (rows, cols)), shape=(ns * n_int, ns * n_int))


def ds_elastic(dim, shear, bulk):
    """DS_elast of CONSTITUTIVE.m:413-422 (IOTA = [1;1;1;zeros(ns-3,1)])."""
    ns = _ns(dim)
    iota = np.concatenate([np.ones(3), np.zeros(ns - 3)])
    VOL = np.outer(iota, iota)
    ident = np.diag([1.0, 1, 1, 0.5, 0.5, 0.5]) if dim == 3 else np.diag([1.0, 1, 0.5])
    DEV = ident - VOL / 3
    return 2 * DEV.reshape(-1, 1, order="F") * shear[None, :] + VOL.reshape(-1, 1, order="F") * bulk[None, :]


class RefPattern:
    def __init__(self, elem, Q, dim, DPhi, WEIGHT, shear, bulk):
        n_n = Q.shape[1]
        self.dim = dim
        self.ns = _ns(dim)
        self.Q_flat = np.flatnonzero(Q.ravel(order="F"))
        B = global_B(elem, DPhi, dim, n_n)
        self.B_Q = B[:, self.Q_flat].tocsc()
        n_int = DPhi[0].shape[1]
        ns = self.ns
        Dpat = block_D(np.ones((ns * ns, n_int)), np.ones(n_int), ns)
        Bp = self.B_Q.copy()
        Bp.data[:] = 1.0
        K = (Bp.T @ Dpat @ Bp)
        K = (K + K.T).tocsc()
        K.sort_indices()
        # column-major find: per column, row indices ascending
        self.ref_J = np.repeat(np.arange(K.shape[1]), np.diff(K.indptr))
        self.ref_I = K.indices.copy()
        self.n_Q = K.shape[0]
        self.nnz = self.ref_I.size
        De = block_D(ds_elastic(dim, shear, bulk), WEIGHT, ns)
        Ke = self.B_Q.T @ De @ self.B_Q
        Ke = ((Ke + Ke.T) / 2).tocsr()
        self.V_elast = np.asarray(Ke[self.ref_I, self.ref_J]).ravel()
        key = self.ref_J.astype(np.int64) * self.n_Q + self.ref_I
        tkey = self.ref_I.astype(np.int64) * self.n_Q + self.ref_J
        order = np.argsort(key)
        self.transpose_perm = order[np.searchsorted(key[order], tkey)]

    def global_tangent(self, DS, WEIGHT):
        Dt = block_D(DS, WEIGHT, self.ns)
        Kt = (self.B_Q.T @ Dt @ self.B_Q).tocsr()
        return np.asarray(Kt[self.ref_I, self.ref_J]).ravel()

    def build_K_r(self, r, V_tang):
        V = r * self.V_elast + (1 - r) * V_tang
        return (V + V[self.transpose_perm]) / 2

    def scatter_map(self, elem, n_n):
        dim = self.dim
        n_p, n_e = elem.shape
        nld = dim * n_p
        free_idx = np.zeros(dim * n_n, dtype=np.int64)
        free_idx[self.Q_flat] = np.arange(1, self.n_Q + 1)
        P = sp.csr_matrix((np.arange(1, self.nnz + 1, dtype=np.float64), (self.ref_I, self.ref_J)),
                          shape=(self.n_Q, self.n_Q))
        smap = np.zeros((nld * nld, n_e), dtype=np.int64)
        for e in range(n_e):
            dofs = (dim * elem[:, e][:, None] + np.arange(dim)[None, :]).ravel()
            fr = free_idx[dofs]
            m = fr > 0
            if not m.any():
                continue
            full = np.zeros((nld, nld))
            act = np.flatnonzero(m)
            full[np.ix_(act, act)] = P[fr[m] - 1][:, fr[m] - 1].toarray()
            smap[:, e] = full.ravel(order="F").astype(np.int64)
        return smap
