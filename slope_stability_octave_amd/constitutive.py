"""Device-resident mirror of the reference's CONSTITUTIVE handle class
(+CONSTITUTIVE_PROBLEM/CONSTITUTIVE.m).

Same method names and meaning: ``reduction``, ``constitutive_problem_stress``,
``constitutive_problem_stress_tangent``, ``build_F``, ``build_F_all``,
``build_F_reduced``, ``build_F_and_DS_all``, ``build_F_and_DS_reduced``,
``init_K_r_pattern``, ``build_K_tangent_QQ_vals``, ``build_K_r_QQ_vals``,
``set_element_data``.  The strain-displacement matrix B is applied
matrix-free from the element gradients (DPhi), so the constructor takes the
element data instead of B.  Every array lives on the GPU; U, F and f are
dim x n_n fields stored node-major (the layout of Octave's U(:)).
"""
from __future__ import annotations

import time

import numpy as np

from . import assembly as A
from . import constitutive_mex as CM
from ._lib import DeviceArray


class CONSTITUTIVE:
    def __init__(self, elem, DPhi, c0, phi, psi, Davis_type, shear, bulk, lame, WEIGHT, n_strain, n_int, dim,
                 friction_pairs=None, materials=None):
        """Host arrays are uploaded; DeviceArrays (the device setup path) are
        used in place -- DPhi then in the device layout (n_p values of a
        point contiguous), and ``friction_pairs`` = (phi, psi) of the
        distinct materials, which the Davis C test needs on the host.
        ``materials`` = (mat_id per element, table n_mat x (c0, phi, psi,
        shear, bulk, lame, ...)) describes the same per-point fields per
        material (see set_material_table)."""
        self.dim = int(dim)
        self.n_strain = int(n_strain)
        self.n_int = int(n_int)
        self.Davis_type = Davis_type
        self.elem = np.asarray(elem, dtype=np.int64)
        self.n_p, self.n_e = self.elem.shape
        self.n_q = self.n_int // self.n_e
        self.n_n = int(self.elem.max()) + 1

        def dev(x):
            return x if isinstance(x, DeviceArray) else DeviceArray.from_host(np.asarray(x, dtype=np.float64).ravel())

        if friction_pairs is None:
            if isinstance(phi, DeviceArray) or isinstance(psi, DeviceArray):
                raise ValueError("device material fields need friction_pairs=(phi, psi) of the materials")
            # the distinct (phi, psi) pairs: Davis C's all-points test
            # (reduction.m:47) gives the same answer on them
            pairs = np.unique(np.stack([np.asarray(phi, dtype=np.float64).ravel(),
                                        np.asarray(psi, dtype=np.float64).ravel()]), axis=1)
            friction_pairs = (pairs[0], pairs[1])
        self._phi_u, self._psi_u = (np.asarray(x, dtype=np.float64) for x in friction_pairs)
        self._c_all = {}
        self.c0, self.phi, self.psi = (dev(x) for x in (c0, phi, psi))
        self.shear, self.bulk, self.lame = (dev(x) for x in (shear, bulk, lame))
        self.WEIGHT = dev(WEIGHT)
        # n_p x n_int, Octave column-major: the n_p values of a point contiguous
        self.DPhi = [d if isinstance(d, DeviceArray) else
                     DeviceArray.from_host(np.asarray(d, dtype=np.float64).ravel(order="F")) for d in DPhi]
        self.c_bar = DeviceArray(self.n_int)
        self.sin_phi = DeviceArray(self.n_int)
        self._lambda = DeviceArray(1)
        ns = self.n_strain
        self.s_rows = 6 if self.dim == 3 else 4
        self.E = DeviceArray((self.n_int, ns))
        self.S = DeviceArray((self.n_int, self.s_rows))
        self.DS = DeviceArray((self.n_int, ns * ns))
        self.pattern = None
        self.pattern_initialized = False
        self.material_fields = None
        self.coord_dev = None  # set_node_coordinates
        if materials is not None:
            self.set_material_table(*materials)
        self.profiler = None
        self.timings = {}
        self.counts = {}

    def _t(self, name, t0):
        self.timings[name] = self.timings.get(name, 0.0) + time.perf_counter() - t0
        self.counts[name] = self.counts.get(name, 0) + 1

    def set_material_table(self, mat_id, table):
        """The per-point material fields as a per-element material id and a
        table (rows c0, phi, psi, shear, bulk, lame per material; the values
        the per-point arrays hold).  The native Newton loops then reduce the
        strength once per material and the return mappings read each point's
        material from its element (the same values, so the same results)."""
        mat_id = np.asarray(mat_id).ravel()
        tab = np.asarray(table, dtype=np.float64)
        if mat_id.size != self.n_e or tab.ndim != 2 or tab.shape[1] < 6:
            raise ValueError("material table: one id per element and rows (c0, phi, psi, shear, bulk, lame)")
        if mat_id.size and (mat_id.min() < 0 or mat_id.max() >= tab.shape[0]):
            raise ValueError("material table: material id out of range")
        n_mat = tab.shape[0]
        mf = dict(mat_id=DeviceArray.from_host(mat_id.astype(np.int32)), n_mat=n_mat,
                  c_bar=DeviceArray(n_mat), sin_phi=DeviceArray(n_mat))
        for k, name in enumerate(("c0", "phi", "psi", "shear", "bulk", "lame")):
            mf[name] = DeviceArray.from_host(np.ascontiguousarray(tab[:, k]))
        self.material_fields = mf

    # -- material reduction (CONSTITUTIVE.m:178-200, reduction.m) -------------
    def reduction(self, lam):
        t0 = time.perf_counter()
        c_all = True
        if self.Davis_type == "C":
            c_all = self._c_all.get(float(lam))
            if c_all is None:
                c_all = bool(np.all(np.arctan(np.tan(self._phi_u) / lam) > self._psi_u))
                self._c_all[float(lam)] = c_all
        self._lambda.set(np.array([float(lam)]))
        CM.reduction_dev(self.n_int, self.Davis_type, self.c0, self.phi, self.psi, self._lambda, False,
                         self.c_bar, self.sin_phi, davis_c_all=c_all)
        mf = self.material_fields
        if mf is not None:
            CM.reduction_dev(mf["n_mat"], self.Davis_type, mf["c0"], mf["phi"], mf["psi"], self._lambda, False,
                             mf["c_bar"], mf["sin_phi"], davis_c_all=c_all)
        self._t("CONSTITUTIVE.reduction", t0)

    # -- stress / tangent (CONSTITUTIVE.m:202-266) -----------------------------
    def set_node_coordinates(self, coord_dev):
        """Node coordinates on the device (n_n x dim, node-major) from which
        DPhi was computed (device_setup.gradients_dev): strain and force then
        compute each point's gradients from them, the same bits as reading
        DPhi, without its 3 x n_p doubles per point of traffic."""
        if not isinstance(coord_dev, DeviceArray) or coord_dev.nbytes < 8 * self.dim * self.n_n:
            raise ValueError("node coordinates: a device array of n_n x dim doubles expected")
        self.coord_dev = coord_dev

    def _strain(self, U):
        if self.coord_dev is not None:
            self.pattern_or_raise().strain_geo(self.n_q, self.coord_dev, U, self.E)
        else:
            self.pattern_or_raise().strain(self.n_q, self.DPhi, U, self.E)

    def constitutive_problem_stress(self, U):
        t0 = time.perf_counter()
        self._strain(U)
        CM.constitutive_dev(self.dim, self.n_int, self.E, self.c_bar, self.sin_phi, self.shear, self.bulk,
                            self.lame, self.S, None)
        self._t("CONSTITUTIVE.stress", t0)

    def constitutive_problem_stress_tangent(self, U):
        t0 = time.perf_counter()
        self._strain(U)
        CM.constitutive_dev(self.dim, self.n_int, self.E, self.c_bar, self.sin_phi, self.shear, self.bulk,
                            self.lame, self.S, self.DS)
        self._t("CONSTITUTIVE.stress_tangent", t0)

    def build_F(self, F=None):
        """F = B' * (WEIGHT .* S(1:n_strain,:)) (CONSTITUTIVE.m:268-285)."""
        t0 = time.perf_counter()
        if F is None:
            F = DeviceArray((self.n_n, self.dim))
        if self.coord_dev is not None:
            self.pattern_or_raise().force_geo(self.n_q, self.coord_dev, self.S, self.s_rows, self.WEIGHT, F)
        else:
            self.pattern_or_raise().force(self.n_q, self.DPhi, self.S, self.s_rows, self.WEIGHT, F)
        self._t("CONSTITUTIVE.build_F", t0)
        return F

    def build_F_all(self, lam, U, F=None):
        self.reduction(lam)
        self.constitutive_problem_stress(U)
        return self.build_F(F)

    def build_F_reduced(self, U, F=None):
        self.constitutive_problem_stress(U)
        return self.build_F(F)

    def build_F_and_DS_all(self, lam, U, F=None):
        self.reduction(lam)
        self.constitutive_problem_stress_tangent(U)
        return self.build_F(F)

    def build_F_and_DS_reduced(self, U, F=None):
        self.constitutive_problem_stress_tangent(U)
        return self.build_F(F)

    # -- potential (CONSTITUTIVE.m:363-387, potential_{2D,3D}.m) ---------------
    def potential_energy(self, U):
        """Integrated potential sum_p WEIGHT_p Psi(E_p(U)) at the current
        reduction (the energy functional of the direct continuation)."""
        t0 = time.perf_counter()
        self._strain(U)
        if not hasattr(self, "_psi_buf"):
            self._psi_buf = DeviceArray(self.n_int)
        CM.potential_dev(self.dim, self.n_int, self.E, self.c_bar, self.sin_phi, self.shear, self.bulk,
                         self.lame, self._psi_buf)
        from .linear_solvers import dot
        val = dot(self.WEIGHT, self._psi_buf)
        self._t("CONSTITUTIVE.potential", t0)
        return val

    # -- K_r(Q,Q) (CONSTITUTIVE.m:393-532) ------------------------------------
    def pattern_or_raise(self):
        if self.pattern is None:
            raise RuntimeError("init_K_r_pattern(Q) must be called first")
        return self.pattern

    def init_K_r_pattern(self, Q, pattern=None):
        """Node-block pattern over all nodes with the Dirichlet mask of Q and
        the elastic values V_elast (CONSTITUTIVE.m:393-464).  ``pattern``: one
        already built (on the device) for this mesh and Q."""
        t0 = time.perf_counter()
        self.Q = np.asarray(Q, dtype=bool)
        self.pattern = pattern if pattern is not None else A.Pattern(self.elem, self.Q, self.dim)
        P = self.pattern
        self.mask = P.d_mask
        ns = self.n_strain
        DSe = DeviceArray((self.n_int, ns * ns))
        A.ds_elastic_dev(self.dim, self.n_int, self.shear, self.bulk, DSe)
        Kraw = DeviceArray(P.nnzb * P.bs)
        P.assemble(self.n_q, self.DPhi, DSe, self.WEIGHT, Kraw)
        # V_elast_QQ is the symmetrised elastic operator (CONSTITUTIVE.m:429)
        self.K_elast = DeviceArray(P.nnzb * P.bs)
        P.regularize(1.0, Kraw, Kraw, self.K_elast)
        self.K_tang = DeviceArray(P.nnzb * P.bs)
        self.K_r = DeviceArray(P.nnzb * P.bs)
        self.n_Q = P.n_free
        self.pattern_initialized = True
        self._t("CONSTITUTIVE.init_K_r_pattern", t0)

    def build_K_tangent_QQ_vals(self):
        t0 = time.perf_counter()
        P = self.pattern_or_raise()
        P.assemble(self.n_q, self.DPhi, self.DS, self.WEIGHT, self.K_tang)
        self._t("CONSTITUTIVE.build_K_tangent_QQ_vals", t0)
        return self.K_tang

    def build_K_r_QQ_vals(self, r):
        """K_r = r K_elast + (1-r) K_tang, symmetrised, Dirichlet-masked; the
        device block values (the pattern is fixed).  One call: the tangent is
        not stored on the way."""
        t0 = time.perf_counter()
        P = self.pattern_or_raise()
        P.assemble_K_r(self.n_q, self.DPhi, self.DS, self.WEIGHT, r, self.K_elast, self.K_r)
        self._t("CONSTITUTIVE.build_K_r_QQ_vals", t0)
        return self.K_r

    def build_K_r_QQ_vals_ref(self, r):
        """(I, J, V) in the reference's order (1-based I, J over free DOFs)."""
        K = self.build_K_r_QQ_vals(r)
        P = self.pattern
        V = DeviceArray(P.nnz_free)
        P.extract_ref(K, V)
        return V.get()
