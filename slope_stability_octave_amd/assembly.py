"""Tangent assembly, strain gather and internal force on the GPU.

``assemble_K_tangent_vals`` / ``assemble_K_tangent_vals_2D`` are drop-ins for
the reference MEX files (assemble_K_tangent_vals.c:38-172,
assemble_K_tangent_vals_2D.c:31-165): same arguments, int64 scatter map,
host arrays in and out.

``Pattern`` is the device-resident replacement of CONSTITUTIVE's pattern and
scatter map (CONSTITUTIVE.m:393-464, 582-635): node-block CSR over all nodes
plus the reference-order extraction map.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import DeviceArray, call, ptr

_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_int = ctypes.c_int
_dbl = ctypes.c_double

_lib.extend_signatures({
    "ssr_assemble_K_tangent_vals": [_int, _i64, _i64, _i64, _int, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp],
    "ssr_dev_assemble_K_tangent_vals": [_int, _i64, _i64, _i64, _int, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp],
    "ssr_pattern_create": [_int, _i64, _i64, _i64, _vp, _vp, _int, _vp],
    "ssr_pattern_destroy": [_vp],
    "ssr_pattern_info": [_vp, _vp],
    "ssr_pattern_get": [_vp, _int, _vp],
    "ssr_pattern_dev_ptr": [_vp, _int, _vp],
    "ssr_dev_assemble_bsr": [_vp, _int, _vp, _vp, _vp, _vp, _int, _vp, _vp],
    "ssr_dev_strain": [_vp, _int, _vp, _vp, _vp, _vp, _vp],
    "ssr_dev_force": [_vp, _int, _vp, _vp, _vp, _vp, _int, _vp, _vp],
    "ssr_dev_strain_geo": [_vp, _int, _vp, _vp, _vp],
    "ssr_dev_force_geo": [_vp, _int, _vp, _vp, _int, _vp, _vp],
    "ssr_dev_regularize": [_vp, _dbl, _vp, _vp, _vp],
    "ssr_dev_assemble_K_r": [_vp, _int, _vp, _vp, _vp, _vp, _int, _vp, _dbl, _vp, _vp],
    "ssr_dev_extract_ref": [_vp, _vp, _vp],
    "ssr_dev_ds_elastic": [_int, _i64, _vp, _vp, _vp],
})


def _assemble_host(dim, DPhi, DS, WEIGHT, scatter_map, n_q, nnz_out):
    _lib.require_gpu()
    if not (isinstance(scatter_map, np.ndarray) and scatter_map.dtype == np.int64):
        raise TypeError(("assemble_K_tangent_vals_2D" if dim == 2 else "assemble_K_tangent_vals")
                        + ":type scatter_map must be int64.")
    D = [np.asfortranarray(np.asarray(x, dtype=np.float64)) for x in DPhi]
    n_p, n_int = D[0].shape
    ns = 6 if dim == 3 else 3
    DSf = np.asfortranarray(np.asarray(DS, dtype=np.float64))
    if DSf.shape != (ns * ns, n_int):
        raise ValueError(f"DS must have size ({ns * ns}, n_int)")
    smap = np.asfortranarray(scatter_map)
    nld = dim * n_p
    if smap.shape[0] != nld * nld:
        raise ValueError("scatter_map row count must be n_local_dof^2.")
    n_e = smap.shape[1]
    W = np.ascontiguousarray(np.asarray(WEIGHT, dtype=np.float64).ravel())
    V = np.zeros(int(nnz_out))
    call("ssr_assemble_K_tangent_vals", dim, n_p, n_int, n_e, int(n_q), ptr(D[0]), ptr(D[1]),
         ptr(D[2]) if dim == 3 else _vp(0), ptr(DSf), ptr(W), ptr(smap), int(nnz_out), ptr(V))
    return V


def assemble_K_tangent_vals(DPhi1, DPhi2, DPhi3, DS, WEIGHT, scatter_map, n_q, nnz_out):
    """V_tang = assemble_K_tangent_vals(DPhi1, DPhi2, DPhi3, DS, WEIGHT, scatter_map, n_q, nnz_out)."""
    return _assemble_host(3, (DPhi1, DPhi2, DPhi3), DS, WEIGHT, scatter_map, n_q, nnz_out)


def assemble_K_tangent_vals_2D(DPhi1, DPhi2, DS, WEIGHT, scatter_map, n_q, nnz_out):
    """V_tang = assemble_K_tangent_vals_2D(DPhi1, DPhi2, DS, WEIGHT, scatter_map, n_q, nnz_out)."""
    return _assemble_host(2, (DPhi1, DPhi2), DS, WEIGHT, scatter_map, n_q, nnz_out)


class _DevRef:
    """A borrowed device address (usable wherever a DeviceArray is)."""

    def __init__(self, addr):
        self.ptr = addr


class Pattern:
    """Node-block CSR pattern of K (all nodes, dim x dim blocks), element
    block map, transpose map and reference-order extraction map."""

    def __init__(self, elem, Q, dim, device=True):
        elem = np.ascontiguousarray(np.asarray(elem, dtype=np.int64).T)  # element-major
        Qm = np.ascontiguousarray(np.asarray(Q, dtype=np.uint8).T)  # node-major
        n_e, n_p = elem.shape
        n_n = Qm.shape[0]
        if device:
            _lib.require_gpu()
        h = _vp(0)
        call("ssr_pattern_create", dim, n_n, n_p, n_e, ptr(elem), ptr(Qm), 1 if device else 0, ctypes.byref(h))
        self._adopt(h)

    @classmethod
    def from_handle(cls, h):
        """Wrap a pattern the library created (e.g. ssr_pattern_create_dev)."""
        self = object.__new__(cls)
        self._adopt(h)
        return self

    def _adopt(self, h):
        self.handle = h
        info = np.zeros(8, dtype=np.int64)
        call("ssr_pattern_info", h, ptr(info))
        self.dim, self.n_n, self.n_p, self.n_e, self.nnzb, self.n_free, self.nnz_free, _ = [int(x) for x in info]
        self.bs = self.dim * self.dim
        self.n_dof = self.dim * self.n_n

    def _dptr(self, which):
        out = _vp(0)
        call("ssr_pattern_dev_ptr", self.handle, which, ctypes.byref(out))
        return _DevRef(out.value)

    @property
    def d_row_ptr(self):
        return self._dptr(0)

    @property
    def d_col(self):
        return self._dptr(1)

    @property
    def d_mask(self):
        return self._dptr(7)

    def table(self, which):
        sizes = {0: (self.n_n + 1, np.int32), 1: (self.nnzb, np.int32), 2: (self.n_e * self.n_p * self.n_p, np.int32),
                 3: (self.nnzb, np.int32), 4: (self.n_n, np.int32), 5: (self.nnz_free, np.int64),
                 6: (self.n_dof, np.uint8)}
        n, dt = sizes[which]
        out = np.empty(n, dtype=dt)
        call("ssr_pattern_get", self.handle, which, ptr(out))
        return out

    @property
    def row_ptr(self):
        return self.table(0)

    @property
    def col(self):
        return self.table(1)

    @property
    def ref_pos(self):
        return self.table(5)

    @property
    def mask(self):
        return self.table(6)

    def __del__(self):
        try:
            if self.handle:
                _lib.load().ssr_pattern_destroy(self.handle)
                self.handle = None
        except Exception:
            pass

    # ---- device operations -------------------------------------------------
    def assemble(self, n_q, DPhi, DS, WEIGHT, out, ds_layout=0):
        """out (nnzb*bs) = sum_e Ke in block values; DS MEX layout (0) or packed symmetric (1)."""
        call("ssr_dev_assemble_bsr", self.handle, int(n_q), ptr(DPhi[0]), ptr(DPhi[1]),
             ptr(DPhi[2]) if self.dim == 3 else _vp(0), ptr(DS), int(ds_layout), ptr(WEIGHT), ptr(out))

    def strain(self, n_q, DPhi, U, E):
        call("ssr_dev_strain", self.handle, int(n_q), ptr(DPhi[0]), ptr(DPhi[1]),
             ptr(DPhi[2]) if self.dim == 3 else _vp(0), ptr(U), ptr(E))

    def force(self, n_q, DPhi, S, s_rows, WEIGHT, F):
        call("ssr_dev_force", self.handle, int(n_q), ptr(DPhi[0]), ptr(DPhi[1]),
             ptr(DPhi[2]) if self.dim == 3 else _vp(0), ptr(S), int(s_rows), ptr(WEIGHT), ptr(F))

    def strain_geo(self, n_q, coord, U, E):
        """strain() with the gradients computed from node coordinates (n_n x dim)."""
        call("ssr_dev_strain_geo", self.handle, int(n_q), ptr(coord), ptr(U), ptr(E))

    def force_geo(self, n_q, coord, S, s_rows, WEIGHT, F):
        """force() with the gradients computed from node coordinates (n_n x dim)."""
        call("ssr_dev_force_geo", self.handle, int(n_q), ptr(coord), ptr(S), int(s_rows), ptr(WEIGHT), ptr(F))

    def regularize(self, r, Kel, Kt, Kr):
        call("ssr_dev_regularize", self.handle, float(r), ptr(Kel), ptr(Kt), ptr(Kr))

    def assemble_K_r(self, n_q, DPhi, DS, WEIGHT, r, Kel, Kr, ds_layout=0):
        """Kr = regularised tangent straight from DS (assemble + regularize in one call)."""
        call("ssr_dev_assemble_K_r", self.handle, int(n_q), ptr(DPhi[0]), ptr(DPhi[1]),
             ptr(DPhi[2]) if self.dim == 3 else _vp(0), ptr(DS), int(ds_layout), ptr(WEIGHT), float(r), ptr(Kel),
             ptr(Kr))

    def extract_ref(self, Kvals, Vref):
        call("ssr_dev_extract_ref", self.handle, ptr(Kvals), ptr(Vref))


def ds_elastic_dev(dim, n_int, shear, bulk, DS):
    call("ssr_dev_ds_elastic", int(dim), int(n_int), ptr(shear), ptr(bulk), ptr(DS))
