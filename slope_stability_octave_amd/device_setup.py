"""Problem setup on the GPU: FE gradients and weights, per-point material
fields, the volume force and the fixed pattern, all built on the device from
the mesh (libssr_hip.so: ssr_dev_fe_gradients, ssr_dev_materials,
ssr_dev_volume_force, ssr_pattern_create_dev).

The host keeps the mesh as the generators or the HDF5 loader give it (coord,
elem, Q) and the small material table; everything per point or per element is
produced in HBM.  Reference counterparts: elastic_stiffness_matrix_{2D,3D}.m
(DPhi, WEIGHT), heterogenous_materials.m (fields), vector_volume_{2D,3D}.m
(f_V), CONSTITUTIVE.init_K_r_pattern / build_scatter_map (pattern).
"""
from __future__ import annotations

import ctypes
import time

import numpy as np

from . import _lib
from ._lib import DeviceArray, call, ptr

_vp, _i64, _int = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
_lib.extend_signatures({
    "ssr_pattern_create_dev": [_int, _i64, _i64, _i64, _vp, _vp, _vp],
    "ssr_dev_fe_gradients": [_int, _i64, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp],
    "ssr_dev_volume_force": [_vp, _int, _vp, _vp, _int, _vp],
    "ssr_dev_materials": [_i64, _int, _int, _int, _vp, _vp, _vp],
})

# field order of the material table (heterogenous_materials.m:64-87)
FIELDS = ("c0", "phi", "psi", "shear", "bulk", "lame", "gamma")


def material_table(mats):
    """n_mat x 7 table: c0, phi, psi (rad), shear, bulk, lame, gamma_sat
    (fully saturated, as the slope drivers use), computed exactly as
    fem.materials_per_point computes each field."""
    rows = []
    for m in mats:
        young, nu = m["young"], m["poisson"]
        shear = young / (2 * (1 + nu))
        bulk = young / (3 * (1 - 2 * nu))
        rows.append([m["c0"], m["phi"] * np.pi / 180, m["psi"] * np.pi / 180, shear, bulk, bulk - 2 * shear / 3,
                     m["gamma_sat"]])
    return np.asarray(rows, dtype=np.float64)


def pattern_dev(elem_dev, Q_dev, dim, n_n, n_p, n_e):
    """Pattern (assembly.Pattern) built on the device from device arrays."""
    from .assembly import Pattern
    h = _vp(0)
    call("ssr_pattern_create_dev", int(dim), int(n_n), int(n_p), int(n_e), ptr(elem_dev), ptr(Q_dev),
         ctypes.byref(h))
    return Pattern.from_handle(h)


def gradients_dev(coord_dev, elem_dev, dim, n_n, n_p, n_e):
    """(DPhi list of DeviceArray n_int x n_p, WEIGHT DeviceArray n_int)."""
    n_q = 11 if dim == 3 else 7
    n_int = n_e * n_q
    DPhi = [DeviceArray((n_int, n_p)) for _ in range(dim)]
    W = DeviceArray(n_int)
    call("ssr_dev_fe_gradients", int(dim), int(n_n), int(n_p), int(n_e), ptr(coord_dev), ptr(elem_dev),
         ptr(DPhi[0]), ptr(DPhi[1]), ptr(DPhi[2]) if dim == 3 else _vp(0), ptr(W))
    return DPhi, W


def materials_dev(table, mat_id_dev, n_e, n_q):
    """dict field -> DeviceArray(n_int) from the material table."""
    tab = DeviceArray.from_host(np.ascontiguousarray(table, dtype=np.float64))
    n_int = n_e * n_q
    out = {k: DeviceArray(n_int) for k in FIELDS}
    ptrs = (ctypes.c_void_p * len(FIELDS))(*[out[k].ptr for k in FIELDS])
    call("ssr_dev_materials", int(n_e), int(n_q), int(table.shape[0]), len(FIELDS), ptr(tab), ptr(mat_id_dev),
         ctypes.cast(ptrs, _vp))
    return out


def volume_force_dev(pattern, n_q, WEIGHT, f_int, f_rows):
    """f_V as DeviceArray (n_n, dim), node-major."""
    f = DeviceArray((pattern.n_n, pattern.dim))
    call("ssr_dev_volume_force", pattern.handle, int(n_q), ptr(WEIGHT), ptr(f_int), int(f_rows), ptr(f))
    return f


def finish_device(coord, elem, Q, dim, mats, mat_id):
    """The problem dict of problems._finish with every per-point and
    per-element array produced on the GPU (``device`` True): DPhi, WEIGHT,
    the material fields and f are DeviceArrays and ``pattern`` is the device
    pattern.  ``setup_s`` holds the wall time of the device part."""
    _lib.require_gpu()
    t0 = time.perf_counter()
    n_p, n_e = elem.shape
    n_n = coord.shape[1]
    n_q = 11 if dim == 3 else 7
    coord_d = DeviceArray.from_host(np.ascontiguousarray(np.asarray(coord, dtype=np.float64).T))
    elem_d = DeviceArray.from_host(np.ascontiguousarray(np.asarray(elem, dtype=np.int32).T))
    Q_d = DeviceArray.from_host(np.ascontiguousarray(np.asarray(Q, dtype=np.uint8).T))
    mid_d = DeviceArray.from_host(np.ascontiguousarray(np.asarray(mat_id, dtype=np.int32)))
    DPhi, W = gradients_dev(coord_d, elem_d, dim, n_n, n_p, n_e)
    table = material_table(mats)
    m = materials_dev(table, mid_d, n_e, n_q)
    pattern = pattern_dev(elem_d, Q_d, dim, n_n, n_p, n_e)
    f = volume_force_dev(pattern, n_q, W, m["gamma"], 1)
    call("ssr_synchronize")
    t = time.perf_counter() - t0
    # coord_dev stays with the problem: the Newton path computes each point's
    # gradients from it (the bits of DPhi) instead of reading DPhi
    return dict(coord=coord, coord_dev=coord_d, elem=elem, Q=Q, dim=dim, DPhi=DPhi, WEIGHT=W, n_q=n_q, mats=m, f=f,
                n_int=n_e * n_q,
                n_dof_free=int(np.asarray(Q).sum()), pattern=pattern, material_table=table,
                mat_id=np.asarray(mat_id), device=True, setup_s=t)
