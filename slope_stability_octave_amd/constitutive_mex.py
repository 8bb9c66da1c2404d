"""Drop-ins for the reference's constitutive MEX functions.

Same argument meaning, shapes and checks as

* ``constitutive_problem_3D_SDS_mex`` (constitutive_problem_3D_SDS_mex.c:29-64)
* ``constitutive_problem_3D_S_mex``   (constitutive_problem_3D_S_mex.c:27-58)
* ``constitutive_problem_2D_SDS_mex`` (constitutive_problem_2D_SDS_mex.c:29-64)
* ``constitutive_problem_2D_S_mex``   (constitutive_problem_2D_S_mex.c:27-59)

Arrays follow Octave's layout: ``E`` is ``n_strain x n_int`` with each point's
components contiguous (numpy: shape ``(n_strain, n_int)``, Fortran order, or
the transposed C-order ``(n_int, n_strain)`` buffer).  Outputs come back the
same way.  The computation runs in the HIP kernels of libssr_hip.so; host
buffers cross PCIe in each call exactly like the MEX boundary.  The device
entry points (``*_dev``) keep everything resident.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from ._lib import DeviceArray, call, ptr


def _pm(E, rows):
    """Point-major float64 copy/view of an Octave-layout (rows x n) array."""
    E = np.asarray(E, dtype=np.float64)
    if E.ndim == 1:
        E = E.reshape(rows, -1, order="F")
    if E.shape[0] != rows:
        raise ValueError(f"E must have {rows} rows")
    return np.asfortranarray(E)


def _vec(x, n):
    x = np.ascontiguousarray(np.asarray(x, dtype=np.float64).ravel())
    if x.size != n:
        raise ValueError(f"material array of length {x.size}, expected {n}")
    return x


def _run(dim, want_ds, E, c_bar, sin_phi, shear, bulk, lame):
    _lib.require_gpu()
    rows = 6 if dim == 3 else 3
    if dim == 2:
        E = np.asarray(E, dtype=np.float64)
        if E.ndim != 2 or E.shape[0] != 3:
            tag = "SDS" if want_ds else "S"
            raise ValueError(f"constitutive_problem_2D_{tag}_mex: E must have 3 rows.")
    E = _pm(E, rows)
    n = E.shape[1]
    mats = [_vec(x, n) for x in (c_bar, sin_phi, shear, bulk, lame)]
    ns, nd = (6, 36) if dim == 3 else (4, 9)
    S = np.empty((ns, n), order="F")
    DS = np.empty((nd, n), order="F") if want_ds else None
    if dim == 3:
        if want_ds:
            call("ssr_constitutive_problem_3D_SDS", n, ptr(E), *[ptr(m) for m in mats], ptr(S), ptr(DS))
        else:
            call("ssr_constitutive_problem_3D_S", n, ptr(E), *[ptr(m) for m in mats], ptr(S))
    else:
        if want_ds:
            call("ssr_constitutive_problem_2D_SDS", n, 3, ptr(E), *[ptr(m) for m in mats], ptr(S), ptr(DS))
        else:
            call("ssr_constitutive_problem_2D_S", n, 3, ptr(E), *[ptr(m) for m in mats], ptr(S))
    return S, DS


def constitutive_problem_3D_SDS_mex(E, c_bar, sin_phi, shear, bulk, lame):
    """[S, DS] = constitutive_problem_3D_SDS_mex(E, c_bar, sin_phi, shear, bulk, lame)."""
    return _run(3, True, E, c_bar, sin_phi, shear, bulk, lame)


def constitutive_problem_3D_S_mex(E, c_bar, sin_phi, shear, bulk, lame):
    """S = constitutive_problem_3D_S_mex(E, c_bar, sin_phi, shear, bulk, lame)."""
    return _run(3, False, E, c_bar, sin_phi, shear, bulk, lame)[0]


def constitutive_problem_2D_SDS_mex(E, c_bar, sin_phi, shear, bulk, lame):
    """[S, DS] = constitutive_problem_2D_SDS_mex(E, ...); S is 4 x n (s11, s22, s12, s33)."""
    return _run(2, True, E, c_bar, sin_phi, shear, bulk, lame)


def constitutive_problem_2D_S_mex(E, c_bar, sin_phi, shear, bulk, lame):
    """S = constitutive_problem_2D_S_mex(E, ...)."""
    return _run(2, False, E, c_bar, sin_phi, shear, bulk, lame)[0]


def constitutive_dev(dim, n, E, c_bar, sin_phi, shear, bulk, lame, S, DS=None, branch=None):
    """Device-resident MEX-layout evaluation on DeviceArrays (no host traffic)."""
    name = "ssr_dev_constitutive_3D" if dim == 3 else "ssr_dev_constitutive_2D"
    call(name, int(n), ptr(E), ptr(c_bar), ptr(sin_phi), ptr(shear), ptr(bulk), ptr(lame), ptr(S),
         ptr(DS), ptr(branch))


def constitutive_mat_dev(dim, n, n_q, E, mat_id, n_mat, c_bar, sin_phi, shear, bulk, lame, S, DS=None,
                         branch=None):
    """S (and DS) with per-material fields: point p takes material
    mat_id[p // n_q] (device arrays)."""
    call("ssr_dev_constitutive_mat", int(dim), int(n), int(n_q), ptr(E), ptr(mat_id), int(n_mat), ptr(c_bar),
         ptr(sin_phi), ptr(shear), ptr(bulk), ptr(lame), ptr(S), ptr(DS) if DS is not None else None,
         ptr(branch) if branch is not None else None)


def reduction_dev(n, davis, c0, phi, psi, lam, lam_per_point, c_bar, sin_phi, davis_c_all=True):
    """Davis reduction (reduction.m:25-58) on device arrays; writes c_bar, sin_phi."""
    call("ssr_dev_reduction", int(n), ord(davis), ptr(c0), ptr(phi), ptr(psi), ptr(lam),
         1 if lam_per_point else 0, 1 if davis_c_all else 0, ptr(c_bar), ptr(sin_phi))


def potential_dev(dim, n, E, c_bar, sin_phi, shear, bulk, lame, Psi):
    """Psi per point (potential_{2D,3D}.m), device arrays."""
    call("ssr_dev_potential", int(dim), int(n), ptr(E), ptr(c_bar), ptr(sin_phi), ptr(shear), ptr(bulk), ptr(lame),
         ptr(Psi))
