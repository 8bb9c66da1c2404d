"""Smoothed-aggregation AMG preconditioner (libssr_hip.so) for the DFGMRES /
DCG solvers: rigid-body near-null space as near_null_space_elasticity_3D.m:46-103
(centred coordinates; 3 modes in 2D, 6 in 3D)."""
from __future__ import annotations

import numpy as np

import ctypes

from ._lib import call, extend_signatures, ptr

_vp, _i64, _int, _cp = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_char_p
extend_signatures({
    "ssr_amg_setup_ijv": [_cp, _vp, _vp, _vp, _i64, _i64, _vp, _int, _int, _vp, _int, _vp],
    "ssr_amg_update_values": [_cp, _vp, _i64, _vp],
    "ssr_amg_apply": [_cp, _vp, _i64, _vp],
    "ssr_amg_apply_dev": [_cp, _vp, _vp],
    "ssr_amg_clear": [_cp],
    "ssr_amg_info": [_vp, _vp],
})

# reuse_hierarchy: the per-step solver copies of a continuation share one AMG
# object, so the hierarchy is built once (then every rebuild_every setups);
# the finest level always applies the current tangent; a failed Newton solve
# of the continuation asks for a rebuild before its retry.
# cheb_degree None: 3 in 3D, 2 in 2D (measured with the node-block smoother of
# the finest level, which is the default)
DEFAULTS = dict(theta=0.0, max_levels=12, max_coarse=600, cheb_degree=None, cheb_lower=1.0 / 30.0, cheb_upper=1.1,
                refresh_every=0, rebuild_every=0, reuse_hierarchy=True)


def rigid_body_modes(coord):
    """Near-null space over all dim*n_n DOFs (node-major), column-major."""
    dim, n_n = coord.shape
    c = coord - coord.mean(axis=1, keepdims=True)
    modes = []
    for t in range(dim):
        u = np.zeros((dim, n_n))
        u[t] = 1.0
        modes.append(u)
    if dim == 2:
        u = np.zeros((2, n_n))
        u[0], u[1] = -c[1], c[0]
        modes.append(u)
    else:
        x, y, z = c
        for (a, b, s) in ((1, 2, (-z, y)), (0, 2, (z, -x)), (0, 1, (-y, x))):
            u = np.zeros((3, n_n))
            u[a], u[b] = s
            modes.append(u)
    return np.stack([m.ravel(order="F") for m in modes], axis=1)


def attach(solver, options=None):
    """Install SA-AMG as the solver's preconditioner.  options may contain
    'coord' (dim x n_n, required) and any key of DEFAULTS."""
    o = dict(DEFAULTS)
    o.update(options or {})
    coord = o.get("coord")
    if coord is None:
        raise ValueError("AMG needs the node coordinates (options['coord'])")
    B = np.asfortranarray(rigid_body_modes(np.asarray(coord, dtype=np.float64)))
    if o["cheb_degree"] is None:
        o["cheb_degree"] = 3 if np.asarray(coord).shape[0] == 3 else 2
    vec = np.array([o["theta"], o["max_levels"], o["max_coarse"], o["cheb_degree"], o["cheb_lower"],
                    o["cheb_upper"], o["refresh_every"], o["rebuild_every"]], dtype=np.float64)
    call("ssr_solver_set_amg", solver.handle, ptr(B), B.shape[1], ptr(vec))


# ---------------------------------------------------------------------------
# id-keyed instances from COO triplets: the LINEAR_SOLVERS.hypre_boomeramg_*
# functions (hypre_boomeramg_setup_ijv.m, _update_values.m, _apply.m, _clear.m)
# ---------------------------------------------------------------------------
def _opts_vector(opts):
    if opts is None:
        return None
    o = dict(DEFAULTS)
    o.update(opts)
    if o["cheb_degree"] is None:
        o["cheb_degree"] = 3
    return np.array([o["theta"], o["max_levels"], o["max_coarse"], o["cheb_degree"], o["cheb_lower"],
                     o["cheb_upper"], o["refresh_every"], o["rebuild_every"]], dtype=np.float64)


def hypre_boomeramg_setup_ijv(I, J, V, n, null_space, opts, instance_id, dofs_per_node=3, force_rebuild=False):
    """info = hypre_boomeramg_setup_ijv(I, J, V, n, null_space, opts, instance_id)
    (1-based I, J); returns a dict like the reference's info struct."""
    I = np.ascontiguousarray(np.asarray(I, dtype=np.float64).ravel())
    J = np.ascontiguousarray(np.asarray(J, dtype=np.float64).ravel())
    V = np.ascontiguousarray(np.asarray(V, dtype=np.float64).ravel())
    if not (I.size == J.size == V.size):
        raise ValueError("hypre:badIJV I, J, V must have equal length.")
    Z = None if null_space is None else np.asfortranarray(np.asarray(null_space, dtype=np.float64))
    nz = 0 if Z is None else (Z.shape[1] if Z.ndim == 2 else 1)
    ov = _opts_vector(opts)
    info = np.zeros(5)
    call("ssr_amg_setup_ijv", str(instance_id).encode(), ptr(I), ptr(J), ptr(V), I.size, int(n),
         ptr(Z) if Z is not None else None, nz, int(dofs_per_node), ptr(ov) if ov is not None else None,
         1 if force_rebuild else 0, ptr(info))
    return dict(instance_id=str(instance_id), n=int(info[0]), nnz=int(info[1]), num_functions=int(info[2]),
                num_interp_vectors=int(info[3]), setup_time_seconds=float(info[4]))


def hypre_boomeramg_update_values(V, instance_id):
    V = np.ascontiguousarray(np.asarray(V, dtype=np.float64).ravel())
    t = ctypes.c_double(0.0)
    call("ssr_amg_update_values", str(instance_id).encode(), ptr(V), V.size, ctypes.byref(t))
    return t.value


def hypre_boomeramg_apply(r, instance_id):
    r = np.ascontiguousarray(np.asarray(r, dtype=np.float64).ravel())
    y = np.empty_like(r)
    call("ssr_amg_apply", str(instance_id).encode(), ptr(r), r.size, ptr(y))
    return y


def hypre_boomeramg_clear(instance_id=None):
    call("ssr_amg_clear", None if instance_id is None else str(instance_id).encode())


def hypre_boomeramg_info():
    c, t = ctypes.c_int64(0), ctypes.c_int64(0)
    call("ssr_amg_info", ctypes.byref(c), ctypes.byref(t))
    return dict(instances=c.value, total_n=t.value)
