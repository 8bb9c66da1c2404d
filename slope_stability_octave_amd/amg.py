"""Smoothed-aggregation AMG preconditioner (libssr_hip.so) for the DFGMRES /
DCG solvers: rigid-body near-null space as near_null_space_elasticity_3D.m:46-103
(centred coordinates; 3 modes in 2D, 6 in 3D)."""
from __future__ import annotations

import numpy as np

from ._lib import call, ptr

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
