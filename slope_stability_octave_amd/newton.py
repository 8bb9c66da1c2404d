"""Semismooth Newton solvers of the SSR path, over device-resident fields.

Mirrors +NEWTON:
* ``newton_ind_SSR`` -- newton_ind_SSR.m:1-206, bordered Newton for
  F_lambda(U) = f, f'U = omega, unknowns U and lambda;
* ``damping_ALG5``   -- damping_ALG5.m:1-60 (residual-decrease halving);
* ``newton``         -- newton.m:1-192, Newton at fixed (already reduced) lambda;
* ``damping``        -- damping.m:1-56 (directional-derivative bisection).

Every vector (U, F, f, V, W, d_U) is a DeviceArray of shape (n_n, dim);
reductions (norms, f'V) return scalars to the host, which owns the control
flow exactly as the reference's Octave loop does.  Restriction to the free
DOFs Q is the Dirichlet mask (zero at fixed DOFs).

``driver="native"`` runs the same loops inside libssr_hip.so
(``ssr_newton_ind_SSR``, ``ssr_newton``, ``ssr_newton_ind_LL``): one library
call per Newton solve instead of a few dozen per iteration, with the same
kernels in the same order, so both drivers give the same iterates.  The
functions here default to the Python loop; the continuations started by
problems.run_ssr / run_ssr_direct / run_ll default to the native one.
"""
from __future__ import annotations

import ctypes
import math
import time

import numpy as np

from . import _lib
from ._lib import DeviceArray, call, ptr
from .linear_solvers import axpby, dot, lincomb

_vp = ctypes.c_void_p
_int = ctypes.c_int
_dbl = ctypes.c_double
_i64 = ctypes.c_int64

_lib.extend_signatures({
    "ssr_newton_work_len": [_vp, _vp],
    "ssr_newton_ind_SSR": [_vp, _vp, _int, _vp, _dbl, _dbl, _int, _int, _dbl, _dbl, _vp, _i64, _vp, _vp, _vp,
                           _vp, _vp, _vp, _vp],
    "ssr_newton": [_vp, _vp, _int, _vp, _int, _int, _dbl, _dbl, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp],
    "ssr_newton_ind_LL": [_vp, _vp, _int, _vp, _dbl, _dbl, _int, _int, _dbl, _dbl, _vp, _i64, _vp, _vp, _vp, _vp,
                          _vp, _vp, _vp],
    "ssr_SSR_indirect_continuation": [_vp, _vp, _int, _int, _vp, _vp, _vp, _vp, _vp, _int, _vp, _vp, _int, _vp,
                                      _vp, _int, _vp, _vp, _vp],
    "ssr_solver_clone": [_vp, _int, _vp],
})

DRIVERS = ("python", "native")
NEWTON_STATUS = {0: "unknown", 1: "converged", 2: "stalled_regularization", 3: "max_iterations",
                 4: "nan_residual"}


class _Constitutive(ctypes.Structure):
    """ctypes image of ssr_constitutive (include/ssr_hip.h)."""

    _fields_ = [("dim", _int), ("n_q", _int), ("davis", _int), ("s_rows", _int), ("n_int", _i64),
                ("pattern", _vp), ("DPhi1", _vp), ("DPhi2", _vp), ("DPhi3", _vp), ("WEIGHT", _vp),
                ("c0", _vp), ("phi", _vp), ("psi", _vp), ("shear", _vp), ("bulk", _vp), ("lame", _vp),
                ("phi_host", _vp), ("psi_host", _vp), ("lambda_", _vp), ("c_bar", _vp), ("sin_phi", _vp),
                ("E", _vp), ("S", _vp), ("DS", _vp), ("K_elast", _vp), ("K_tang", _vp), ("K_r", _vp),
                ("mask", _vp), ("f", _vp), ("n_friction", _i64),
                ("mat_id", _vp), ("n_mat", _i64), ("c0_m", _vp), ("phi_m", _vp), ("psi_m", _vp),
                ("shear_m", _vp), ("bulk_m", _vp), ("lame_m", _vp), ("c_bar_m", _vp), ("sin_phi_m", _vp),
                ("coord", _vp)]


def native_description(cmb, f):
    """The ssr_constitutive view of a CONSTITUTIVE object and load vector f.
    The arrays stay owned by ``cmb`` and ``f``; the host copies of phi and psi
    are kept on ``cmb`` for as long as it lives."""
    if not cmb.pattern_initialized:
        raise RuntimeError("init_K_r_pattern(Q) must be called first")
    n, ns, P = cmb.n_int, cmb.n_strain, cmb.pattern
    need = dict(WEIGHT=n, c0=n, phi=n, psi=n, shear=n, bulk=n, lame=n, c_bar=n, sin_phi=n, E=n * ns,
                S=n * cmb.s_rows, DS=n * ns * ns, K_elast=P.nnzb * P.bs, K_tang=P.nnzb * P.bs, K_r=P.nnzb * P.bs)
    for name, count in need.items():
        if getattr(cmb, name).nbytes < 8 * count:
            raise ValueError(f"native Newton driver: {name} holds fewer than {count} doubles")
    if any(d.nbytes < 8 * cmb.n_p * n for d in cmb.DPhi) or f.nbytes < 8 * P.n_dof:
        raise ValueError("native Newton driver: DPhi or f too small for the mesh")
    host = getattr(cmb, "_native_host", None)
    if host is None:
        # the (phi, psi) of the distinct materials: the Davis C test needs no more
        host = cmb._native_host = (np.ascontiguousarray(cmb._phi_u, dtype=np.float64).ravel(),
                                   np.ascontiguousarray(cmb._psi_u, dtype=np.float64).ravel())
    if host[0].size != host[1].size or host[0].size == 0:
        raise ValueError("native Newton driver: phi and psi of the materials must pair up")
    d = _Constitutive()
    d.dim, d.n_q, d.davis, d.s_rows, d.n_int = cmb.dim, cmb.n_q, ord(cmb.Davis_type), cmb.s_rows, cmb.n_int
    d.pattern = cmb.pattern.handle.value
    d.DPhi1, d.DPhi2 = ptr(cmb.DPhi[0]).value, ptr(cmb.DPhi[1]).value
    d.DPhi3 = ptr(cmb.DPhi[2]).value if cmb.dim == 3 else None
    for name in ("WEIGHT", "c0", "phi", "psi", "shear", "bulk", "lame", "c_bar", "sin_phi", "E", "S", "DS",
                 "K_elast", "K_tang", "K_r", "mask"):
        setattr(d, name, ptr(getattr(cmb, name)).value)
    d.lambda_ = ptr(cmb._lambda).value
    d.phi_host, d.psi_host = host[0].ctypes.data, host[1].ctypes.data
    d.n_friction = host[0].size
    mt = getattr(cmb, "material_fields", None)
    if mt is not None:
        # per-material fields: one reduction per material, the kernels read each element's
        d.mat_id = ptr(mt["mat_id"]).value
        d.n_mat = mt["n_mat"]
        for name in ("c0", "phi", "psi", "shear", "bulk", "lame", "c_bar", "sin_phi"):
            setattr(d, name + "_m", ptr(mt[name]).value)
    if getattr(cmb, "coord_dev", None) is not None:
        d.coord = ptr(cmb.coord_dev).value
    d.f = ptr(f).value
    return d


class Workspace:
    """Scratch device vectors reused across Newton iterations."""

    def __init__(self, n_n, dim):
        shape = (n_n, dim)
        self.shape = shape
        self.F = DeviceArray(shape)
        self.F_eps = DeviceArray(shape)
        self.G = DeviceArray(shape)
        self.rhs = DeviceArray(shape)
        self.V = DeviceArray(shape)
        self.W = DeviceArray(shape)
        self.dU = DeviceArray(shape)
        self.Ua = DeviceArray(shape)
        self.Fa = DeviceArray(shape)
        self.tmp = DeviceArray(shape)
        self._native = None

    def native_work(self, n):
        """Device scratch of the native Newton driver (n doubles, kept)."""
        if self._native is None or self._native.shape[0] < n:
            self._native = DeviceArray(n)
        return self._native


def norm_Q(a, b, mask, tmp):
    """||a(Q) - b(Q)||."""
    lincomb(1.0, a, -1.0, b, tmp)
    return math.sqrt(max(dot(tmp, tmp, mask), 0.0))


def damping_ALG5(it_damp_max, U_it, lambda_it, d_U, d_l, f, criterion, mask, cmb, ws):
    if math.isnan(d_l):
        return 0.0
    alpha = 1.0
    it_damp = 0
    while True:
        it_damp += 1
        lincomb(1.0, U_it, alpha, d_U, ws.Ua)
        cmb.build_F_all(lambda_it + alpha * d_l, ws.Ua, ws.Fa)
        crit_alpha = norm_Q(ws.Fa, f, mask, ws.tmp)
        if crit_alpha >= criterion or math.isnan(crit_alpha):
            alpha = alpha / 2
        else:
            break
        if it_damp >= it_damp_max:
            alpha = 0.0
            break
    return alpha


def newton_ind_SSR(U_ini, omega, lambda_ini, it_newt_max, it_damp_max, tol, r_min, K_elast, Q, f, cmb,
                   linear_solver, ws=None, log=None, driver="python"):
    """Returns (U_it, lambda_it, flag_N, it, history)."""
    if not cmb.pattern_initialized:
        cmb.init_K_r_pattern(Q)
    if driver == "native":
        return _newton_ind_SSR_native(U_ini, omega, lambda_ini, it_newt_max, it_damp_max, tol, r_min, f, cmb,
                                      linear_solver, ws, log)
    if driver != "python":
        raise ValueError(f"unknown Newton driver {driver!r} (expected one of {DRIVERS})")
    mask = cmb.mask
    if ws is None:
        ws = Workspace(cmb.n_n, cmb.dim)
    U_it = DeviceArray(U_ini.shape)
    lincomb(1.0, U_ini, 0.0, U_ini, U_it)
    lambda_it = float(lambda_ini)
    eps = tol / 1000
    norm_f = math.sqrt(dot(f, f, mask))
    fQ_f = None
    flag_N = 0
    compute_diffs = True
    r = r_min
    it = 0
    hist = dict(residual=[], r=[], alpha=[], linear_iters_W=[], linear_iters_V=[])
    rel_resid = float("nan")
    criterion = float("nan")
    status = "unknown"
    t_start = time.perf_counter()
    F = ws.F
    while True:
        it += 1
        if compute_diffs:
            cmb.build_F_and_DS_all(lambda_it, U_it, F)
            criterion = norm_Q(F, f, mask, ws.tmp)
            rel_resid = criterion / norm_f
            hist["residual"].append(rel_resid)
            if rel_resid < tol and it > 1:
                status = "converged"
                break
        else:
            hist["residual"].append(rel_resid)
        hist["r"].append(r)
        K_r = cmb.build_K_r_QQ_vals(r)
        # G = (F(lambda+eps) - F)/eps  (one stress-only evaluation)
        cmb.build_F_all(lambda_it + eps, U_it, ws.F_eps)
        lincomb(1.0 / eps, ws.F_eps, -1.0 / eps, F, ws.G)
        linear_solver.setup_preconditioner(K_r)
        linear_solver.A_orthogonalize(K_r)
        # W = K \ (-G(Q)),  V = K \ (f(Q) - F(Q))
        lincomb(-1.0, ws.G, 0.0, ws.G, ws.rhs)
        _mask_inplace(ws.rhs, mask)
        _, lin_w = linear_solver.solve(K_r, ws.rhs, ws.W)
        lincomb(1.0, f, -1.0, F, ws.rhs)
        _mask_inplace(ws.rhs, mask)
        _, lin_v = linear_solver.solve(K_r, ws.rhs, ws.V)
        hist["linear_iters_W"].append(lin_w)
        hist["linear_iters_V"].append(lin_v)
        fV = dot(f, ws.V, mask)
        fW = dot(f, ws.W, mask)
        d_l = -fV / fW if fW != 0 else float("nan")
        lincomb(1.0, ws.V, d_l, ws.W, ws.dU)
        alpha = damping_ALG5(it_damp_max, U_it, lambda_it, ws.dU, d_l, f, criterion, mask, cmb, ws)
        hist["alpha"].append(alpha)
        compute_diffs = True
        if alpha < 1e-1:
            if alpha == 0:
                compute_diffs = False
                r = r * 2
            else:
                r = r * 2 ** (1 / 4)
        else:
            linear_solver.expand_deflation_basis(ws.W)
            linear_solver.expand_deflation_basis(ws.V)
            if alpha > 0.5:
                r = max(r / math.sqrt(2), r_min)
        if alpha == 0 and r > 1:
            status = "stalled_regularization"
            if rel_resid > 10 * tol:
                flag_N = 1
            break
        # U = U + alpha dU; U = omega U / (f(Q)'U(Q)); lambda += alpha d_l
        axpby(alpha, ws.dU, 1.0, U_it)
        fU = dot(f, U_it, mask)
        axpby(0.0, U_it, omega / fU, U_it)
        lambda_it = lambda_it + alpha * d_l
        if log:
            log(f"  newton_ind_SSR it={it}  rel_resid={rel_resid:.2e}  alpha={alpha:.2f}  r={r:.3g}  "
                f"lin_it=[W:{lin_w},V:{lin_v}]")
        if math.isnan(rel_resid) or it == it_newt_max:
            if rel_resid > 10 * tol or math.isnan(rel_resid):
                flag_N = 1
            status = "nan_residual" if math.isnan(rel_resid) else "max_iterations"
            break
    hist["status"] = status
    hist["wall_time"] = time.perf_counter() - t_start
    return U_it, lambda_it, flag_N, it, hist


def _newton_ind_SSR_native(U_ini, omega, lambda_ini, it_newt_max, it_damp_max, tol, r_min, f, cmb, linear_solver,
                           ws, log):
    """newton_ind_SSR.m:1-206 as one ssr_newton_ind_SSR call; the history and
    the solver's counters are filled from the per-iteration record it returns."""
    if ws is None:
        ws = Workspace(cmb.n_n, cmb.dim)
    desc = native_description(cmb, f)
    wl = _i64(0)
    call("ssr_newton_work_len", ctypes.byref(desc), ctypes.byref(wl))
    work = ws.native_work(wl.value)
    U_it = DeviceArray(U_ini.shape)
    rec = np.full((int(it_newt_max), 6), np.nan)
    times = np.zeros(4)
    lam, flag, it, st = _dbl(0.0), _int(0), _int(0), _int(0)
    method = 1 if getattr(linear_solver, "method", "dfgmres") == "dcg" else 0
    t_start = time.perf_counter()
    call("ssr_newton_ind_SSR", ctypes.byref(desc), linear_solver.handle, method, ptr(U_ini), float(omega),
         float(lambda_ini), int(it_newt_max), int(it_damp_max), float(tol), float(r_min), ptr(work), wl.value,
         ptr(U_it), ctypes.byref(lam), ctypes.byref(flag), ctypes.byref(it), ctypes.byref(st), ptr(rec), ptr(times))
    wall = time.perf_counter() - t_start
    linear_solver._K = cmb.K_r
    n_it = it.value
    hist = dict(residual=[float(x) for x in rec[:n_it, 0]], r=[], alpha=[], linear_iters_W=[], linear_iters_V=[])
    for i in range(n_it):
        row = rec[i]
        if math.isnan(row[1]):  # stopped at the convergence test
            continue
        hist["r"].append(float(row[1]))
        hist["alpha"].append(float(row[2]))
        lin_w, lin_v = int(row[3]), int(row[4])
        hist["linear_iters_W"].append(lin_w)
        hist["linear_iters_V"].append(lin_v)
        _count_native(linear_solver, int(row[5]), [lin_w, lin_v])
        if log and not (i == n_it - 1 and st.value == 2):
            log(f"  newton_ind_SSR it={i + 1}  rel_resid={row[0]:.2e}  alpha={row[2]:.2f}  r={row[1]:.3g}  "
                f"lin_it=[W:{lin_w},V:{lin_v}]")
    _record_native(linear_solver, times, wall)
    hist["status"] = NEWTON_STATUS.get(st.value, "unknown")
    hist["wall_time"] = wall
    return U_it, lam.value, flag.value, n_it, hist


def _record_native(linear_solver, times, wall):
    tt = linear_solver.timings
    if tt is not None:
        for key, v in zip(("DFGMRES.solve", "DFGMRES.A_orthogonalize", "DFGMRES.update_preconditioner",
                           "newton_ind_SSR.native_other"), times):
            tt[key] = tt.get(key, 0.0) + float(v)
        tt["newton_ind_SSR.native"] = tt.get("newton_ind_SSR.native", 0.0) + wall


def _count_native(linear_solver, k, lin_its):
    c = linear_solver.counters
    if c is None:
        return
    c["DFGMRES.deflation_size.max"] = max(c.get("DFGMRES.deflation_size.max", 0), k)
    c["DFGMRES.deflation_size.sum"] = c.get("DFGMRES.deflation_size.sum", 0) + k
    c["DFGMRES.solve.n_iters"] = c.get("DFGMRES.solve.n_iters", 0) + sum(lin_its)
    c["DFGMRES.solve.calls"] = c.get("DFGMRES.solve.calls", 0) + len(lin_its)
    at_max = sum(1 for x in lin_its if x >= linear_solver.max_iterations)
    if at_max:
        c["DFGMRES.solve.at_maxit"] = c.get("DFGMRES.solve.at_maxit", 0) + at_max


def _newton_native(U_ini, tol, it_newt_max, it_damp_max, r_min, f, cmb, linear_solver, ws, log):
    """newton.m:1-192 as one ssr_newton call (material reduced beforehand)."""
    if ws is None:
        ws = Workspace(cmb.n_n, cmb.dim)
    desc = native_description(cmb, f)
    wl = _i64(0)
    call("ssr_newton_work_len", ctypes.byref(desc), ctypes.byref(wl))
    work = ws.native_work(wl.value)
    U_it = DeviceArray(U_ini.shape)
    rec = np.full((int(it_newt_max), 5), np.nan)
    times = np.zeros(4)
    flag, it, st = _int(0), _int(0), _int(0)
    method = 1 if getattr(linear_solver, "method", "dfgmres") == "dcg" else 0
    t_start = time.perf_counter()
    call("ssr_newton", ctypes.byref(desc), linear_solver.handle, method, ptr(U_ini), int(it_newt_max),
         int(it_damp_max), float(tol), float(r_min), ptr(work), wl.value, ptr(U_it), ctypes.byref(flag),
         ctypes.byref(it), ctypes.byref(st), ptr(rec), ptr(times))
    wall = time.perf_counter() - t_start
    linear_solver._K = cmb.K_r
    n_it = it.value
    hist = dict(residual=[], r=[], alpha=[], linear_iters=[])
    for i in range(n_it):
        row = rec[i]
        if i == 0 or (hist["alpha"] and hist["alpha"][-1] != 0):  # a fresh residual was computed
            hist["residual"].append(float(row[0]))
        if math.isnan(row[1]):  # stopped at the convergence test
            continue
        hist["r"].append(float(row[1]))
        hist["alpha"].append(float(row[2]))
        hist["linear_iters"].append(int(row[3]))
        _count_native(linear_solver, int(row[4]), [int(row[3])])
        if log and not (i == n_it - 1 and st.value == 2):
            log(f"  newton it={i + 1}  rel_resid={row[0]:.2e}  alpha={row[2]:.2f}  r={row[1]:.3g}  "
                f"lin_it={int(row[3])}")
    _record_native(linear_solver, times, wall)
    hist["status"] = NEWTON_STATUS.get(st.value, "unknown")
    return U_it, flag.value, n_it, hist


def _mask_inplace(x, mask):
    from ._lib import call, ptr
    call("ssr_dev_mul", int(np.prod(x.shape)), ptr(mask), ptr(x))


def damping(it_damp_max, U_it, dU, F, f, cmb, ws):
    """Bisection on the directional derivative (damping.m:24-55)."""
    lincomb(1.0, F, -1.0, f, ws.tmp)
    decrease = dot(ws.tmp, dU)
    ndu = math.sqrt(dot(dU, dU))
    if math.isnan(decrease) or math.isinf(ndu) or decrease >= 0:
        return 0.0
    alpha, a_min, a_max = 1.0, 0.0, 1.0
    it_damp = 0
    while it_damp < it_damp_max:
        it_damp += 1
        lincomb(1.0, U_it, alpha, dU, ws.Ua)
        cmb.build_F_reduced(ws.Ua, ws.Fa)
        lincomb(1.0, ws.Fa, -1.0, f, ws.tmp)
        decrease = dot(ws.tmp, dU)
        if decrease < 0:
            if alpha == 1:
                break
            a_min = alpha
        else:
            a_max = alpha
        alpha = (a_min + a_max) / 2
    return alpha


def newton(U_ini, tol, it_newt_max, it_damp_max, r_min, K_elast, Q, f, cmb, linear_solver, ws=None, log=None,
           driver="python"):
    """Newton at the current (reduced) material; returns (U, flag_N, it, history)."""
    if not cmb.pattern_initialized:
        cmb.init_K_r_pattern(Q)
    if driver == "native":
        return _newton_native(U_ini, tol, it_newt_max, it_damp_max, r_min, f, cmb, linear_solver, ws, log)
    if driver != "python":
        raise ValueError(f"unknown Newton driver {driver!r} (expected one of {DRIVERS})")
    mask = cmb.mask
    if ws is None:
        ws = Workspace(cmb.n_n, cmb.dim)
    U_it = DeviceArray(U_ini.shape)
    lincomb(1.0, U_ini, 0.0, U_ini, U_it)
    it = 0
    flag_N = 0
    r = r_min
    compute_diffs = True
    norm_f = math.sqrt(dot(f, f, mask))
    rel_resid = float("nan")
    hist = dict(residual=[], r=[], alpha=[], linear_iters=[])
    status = "unknown"
    F = ws.F
    while True:
        it += 1
        if compute_diffs:
            cmb.build_F_and_DS_reduced(U_it, F)
            criterion = norm_Q(F, f, mask, ws.tmp)
            rel_resid = criterion / norm_f
            hist["residual"].append(rel_resid)
            if rel_resid < tol:
                status = "converged"
                break
        hist["r"].append(r)
        K_r = cmb.build_K_r_QQ_vals(r)
        linear_solver.setup_preconditioner(K_r)
        linear_solver.A_orthogonalize(K_r)
        lincomb(1.0, f, -1.0, F, ws.rhs)
        _mask_inplace(ws.rhs, mask)
        _, lin_it = linear_solver.solve(K_r, ws.rhs, ws.dU)
        hist["linear_iters"].append(lin_it)
        alpha = damping(it_damp_max, U_it, ws.dU, F, f, cmb, ws)
        hist["alpha"].append(alpha)
        compute_diffs = True
        if alpha < 1e-1:
            if alpha == 0:
                compute_diffs = False
                r = r * 2
            else:
                r = r * 2 ** (1 / 4)
        else:
            linear_solver.expand_deflation_basis(ws.dU)
            if alpha > 0.5:
                r = max(r / math.sqrt(2), r_min)
        if alpha == 0 and r > 1:
            status = "stalled_regularization"
            if rel_resid > 10 * tol:
                flag_N = 1
            break
        axpby(alpha, ws.dU, 1.0, U_it)
        if log:
            log(f"  newton it={it}  rel_resid={rel_resid:.2e}  alpha={alpha:.2f}  r={r:.3g}  lin_it={lin_it}")
        if math.isnan(rel_resid) or it == it_newt_max:
            if rel_resid > 10 * tol or math.isnan(rel_resid):
                flag_N = 1
            status = "nan_residual" if math.isnan(rel_resid) else "max_iterations"
            break
    hist["status"] = status
    return U_it, flag_N, it, hist


def newton_ind_LL(U_ini, t_ini, omega, it_newt_max, it_damp_max, tol, r_min, K_elast, Q, f, cmb, linear_solver,
                  ws=None, log=None, driver="python"):
    """newton_ind_LL.m:1-198: bordered Newton for F(U) = t f, f'U = omega at
    the unreduced material (unknowns U and the load factor t), with the
    directional-derivative damping of damping.m against f = 0 (f'dU = 0 by
    construction of d_t).  As in the reference, t takes the full d_t even
    when U is damped, and reaching it_newt_max is a failure.
    Returns (U_it, t_it, flag_N, it)."""
    if not cmb.pattern_initialized:
        cmb.init_K_r_pattern(Q)
    if driver == "native":
        return _newton_ind_LL_native(U_ini, t_ini, omega, it_newt_max, it_damp_max, tol, r_min, f, cmb,
                                     linear_solver, ws, log)
    if driver != "python":
        raise ValueError(f"unknown Newton driver {driver!r} (expected one of {DRIVERS})")
    mask = cmb.mask
    if ws is None:
        ws = Workspace(cmb.n_n, cmb.dim)
    U_it = DeviceArray(U_ini.shape)
    lincomb(1.0, U_ini, 0.0, U_ini, U_it)
    t_it = float(t_ini)
    norm_f = math.sqrt(dot(f, f, mask))
    zero = DeviceArray(U_ini.shape, zero=True)
    flag_N = 0
    r = r_min
    compute_diffs = True
    it = 0
    rel_resid = float("nan")
    F = ws.F
    while True:
        it += 1
        if compute_diffs:
            cmb.build_F_and_DS_reduced(U_it, F)
        lincomb(1.0, F, -t_it, f, ws.tmp)
        criterion = math.sqrt(max(dot(ws.tmp, ws.tmp, mask), 0.0))
        rel_resid = criterion / norm_f
        if criterion < tol * norm_f and it > 1:
            break
        K_r = cmb.build_K_r_QQ_vals(r)
        linear_solver.setup_preconditioner(K_r)
        linear_solver.A_orthogonalize(K_r)
        lincomb(1.0, f, 0.0, f, ws.rhs)
        _mask_inplace(ws.rhs, mask)
        _, lin_w = linear_solver.solve(K_r, ws.rhs, ws.W)
        lincomb(t_it, f, -1.0, F, ws.rhs)
        _mask_inplace(ws.rhs, mask)
        _, lin_v = linear_solver.solve(K_r, ws.rhs, ws.V)
        fV = dot(f, ws.V, mask)
        fW = dot(f, ws.W, mask)
        d_t = -fV / fW if fW != 0 else float("nan")
        lincomb(1.0, ws.V, d_t, ws.W, ws.dU)
        alpha = damping(it_damp_max, U_it, ws.dU, F, zero, cmb, ws)
        compute_diffs = True
        if alpha < 1e-1:
            if alpha == 0:
                compute_diffs = False
                r = r * 2
            else:
                r = r * 2 ** (1 / 4)
        else:
            linear_solver.expand_deflation_basis(ws.W)
            linear_solver.expand_deflation_basis(ws.V)
            if alpha > 0.5:
                r = max(r / math.sqrt(2), r_min)
        if alpha == 0 and r > 1:
            flag_N = 1
            break
        axpby(alpha, ws.dU, 1.0, U_it)
        fU = dot(f, U_it, mask)
        axpby(0.0, U_it, omega / fU, U_it)
        t_it = t_it + d_t
        if log:
            log(f"  newton_ind_LL it={it}  rel_resid={rel_resid:.2e}  alpha={alpha:.2f}  r={r:.3g}  "
                f"lin_it=[W:{lin_w},V:{lin_v}]")
        if math.isnan(rel_resid) or it == it_newt_max:
            flag_N = 1
            break
    return U_it, t_it, flag_N, it


def _newton_ind_LL_native(U_ini, t_ini, omega, it_newt_max, it_damp_max, tol, r_min, f, cmb, linear_solver, ws,
                          log):
    """newton_ind_LL.m:1-198 as one ssr_newton_ind_LL call."""
    if ws is None:
        ws = Workspace(cmb.n_n, cmb.dim)
    desc = native_description(cmb, f)
    wl = _i64(0)
    call("ssr_newton_work_len", ctypes.byref(desc), ctypes.byref(wl))
    work = ws.native_work(wl.value)
    U_it = DeviceArray(U_ini.shape)
    rec = np.full((int(it_newt_max), 6), np.nan)
    times = np.zeros(4)
    t, flag, it, st = _dbl(0.0), _int(0), _int(0), _int(0)
    method = 1 if getattr(linear_solver, "method", "dfgmres") == "dcg" else 0
    t_start = time.perf_counter()
    call("ssr_newton_ind_LL", ctypes.byref(desc), linear_solver.handle, method, ptr(U_ini), float(t_ini),
         float(omega), int(it_newt_max), int(it_damp_max), float(tol), float(r_min), ptr(work), wl.value, ptr(U_it),
         ctypes.byref(t), ctypes.byref(flag), ctypes.byref(it), ctypes.byref(st), ptr(rec), ptr(times))
    wall = time.perf_counter() - t_start
    linear_solver._K = cmb.K_r
    n_it = it.value
    for i in range(n_it):
        row = rec[i]
        if math.isnan(row[1]):  # stopped at the convergence test
            continue
        lin_w, lin_v = int(row[3]), int(row[4])
        _count_native(linear_solver, int(row[5]), [lin_w, lin_v])
        if log and not (i == n_it - 1 and st.value == 2):
            log(f"  newton_ind_LL it={i + 1}  rel_resid={row[0]:.2e}  alpha={row[2]:.2f}  r={row[1]:.3g}  "
                f"lin_it=[W:{lin_w},V:{lin_v}]")
    _record_native(linear_solver, times, wall)
    return U_it, t.value, flag.value, n_it
