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
"""
from __future__ import annotations

import math
import time

import numpy as np

from ._lib import DeviceArray
from .linear_solvers import axpby, dot, lincomb


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
                   linear_solver, ws=None, log=None):
    """Returns (U_it, lambda_it, flag_N, it, history)."""
    if not cmb.pattern_initialized:
        cmb.init_K_r_pattern(Q)
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


def newton(U_ini, tol, it_newt_max, it_damp_max, r_min, K_elast, Q, f, cmb, linear_solver, ws=None, log=None):
    """Newton at the current (reduced) material; returns (U, flag_N, it, history)."""
    if not cmb.pattern_initialized:
        cmb.init_K_r_pattern(Q)
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
                  ws=None, log=None):
    """newton_ind_LL.m:1-198: bordered Newton for F(U) = t f, f'U = omega at
    the unreduced material (unknowns U and the load factor t), with the
    directional-derivative damping of damping.m against f = 0 (f'dU = 0 by
    construction of d_t).  As in the reference, t takes the full d_t even
    when U is damped, and reaching it_newt_max is a failure.
    Returns (U_it, t_it, flag_N, it)."""
    if not cmb.pattern_initialized:
        cmb.init_K_r_pattern(Q)
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
