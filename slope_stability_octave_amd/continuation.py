"""Indirect SSR continuation (the time-to-FoS harness).

Mirrors +CONTINUATION/SSR_indirect_continuation.m:1-237 and
init_phase_SSR_indirect_continuation.m:1-143: omega-stepping with the secant
predictor U_ini = d_omega (U - U_old)/(omega - omega_old) + U (:119), retries
with halved d_omega (:164-167), d_omega doubling when the lambda increment
shrinks (:214-216) and the stop rules (:203-211, :220-227).  The factor of
safety is lambda_hist[-1].

``predictor`` (optional) replaces the secant extrapolation: a callable
(state) -> U_ini, (U_ini, lambda_ini) or None; None falls back to the secant
step.  ``state`` holds the last converged fields (``U_hist``, newest first,
at most four), the histories and the target omega.  ``recorder`` (optional)
is called after every accepted step as recorder(state, U_new, lambda_new)
with the state the step started from (training-data capture).
"""
from __future__ import annotations

import ctypes
import math
import time

import numpy as np

from ._lib import DeviceArray, call, ptr
from .linear_solvers import dot, lincomb
from .newton import Workspace, newton, newton_ind_SSR


class ContinuationError(RuntimeError):
    pass



def _rebuild_after_failure(solver):
    """A failed Newton solve asks for a fresh preconditioner before the retry:
    with one AMG hierarchy shared by the steps' solver copies, the retry's
    first setup rebuilds it from the current tangent."""
    inv = getattr(solver, "invalidate_preconditioner", None)
    if inv is not None:
        inv()

def _umax(U):
    """max over nodes of |U(:, node)| (SSR_indirect_continuation.m:92-93,182), on the device."""
    out = ctypes.c_double(0.0)
    call("ssr_node_norm_max", int(U.shape[1]), int(U.shape[0]), ptr(U), ctypes.byref(out))
    return out.value


def init_phase_SSR_indirect_continuation(lambda_init, d_lambda_init, d_lambda_min, it_newt_max, it_damp_max,
                                         tol, r_min, K_elast, Q, f, cmb, solver, ws, log=None, driver="python"):
    if not cmb.pattern_initialized:
        cmb.init_K_r_pattern(Q)
    mask = cmb.mask
    U_ini = DeviceArray(f.shape, zero=True)
    if log:
        log(f"Init step 1: lambda={lambda_init:.6g}")
    cmb.reduction(lambda_init)
    its = []
    U_it, flag_N, it_newt, _ = newton(U_ini, tol, it_newt_max, it_damp_max, r_min, K_elast, Q, f, cmb, solver,
                                      ws, log, driver=driver)
    its.append(it_newt)
    if flag_N == 1:
        raise ContinuationError("Initial choice of lambda seems to be too large.")
    U1 = U_it
    omega1 = dot(f, U1, mask)
    d_lambda = d_lambda_init
    lambda1 = lambda_init
    solver.expand_deflation_basis(U_it)
    while True:
        lambda_it = lambda1 + d_lambda
        cmb.reduction(lambda_it)
        U_it, flag_N, it_newt, _ = newton(U1, tol, it_newt_max, it_damp_max, r_min, K_elast, Q, f, cmb, solver,
                                          ws, log, driver=driver)
        its.append(it_newt)
        if flag_N == 1:
            d_lambda = d_lambda / 2
        else:
            U2 = U_it
            omega2 = dot(f, U2, mask)
            lambda2 = lambda_it
            if (omega2 - omega1) / max(1.0, omega1) < 1e-5:
                U1, lambda1, omega1 = U2, lambda2, omega2
            else:
                break
        if d_lambda < d_lambda_min:
            raise ContinuationError("It seems that the FoS is equal to lambda_init.")
        if lambda1 > 10:
            raise ContinuationError("It seems that the FoS is greater than 10.")
    if log:
        log(f"   lambda1 = {lambda1:.6g}, lambda2 = {lambda2:.6g}, omega1 = {omega1:.6g}, omega2 = {omega2:.6g}")
    return U1, U2, omega1, omega2, lambda1, lambda2, its


_STOP_REASONS = {1: "minimal increase of lambda achieved", 2: "maximal omega achieved",
                 3: "impossible to increase omega", 4: "maximal number of steps achieved"}


def _shares_hierarchy(solver):
    """Whether solver.copy() shares the solver's preconditioner object (AMG
    with reuse_hierarchy) or builds a node-block Jacobi of its own; None when
    the copy would build a new AMG (not done by the library's continuation)."""
    from .linear_solvers import _amg_option
    name = getattr(solver, "preconditioner_name", None)
    if name == "amg":
        return True if _amg_option(solver.amg_options, "reuse_hierarchy") else None
    return False if name == "block_jacobi" else None


def _SSR_indirect_continuation_native(lambda_init, d_lambda_init, d_lambda_min, d_lambda_diff_scaled_min, step_max,
                                      omega_max_stop, it_newt_max, it_damp_max, tol, r_min, Q, f, cmb,
                                      linear_system_solver, log=None):
    """The whole continuation as one ssr_SSR_indirect_continuation call (the
    same steps as the host loop below, over the native Newton loops)."""
    from . import _lib
    from .newton import native_description
    t_total = time.perf_counter()
    if not cmb.pattern_initialized:
        cmb.init_K_r_pattern(Q)
    share = _shares_hierarchy(linear_system_solver)
    if share is None:
        raise ValueError("the library's continuation needs a solver whose copies share its AMG hierarchy "
                         "(reuse_hierarchy) or use node-block Jacobi")
    desc = native_description(cmb, f)
    params = np.array([lambda_init, d_lambda_init, d_lambda_min, d_lambda_diff_scaled_min, step_max, omega_max_stop,
                       it_newt_max, it_damp_max, tol, r_min], dtype=np.float64)
    cap = int(step_max) + 8
    lam_h, om_h, um_h = np.zeros(cap), np.zeros(cap), np.zeros(cap)
    acap, icap = 6 * cap, 4096
    att, ini = np.zeros(acap, dtype=np.int32), np.zeros(icap, dtype=np.int32)
    nh, na, ni, stop = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
    kry = ctypes.c_int64(0)
    U = DeviceArray(f.shape)
    method = 1 if getattr(linear_system_solver, "method", "dfgmres") == "dcg" else 0
    lib = _lib.load()
    rc = lib.ssr_SSR_indirect_continuation(
        ctypes.byref(desc), linear_system_solver.handle, method, 1 if share else 0, ptr(params), ptr(U), ptr(lam_h),
        ptr(om_h), ptr(um_h), cap, ctypes.byref(nh), ptr(att), acap, ctypes.byref(na), ptr(ini), icap,
        ctypes.byref(ni), ctypes.byref(kry), ctypes.byref(stop))
    if rc != 0:
        msg = lib.ssr_last_error().decode(errors="replace")
        if stop.value < 0:
            raise ContinuationError(msg)
        raise _lib.SSRError(f"ssr_SSR_indirect_continuation: {msg} (status {rc})")
    linear_system_solver._K = cmb.K_r
    n = nh.value
    if log:
        for k in range(2, n):
            log(f"  accepted: lambda={lam_h[k]:.6g}, omega={om_h[k]:.6g}, U_max={um_h[k]:.6g}")
        log(f"Stop: {_STOP_REASONS.get(stop.value, 'unknown')}.")
    c = linear_system_solver.counters
    if c is not None:
        c["DFGMRES.solve.n_iters"] = c.get("DFGMRES.solve.n_iters", 0) + int(kry.value)
    stats = dict(attempt_newton_iterations=[int(v) for v in att[:na.value]],
                 init_newton_iterations=[int(v) for v in ini[:ni.value]],
                 step_lambda=[float(v) for v in lam_h[2:n]], step_omega=[float(v) for v in om_h[2:n]],
                 predictor_used=[False] * max(0, na.value), stop_reason=_STOP_REASONS.get(stop.value, "unknown"),
                 krylov_iterations=int(kry.value), total_wall_time=time.perf_counter() - t_total)
    return U, lam_h[:n].copy(), om_h[:n].copy(), um_h[:n].copy(), stats


def SSR_indirect_continuation(lambda_init, d_lambda_init, d_lambda_min, d_lambda_diff_scaled_min, step_max,
                              omega_max_stop, it_newt_max, it_damp_max, tol, r_min, K_elast, Q, f, cmb,
                              linear_system_solver, log=None, predictor=None, recorder=None, driver="python"):
    """Returns (U, lambda_hist, omega_hist, Umax_hist, stats).  ``driver``
    picks the Newton loops (newton_ind_SSR and the initial phase's newton):
    "python" (newton.py) or "native" (the same loops in libssr_hip.so);
    "library" runs the whole continuation in libssr_hip.so
    (ssr_SSR_indirect_continuation: the secant predictor, no predictor or
    recorder hooks) and gives the "native" driver's numbers."""
    if driver == "library":
        if predictor is not None or recorder is not None:
            raise ValueError("the library's continuation has no predictor or recorder hooks")
        return _SSR_indirect_continuation_native(lambda_init, d_lambda_init, d_lambda_min, d_lambda_diff_scaled_min,
                                                 step_max, omega_max_stop, it_newt_max, it_damp_max, tol, r_min, Q,
                                                 f, cmb, linear_system_solver, log)
    t_total = time.perf_counter()
    ws = Workspace(cmb.n_n, cmb.dim)
    mask_ready = cmb.pattern_initialized or cmb.init_K_r_pattern(Q) or True  # noqa: F841
    mask = cmb.mask
    stats = dict(attempt_step=[], attempt_success=[], attempt_wall_time=[], attempt_newton_iterations=[],
                 attempt_omega_target=[], attempt_lambda_before=[], attempt_lambda_after=[],
                 step_lambda=[], step_omega=[], step_newton_iterations=[], newton_histories=[],
                 predictor_used=[])
    U_old, U, omega_old, omega, lambda_old, lam, init_its = init_phase_SSR_indirect_continuation(
        lambda_init, d_lambda_init, d_lambda_min, it_newt_max, it_damp_max, tol, r_min, K_elast, Q, f, cmb,
        linear_system_solver.copy(), ws, log, driver=driver)
    stats["init_newton_iterations"] = init_its
    linear_system_solver.expand_deflation_basis(U_old)
    lambda_hist = [lambda_old, lam]
    omega_hist = [omega_old, omega]
    Umax_hist = [_umax(U_old), _umax(U)]
    d_omega = omega - omega_old
    if omega_max_stop < omega + d_omega:
        raise ContinuationError("Too small value of omega_max_stop. It is necessary to increase this input quantinty.")
    step = 2
    n_omega, n_omega_max = 0, 5
    U_ini = DeviceArray(U.shape)
    history = dict(lambda_hist=lambda_hist, omega_hist=omega_hist)
    U_hist = [U, U_old]  # converged fields, newest first (at most four kept)
    while True:
        omega_it = min(omega + d_omega, omega_max_stop)
        d_omega = omega_it - omega
        if log:
            log(f"Step {step + 1}: omega_target={omega_it:.6g}, d_omega={d_omega:.6g}")
        used = False
        lambda_ini = lam
        state = dict(U=U, U_old=U_old, omega=omega, omega_old=omega_old, omega_it=omega_it, lam=lam,
                     history=history, U_hist=list(U_hist), lambda_hist=list(lambda_hist),
                     omega_hist=list(omega_hist))
        if predictor is not None:
            guess = predictor(state)
            if guess is not None:
                if isinstance(guess, tuple):
                    guess, lambda_ini = guess
                lincomb(1.0, guess, 0.0, guess, U_ini)
                used = True
        if not used:
            a = d_omega / (omega - omega_old)
            # U_ini = a (U - U_old) + U
            lincomb(1.0 + a, U, -a, U_old, U_ini)
        stats["predictor_used"].append(used)
        t0 = time.perf_counter()
        U_it, lambda_it, flag, it_newt, hist = newton_ind_SSR(
            U_ini, omega_it, lambda_ini, it_newt_max, it_damp_max, tol, r_min, K_elast, Q, f, cmb,
            linear_system_solver.copy(), ws, log, driver=driver)
        wall = time.perf_counter() - t0
        stats["attempt_step"].append(step + 1)
        stats["attempt_success"].append(flag == 0)
        stats["attempt_wall_time"].append(wall)
        stats["attempt_newton_iterations"].append(it_newt)
        stats["attempt_omega_target"].append(omega_it)
        stats["attempt_lambda_before"].append(lam)
        stats["attempt_lambda_after"].append(lambda_it if flag == 0 else float("nan"))
        stats["newton_histories"].append(hist)
        if flag == 1:
            if log:
                log(f"  retry: reducing d_omega from {d_omega:.6g} to {d_omega / 2:.6g}")
            _rebuild_after_failure(linear_system_solver)
            d_omega = d_omega / 2
            n_omega += 1
        else:
            if recorder is not None:
                recorder(state, U_it, lambda_it)
            step += 1
            U_old, U = U, U_it
            U_hist = [U_it] + U_hist[:3]
            linear_system_solver.expand_deflation_basis(U)
            omega_old, omega = omega, omega_it
            d_lambda = lambda_it - lam
            lam = lambda_it
            n_omega = 0
            omega_hist.append(omega)
            lambda_hist.append(lam)
            Umax_hist.append(_umax(U))
            stats["step_lambda"].append(lam)
            stats["step_omega"].append(omega)
            stats["step_newton_iterations"].append(it_newt)
            scaled = d_lambda / d_omega * (omega_hist[-1] - omega_hist[0])
            if log:
                log(f"  accepted: lambda={lam:.6g}, d_lambda={d_lambda:.6g}, d_lambda_diff_scaled={scaled:.6g}, "
                    f"omega={omega:.6g}, d_omega={d_omega:.6g}, U_max={Umax_hist[-1]:.6g}")
            if scaled < d_lambda_diff_scaled_min:
                if log:
                    log("Stop: minimal increase of lambda achieved.")
                break
            if omega >= omega_max_stop:
                if log:
                    log("Stop: maximal omega achieved.")
                break
            if (lambda_hist[-1] - lambda_hist[-2]) < 0.9 * (lambda_hist[-2] - lambda_hist[-3]):
                d_omega = 2 * d_omega
        if n_omega >= n_omega_max:
            if log:
                log("Stop: impossible to increase omega.")
            break
        if step >= step_max:
            if log:
                log("Stop: maximal number of steps achieved.")
            break
    stats["total_wall_time"] = time.perf_counter() - t_total
    return U, np.array(lambda_hist), np.array(omega_hist), np.array(Umax_hist), stats


# ---------------------------------------------------------------------------
# Direct SSR continuation (lambda-stepping with omega = -dJ/dlambda)
# ---------------------------------------------------------------------------
def omega_SSR_direct_continuation(lambda_ini, U_ini, eps, d_lambda, it_newt_max, it_damp_max, tol, r_min, K_elast,
                                  Q, f, cmb, solver, ws, log=None, driver="python"):
    """omega_SSR_direct_continuation.m:49-81: U at lambda, and omega as the
    backward difference of the energy J = Psi - f'U over eps.  Returns
    (U, omega, flag, newton iterations)."""
    mask = cmb.mask
    cmb.reduction(lambda_ini)
    U, flag, it1, _ = newton(U_ini, tol, it_newt_max, it_damp_max, r_min, K_elast, Q, f, cmb, solver, ws, log,
                             driver=driver)
    if flag == 1:
        return U, 0.0, 1, it1
    J = cmb.potential_energy(U) - dot(f, U, mask)
    beta = min(1.0, eps / d_lambda)
    U_beta = DeviceArray(U.shape)
    lincomb(beta, U_ini, 1.0 - beta, U, U_beta)
    cmb.reduction(lambda_ini - eps)
    U_eps, flag, it2, _ = newton(U_beta, tol, it_newt_max, it_damp_max, r_min, K_elast, Q, f, cmb, solver, ws, log,
                                 driver=driver)
    if flag == 1:
        return U, 0.0, 1, it1 + it2
    J_eps = cmb.potential_energy(U_eps) - dot(f, U_eps, mask)
    return U, (J_eps - J) / eps, 0, it1 + it2


def init_phase_SSR_direct_continuation(lambda_init, d_lambda_init, d_lambda_min, it_newt_max, it_damp_max, tol,
                                       eps, r_min, K_elast, Q, f, cmb, solver, ws, log=None, driver="python"):
    """init_phase_SSR_direct_continuation.m:1-143."""
    U0 = DeviceArray(f.shape, zero=True)
    if log:
        log(f"Init step 1: lambda={lambda_init:.6g}")
    U1, omega1, flag, _ = omega_SSR_direct_continuation(lambda_init, U0, eps, 1000.0, it_newt_max, it_damp_max, tol,
                                                        r_min, K_elast, Q, f, cmb, solver.copy(), ws, log,
                                                        driver=driver)
    if flag == 1:
        raise ContinuationError("Initial choice of lambda seems to be too large.")
    d_lambda = d_lambda_init
    lambda1 = lambda_init
    while True:
        lambda_it = lambda1 + d_lambda
        solver.expand_deflation_basis(U1)
        U_it, omega_it, flag, _ = omega_SSR_direct_continuation(lambda_it, U1, eps, d_lambda, it_newt_max,
                                                                it_damp_max, tol, r_min, K_elast, Q, f, cmb,
                                                                solver.copy(), ws, log, driver=driver)
        if flag == 1:
            _rebuild_after_failure(solver)
            d_lambda = d_lambda / 2
        else:
            U2, omega2, lambda2 = U_it, omega_it, lambda_it
            if (omega2 - omega1) / max(1.0, omega1) < 1e-5:
                U1, lambda1, omega1 = U2, lambda2, omega2
            else:
                break
        if d_lambda < d_lambda_min:
            raise ContinuationError("It seems that FoS is equal to lambda_init.")
        if lambda1 > 10:
            raise ContinuationError("It seems that the FoS is greater than 10.")
    if log:
        log(f"   lambda1 = {lambda1:.6g}, lambda2 = {lambda2:.6g}, omega1 = {omega1:.6g}, omega2 = {omega2:.6g}")
    return U1, U2, omega1, omega2, lambda1, lambda2


def SSR_direct_continuation(lambda_init, d_lambda_init, d_lambda_min, d_lambda_diff_scaled_min, step_max,
                            it_newt_max, it_damp_max, tol, r_min, K_elast, Q, f, cmb, linear_system_solver, log=None,
                            driver="python"):
    """SSR_direct_continuation.m:47-146: lambda steps, halved on failure or a
    decreasing omega, halved when omega grows by more than 1.5x; stops on the
    scaled lambda increase, the minimal increment or the step limit.
    Returns (U, lambda_hist, omega_hist, Umax_hist, work_hist)."""
    if not cmb.pattern_initialized:
        cmb.init_K_r_pattern(Q)
    mask = cmb.mask
    ws = Workspace(cmb.n_n, cmb.dim)
    eps = tol * 100
    U_old, U, omega_old, omega, lambda_old, lam = init_phase_SSR_direct_continuation(
        lambda_init, d_lambda_init, d_lambda_min, it_newt_max, it_damp_max, tol, eps, r_min, K_elast, Q, f, cmb,
        linear_system_solver, ws, log, driver=driver)
    linear_system_solver.expand_deflation_basis(U)
    lambda_hist, omega_hist = [lambda_old, lam], [omega_old, omega]
    Umax_hist = [_umax(U), _umax(U)]
    work_hist = [dot(U_old, f), dot(U, f)]
    d_omega = omega - omega_old
    d_lambda = lam - lambda_init
    step = 2
    while True:
        lambda_it = lam + d_lambda
        if log:
            log(f"Step {step + 1}: lambda_target={lambda_it:.6g}")
        U_it, omega_it, flag, _ = omega_SSR_direct_continuation(lambda_it, U, eps, d_lambda, it_newt_max,
                                                                it_damp_max, tol, r_min, K_elast, Q, f, cmb,
                                                                linear_system_solver.copy(), ws, log, driver=driver)
        d_omega_test = omega_it - omega_hist[-1]
        if flag == 1 or d_omega_test < 0:
            if log:
                log(f"  retry: reducing d_lambda from {d_lambda:.6g} to {d_lambda / 2:.6g}")
            if flag == 1:
                _rebuild_after_failure(linear_system_solver)
            d_lambda = d_lambda / 2
        else:
            U, omega, lam = U_it, omega_it, lambda_it
            step += 1
            lambda_hist.append(lam)
            omega_hist.append(omega)
            Umax_hist.append(_umax(U))
            work_hist.append(dot(U, f))
            linear_system_solver.expand_deflation_basis(U)
            scaled = d_lambda / d_omega_test * (omega_hist[-1] - omega_hist[0])
            if log:
                log(f"  accepted: lambda={lam:.6g}, d_lambda={d_lambda:.6g}, omega={omega:.6g}, "
                    f"d_omega={d_omega:.6g}, d_lambda_diff_scaled={scaled:.6g}, U_max={Umax_hist[-1]:.6g}")
            if scaled < d_lambda_diff_scaled_min:
                if log:
                    log("Stop: minimal increase of lambda achieved.")
                break
            if d_omega_test > 1.5 * d_omega:
                d_lambda = d_lambda / 2
            d_omega = d_omega_test
        if d_lambda < d_lambda_min:
            if log:
                log("Stop: minimal increment of lambda achieved.")
            break
        if step >= step_max:
            if log:
                log("Stop: maximal number of steps achieved.")
            break
    return U, np.array(lambda_hist), np.array(omega_hist), np.array(Umax_hist), np.array(work_hist)


# ---------------------------------------------------------------------------
# Limit-load (LL) indirect continuation
# ---------------------------------------------------------------------------
def LL_indirect_continuation(d_omega_ini, d_t_min, step_max, omega_max, it_newt_max, it_damp_max, tol, r_min,
                             K_elast, U_elast, Q, f, cmb, linear_system_solver, log=None, driver="python"):
    """LL_indirect_continuation.m:1-160: omega-stepping of the load factor t
    from the scaled elastic field, secant predictor after the first step,
    halved d_omega on failure (at most five times in a row), doubled when the
    t increments shrink and Newton took fewer than 20 iterations; stops on
    d_t < d_t_min, omega_max or step_max.  Returns (U, t_hist, omega_hist,
    Umax_hist, newton_iterations).  ``driver``: "python" or "native" Newton
    loop (newton.newton_ind_LL)."""
    from .newton import newton_ind_LL
    if not cmb.pattern_initialized:
        cmb.init_K_r_pattern(Q)
    ws = Workspace(cmb.n_n, cmb.dim)
    omega, omega_old, d_omega, t = 0.0, 0.0, float(d_omega_ini), 0.0
    U = U_elast
    U_old = U
    U_ini = DeviceArray(U.shape)
    lincomb(1.0, U, 0.0, U, U_ini)
    n_omega, n_omega_max = 0, 5
    step = 1
    t_hist, omega_hist, Umax_hist = [0.0], [0.0], [0.0]
    newton_total = 0
    while True:
        omega_it = omega + d_omega
        if omega_it > omega_max:
            omega_it = omega_max
            d_omega = omega_it - omega
        if log:
            log(f"Step {step + 1}: omega_target={omega_it:.6g}, d_omega={d_omega:.6g}")
        if step > 1:
            a = d_omega / (omega - omega_old)
            lincomb(1.0 + a, U, -a, U_old, U_ini)
        U_it, t_it, flag, its = newton_ind_LL(U_ini, t, omega_it, it_newt_max, it_damp_max, tol, r_min, K_elast, Q,
                                              f, cmb, linear_system_solver.copy(), ws, log, driver=driver)
        newton_total += its
        if flag == 1:
            if log:
                log(f"  retry: reducing d_omega from {d_omega:.6g} to {d_omega / 2:.6g}")
            _rebuild_after_failure(linear_system_solver)
            d_omega = d_omega / 2
            n_omega += 1
        else:
            step += 1
            U_old, U = U, U_it
            linear_system_solver.expand_deflation_basis(U)
            omega_old, omega = omega, omega_it
            d_t = t_it - t
            t = t_it
            n_omega = 0
            omega_hist.append(omega)
            t_hist.append(t)
            Umax_hist.append(_umax(U))
            if log:
                log(f"  accepted: t={t:.6g}, d_t={d_t:.6g}, omega={omega:.6g}, d_omega={d_omega:.6g}, "
                    f"U_max={Umax_hist[-1]:.6g}")
            if d_t < d_t_min:
                if log:
                    log("Stop: minimal increment of t achieved.")
                break
            if its < 20 and step > 2 and (t_hist[-1] - t_hist[-2]) < 0.9 * (t_hist[-2] - t_hist[-3]):
                d_omega = 2 * d_omega
        if n_omega >= n_omega_max:
            if log:
                log("Stop: impossible to increase omega further (max reductions reached).")
            break
        if omega >= omega_max:
            if log:
                log("Stop: maximal omega achieved.")
            break
        if step >= step_max:
            if log:
                log("Stop: maximal number of steps reached.")
            break
    return U, np.array(t_hist), np.array(omega_hist), np.array(Umax_hist), newton_total
