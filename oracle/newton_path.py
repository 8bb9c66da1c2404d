"""CPU restatement of the SSR Newton/continuation path (test infrastructure only).

FP64, numpy/scipy, over the free DOFs Q exactly as the reference's Octave code:

* ``ConstitutiveCPU``  -- CONSTITUTIVE.m (reduction, stress, stress+tangent via
  the C oracle, build_F via the global B, K_r(Q,Q) via the element-level
  oracle assembly, V_elast and transpose_perm from ``RefPattern``);
* ``newton_ind_SSR`` / ``damping_ALG5`` -- newton_ind_SSR.m:68-175, damping_ALG5.m;
* ``newton`` / ``damping``               -- newton.m:74-165, damping.m;
* ``init_phase`` / ``SSR_indirect_continuation`` -- the continuation files.

The linear systems are solved directly (sparse LU of K_r(Q,Q)) instead of by
DFGMRES+BoomerAMG at tolerance 1e-1: the converged Newton states satisfy the
same nonlinear tolerance, so the factor of safety is comparable.
"""
from __future__ import annotations

import math

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import fe_oracle as FO
from . import oracle as O


class ConstitutiveCPU:
    def __init__(self, coord, elem, Q, dim, DPhi, WEIGHT, mats, Davis_type="B"):
        self.dim = dim
        self.ns = 6 if dim == 3 else 3
        self.elem = elem
        self.n_n = coord.shape[1]
        self.n_q = WEIGHT.size // elem.shape[1]
        self.DPhi = DPhi
        self.WEIGHT = WEIGHT
        self.m = mats
        self.Davis_type = Davis_type
        self.Q = Q
        self.B = FO.global_B(elem, DPhi, dim, self.n_n)
        self.R = FO.RefPattern(elem, Q, dim, DPhi, WEIGHT, mats["shear"], mats["bulk"])
        self.smap = self.R.scatter_map(elem, self.n_n)
        self.qf = self.R.Q_flat
        self.DS = None

    def reduction(self, lam):
        self.c_bar, self.sin_phi = O.reduction(self.m["c0"], self.m["phi"], self.m["psi"], lam, self.Davis_type)

    def _E(self, U):
        return (self.B @ U.ravel(order="F")).reshape(self.ns, -1, order="F")

    def _mc(self, E, want_ds):
        f = O.mc3d if self.dim == 3 else O.mc2d
        S, DS, _ = f(E, self.c_bar, self.sin_phi, self.m["shear"], self.m["bulk"], self.m["lame"], want_ds=want_ds)
        return S, DS

    def build_F(self, S):
        s = (self.WEIGHT[None, :] * S[: self.ns]).ravel(order="F")
        return (self.B.T @ s).reshape(self.dim, -1, order="F")

    def build_F_all(self, lam, U):
        self.reduction(lam)
        S, _ = self._mc(self._E(U), False)
        return self.build_F(S)

    def build_F_reduced(self, U):
        S, _ = self._mc(self._E(U), False)
        return self.build_F(S)

    def build_F_and_DS_all(self, lam, U):
        self.reduction(lam)
        return self.build_F_and_DS_reduced(U)

    def build_F_and_DS_reduced(self, U):
        S, DS = self._mc(self._E(U), True)
        self.DS = np.asfortranarray(DS)
        return self.build_F(S)

    def build_K_r(self, r):
        Vt = O.assemble(self.dim, self.DPhi, self.DS, self.WEIGHT, self.smap, self.n_q, self.R.nnz)
        V = self.R.build_K_r(r, Vt)
        return sp.csc_matrix((V, (self.R.ref_I, self.R.ref_J)), shape=(self.R.n_Q, self.R.n_Q))


def _nQ(x, qf):
    return x.ravel(order="F")[qf]


def damping_ALG5(it_damp_max, U_it, lam, d_U, d_l, f, criterion, cmb):
    if math.isnan(d_l):
        return 0.0
    alpha, it = 1.0, 0
    fq = _nQ(f, cmb.qf)
    while True:
        it += 1
        Fa = cmb.build_F_all(lam + alpha * d_l, U_it + alpha * d_U)
        crit = np.linalg.norm(_nQ(Fa, cmb.qf) - fq)
        if crit >= criterion or math.isnan(crit):
            alpha /= 2
        else:
            break
        if it >= it_damp_max:
            return 0.0
    return alpha


def _scatter(vq, cmb):
    out = np.zeros(cmb.dim * cmb.n_n)
    out[cmb.qf] = vq
    return out.reshape(cmb.dim, -1, order="F")


def newton_ind_SSR(U_ini, omega, lambda_ini, it_newt_max, it_damp_max, tol, r_min, f, cmb):
    U_it = U_ini.copy()
    lam = float(lambda_ini)
    eps = tol / 1000
    fq = _nQ(f, cmb.qf)
    norm_f = np.linalg.norm(fq)
    flag, compute, r, it = 0, True, r_min, 0
    rel = float("nan")
    while True:
        it += 1
        if compute:
            F = cmb.build_F_and_DS_all(lam, U_it)
            criterion = np.linalg.norm(_nQ(F, cmb.qf) - fq)
            rel = criterion / norm_f
            if rel < tol and it > 1:
                break
        K = cmb.build_K_r(r)
        G = (cmb.build_F_all(lam + eps, U_it) - F) / eps
        lu = spla.splu(K)
        W = _scatter(lu.solve(-_nQ(G, cmb.qf)), cmb)
        V = _scatter(lu.solve(fq - _nQ(F, cmb.qf)), cmb)
        d_l = -(fq @ _nQ(V, cmb.qf)) / (fq @ _nQ(W, cmb.qf))
        d_U = V + d_l * W
        alpha = damping_ALG5(it_damp_max, U_it, lam, d_U, d_l, f, criterion, cmb)
        compute = True
        if alpha < 1e-1:
            if alpha == 0:
                compute = False
                r *= 2
            else:
                r *= 2 ** 0.25
        elif alpha > 0.5:
            r = max(r / math.sqrt(2), r_min)
        if alpha == 0 and r > 1:
            if rel > 10 * tol:
                flag = 1
            break
        U_it = U_it + alpha * d_U
        U_it = omega * U_it / (fq @ _nQ(U_it, cmb.qf))
        lam = lam + alpha * d_l
        if math.isnan(rel) or it == it_newt_max:
            if rel > 10 * tol or math.isnan(rel):
                flag = 1
            break
    return U_it, lam, flag, it


def damping(it_damp_max, U_it, dU, F, f, cmb):
    dec = (F - f).ravel() @ dU.ravel()
    if math.isnan(dec) or np.isinf(np.linalg.norm(dU)) or dec >= 0:
        return 0.0
    alpha, a0, a1, it = 1.0, 0.0, 1.0, 0
    while it < it_damp_max:
        it += 1
        Fa = cmb.build_F_reduced(U_it + alpha * dU)
        dec = (Fa - f).ravel() @ dU.ravel()
        if dec < 0:
            if alpha == 1:
                break
            a0 = alpha
        else:
            a1 = alpha
        alpha = (a0 + a1) / 2
    return alpha


def newton(U_ini, tol, it_newt_max, it_damp_max, r_min, f, cmb):
    U_it = U_ini.copy()
    fq = _nQ(f, cmb.qf)
    norm_f = np.linalg.norm(fq)
    flag, compute, r, it = 0, True, r_min, 0
    rel = float("nan")
    while True:
        it += 1
        if compute:
            F = cmb.build_F_and_DS_reduced(U_it)
            rel = np.linalg.norm(_nQ(F, cmb.qf) - fq) / norm_f
            if rel < tol:
                break
        K = cmb.build_K_r(r)
        dU = _scatter(spla.splu(K).solve(fq - _nQ(F, cmb.qf)), cmb)
        alpha = damping(it_damp_max, U_it, dU, F, f, cmb)
        compute = True
        if alpha < 1e-1:
            if alpha == 0:
                compute = False
                r *= 2
            else:
                r *= 2 ** 0.25
        elif alpha > 0.5:
            r = max(r / math.sqrt(2), r_min)
        if alpha == 0 and r > 1:
            if rel > 10 * tol:
                flag = 1
            break
        U_it = U_it + alpha * dU
        if math.isnan(rel) or it == it_newt_max:
            if rel > 10 * tol or math.isnan(rel):
                flag = 1
            break
    return U_it, flag, it


def SSR_indirect_continuation(lambda_init, d_lambda_init, d_lambda_min, d_lambda_diff_scaled_min, step_max,
                              omega_max_stop, it_newt_max, it_damp_max, tol, r_min, f, cmb, log=None):
    fq = _nQ(f, cmb.qf)
    # init phase
    cmb.reduction(lambda_init)
    U, flag, _ = newton(np.zeros_like(f), tol, it_newt_max, it_damp_max, r_min, f, cmb)
    if flag:
        raise RuntimeError("Initial choice of lambda seems to be too large.")
    U1, omega1, lambda1, d_lambda = U, fq @ _nQ(U, cmb.qf), lambda_init, d_lambda_init
    while True:
        lam_it = lambda1 + d_lambda
        cmb.reduction(lam_it)
        U_it, flag, _ = newton(U1, tol, it_newt_max, it_damp_max, r_min, f, cmb)
        if flag:
            d_lambda /= 2
        else:
            U2, omega2, lambda2 = U_it, fq @ _nQ(U_it, cmb.qf), lam_it
            if (omega2 - omega1) / max(1, omega1) < 1e-5:
                U1, lambda1, omega1 = U2, lambda2, omega2
            else:
                break
        if d_lambda < d_lambda_min:
            raise RuntimeError("It seems that the FoS is equal to lambda_init.")
        if lambda1 > 10:
            raise RuntimeError("It seems that the FoS is greater than 10.")
    U_old, U, omega_old, omega, lam = U1, U2, omega1, omega2, lambda2
    lam_h, om_h = [lambda1, lambda2], [omega1, omega2]
    d_omega = omega - omega_old
    step, n_omega = 2, 0
    while True:
        omega_it = min(omega + d_omega, omega_max_stop)
        d_omega = omega_it - omega
        U_ini = d_omega * (U - U_old) / (omega - omega_old) + U
        U_it, lam_it, flag, it = newton_ind_SSR(U_ini, omega_it, lam, it_newt_max, it_damp_max, tol, r_min, f, cmb)
        if flag:
            d_omega /= 2
            n_omega += 1
        else:
            step += 1
            U_old, U = U, U_it
            omega_old, omega = omega, omega_it
            d_lam = lam_it - lam
            lam = lam_it
            n_omega = 0
            om_h.append(omega)
            lam_h.append(lam)
            if log:
                log(f"  cpu step {step}: lambda={lam:.8f} omega={omega:.6g} newton_it={it}")
            if d_lam / d_omega * (om_h[-1] - om_h[0]) < d_lambda_diff_scaled_min:
                break
            if omega >= omega_max_stop:
                break
            if lam_h[-1] - lam_h[-2] < 0.9 * (lam_h[-2] - lam_h[-3]):
                d_omega *= 2
        if n_omega >= 5 or step >= step_max:
            break
    return U, np.array(lam_h), np.array(om_h)
