"""Learned continuation predictor: a small MLP (HIP kernels in csrc/mlp.hip)
that proposes the next displacement as coefficients over the last four
converged fields and the next lambda, trained on trajectories the plain
solver generates.

Baseline it corrects: the reference's secant step
(SSR_indirect_continuation.m:119), U_sec = (1+a) U_k - a U_{k-1} with
a = d_omega / (omega_k - omega_{k-1}), and lambda_sec = lambda_k + a
(lambda_k - lambda_{k-1}).  The network outputs delta (4) and t (1):
U_pred = sum_i (c_sec,i + delta_i) U_{k-i}, lambda_pred = lambda_sec +
t (lambda_k - lambda_{k-1}).  At run time Newton's residual is the judge:
the prediction is used only when ||F_lambda(U) - f||_Q at (U_pred,
lambda_pred) is below the residual of the secant start (U_sec, lambda_k);
otherwise the secant step runs unchanged.

Output coordinates.  The four fields F = [U_k .. U_{k-3}] are nearly
collinear (the Gram matrix M = F'F has condition numbers around 1e8), so raw
coefficients are a hopeless regression target.  The network instead outputs
w in the orthonormal (Gram-Schmidt) basis of the span, scaled by the last
step length: with M = R'R (Cholesky) and |D1| = |U_k - U_{k-1}|,
    delta = R^-1 w |D1|,   F delta = Q w |D1|  (Q orthonormal),
so an error in w maps isometrically to the displacement error in units of
the last step.  M is known at run time (ten inner products of the fields).

Training sample of step k: features of the (lambda, omega) history, of the
field geometry (step-length ratios and angles between successive steps, from
M) and of the material; target w* = R delta* / |D1| with delta* the least
squares correction of the secant over the span; and the weight
rho = |D1|^2 / s, s = |U_sec - U_{k+1}|^2.  The normalised error of a
prediction is then
    L(w) = rho |w - w*|^2 + L_opt,   L(0) = 1 (the secant),
L_opt the best the span allows; the kernels minimise the mean of
rho |w - w*|^2 + w_lam (t - t*)^2.
"""
from __future__ import annotations

import ctypes
import math
import time

import numpy as np

from ._lib import DeviceArray, call, extend_signatures, ptr
from .linear_solvers import dot, lincomb

_vp, _int, _i64, _dbl, _flt, _u64 = (ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double,
                                      ctypes.c_float, ctypes.c_uint64)
extend_signatures({
    "ssr_mlp_create": [_int, _int, _int, _int, _u64, _dbl, _vp],
    "ssr_mlp_destroy": [_vp],
    "ssr_mlp_info": [_vp, _vp, _int],
    "ssr_mlp_get_params": [_vp, _vp],
    "ssr_mlp_set_params": [_vp, _vp],
    "ssr_mlp_forward": [_vp, _int, _vp, _vp],
    "ssr_mlp_backward": [_vp, _int, _vp, _vp, _vp],
    "ssr_mlp_predictor_loss": [_int, _int, _vp, _vp, _vp, _vp, _vp, _flt, _vp, _vp],
    "ssr_mlp_train_step": [_vp, _int, _vp, _vp, _vp, _vp, _vp, _flt, _flt, _flt, _vp],
})

N_FIELDS = 4
N_FEATURES = 17
N_OUT = N_FIELDS + 1
_D1 = np.array([1.0, -1.0, 0.0, 0.0])
_DIFF = np.array([[1.0, -1.0, 0.0, 0.0], [0.0, 1.0, -1.0, 0.0], [0.0, 0.0, 1.0, -1.0]])


class MLP:
    """Device network; widths {n_in, hidden x (layers-1), n_out}."""

    def __init__(self, n_in=N_FEATURES, n_hidden=128, n_out=N_OUT, n_layers=3, seed=0, out_scale=1e-2):
        h = _vp(0)
        call("ssr_mlp_create", n_in, n_hidden, n_out, n_layers, seed, float(out_scale), ctypes.byref(h))
        self.handle = h
        info = np.zeros(2 + n_layers + 1, dtype=np.int64)
        call("ssr_mlp_info", self.handle, ptr(info), info.size)
        self.n_params = int(info[0])
        self.widths = [int(w) for w in info[2:]]

    def __del__(self):
        try:
            if self.handle:
                from . import _lib
                _lib.load().ssr_mlp_destroy(self.handle)
                self.handle = None
        except Exception:
            pass

    def get_params(self):
        out = np.empty(self.n_params, dtype=np.float32)
        call("ssr_mlp_get_params", self.handle, ptr(out))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.n_params
        call("ssr_mlp_set_params", self.handle, ptr(p))

    def forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        dX = DeviceArray.from_host(X)
        dY = DeviceArray((X.shape[0], self.widths[-1]), dtype=np.float32)
        call("ssr_mlp_forward", self.handle, X.shape[0], ptr(dX), ptr(dY))
        return dY.get()

    def backward(self, X, dY):
        X = np.ascontiguousarray(X, dtype=np.float32)
        dX = DeviceArray.from_host(X)
        ddY = DeviceArray.from_host(np.ascontiguousarray(dY, dtype=np.float32))
        g = DeviceArray(self.n_params, dtype=np.float32)
        call("ssr_mlp_backward", self.handle, X.shape[0], ptr(dX), ptr(ddY), ptr(g))
        return g.get()


# ---------------------------------------------------------------------------
# features and training data
# ---------------------------------------------------------------------------
def _slog(x, lo=-7.0, hi=7.0):
    return float(np.clip(math.log(max(abs(x), 1e-12)), lo, hi))


def secant_coefficients(omega_hist, omega_target):
    a = (omega_target - omega_hist[-1]) / (omega_hist[-1] - omega_hist[-2])
    return np.array([1.0 + a, -a, 0.0, 0.0]), a


def gram_features(M):
    """Step-length ratios and cosines between the last three steps of the
    fields (from their Gram matrix)."""
    MD = _DIFF @ M @ _DIFF.T
    ln = np.sqrt(np.maximum(np.diag(MD), 1e-300))
    return [_slog(ln[1] / ln[0]), _slog(ln[2] / ln[0]), MD[0, 1] / (ln[0] * ln[1]), MD[1, 2] / (ln[1] * ln[2]),
            MD[0, 2] / (ln[0] * ln[2])]


def features(lambda_hist, omega_hist, omega_target, material, M):
    """17 scale-free features: the last four (lambda, omega) pairs and the
    target omega (10), the geometry of the last three field steps (5), the
    material (slope angle deg, c0 kPa, phi deg)."""
    lam = np.asarray(lambda_hist[-4:], dtype=np.float64)
    om = np.asarray(omega_hist[-4:], dtype=np.float64)
    d_om = omega_target - om[-1]
    dw = np.diff(om)[::-1]   # newest first: om_k - om_{k-1}, ...
    dl = np.diff(lam)[::-1]
    f = [_slog(dw[i] / d_om) for i in range(3)]
    f += [_slog(dl[i] / dl[0]) if dl[0] != 0 else 0.0 for i in (1, 2)]
    s = dl / dw
    f += [_slog(s[i] / s[0]) if s[0] != 0 else 0.0 for i in (1, 2)]
    f += [lam[-1], math.log(om[-1] / om[0])]
    f += gram_features(M)
    beta, c0, phi = material
    f += [beta / 90.0, c0 / 30.0, phi / 45.0]
    return np.array(f, dtype=np.float64)


def field_gram(F, mask):
    M = np.zeros((N_FIELDS, N_FIELDS))
    for i in range(N_FIELDS):
        for j in range(i, N_FIELDS):
            M[i, j] = M[j, i] = dot(F[i], F[j], mask)
    return M


def whiten(M):
    """(R, |D1|) with M = R'R; None if the fields are rank deficient."""
    try:
        R = np.linalg.cholesky(M).T
    except np.linalg.LinAlgError:
        return None
    d1 = float(_D1 @ M @ _D1)
    if not (d1 > 0) or not np.all(np.isfinite(R)):
        return None
    return R, math.sqrt(d1)


class TrajectoryRecorder:
    """Continuation hook: after each accepted step with four earlier fields,
    stores (features, G, g, t) of the prediction problem it just solved."""

    def __init__(self, mask, material):
        self.mask = mask
        self.material = material
        self.X, self.G, self.g, self.h, self.t = [], [], [], [], []
        self.L_opt = []  # best normalised error within the span (secant = 1)
        self.raw = []  # last four (lambda, omega), target omega, new lambda

    def __call__(self, state, U_new, lam_new):
        Uh = state["U_hist"]
        lh, oh = state["lambda_hist"], state["omega_hist"]
        if len(Uh) < N_FIELDS or len(lh) < N_FIELDS:
            return
        om_t = state["omega_it"]
        c, a = secant_coefficients(oh, om_t)
        F = Uh[:N_FIELDS]
        m = self.mask
        M = field_gram(F, m)
        Fu = np.array([dot(F[i], U_new, m) for i in range(N_FIELDS)])
        uu = dot(U_new, U_new, m)
        s = float(c @ M @ c - 2 * c @ Fu + uu)  # |U_sec - U*|^2
        wh = whiten(M)
        if wh is None or not (s > 0) or not math.isfinite(s):
            return
        R, d1 = wh
        delta = np.linalg.solve(M, Fu - M @ c)  # least-squares correction of the secant
        w = R @ delta / d1
        rho = d1 * d1 / s
        self.X.append(features(lh, oh, om_t, self.material, M))
        self.G.append(rho * np.eye(N_FIELDS))
        self.g.append(rho * w)
        self.h.append(rho * float(w @ w))
        self.L_opt.append(1.0 - rho * float(w @ w))
        dl1 = lh[-1] - lh[-2]
        lam_sec = lh[-1] + a * dl1
        self.t.append((lam_new - lam_sec) / dl1 if dl1 != 0 else 0.0)
        self.raw.append(np.r_[np.asarray(lh[-4:], float), np.asarray(oh[-4:], float), om_t, lam_new])

    def arrays(self):
        return (np.array(self.X), np.array(self.G), np.array(self.g), np.array(self.t), np.array(self.h))


class Normaliser:
    def __init__(self, mean, std):
        self.mean = np.asarray(mean, dtype=np.float64)
        self.std = np.asarray(std, dtype=np.float64)

    @classmethod
    def fit(cls, X):
        return cls(X.mean(axis=0), np.maximum(X.std(axis=0), 1e-3))

    def __call__(self, X):
        return ((np.atleast_2d(X) - self.mean) / self.std).astype(np.float32)


def train(net, data, epochs=2000, lr=2e-3, wd=1e-4, w_lam=0.5, log=None, val=None):
    """Full-batch AdamW on the GPU.  data = (X, G, g, t[, h]) with X already
    normalised (h defaults to 1, the normalised secant error).  Returns the
    loss history (and validation history)."""
    X, G, g, t = data[:4]
    h = data[4] if len(data) > 4 else np.ones(X.shape[0])
    B = X.shape[0]
    dX = DeviceArray.from_host(np.ascontiguousarray(X, dtype=np.float32))
    dG = DeviceArray.from_host(np.ascontiguousarray(G.reshape(B, -1), dtype=np.float32))
    dg = DeviceArray.from_host(np.ascontiguousarray(g, dtype=np.float32))
    dh = DeviceArray.from_host(np.ascontiguousarray(h, dtype=np.float32))
    dt = DeviceArray.from_host(np.ascontiguousarray(t, dtype=np.float32))
    hist, vhist = [], []
    loss = ctypes.c_float(0.0)
    best = (float("inf"), None)
    for ep in range(epochs):
        call("ssr_mlp_train_step", net.handle, B, ptr(dX), ptr(dG), ptr(dg), ptr(dh), ptr(dt), float(w_lam),
             float(lr), float(wd), ctypes.byref(loss))
        hist.append(loss.value)
        if val is not None and (ep % 25 == 0 or ep == epochs - 1):
            ev = evaluate(net, val, w_lam)
            vhist.append((ep, ev))
            score = ev["L_U"] + w_lam * ev["L_lambda"]
            if score < best[0]:
                best = (score, net.get_params())
        if log and (ep % 200 == 0 or ep == epochs - 1):
            log(f"[mlp] epoch {ep} loss {loss.value:.4f}" + (f" val {vhist[-1][1]}" if vhist else ""))
    if best[1] is not None:  # early stopping on the validation set
        net.set_params(best[1])
    return hist, vhist


def evaluate(net, data, w_lam=0.5):
    """Mean normalised displacement error L_U (secant = 1) and lambda error
    of the network on data, with the secant's lambda error for reference."""
    X, G, g, t = data[:4]
    Y = net.forward(X).astype(np.float64)
    B = X.shape[0]
    d = Y[:, :N_FIELDS]
    LU = np.einsum("si,sij,sj->s", d, G, d) - 2 * (d * g).sum(1) + 1.0
    Ll = (Y[:, N_FIELDS] - t) ** 2
    return dict(L_U=float(LU.mean()), L_lambda=float(Ll.mean()), secant_L_lambda=float((t ** 2).mean()), n=B)


# ---------------------------------------------------------------------------
# run-time predictor
# ---------------------------------------------------------------------------
class LearnedPredictor:
    """Callable for SSR_indirect_continuation(predictor=...)."""

    def __init__(self, net, normaliser, material, cmb, f):
        self.net = net
        self.norm = normaliser
        self.material = material
        self.cmb = cmb
        self.f = f
        self.used = 0
        self.rejected = 0
        self.F = None
        self.start_residuals = []  # (network start, secant start) per judged step

    def _residual(self, lam, U):
        from .newton import norm_Q
        if self.F is None:
            self.F = DeviceArray(U.shape)
            self.tmp = DeviceArray(U.shape)
        self.cmb.build_F_all(lam, U, self.F)
        return norm_Q(self.F, self.f, self.cmb.mask, self.tmp)

    def __call__(self, state):
        Uh = state["U_hist"]
        lh, oh = state["history"]["lambda_hist"], state["history"]["omega_hist"]
        if len(Uh) < N_FIELDS or len(lh) < N_FIELDS:
            return None
        om_t = state["omega_it"]
        c, a = secant_coefficients(oh, om_t)
        M = field_gram(Uh[:N_FIELDS], self.cmb.mask)
        wh = whiten(M)
        if wh is None:
            self.rejected += 1
            return None
        R, d1 = wh
        x = self.norm(features(lh, oh, om_t, self.material, M))
        y = self.net.forward(x)[0].astype(np.float64)
        coef = c + np.linalg.solve(R, y[:N_FIELDS]) * d1
        dl1 = lh[-1] - lh[-2]
        lam_pred = lh[-1] + a * dl1 + y[N_FIELDS] * dl1
        U = DeviceArray(Uh[0].shape)
        lincomb(coef[0], Uh[0], coef[1], Uh[1], U)
        lincomb(1.0, U, coef[2], Uh[2], U)
        lincomb(1.0, U, coef[3], Uh[3], U)
        fU = dot(self.f, U, self.cmb.mask)
        if not (abs(fU) > 0) or not math.isfinite(fU):
            self.rejected += 1
            return None
        lincomb(om_t / fU, U, 0.0, U, U)
        # Newton's residual judges: the learned start must beat the secant start
        U_sec = DeviceArray(Uh[0].shape)
        lincomb(c[0], Uh[0], c[1], Uh[1], U_sec)
        r_pred = self._residual(lam_pred, U)
        r_sec = self._residual(lh[-1], U_sec)
        self.start_residuals.append((float(r_pred), float(r_sec)))
        if not (r_pred < r_sec):
            self.rejected += 1
            return None
        self.used += 1
        return U, lam_pred


def save(path, net, normaliser, meta=None):
    np.savez(path, params=net.get_params(), widths=np.array(net.widths), mean=normaliser.mean, std=normaliser.std,
             **(meta or {}))


def load(path):
    with np.load(path) as z:
        widths = [int(w) for w in z["widths"]]
        net = MLP(widths[0], widths[1], widths[-1], len(widths) - 1)
        net.set_params(z["params"])
        return net, Normaliser(z["mean"], z["std"])


def predictor_factory(net, normaliser, material):
    """For problems.run_ssr(predictor=...): builds the LearnedPredictor once
    the device problem exists."""
    def make(cmb, f):
        return LearnedPredictor(net, normaliser, material, cmb, f)
    make.is_factory = True
    return make


def recorder_factory(material):
    def make(cmb):
        return TrajectoryRecorder(cmb.mask, material)
    make.is_factory = True
    return make


# ---------------------------------------------------------------------------
# end-to-end: trajectories -> training -> use
# ---------------------------------------------------------------------------
def collect(cases, preconditioner="amg", log=None):
    """Run the plain solver on each (name, problem, material, params) case and
    collect the training samples of every trajectory.  Returns (X, G, g, t, h)
    raw; cases whose continuation fails are skipped."""
    from . import problems
    from .continuation import ContinuationError
    Xs, Gs, gs, ts, hs = [], [], [], [], []
    for name, prob, mat, p in cases:
        try:
            res = problems.run_ssr(prob, p, preconditioner=preconditioner, recorder=recorder_factory(mat))
        except ContinuationError as exc:
            if log:
                log(f"[mlp] {name} skipped: {exc}")
            continue
        X, G, g, t, hh = res["recorder"].arrays()
        if log:
            lo = np.mean(res["recorder"].L_opt) if len(X) else float("nan")
            log(f"[mlp] {name} material {tuple(round(x, 2) for x in mat)} FoS {res['FoS']:.5f} "
                f"samples {len(X)} span optimum L {lo:.3f} time {res['time_to_fos_s']:.2f}s")
        if len(X):
            Xs.append(X), Gs.append(G), gs.append(g), ts.append(t), hs.append(hh)
    return (np.concatenate(Xs), np.concatenate(Gs), np.concatenate(gs), np.concatenate(ts), np.concatenate(hs))


def generate(seeds, h=1.0, params=None, preconditioner="amg", log=None, linear_tol=1e-2):
    """Training samples from the plain solver on the seeded 2D slope family."""
    from . import problems

    def cases():
        for seed in seeds:
            prob, mat = problems.slope_2d_family(seed, h=h)
            p = dict(problems.PARAMS_FAMILY)
            p.update(params or {})
            p.setdefault("linear_tol", linear_tol)
            yield f"seed {seed}", prob, mat, p
    return collect(cases(), preconditioner=preconditioner, log=log)


def fit(train_raw, val_raw=None, epochs=3000, lr=1e-3, wd=1e-3, w_lam=0.5, seed=0, log=None):
    """Normalise the features, train a 3-layer, 128-unit network on the GPU;
    returns (net, normaliser, report)."""
    norm = Normaliser.fit(train_raw[0])
    tr = (norm(train_raw[0]),) + tuple(train_raw[1:])
    va = None if val_raw is None else (norm(val_raw[0]),) + tuple(val_raw[1:])
    net = MLP(N_FEATURES, 128, N_OUT, 3, seed=seed)
    t0 = time.perf_counter()
    hist, vhist = train(net, tr, epochs=epochs, lr=lr, wd=wd, w_lam=w_lam, log=log, val=va)
    rep = dict(train_time_s=time.perf_counter() - t0, train_loss_first=hist[0], train_loss_last=hist[-1],
               train=evaluate(net, tr, w_lam), samples=int(tr[0].shape[0]))
    if va is not None:
        rep["val"] = evaluate(net, va, w_lam)
        rep["val_samples"] = int(va[0].shape[0])
    return net, norm, rep
