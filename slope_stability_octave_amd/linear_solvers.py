"""Deflated Krylov solvers on the GPU, mirroring +LINEAR_SOLVERS.

``DFGMRES`` keeps the method names and meaning of DFGMRES.m:77-221
(setup_preconditioner, update_preconditioner_values, A_orthogonalize, solve,
expand_deflation_basis, copy); ``DCG`` swaps the Krylov core for the deflated
CG of dcg_solver.m, as DCG.m does.  The operator is the node-block matrix of a
``Pattern`` (device values); vectors are device arrays over all dim*n_n DOFs
with Dirichlet entries zero.  The solver loops run in libssr_hip.so.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np

from . import _lib
from ._lib import DeviceArray, call, ptr

_vp = ctypes.c_void_p
_int = ctypes.c_int
_dbl = ctypes.c_double
_i64 = ctypes.c_int64

_lib.extend_signatures({
    "ssr_solver_create": [_vp, _int, _dbl, _dbl, _int, _vp],
    "ssr_solver_destroy": [_vp],
    "ssr_solver_set_matrix": [_vp, _vp, _int],
    "ssr_solver_deflation_size": [_vp, _vp],
    "ssr_solver_clear_deflation": [_vp],
    "ssr_solver_expand_deflation": [_vp, _vp],
    "ssr_solver_copy_deflation": [_vp, _vp],
    "ssr_solver_share_preconditioner": [_vp, _vp],
    "ssr_solver_invalidate_preconditioner": [_vp],
    "ssr_solver_get_deflation": [_vp, _int, _vp],
    "ssr_solver_A_orthogonalize": [_vp],
    "ssr_solver_dfgmres": [_vp, _vp, _vp, _vp, _vp],
    "ssr_solver_dcg": [_vp, _vp, _vp, _vp, _vp],
    "ssr_solver_precond_apply": [_vp, _vp, _vp],
    "ssr_solver_stats": [_vp, _vp],
    "ssr_solver_set_amg": [_vp, _vp, _int, _vp],
    "ssr_solver_set_precision": [_vp, _int],
    "ssr_solver_set_symmetric": [_vp, _int],
    "ssr_solver_get_precision": [_vp, _vp],
    "ssr_dev_spmv_bsr": [_int, _i64, _vp, _vp, _vp, _vp, _vp, _dbl, _dbl],
    "ssr_dev_dot": [_i64, _vp, _vp, _vp, _vp],
    "ssr_dev_spmv_bsr_f32": [_int, _i64, _vp, _vp, _vp, _vp, _vp, _dbl, _dbl],
    "ssr_dev_spmv_bsr_sym": [_int, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _dbl, _dbl],
    "ssr_dev_spmv_bsr_sym_f32": [_int, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _dbl, _dbl],
    "ssr_dev_convert_f64_f32": [_i64, _vp, _vp],
    "ssr_dot": [_i64, _vp, _vp, _vp, _vp],
    "ssr_node_norm_max": [_int, _i64, _vp, _vp],
    "ssr_dev_multidot": [_i64, _int, _vp, _i64, _vp, _vp],
    "ssr_dev_multidot_noalloc": [_i64, _int, _vp, _i64, _vp, _vp],
    "ssr_dev_gemv_update": [_i64, _int, _vp, _i64, _vp, _dbl, _dbl, _vp],
    "ssr_dev_multidot_f32": [_i64, _int, _vp, _i64, _vp, _vp],
    "ssr_dev_gemv_update_f32": [_i64, _int, _vp, _i64, _vp, _dbl, _dbl, _vp],
    "ssr_dev_axpby": [_i64, _dbl, _vp, _dbl, _vp],
    "ssr_dev_lincomb": [_i64, _dbl, _vp, _dbl, _vp, _vp],
    "ssr_dev_mul": [_i64, _vp, _vp],
    "ssr_dev_fill": [_i64, _dbl, _vp],
    "ssr_dev_block_jacobi_setup": [_int, _i64, _vp, _vp, _vp],
    "ssr_dev_block_jacobi_apply": [_int, _i64, _vp, _vp, _vp, _vp],
    "ssr_dev_csr_spmv": [_i64, _vp, _vp, _vp, _vp, _vp, _dbl, _dbl],
    "ssr_dev_ewmul": [_i64, _vp, _vp, _vp],
    "ssr_dev_cheb_step": [_i64, _dbl, _dbl, _vp, _vp, _vp, _vp],
    "ssr_dev_dense_gemv": [_int, _vp, _vp, _vp],
})

PRECONDITIONERS = {"block_jacobi": 0}


def _amg_option(options, key):
    from . import amg
    return (options or {}).get(key, amg.DEFAULTS[key])


def dot(a, b, mask=None) -> float:
    out = ctypes.c_double(0.0)
    call("ssr_dot", int(np.prod(a.shape)), ptr(a), ptr(b), ptr(mask), ctypes.byref(out))
    return out.value


def lincomb(a, x, b, y, z):
    """z = a*x + b*y on device."""
    call("ssr_dev_lincomb", int(np.prod(z.shape)), float(a), ptr(x), float(b), ptr(y), ptr(z))


def fill(x, v=0.0):
    """x[:] = v on the device (write-only kernel)."""
    call("ssr_dev_fill", int(np.prod(x.shape)), float(v), ptr(x))


def axpby(a, x, b, y):
    """y = a*x + b*y on device."""
    call("ssr_dev_axpby", int(np.prod(y.shape)), float(a), ptr(x), float(b), ptr(y))


def spmv(pattern, K, x, y, alpha=1.0, beta=0.0):
    call("ssr_dev_spmv_bsr", pattern.dim, pattern.n_n, ptr(pattern.d_row_ptr), ptr(pattern.d_col), ptr(K),
         ptr(x), ptr(y), float(alpha), float(beta))


class DFGMRES:
    """Deflated flexible GMRES with a GPU preconditioner (DFGMRES.m).

    ``symmetric`` (default True: every operator of the SSR path, K_r, is
    symmetrised by construction) lets the deflation projector use the A W kept
    by A_orthogonalize; pass False for a non-symmetric operator."""

    method = "dfgmres"

    def __init__(self, pattern, tolerance=1e-1, max_iterations=100, tolerance_deflation_basis=1e-3,
                 preconditioner="block_jacobi", verbose=0, amg_options=None, precision=64, symmetric=True, *,
                 share_preconditioner_of=None):
        self.pattern = pattern
        self.precision = int(precision)
        self.tolerance = float(tolerance)
        self.max_iterations = int(max_iterations)
        self.tolerance_deflation_basis = float(tolerance_deflation_basis)
        self.preconditioner_name = preconditioner
        self.amg_options = amg_options
        self.verbose = verbose
        self.profiler = None
        self.timings = {}
        self.counters = {}
        h = _vp(0)
        call("ssr_solver_create", pattern.handle, self.max_iterations, self.tolerance,
             self.tolerance_deflation_basis, PRECONDITIONERS.get(preconditioner, 0), ctypes.byref(h))
        self.handle = h
        if self.precision != 64:
            call("ssr_solver_set_precision", self.handle, self.precision)
        # the operators of the Newton solves (K_r) are symmetric
        self.symmetric = bool(symmetric)
        call("ssr_solver_set_symmetric", self.handle, 1 if self.symmetric else 0)
        if share_preconditioner_of is not None:
            # the other solver's preconditioner object (no AMG of its own is
            # built: that would be a host near-null-space and an upload per copy)
            call("ssr_solver_share_preconditioner", self.handle, share_preconditioner_of.handle)
        elif preconditioner == "amg":
            from . import amg
            amg.attach(self, amg_options)
        self._K = None
        self.last_iterations = 0
        self.last_relres = 0.0

    def __del__(self):
        try:
            if self.handle:
                _lib.load().ssr_solver_destroy(self.handle)
                self.handle = None
        except Exception:
            pass

    def _t(self, name, t0):
        if self.timings is not None:
            self.timings[name] = self.timings.get(name, 0.0) + time.perf_counter() - t0

    # -- preconditioner -------------------------------------------------------
    def setup_preconditioner(self, K):
        t0 = time.perf_counter()
        self._K = K
        self._fresh = True  # the operator was just handed over (A_orthogonalize need not again)
        call("ssr_solver_set_matrix", self.handle, ptr(K), 1)
        self._t("DFGMRES.update_preconditioner", t0)

    def update_preconditioner_values(self, K):
        self.setup_preconditioner(K)

    def invalidate_preconditioner(self):
        """The next setup rebuilds the preconditioner in full (also for every
        copy sharing it)."""
        call("ssr_solver_invalidate_preconditioner", self.handle)
        if self.counters is not None:
            self.counters["AMG.rebuilds_requested"] = self.counters.get("AMG.rebuilds_requested", 0) + 1

    def set_operator(self, K):
        self._K = K
        self._fresh = False
        call("ssr_solver_set_matrix", self.handle, ptr(K), 0)

    # -- deflation ------------------------------------------------------------
    @property
    def deflation_size(self) -> int:
        k = ctypes.c_int(0)
        call("ssr_solver_deflation_size", self.handle, ctypes.byref(k))
        return k.value

    def expand_deflation_basis(self, v):
        call("ssr_solver_expand_deflation", self.handle, ptr(v))

    def A_orthogonalize(self, K):
        t0 = time.perf_counter()
        if not (K is self._K and getattr(self, "_fresh", False)):
            self.set_operator(K)
        self._fresh = False
        call("ssr_solver_A_orthogonalize", self.handle)
        self._t("DFGMRES.A_orthogonalize", t0)
        if self.counters is not None:
            k = self.deflation_size
            c = self.counters
            c["DFGMRES.deflation_size.max"] = max(c.get("DFGMRES.deflation_size.max", 0), k)
            c["DFGMRES.deflation_size.sum"] = c.get("DFGMRES.deflation_size.sum", 0) + k

    def deflation_basis(self):
        """Host copy of the basis, columns over all DOFs."""
        k = self.deflation_size
        out = np.empty((k, self.pattern.n_dof))
        tmp = DeviceArray(self.pattern.n_dof)
        for c in range(k):
            call("ssr_solver_get_deflation", self.handle, c, ptr(tmp))
            out[c] = tmp.get()
        return out.T

    def copy(self):
        """New solver with the same settings and a copy of the basis; the
        preconditioner is rebuilt on the next setup (DFGMRES.m:204-221), or,
        with amg_options['reuse_hierarchy'] (the default), the copy shares
        this solver's AMG object so one hierarchy serves every copy."""
        share = self.preconditioner_name == "amg" and _amg_option(self.amg_options, "reuse_hierarchy")
        new = type(self)(self.pattern, self.tolerance, self.max_iterations, self.tolerance_deflation_basis,
                         self.preconditioner_name, self.verbose, self.amg_options, self.precision, self.symmetric,
                         share_preconditioner_of=self if share else None)
        call("ssr_solver_copy_deflation", new.handle, self.handle)
        new.profiler = self.profiler
        new.timings = self.timings  # shared, like the reference's profiler handle
        new.counters = self.counters
        return new

    # -- solve ----------------------------------------------------------------
    def solve(self, K, b, x=None):
        """Return (x, iterations) for K x = b (device arrays)."""
        t0 = time.perf_counter()
        if K is not self._K:
            self.set_operator(K)
        self._fresh = False
        if x is None:
            x = DeviceArray(b.shape)
        it = ctypes.c_int(0)
        rr = ctypes.c_double(0.0)
        name = "ssr_solver_dfgmres" if self.method == "dfgmres" else "ssr_solver_dcg"
        call(name, self.handle, ptr(b), ptr(x), ctypes.byref(it), ctypes.byref(rr))
        self.last_iterations = it.value
        self.last_relres = rr.value
        self._t("DFGMRES.solve", t0)
        if self.counters is not None:
            self.counters["DFGMRES.solve.n_iters"] = self.counters.get("DFGMRES.solve.n_iters", 0) + it.value
            self.counters["DFGMRES.solve.calls"] = self.counters.get("DFGMRES.solve.calls", 0) + 1
            if it.value >= self.max_iterations:
                self.counters["DFGMRES.solve.at_maxit"] = self.counters.get("DFGMRES.solve.at_maxit", 0) + 1
        return x, it.value

    def stats(self):
        out = np.zeros(4, dtype=np.int64)
        call("ssr_solver_stats", self.handle, ptr(out))
        return dict(iterations=int(out[0]), solves=int(out[1]), precond_applies=int(out[2]), spmv=int(out[3]))


class DCG(DFGMRES):
    """Deflated preconditioned CG (DCG.m / dcg_solver.m)."""

    method = "dcg"
