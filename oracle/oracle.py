"""CPU oracle for the SSR Newton-step hot path (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module, and only as the checker or as the timed
CPU baseline.  The product path (``slope_stability_octave_amd``) never does.

Two native libraries back it:

* ``liboracle.so`` -- this build's C restatement (``mc_oracle.c``) of the
  reference kernels; every routine cites the reference lines it follows.
* ``_ref/libref.so`` -- the reference's own per-point kernels
  (``constitutive_{2D,3D}_kernel.h``) compiled from their sources by
  ``oracle/Makefile``; present only where it was built.

The Newton-level restatement (``newton_path.py``) is numpy/scipy.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_ORC_PATH = os.path.join(_HERE, "liboracle.so")
_REF_PATH = os.path.join(_HERE, "_ref", "libref.so")

_dp = ctypes.POINTER(ctypes.c_double)
_i64 = ctypes.c_int64


def build(force: bool = False) -> None:
    """Compile the oracle (and the reference harness where the sources exist)."""
    if force or not os.path.exists(_ORC_PATH):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _ptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(_dp)


def _load(path):
    lib = ctypes.CDLL(path)
    return lib


_orc = None
_ref = None


def lib():
    global _orc
    if _orc is None:
        if not os.path.exists(_ORC_PATH):
            build()
        _orc = _load(_ORC_PATH)
        for name in ("orc_mc3d", "orc_mc2d"):
            f = getattr(_orc, name)
            f.restype = None
            f.argtypes = [_i64, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_void_p]
        _orc.orc_reduction.restype = ctypes.c_int
        _orc.orc_reduction.argtypes = [ctypes.c_int, _i64, _dp, _dp, _dp, _dp, ctypes.c_int, _dp, _dp]
        _orc.orc_assemble3d.restype = None
        _orc.orc_assemble3d.argtypes = [_i64, _i64, ctypes.c_int, _dp, _dp, _dp, _dp, _dp,
                                        ctypes.c_void_p, _i64, _dp]
        _orc.orc_assemble2d.restype = None
        _orc.orc_assemble2d.argtypes = [_i64, _i64, ctypes.c_int, _dp, _dp, _dp, _dp,
                                        ctypes.c_void_p, _i64, _dp]
        _orc.orc_strain.restype = None
        _orc.orc_strain.argtypes = [ctypes.c_int, _i64, _i64, ctypes.c_int, ctypes.c_void_p,
                                    _dp, _dp, _dp, _dp, _dp]
        _orc.orc_force.restype = None
        _orc.orc_force.argtypes = [ctypes.c_int, _i64, _i64, ctypes.c_int, _i64, ctypes.c_void_p,
                                   _dp, _dp, _dp, _dp, ctypes.c_int, _dp, _dp]
        _orc.orc_num_threads.restype = ctypes.c_int
    return _orc


def ref_available() -> bool:
    return os.path.exists(_REF_PATH)


def ref_lib():
    global _ref
    if _ref is None:
        if not ref_available():
            raise FileNotFoundError("oracle/_ref/libref.so not built (needs the reference checkout)")
        _ref = _load(_REF_PATH)
        for name in ("ref_mc3d", "ref_mc2d"):
            f = getattr(_ref, name)
            f.restype = None
            f.argtypes = [_i64, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
    return _ref


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def num_threads() -> int:
    return int(lib().orc_num_threads())


# --------------------------------------------------------------------------
# constitutive
# --------------------------------------------------------------------------
def mc3d(E, c_bar, sin_phi, shear, bulk, lame, want_ds=True, use_ref=False):
    """Return (S 6xn, DS 36xn or None, branch n) for engineering strains E (6xn,
    stored point-major as in the MEX interface)."""
    E = _c(E)
    n = E.shape[1] if E.ndim == 2 else E.size // 6
    args = [_c(x) for x in (c_bar, sin_phi, shear, bulk, lame)]
    S = np.empty((n, 6))
    DS = np.empty((n, 36)) if want_ds else None
    Et = np.ascontiguousarray(E.T) if E.ndim == 2 else E
    if use_ref:
        ref_lib().ref_mc3d(n, _ptr(Et), *[_ptr(a) for a in args], _ptr(S), _ptr(DS))
        br = None
    else:
        br = np.empty(n, dtype=np.int8)
        lib().orc_mc3d(n, _ptr(Et), *[_ptr(a) for a in args], _ptr(S), _ptr(DS),
                       br.ctypes.data_as(ctypes.c_void_p))
    return S.T, (DS.T if DS is not None else None), br


def mc2d(E, c_bar, sin_phi, shear, bulk, lame, want_ds=True, use_ref=False):
    E = _c(E)
    n = E.shape[1]
    args = [_c(x) for x in (c_bar, sin_phi, shear, bulk, lame)]
    S = np.empty((n, 4))
    DS = np.empty((n, 9)) if want_ds else None
    Et = np.ascontiguousarray(E.T)
    if use_ref:
        ref_lib().ref_mc2d(n, _ptr(Et), *[_ptr(a) for a in args], _ptr(S), _ptr(DS))
        br = None
    else:
        br = np.empty(n, dtype=np.int8)
        lib().orc_mc2d(n, _ptr(Et), *[_ptr(a) for a in args], _ptr(S), _ptr(DS),
                       br.ctypes.data_as(ctypes.c_void_p))
    return S.T, (DS.T if DS is not None else None), br


def reduction(c0, phi, psi, lam, davis="B"):
    c0, phi, psi = _c(c0), _c(phi), _c(psi)
    lam = np.atleast_1d(_c(lam))
    n = c0.size
    cb = np.empty(n)
    sp = np.empty(n)
    rc = lib().orc_reduction(ord(davis), n, _ptr(c0), _ptr(phi), _ptr(psi), _ptr(lam),
                             1 if lam.size == n and n > 1 else 0, _ptr(cb), _ptr(sp))
    if rc != 0:
        raise ValueError("Incorrect choice of the Davis approach.")
    return cb, sp


# --------------------------------------------------------------------------
# element-level operators
# --------------------------------------------------------------------------
def assemble(dim, DPhi, DS, WEIGHT, smap, n_q, nnz):
    """V_tang (nnz) from DPhi (tuple of n_p x n_int arrays), DS (ns^2 x n_int),
    WEIGHT (n_int), int64 scatter map (n_ld^2 x n_e, 1-based, 0 = skip)."""
    D = [np.asfortranarray(_c(x)) for x in DPhi]
    n_p = D[0].shape[0]
    smap = np.asfortranarray(smap, dtype=np.int64)
    n_e = smap.shape[1]
    DSf = np.asfortranarray(_c(DS))
    W = _c(WEIGHT).ravel()
    V = np.zeros(int(nnz))
    sm = smap.ctypes.data_as(ctypes.c_void_p)
    if dim == 3:
        lib().orc_assemble3d(n_p, n_e, int(n_q), _ptr(D[0]), _ptr(D[1]), _ptr(D[2]), _ptr(DSf), _ptr(W),
                             sm, int(nnz), _ptr(V))
    else:
        lib().orc_assemble2d(n_p, n_e, int(n_q), _ptr(D[0]), _ptr(D[1]), _ptr(DSf), _ptr(W),
                             sm, int(nnz), _ptr(V))
    return V


def strain(dim, elem, DPhi, U, n_q):
    """E = B*U, elem n_p x n_e (0-based), U dim x n_n -> E ns x n_int."""
    elem = np.asfortranarray(elem, dtype=np.int64)
    n_p, n_e = elem.shape
    D = [np.asfortranarray(_c(x)) for x in DPhi]
    if dim == 2:
        D = D + [D[0]]
    Uf = np.asfortranarray(_c(U))
    ns = 6 if dim == 3 else 3
    E = np.empty((ns, n_e * n_q), order="F")
    lib().orc_strain(dim, n_p, n_e, int(n_q), elem.ctypes.data_as(ctypes.c_void_p),
                     _ptr(D[0]), _ptr(D[1]), _ptr(D[2]), _ptr(Uf), E.ctypes.data_as(_dp))
    return E


def force(dim, elem, DPhi, S, WEIGHT, n_q, n_n):
    """F = B'*(WEIGHT.*S(1:ns,:)), returned dim x n_n."""
    elem = np.asfortranarray(elem, dtype=np.int64)
    n_p, n_e = elem.shape
    D = [np.asfortranarray(_c(x)) for x in DPhi]
    if dim == 2:
        D = D + [D[0]]
    Sf = np.asfortranarray(_c(S))
    F = np.empty((dim, n_n), order="F")
    lib().orc_force(dim, n_p, n_e, int(n_q), int(n_n), elem.ctypes.data_as(ctypes.c_void_p),
                    _ptr(D[0]), _ptr(D[1]), _ptr(D[2]), Sf.ctypes.data_as(_dp), Sf.shape[0],
                    _ptr(_c(WEIGHT).ravel()), F.ctypes.data_as(_dp))
    return F
