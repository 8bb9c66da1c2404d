"""Pins the CPU oracle of the return mapping and the Davis reduction.

Checks: golden vectors written by the reference's own compiled kernels,
closed forms (elastic, apex), the yield-surface invariant, symmetry of DS,
DS against a central finite difference of S, and Davis A/B/C against an
independent numpy evaluation.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from slope_stability_octave_amd import workloads as W
from tests.helpers import pointwise_rel, yield_value_2d, yield_value_3d


def _mats(g):
    return g["c_bar"], g["sin_phi"], g["shear"], g["bulk"], g["lame"]


def test_oracle_matches_reference_golden_3d(golden3):
    S, DS, br = O.mc3d(golden3["E"].T, *_mats(golden3))
    assert np.array_equal(br, golden3["branch"])
    assert pointwise_rel(S, golden3["S"].T).max() <= 1e-15
    assert pointwise_rel(DS, golden3["DS"].T).max() <= 1e-15
    assert set(np.unique(br)) == {0, 1, 2, 3, 4}


def test_oracle_matches_reference_golden_2d(golden2):
    S, DS, br = O.mc2d(golden2["E"].T, *_mats(golden2))
    assert np.array_equal(br, golden2["branch"])
    assert pointwise_rel(S, golden2["S"].T).max() <= 1e-15
    assert pointwise_rel(DS, golden2["DS"].T).max() <= 1e-15
    assert set(np.unique(br)) == {0, 1, 2, 3, 4}


@pytest.mark.skipif(not O.ref_available(), reason="reference harness not built here")
def test_oracle_bitwise_equals_reference_kernels():
    d = W.config2_points(50000)
    cb, sp = O.reduction(d["c0"], d["phi"], d["psi"], d["lam"], "B")
    for f in (O.mc3d,):
        S1, D1, _ = f(d["E"].T, cb, sp, d["shear"], d["bulk"], d["lame"])
        S2, D2, _ = f(d["E"].T, cb, sp, d["shear"], d["bulk"], d["lame"], use_ref=True)
        assert np.array_equal(S1, S2) and np.array_equal(D1, D2)
    d2 = W.config2_points(50000, dim=2)
    cb, sp = O.reduction(d2["c0"], d2["phi"], d2["psi"], d2["lam"], "B")
    S1, D1, _ = O.mc2d(d2["E"].T, cb, sp, d2["shear"], d2["bulk"], d2["lame"])
    S2, D2, _ = O.mc2d(d2["E"].T, cb, sp, d2["shear"], d2["bulk"], d2["lame"], use_ref=True)
    assert np.array_equal(S1, S2) and np.array_equal(D1, D2)


def _sample(n=20000, dim=3, seed=11):
    d = W.config2_points(n, seed_e=seed, seed_mask=seed + 1, dim=dim)
    cb, sp = O.reduction(d["c0"], d["phi"], d["psi"], d["lam"], "B")
    return d, cb, sp


def test_closed_forms_elastic_and_apex_3d():
    d, cb, sp = _sample()
    G, K, L = W.moduli()
    S, DS, br = O.mc3d(d["E"].T, cb, sp, d["shear"], d["bulk"], d["lame"])
    E = d["E"].T
    t = E.copy()
    t[3:] *= 0.5
    I1 = E[0] + E[1] + E[2]
    el = br == 0
    S_el = L * I1 * np.array([1, 1, 1, 0, 0, 0])[:, None] + 2 * G * t
    assert np.allclose(S[:, el], S_el[:, el], rtol=1e-14, atol=0)
    D_el = (L * np.outer([1, 1, 1, 0, 0, 0], [1, 1, 1, 0, 0, 0]) + 2 * G * np.diag([1, 1, 1, .5, .5, .5]))
    assert np.allclose(DS[:, el], D_el.reshape(-1, order="F")[:, None], rtol=1e-15, atol=0)
    ap = br == 4
    assert ap.any()
    sa = cb[ap] / (2 * sp[ap])
    assert np.array_equal(S[:3, ap], np.vstack([sa, sa, sa])) and not S[3:, ap].any()
    assert not DS[:, ap].any()


def test_yield_surface_and_symmetry_3d():
    d, cb, sp = _sample()
    S, DS, br = O.mc3d(d["E"].T, cb, sp, d["shear"], d["bulk"], d["lame"])
    f = yield_value_3d(S, sp, cb)
    pl = br > 0
    assert np.all(np.abs(f[pl]) <= 1e-9 * np.maximum(cb[pl], np.abs(S[:, pl]).max(axis=0)))
    assert np.all(f[~pl] <= 1e-9 * cb[~pl])
    D = DS.reshape(6, 6, -1, order="F")
    assert np.array_equal(D, D.transpose(1, 0, 2))


def test_yield_surface_2d():
    d, cb, sp = _sample(dim=2)
    S, DS, br = O.mc2d(d["E"].T, cb, sp, d["shear"], d["bulk"], d["lame"])
    f = yield_value_2d(S, sp, cb)
    pl = br > 0
    assert np.all(np.abs(f[pl]) <= 1e-9 * np.maximum(cb[pl], np.abs(S[:, pl]).max(axis=0)))
    D = DS.reshape(3, 3, -1, order="F")
    assert np.allclose(D, D.transpose(1, 0, 2), rtol=1e-13, atol=1e-9)


@pytest.mark.parametrize("dim", [3, 2])
def test_tangent_is_derivative_of_stress(dim):
    """DS . dE equals the central difference of S away from branch switches."""
    d, cb, sp = _sample(n=4000, dim=dim, seed=23)
    f = O.mc3d if dim == 3 else O.mc2d
    ns = 6 if dim == 3 else 3
    E = d["E"].T
    mats = (cb, sp, d["shear"], d["bulk"], d["lame"])
    S0, DS, br = f(E, *mats)
    rng = np.random.default_rng(5)
    dE = rng.normal(size=E.shape)
    h = 1e-7 * np.abs(E).max(axis=0)
    Sp, _, bp = f(E + h * dE, *mats)
    Sm, _, bm = f(E - h * dE, *mats)
    keep = (bp == br) & (bm == br)
    fd = (Sp[:ns] - Sm[:ns]) / (2 * h)
    D = DS.reshape(ns, ns, -1, order="F")
    lin = np.einsum("ijn,jn->in", D, dE)
    scale = np.abs(D).max(axis=(0, 1)) * np.abs(dE).max(axis=0) + 1e-30
    err = np.abs(fd - lin).max(axis=0) / scale
    assert keep.mean() > 0.95
    assert np.quantile(err[keep], 0.99) < 1e-5
    for b in range(1, 4):
        m = keep & (br == b)
        if m.any():
            assert np.median(err[m]) < 1e-6


def _numpy_davis(c0, phi, psi, lam, kind):
    if kind == "A":
        beta = np.cos(phi) * np.cos(psi) / (1 - np.sin(phi) * np.sin(psi))
        cl = beta * c0 / lam
        pl = np.arctan(beta * np.tan(phi) / lam)
    elif kind == "B":
        phi1 = np.arctan(np.tan(phi) / lam)
        psi1 = np.arctan(np.tan(psi) / lam)
        beta = np.cos(phi1) * np.cos(psi1) / (1 - np.sin(phi1) * np.sin(psi1))
        cl = beta * c0 / lam
        pl = np.arctan(beta * np.tan(phi1))
    else:
        phi1 = np.arctan(np.tan(phi) / lam)
        beta = np.cos(phi1) * np.cos(psi) / (1 - np.sin(phi1) * np.sin(psi)) if np.all(phi1 > psi) else 1.0
        cl = beta * c0 / lam
        pl = np.arctan(beta * np.tan(phi1))
    return 2 * cl * np.cos(pl), np.sin(pl)


@pytest.mark.parametrize("kind", ["A", "B", "C"])
def test_davis_reduction(kind):
    rng = np.random.default_rng(3)
    n = 1000
    c0 = rng.uniform(5, 30, n)
    phi = np.deg2rad(rng.uniform(20, 45, n))
    psi = np.deg2rad(rng.uniform(0, 10, n))
    lam = rng.uniform(0.8, 2.0, n)
    cb, sp = O.reduction(c0, phi, psi, lam, kind)
    cb2, sp2 = _numpy_davis(c0, phi, psi, lam, kind)
    assert np.allclose(cb, cb2, rtol=1e-14) and np.allclose(sp, sp2, rtol=1e-14)


def test_davis_b_closed_form_at_unit_lambda():
    phi = math.radians(30.0)
    cb, sp = O.reduction(np.array([15.0]), np.array([phi]), np.array([0.0]), 1.0, "B")
    # psi = 0, lambda = 1: beta = cos(phi), phi_l = atan(sin(phi))
    phl = math.atan(math.sin(phi))
    assert abs(cb[0] - 2 * 15.0 * math.cos(phi) * math.cos(phl)) < 1e-13
    assert abs(sp[0] - math.sin(phl)) < 1e-15


def test_davis_c_vector_if_semantics():
    # one point with phi1 <= psi turns beta into 1 for every point
    c0 = np.array([10.0, 10.0])
    phi = np.deg2rad([30.0, 5.0])
    psi = np.deg2rad([0.0, 10.0])
    cb, sp = O.reduction(c0, phi, psi, np.array([1.0, 1.0]), "C")
    assert np.allclose(sp, np.sin(phi)) and np.allclose(cb, 2 * c0 * np.cos(phi))
