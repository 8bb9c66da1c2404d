"""The potential oracle (potential_{2D,3D}.m restated) against what fixes it:
the stress of the (reference-pinned) return-mapping oracle is its gradient,
S = dPsi/dE, checked by central differences on every branch; and the elastic
energy closed form."""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import potential_oracle as PO
from slope_stability_octave_amd import workloads as W


def _material(d):
    cb, sp = O.reduction(d["c0"], d["phi"], d["psi"], d["lam"], "B")
    return cb, sp, d["shear"], d["bulk"], d["lame"]


@pytest.mark.parametrize("dim", [3, 2])
def test_stress_is_gradient_of_potential(dim):
    n = 4000
    d = W.config2_points(n, dim=dim)
    cb, sp, G, K, L = _material(d)
    E = np.ascontiguousarray(d["E"].T)  # rows x n
    fn = PO.potential_3d if dim == 3 else PO.potential_2d
    mc = O.mc3d if dim == 3 else O.mc2d
    S, _, br = mc(E, cb, sp, G, K, L)
    psi, brp = fn(E, cb, sp, G, K, L)
    assert (brp == br).mean() > 0.999
    nr = E.shape[0]
    grad = np.zeros((nr, n))
    h = 1e-7 * np.maximum(np.abs(E).max(axis=0), 1e-6)
    for k in range(nr):
        Ep, Em = E.copy(), E.copy()
        Ep[k] += h
        Em[k] -= h
        grad[k] = (fn(Ep, cb, sp, G, K, L)[0] - fn(Em, cb, sp, G, K, L)[0]) / (2 * h)
    # points whose perturbation stays on one branch (the potential is C1, its
    # second derivative jumps across branch boundaries)
    same = np.ones(n, dtype=bool)
    for k in range(nr):
        for sgn in (1, -1):
            Ek = E.copy()
            Ek[k] += sgn * 10 * h
            same &= fn(Ek, cb, sp, G, K, L)[1] == brp
    scale = np.abs(S[:nr]).max(axis=0)
    err = np.abs(grad - S[:nr]).max(axis=0) / scale
    assert same.mean() > 0.95
    assert np.all(err[same] < 1e-5), err[same].max()
    # every branch is represented
    assert set(np.unique(brp[same])) == {0, 1, 2, 3, 4}


def test_elastic_energy_closed_form():
    rng = np.random.default_rng(0)
    n = 100
    E = rng.normal(0, 1e-6, size=(6, n))
    G = np.full(n, 7692.3077)
    L = np.full(n, 11538.462)
    K = L + 2 * G / 3
    psi, br = PO.potential_3d(E, np.full(n, 1e3), np.full(n, 0.5), G, K, L)
    assert np.all(br == 0)
    tr = E[:3].sum(0)
    ref = 0.5 * L * tr ** 2 + G * ((E[:3] ** 2).sum(0) + 0.5 * (E[3:] ** 2).sum(0))
    assert np.allclose(psi, ref, rtol=1e-12, atol=0)
    psi2, br2 = PO.potential_2d(E[[0, 1, 3]], np.full(n, 1e3), np.full(n, 0.5), G, K, L)
    tr2 = E[0] + E[1]
    ref2 = 0.5 * L * tr2 ** 2 + G * (E[0] ** 2 + E[1] ** 2 + 0.5 * E[3] ** 2)
    assert np.all(br2 == 0) and np.allclose(psi2, ref2, rtol=1e-12, atol=0)
