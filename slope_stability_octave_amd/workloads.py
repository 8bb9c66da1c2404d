"""Synthetic inputs for the constitutive + assembly microbenchmark
(3D box 40x20x20 m, 100x50x40 hexes x 6 P2 tetrahedra, 11 points each).

Strains: N(0, 1e-3^2) i.i.d. per Voigt component; a seeded subset of exactly
floor(0.3 n) points (seeded permutation) is scaled by 5 to reach the edge and
apex branches.  Material: c0 = 15 kPa, phi = 30 deg, psi = 0, E = 20 MPa
(20000 kPa, the reference's units), nu = 0.3; strength reduction lambda ~ U[1,2]
per point, Davis B.  Moduli follow heterogenous_materials.m:64-66.
"""
from __future__ import annotations

import math

import numpy as np

SEED_E = 20261203
SEED_MASK = 20261204

C0 = 15.0
PHI_DEG = 30.0
PSI_DEG = 0.0
YOUNG = 20000.0
POISSON = 0.3


def moduli(young: float = YOUNG, poisson: float = POISSON):
    """(shear, bulk, lame) exactly as heterogenous_materials.m:64-66."""
    shear = young / (2 * (1 + poisson))
    bulk = young / (3 * (1 - 2 * poisson))
    lame = bulk - 2 * shear / 3
    return shear, bulk, lame


def config2_points(n: int, seed_e: int = SEED_E, seed_mask: int = SEED_MASK, dim: int = 3):
    """Return dict with E (n x ns, point-major), c0, phi, psi, lam (n each),
    shear, bulk, lame (n each)."""
    ns = 6 if dim == 3 else 3
    rng = np.random.default_rng(seed_e)
    E = rng.normal(0.0, 1e-3, size=(n, ns))
    rng2 = np.random.default_rng(seed_mask)
    perm = rng2.permutation(n)
    k = int(math.floor(0.3 * n))
    E[perm[:k]] *= 5.0
    lam = rng2.uniform(1.0, 2.0, size=n)
    G, K, L = moduli()
    out = {
        "E": E,
        "c0": np.full(n, C0),
        "phi": np.full(n, PHI_DEG * math.pi / 180),
        "psi": np.full(n, PSI_DEG * math.pi / 180),
        "lam": lam,
        "shear": np.full(n, G),
        "bulk": np.full(n, K),
        "lame": np.full(n, L),
    }
    return out
