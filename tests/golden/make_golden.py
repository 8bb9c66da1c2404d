"""Write the constitutive golden vectors from the reference's own kernels.

Run in the development container, where the reference checkout exists:

    make -C oracle            # builds oracle/_ref/libref.so from
                              # constitutive_{2D,3D}_kernel.h (reference sources)
    python tests/golden/make_golden.py

Inputs are seeded, stratified by return branch (elastic, smooth, left edge,
right edge, apex) and include hand-picked special strains.  Two material sets:
the microbenchmark material (c0=15 kPa, phi=30 deg, E=20 MPa, nu=0.3) and the
2D homogeneous slope material (c0=6 kPa, phi=45 deg, E=40 MPa, nu=0.3), both
with Davis B reduction at per-point lambda.  The expected S/DS come from the
compiled reference kernels (oracle/_ref); the branch labels from the oracle.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
PER_BRANCH = 100


def moduli(young, nu):
    G = young / (2 * (1 + nu))
    K = young / (3 * (1 - 2 * nu))
    return G, K, K - 2 * G / 3


def special_strains(dim):
    ns = 6 if dim == 3 else 3
    rows = [np.zeros(ns)]
    a = 1e-3
    if dim == 3:
        rows += [
            [a, a, a, 0, 0, 0], [-a, -a, -a, 0, 0, 0], [20 * a, 20 * a, 20 * a, 0, 0, 0],
            [a, 0, 0, 0, 0, 0], [-5 * a, 0, 0, 0, 0, 0], [0, 0, 0, 2 * a, 0, 0],
            [0, 0, 0, 0, 8 * a, 0], [3 * a, -a, 0.5 * a, a, -2 * a, 0.7 * a],
            [5 * a, 5 * a, -3 * a, 0, 0, 0], [-3 * a, 5 * a, 5 * a, 0, 0, 0],
            [1e-9, -2e-9, 3e-9, 0, 0, 0],
        ]
    else:
        rows += [
            [a, a, 0], [-a, -a, 0], [20 * a, 20 * a, 0], [a, 0, 0], [-5 * a, 0, 0],
            [0, 0, 2 * a], [0, 0, 8 * a], [3 * a, -a, a], [5 * a, -3 * a, 0], [-3 * a, 5 * a, 4 * a],
            [1e-9, -2e-9, 3e-9],
        ]
    return np.array(rows, dtype=np.float64)


def make(dim, seed):
    ns = 6 if dim == 3 else 3
    rng = np.random.default_rng(seed)
    mats = [(15.0, 30.0, 20000.0, 0.3, (1.0, 2.0)), (6.0, 45.0, 40000.0, 0.3, (0.9, 1.6))]
    sel = {k: [] for k in ("E", "c_bar", "sin_phi", "shear", "bulk", "lame", "branch")}
    for (c0, phi, young, nu, lr) in mats:
        n = 40000
        E = rng.normal(0, 1e-3, size=(n, ns))
        E[rng.permutation(n)[: int(0.3 * n)]] *= 5.0
        E = np.vstack([E, special_strains(dim)])
        n = E.shape[0]
        lam = rng.uniform(*lr, size=n)
        cb, sp = O.reduction(np.full(n, c0), np.full(n, np.deg2rad(phi)), np.zeros(n), lam, "B")
        G, K, L = moduli(young, nu)
        f = O.mc3d if dim == 3 else O.mc2d
        _, _, br = f(E.T, cb, sp, np.full(n, G), np.full(n, K), np.full(n, L))
        idx = list(range(n - special_strains(dim).shape[0], n))
        for b in range(5):
            cand = np.flatnonzero(br[: n - len(idx)] == b)
            idx += list(rng.choice(cand, size=min(PER_BRANCH // 2, cand.size), replace=False))
        idx = np.array(idx)
        sel["E"].append(E[idx]); sel["c_bar"].append(cb[idx]); sel["sin_phi"].append(sp[idx])
        sel["shear"].append(np.full(idx.size, G)); sel["bulk"].append(np.full(idx.size, K))
        sel["lame"].append(np.full(idx.size, L)); sel["branch"].append(br[idx])
    d = {k: np.concatenate(v) for k, v in sel.items()}
    f = O.mc3d if dim == 3 else O.mc2d
    S, DS, _ = f(d["E"].T, d["c_bar"], d["sin_phi"], d["shear"], d["bulk"], d["lame"], use_ref=True)
    keep = np.all(np.isfinite(S), axis=0) & np.all(np.isfinite(DS), axis=0)
    out = {k: v[keep] for k, v in d.items()}
    out["S"] = np.ascontiguousarray(S.T[keep])
    out["DS"] = np.ascontiguousarray(DS.T[keep])
    path = os.path.join(OUT, f"golden_mc{dim}d.npz")
    np.savez_compressed(path, **out)
    print(path, out["E"].shape[0], "points, branches", np.bincount(out["branch"], minlength=5))


if __name__ == "__main__":
    if not O.ref_available():
        sys.exit("oracle/_ref/libref.so is not built: run `make -C oracle` where the reference exists")
    make(3, 7001)
    make(2, 7002)
