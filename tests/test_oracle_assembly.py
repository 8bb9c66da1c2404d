"""Pins the assembly oracle (element-level, scatter map) against the global
B'*D*B formulation of the reference, and the build's pattern tables against
the reference's pattern order.  CPU only."""
import numpy as np
import pytest

from oracle import fe_oracle as FO
from oracle import oracle as O
from slope_stability_octave_amd import fem, mesh
from slope_stability_octave_amd import workloads as W
from slope_stability_octave_amd.assembly import Pattern


def small_problem(dim, seed=0):
    if dim == 2:
        coord, elem, Q = mesh.mesh_P2_2D(2.5, 5, 5, 5, 5, 5)
    else:
        coord, elem, Q = mesh.mesh_P2_3D(1, 1, 1, 1, 1, 1, 2.0, 2.0, 2.0, 2.0, 2.0, 2.0)
    DPhi, WEIGHT, HatP, WF = fem.gradients(coord, elem, dim)
    n_q = WF.size
    n_int = WEIGHT.size
    G, K, L = W.moduli()
    rng = np.random.default_rng(seed)
    ns = 6 if dim == 3 else 3
    E = rng.normal(0, 2e-3, size=(ns, n_int))
    cb, sp = O.reduction(np.full(n_int, 15.0), np.full(n_int, np.pi / 6), np.zeros(n_int), 1.3, "B")
    f = O.mc3d if dim == 3 else O.mc2d
    S, DS, br = f(E, cb, sp, np.full(n_int, G), np.full(n_int, K), np.full(n_int, L))
    return dict(coord=coord, elem=elem, Q=Q, DPhi=DPhi, WEIGHT=WEIGHT, n_q=n_q, DS=DS, S=S,
                shear=np.full(n_int, G), bulk=np.full(n_int, K), branch=br)


@pytest.mark.parametrize("dim", [2, 3])
def test_element_assembly_equals_global_BtDB(dim):
    p = small_problem(dim)
    n_n = p["coord"].shape[1]
    R = FO.RefPattern(p["elem"], p["Q"], dim, p["DPhi"], p["WEIGHT"], p["shear"], p["bulk"])
    smap = R.scatter_map(p["elem"], n_n)
    V = O.assemble(dim, p["DPhi"], p["DS"], p["WEIGHT"], smap, p["n_q"], R.nnz)
    Vg = R.global_tangent(p["DS"], p["WEIGHT"])
    assert np.linalg.norm(V - Vg) <= 1e-13 * np.linalg.norm(Vg)
    assert len(set(p["branch"].tolist())) >= 3  # a plastic mix, not only elastic


@pytest.mark.parametrize("dim", [2, 3])
def test_elastic_tangent_equals_V_elast(dim):
    p = small_problem(dim)
    n_n = p["coord"].shape[1]
    R = FO.RefPattern(p["elem"], p["Q"], dim, p["DPhi"], p["WEIGHT"], p["shear"], p["bulk"])
    smap = R.scatter_map(p["elem"], n_n)
    DSe = FO.ds_elastic(dim, p["shear"], p["bulk"])
    V = O.assemble(dim, p["DPhi"], DSe, p["WEIGHT"], smap, p["n_q"], R.nnz)
    Vs = (V + V[R.transpose_perm]) / 2
    assert np.allclose(Vs, R.V_elast, rtol=1e-12, atol=1e-12 * np.abs(R.V_elast).max())


def test_rigid_body_modes_in_kernel_3d():
    """With no Dirichlet DOFs, K annihilates the six rigid-body modes."""
    p = small_problem(3)
    n_n = p["coord"].shape[1]
    Qall = np.ones_like(p["Q"])
    R = FO.RefPattern(p["elem"], Qall, 3, p["DPhi"], p["WEIGHT"], p["shear"], p["bulk"])
    Vg = R.global_tangent(p["DS"], p["WEIGHT"])
    import scipy.sparse as sp
    K = sp.csr_matrix((Vg, (R.ref_I, R.ref_J)), shape=(R.n_Q, R.n_Q))
    x, y, z = p["coord"]
    modes = []
    for t in range(3):
        m = np.zeros((3, n_n)); m[t] = 1; modes.append(m)
    for (a, b) in ((0, 1), (1, 2), (0, 2)):
        m = np.zeros((3, n_n)); m[a] = -p["coord"][b]; m[b] = p["coord"][a]; modes.append(m)
    for m in modes:
        v = m.ravel(order="F")
        assert np.linalg.norm(K @ v) <= 1e-9 * abs(Vg).max() * np.linalg.norm(v)


@pytest.mark.parametrize("dim", [2, 3])
def test_pattern_order_matches_reference(dim):
    p = small_problem(dim)
    R = FO.RefPattern(p["elem"], p["Q"], dim, p["DPhi"], p["WEIGHT"], p["shear"], p["bulk"])
    P = Pattern(p["elem"], p["Q"], dim, device=False)
    assert P.n_free == R.n_Q and P.nnz_free == R.nnz
    rp = P.ref_pos
    row_ptr, col = P.row_ptr, P.col
    brow = np.repeat(np.arange(P.n_n), np.diff(row_ptr))
    blk = rp // P.bs
    ri, cj = np.divmod(rp % P.bs, dim)
    I = dim * brow[blk] + ri
    J = dim * col[blk] + cj
    free = np.full(P.n_dof, -1)
    free[R.Q_flat] = np.arange(R.n_Q)
    assert np.array_equal(free[I], R.ref_I) and np.array_equal(free[J], R.ref_J)


@pytest.mark.parametrize("dim", [2, 3])
def test_strain_and_force_oracle_vs_global_B(dim):
    p = small_problem(dim)
    n_n = p["coord"].shape[1]
    B = FO.global_B(p["elem"], p["DPhi"], dim, n_n)
    rng = np.random.default_rng(4)
    U = rng.normal(size=(dim, n_n))
    E = O.strain(dim, p["elem"], p["DPhi"], U, p["n_q"])
    Eg = (B @ U.ravel(order="F")).reshape(E.shape, order="F")
    assert np.allclose(E, Eg, rtol=1e-13, atol=1e-13 * np.abs(Eg).max())
    ns = 6 if dim == 3 else 3
    F = O.force(dim, p["elem"], p["DPhi"], p["S"], p["WEIGHT"], p["n_q"], n_n)
    s = (p["WEIGHT"][None, :] * p["S"][:ns]).ravel(order="F")
    Fg = (B.T @ s).reshape((dim, n_n), order="F")
    assert np.allclose(F, Fg, rtol=1e-12, atol=1e-12 * np.abs(Fg).max())
