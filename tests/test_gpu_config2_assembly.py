"""Tangent assembly on the constitutive + assembly microbenchmark geometry:
a slice of the 40x20x20 m box at its element size (0.4 x 0.4 x 0.5 m hexes,
6 P2 tetrahedra each, 11 points), with the benchmark's trial strains (N(0,
1e-3^2), a seeded 30% scaled by 5), c0 = 15 kPa, phi = 30 deg, E = 20 MPa,
nu = 0.3, lambda ~ U[1,2] per point.

* values in the reference's (ref_I, ref_J) order against the oracle's
  element assembly, 1e-11 relative (normwise);
* size-independent properties: the tangent of symmetric DS is symmetric
  (block (v,w) = block (w,v)^T), the elastic tangent equals the elastic
  operator V_elast, and with no Dirichlet DOFs K annihilates the six
  rigid-body modes.
"""
import numpy as np
import pytest

from oracle import fe_oracle as FO
from oracle import oracle as O
from slope_stability_octave_amd import assembly as A
from slope_stability_octave_amd import fem, mesh
from slope_stability_octave_amd import workloads as W
from slope_stability_octave_amd._lib import DeviceArray

pytestmark = pytest.mark.gpu


def box_slice(nx=10, ny=5, nz=4, all_free=False):
    coord, elem, Q = mesh.mesh_P2_box(nx, ny, nz, 0.4 * nx, 0.4 * ny, 0.5 * nz)
    if all_free:
        Q = np.ones_like(Q)
    DPhi, WEIGHT, HatP, WF = fem.gradients(coord, elem, 3)
    n_int = WEIGHT.size
    d = W.config2_points(n_int)
    cb, sp = O.reduction(d["c0"], d["phi"], d["psi"], d["lam"], "B")
    S, DS, br = O.mc3d(np.ascontiguousarray(d["E"].T), cb, sp, d["shear"], d["bulk"], d["lame"])
    return dict(coord=coord, elem=elem, Q=Q, DPhi=DPhi, WEIGHT=WEIGHT, n_q=WF.size, DS=DS, branch=br,
                shear=d["shear"], bulk=d["bulk"])


def assemble_gpu(p, DS):
    P = A.Pattern(p["elem"], p["Q"], 3)
    dD = [DeviceArray.from_host(x.ravel(order="F")) for x in p["DPhi"]]
    dW = DeviceArray.from_host(p["WEIGHT"])
    dDS = DeviceArray.from_host(np.asfortranarray(DS).T.copy())
    Kt = DeviceArray(P.nnzb * P.bs)
    P.assemble(p["n_q"], dD, dDS, dW, Kt)
    Vt = DeviceArray(P.nnz_free)
    P.extract_ref(Kt, Vt)
    return P, Kt.get().reshape(-1, 3, 3), Vt.get()


def test_config2_slice_tangent_matches_oracle():
    p = box_slice()
    n_e = p["elem"].shape[1]
    print(f"config-2 slice: {n_e} elements, branches {np.bincount(p['branch'], minlength=5).tolist()}", flush=True)
    assert len(set(p["branch"].tolist())) == 5  # every return branch present
    R = FO.RefPattern(p["elem"], p["Q"], 3, p["DPhi"], p["WEIGHT"], p["shear"], p["bulk"])
    smap = R.scatter_map(p["elem"], p["coord"].shape[1])
    V0 = O.assemble(3, p["DPhi"], p["DS"], p["WEIGHT"], smap, p["n_q"], R.nnz)
    P, Kb, V = assemble_gpu(p, p["DS"])
    rel = np.abs(V - V0).max() / np.abs(V0).max()
    print(f"max rel {rel:.2e}", flush=True)
    assert rel <= 1e-11
    # symmetric DS (associated flow) gives a symmetric tangent: block (v,w) = block (w,v)^T, bit for
    # bit off the diagonal (each element pair block serves both), to rounding on diagonal blocks
    tblk = P.table(3)
    KbT = np.transpose(Kb[tblk], (0, 2, 1))
    off = tblk != np.arange(tblk.size)
    assert np.array_equal(Kb[off], KbT[off])
    assert np.abs(Kb - KbT).max() <= 1e-13 * np.abs(Kb).max()
    # elastic DS gives the elastic operator V_elast of the reference's pattern
    DSe = FO.ds_elastic(3, p["shear"], p["bulk"])
    _, _, Ve = assemble_gpu(p, DSe)
    Vs = (Ve + Ve[R.transpose_perm]) / 2
    assert np.abs(Vs - R.V_elast).max() <= 1e-11 * np.abs(R.V_elast).max()


def test_config2_slice_rigid_body_modes():
    p = box_slice(6, 3, 3, all_free=True)
    P, Kb, _ = assemble_gpu(p, p["DS"])
    x, y, z = p["coord"]
    n_n = x.size
    rows = np.repeat(np.arange(P.n_n), np.diff(P.row_ptr))
    modes = [np.stack([np.ones(n_n), 0 * x, 0 * x]), np.stack([0 * x, np.ones(n_n), 0 * x]),
             np.stack([0 * x, 0 * x, np.ones(n_n)]), np.stack([-y, x, 0 * x]), np.stack([0 * x, -z, y]),
             np.stack([-z, 0 * x, x])]
    scale = np.abs(Kb).max()
    for m in modes:
        u = m.T  # node-major
        Ku = np.zeros_like(u)
        np.add.at(Ku, rows, np.einsum("bij,bj->bi", Kb, u[P.col]))
        assert np.abs(Ku).max() <= 1e-9 * scale * np.abs(u).max()
