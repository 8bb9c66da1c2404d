"""Setup path on the GPU against the host restatements.

* every pattern table built on the device equals the host builder's bit for
  bit (integer work), on 2D and 3D meshes, with Dirichlet masks, on a mesh
  whose node numbering is shuffled (element-local order unrelated to the
  global one) and on the configuration-sized 3D slope of the time-to-FoS
  workload;
* DPhi and WEIGHT within 1e-13 of fem.gradients, and pinned by invariants:
  the gradients of the P2 partition of unity sum to zero, linear fields are
  differentiated exactly, and the weights integrate the domain's volume
  (closed forms of the slope geometry);
* f_V within 1e-12 of fem.volume_force; its total equals -gamma * volume;
* material fields equal fem.materials_per_point bit for bit;
* a continuation set up on the device reaches the host-setup FoS.
"""
import ctypes

import numpy as np
import pytest

from slope_stability_octave_amd import assembly as A
from slope_stability_octave_amd import device_setup as DSU
from slope_stability_octave_amd import fem, mesh, problems
from slope_stability_octave_amd._lib import DeviceArray, call

pytestmark = pytest.mark.gpu


def _dev_table(P, which, n, dtype):
    out = np.empty(n, dtype=dtype)
    if n:
        src = P._dptr(which).ptr
        call("ssr_dtoh", ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(src), out.nbytes)
    return out


def _all_tables(P):
    rp = P.table(0)
    col = P.table(1)
    brow = np.repeat(np.arange(P.n_n, dtype=np.int32), np.diff(rp))
    n_ub = int((col >= brow).sum())
    npair = P.n_p * (P.n_p + 1) // 2
    t = dict(row_ptr=rp, col=col, emap=P.table(2), tblk=P.table(3), diag=P.table(4), ref_pos=P.table(5),
             mask=P.table(6))
    t["brow"] = _dev_table(P, 8, P.nnzb, np.int32)
    t["ublk"] = _dev_table(P, 9, n_ub, np.int32)
    t["uptr"] = _dev_table(P, 10, n_ub + 1, np.int32)
    t["pslot"] = _dev_table(P, 11, P.n_e * npair, np.int32)
    t["cptr"] = _dev_table(P, 12, P.nnzb + 1, np.int32)
    t["ccode"] = _dev_table(P, 13, P.n_e * P.n_p * P.n_p, np.int32)
    t["nptr"] = _dev_table(P, 14, P.n_n + 1, np.int32)
    t["ncode"] = _dev_table(P, 15, P.n_e * P.n_p, np.int32)
    t["elem"] = _dev_table(P, 5, P.n_e * P.n_p, np.int32)
    t["dmask"] = _dev_table(P, 7, P.n_dof, np.float64)
    return t


def _device_pattern(elem, Q, dim):
    n_p, n_e = elem.shape
    elem_d = DeviceArray.from_host(np.ascontiguousarray(elem.T.astype(np.int32)))
    Q_d = DeviceArray.from_host(np.ascontiguousarray(np.asarray(Q, dtype=np.uint8).T))
    return DSU.pattern_dev(elem_d, Q_d, dim, Q.shape[1], n_p, n_e)


def _shuffled(coord, elem, Q, seed):
    rng = np.random.default_rng(seed)
    perm = rng.permutation(coord.shape[1])  # new id of old node i = perm[i]
    inv = np.argsort(perm)
    return coord[:, inv], perm[elem], Q[:, inv]


MESHES = {
    "2d_h1": lambda: mesh.mesh_P2_2D(1.0, 15, 10, 15, 10, 10),
    "3d_small": lambda: mesh.mesh_P2_3D(2, 2, 2, 1, 2, 2, 20.0, 20.0, 20.0, 10.0, 20.0, 20.0),
    "3d_box": lambda: mesh.mesh_P2_3D(3, 2, 2, 2, 1, 2, 3.0, 2.0, 2.0, 2.0, 1.0, 2.0),
}


@pytest.mark.parametrize("name,shuffle", [("2d_h1", False), ("3d_small", False), ("3d_box", False),
                                          ("2d_h1", True), ("3d_small", True)])
def test_device_pattern_equals_host(name, shuffle):
    coord, elem, Q = MESHES[name]()
    if shuffle:
        coord, elem, Q = _shuffled(coord, elem, Q, 7)
    dim = coord.shape[0]
    print(f"pattern {name} shuffle={shuffle}: {elem.shape[1]} elements, {coord.shape[1]} nodes", flush=True)
    H = A.Pattern(elem, Q, dim)
    D = _device_pattern(elem, Q, dim)
    assert (D.nnzb, D.n_free, D.nnz_free) == (H.nnzb, H.n_free, H.nnz_free)
    th, td = _all_tables(H), _all_tables(D)
    for k in th:
        assert th[k].shape == td[k].shape, k
        assert np.array_equal(th[k], td[k]), k


def test_device_pattern_time_to_fos_mesh():
    """The ~1.5M-unknown slope (the time-to-FoS workload): the tables the
    kernels use equal the host's."""
    coord, elem, Q = mesh.mesh_P2_3D(26, 26, 26, 13, 26, 26, 20.0, 20.0, 20.0, 10.0, 20.0, 20.0)
    print(f"pattern 1.5M: {elem.shape[1]} elements", flush=True)
    H = A.Pattern(elem, Q, 3)
    D = _device_pattern(elem, Q, 3)
    assert (D.nnzb, D.n_free, D.nnz_free) == (H.nnzb, H.n_free, H.nnz_free)
    th, td = _all_tables(H), _all_tables(D)
    for k in ("row_ptr", "col", "ref_pos", "uptr", "pslot", "ublk", "nptr", "ncode", "tblk", "diag"):
        assert np.array_equal(th[k], td[k]), k


def _fe_device(coord, elem, dim):
    n_p, n_e = elem.shape
    coord_d = DeviceArray.from_host(np.ascontiguousarray(coord.T))
    elem_d = DeviceArray.from_host(np.ascontiguousarray(elem.T.astype(np.int32)))
    DPhi, W = DSU.gradients_dev(coord_d, elem_d, dim, coord.shape[1], n_p, n_e)
    # device layout: n_p values of a point contiguous -> n_p x n_int
    return [d.get().reshape(-1, n_p).T for d in DPhi], W.get()


@pytest.mark.parametrize("name", ["2d_h1", "3d_small", "3d_box"])
def test_fe_gradients(name):
    coord, elem, Q = MESHES[name]()
    dim = coord.shape[0]
    DPhi0, W0, HatP, WF = fem.gradients(coord, elem, dim)
    DPhi, W = _fe_device(coord, elem, dim)
    for k in range(dim):
        assert np.abs(DPhi[k] - DPhi0[k]).max() <= 1e-13 * np.abs(DPhi0[k]).max()
    assert np.abs(W - W0).max() <= 1e-13 * np.abs(W0).max()
    # partition of unity: sum_a dphi_a = 0; linear fields exactly
    n_q = WF.size
    X = np.repeat(coord[:, elem], n_q, axis=2)  # dim x n_p x n_int
    for k in range(dim):
        scale = np.abs(DPhi[k]).max()
        assert np.abs(DPhi[k].sum(axis=0)).max() <= 1e-12 * scale
        for j in range(dim):
            g = (DPhi[k] * X[j]).sum(axis=0)
            assert np.abs(g - (1.0 if j == k else 0.0)).max() <= 1e-12
    # volume of the domain: slope profile area (x extrusion in 3D)
    if name == "2d_h1":
        # mesh_P2_2D puts the crest over x1: foundation + x1*y2 + x2*y2/2
        area = (15 + 10 + 15) * 10 + (15 * 10 + 10 * 10 / 2)
        assert abs(W.sum() - area) <= 1e-10 * area
    elif name == "3d_small":
        vol = ((20 + 20 + 20) * 10 + (20 * 20 + 20 * 20 / 2)) * 20
        assert abs(W.sum() - vol) <= 1e-10 * vol
    else:
        # mesh_P2_3D puts the crest over x3: body above the foundation x3*y2 + x2*y2/2
        vol = ((3 + 2 + 2) * 2 + (2 * 1 + 2 * 1 / 2)) * 2
        assert abs(W.sum() - vol) <= 1e-10 * vol


@pytest.mark.parametrize("name", ["2d_h1", "3d_small"])
def test_materials_and_volume_force(name):
    coord, elem, Q = MESHES[name]()
    dim = coord.shape[0]
    mats = problems.HETERO_3D_LAYERS if dim == 3 else problems.HOMO_2D
    rng = np.random.default_rng(3)
    mat_id = rng.integers(0, len(mats), size=elem.shape[1])
    host = problems._finish(coord, elem, Q, dim, mats, mat_id)
    dev = problems._finish(coord, elem, Q, dim, mats, mat_id, device=True)
    for k in DSU.FIELDS:
        assert np.array_equal(dev["mats"][k].get(), host["mats"][k]), k
    f0 = host["f"].T
    f = dev["f"].get()
    assert np.abs(f - f0).max() <= 1e-12 * np.abs(f0).max()
    # total load = -sum_p gamma_p WEIGHT_p along y
    W = dev["WEIGHT"].get()
    g = dev["mats"]["gamma"].get()
    assert abs(f[:, 1].sum() + (g * W).sum()) <= 1e-10 * abs((g * W).sum())
    assert np.abs(f[:, 0].sum()) <= 1e-10 * abs((g * W).sum())


def test_continuation_with_device_setup():
    """2D slope at h = 1: the continuation set up on the device reaches the
    host-setup FoS and follows the same lambda(omega) curve."""
    params = dict(linear_tol=1e-10)
    host = problems.run_ssr(problems.slope_2d(1.0), params=params, preconditioner="amg")
    prob = problems.slope_2d(1.0, device=True)
    print(f"device setup {prob['setup_s']:.3f} s", flush=True)
    dev = problems.run_ssr(prob, params=params, preconditioner="amg")
    assert abs(dev["FoS"] - host["FoS"]) <= 1e-8
    n = min(len(dev["lambda_hist"]), len(host["lambda_hist"]))
    assert np.abs(np.asarray(dev["lambda_hist"][:n]) - np.asarray(host["lambda_hist"][:n])).max() <= 1e-8
