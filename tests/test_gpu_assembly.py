"""HIP tangent assembly, strain gather and internal force vs the oracle.

Tolerance: tangent values within 1e-11 relative, measured normwise
(max |V - V_ref| <= 1e-11 * max |V_ref|); strains/forces within 1e-12 of the
largest entry.
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

TOL_K = 1e-11


def problem(dim, size="small", seed=0):
    if dim == 2:
        coord, elem, Q = mesh.mesh_P2_2D(2.5 if size == "small" else 1.0, 5, 5, 5, 5, 5)
    elif size == "small":
        coord, elem, Q = mesh.mesh_P2_3D(1, 1, 1, 1, 1, 1, 2.0, 2.0, 2.0, 2.0, 2.0, 2.0)
    else:
        coord, elem, Q = mesh.mesh_P2_3D(2, 2, 3, 2, 2, 3, 2.0, 2.0, 3.0, 2.0, 2.0, 3.0)
    DPhi, WEIGHT, HatP, WF = fem.gradients(coord, elem, dim)
    n_int = WEIGHT.size
    G, K, L = W.moduli()
    rng = np.random.default_rng(seed)
    ns = 6 if dim == 3 else 3
    E = rng.normal(0, 2e-3, size=(ns, n_int))
    cb, sp = O.reduction(np.full(n_int, 15.0), np.full(n_int, np.pi / 6), np.zeros(n_int), 1.3, "B")
    f = O.mc3d if dim == 3 else O.mc2d
    S, DS, br = f(E, cb, sp, np.full(n_int, G), np.full(n_int, K), np.full(n_int, L))
    R = FO.RefPattern(elem, Q, dim, DPhi, WEIGHT, np.full(n_int, G), np.full(n_int, K))
    return dict(coord=coord, elem=elem, Q=Q, DPhi=DPhi, WEIGHT=WEIGHT, n_q=WF.size, DS=DS, S=S, R=R,
                shear=np.full(n_int, G), bulk=np.full(n_int, K))


def relK(V, V0):
    return np.abs(V - V0).max() / np.abs(V0).max()


@pytest.mark.parametrize("dim,size", [(3, "small"), (2, "small"), (3, "medium"), (2, "medium")])
def test_mex_dropin_scatter_map(dim, size):
    p = problem(dim, size)
    n_n = p["coord"].shape[1]
    R = p["R"]
    smap = R.scatter_map(p["elem"], n_n)
    V0 = O.assemble(dim, p["DPhi"], p["DS"], p["WEIGHT"], smap, p["n_q"], R.nnz)
    if dim == 3:
        V = A.assemble_K_tangent_vals(*p["DPhi"], p["DS"], p["WEIGHT"], smap, np.int32(p["n_q"]), np.int32(R.nnz))
    else:
        V = A.assemble_K_tangent_vals_2D(*p["DPhi"], p["DS"], p["WEIGHT"], smap, np.int32(p["n_q"]), np.int32(R.nnz))
    assert relK(V, V0) <= TOL_K


def test_mex_dropin_rejects_non_int64_map():
    p = problem(3)
    with pytest.raises(TypeError, match="int64"):
        A.assemble_K_tangent_vals(*p["DPhi"], p["DS"], p["WEIGHT"], np.zeros((900, 1), dtype=np.int32), 11, 10)


@pytest.mark.parametrize("dim,size", [(3, "small"), (2, "small"), (3, "medium")])
def test_block_assembly_regularize_extract(dim, size):
    p = problem(dim, size)
    R = p["R"]
    P = A.Pattern(p["elem"], p["Q"], dim)
    dD = [DeviceArray.from_host(x.ravel(order="F")) for x in p["DPhi"]]
    dW = DeviceArray.from_host(p["WEIGHT"])
    dDS = DeviceArray.from_host(np.asfortranarray(p["DS"]).T.copy())  # point-major
    Kt = DeviceArray(P.nnzb * P.bs)
    P.assemble(p["n_q"], dD, dDS, dW, Kt)
    Vt = DeviceArray(P.nnz_free)
    P.extract_ref(Kt, Vt)
    V0 = R.global_tangent(p["DS"], p["WEIGHT"])
    assert relK(Vt.get(), V0) <= TOL_K
    # elastic operator and the regularised, symmetrised K_r
    n_int = p["WEIGHT"].size
    ns = 6 if dim == 3 else 3
    dDSe = DeviceArray((n_int, ns * ns))
    A.ds_elastic_dev(dim, n_int, DeviceArray.from_host(p["shear"]), DeviceArray.from_host(p["bulk"]), dDSe)
    Kel = DeviceArray(P.nnzb * P.bs)
    P.assemble(p["n_q"], dD, dDSe, dW, Kel)
    r = 0.37
    Kr = DeviceArray(P.nnzb * P.bs)
    P.regularize(r, Kel, Kt, Kr)
    Vr = DeviceArray(P.nnz_free)
    P.extract_ref(Kr, Vr)
    Vr0 = R.build_K_r(r, V0)
    assert relK(Vr.get(), Vr0) <= TOL_K
    # Dirichlet rows/columns masked with a unit diagonal, free part symmetric
    Kh = Kr.get().reshape(-1, dim, dim)
    rows = np.repeat(np.arange(P.n_n), np.diff(P.row_ptr))
    mask = P.mask.reshape(-1, dim)
    col = P.col
    for b in np.flatnonzero((mask[rows] == 0).any(axis=1))[:200]:
        for i in range(dim):
            if mask[rows[b], i] == 0:
                expect = np.zeros(dim)
                if rows[b] == col[b]:
                    expect[i] = 1.0
                assert np.array_equal(Kh[b, i], expect)


@pytest.mark.parametrize("dim,size", [(3, "small"), (2, "small"), (3, "medium"), (2, "medium")])
def test_gather_and_atomic_assembly_forms(dim, size):
    """The default form (element blocks, then a per-block gather in element
    order) against the oracle and against the FP64-atomic form; the gather
    form gives the same bits on every run."""
    from slope_stability_octave_amd import _lib
    p = problem(dim, size, seed=3)
    P = A.Pattern(p["elem"], p["Q"], dim)
    dD = [DeviceArray.from_host(x.ravel(order="F")) for x in p["DPhi"]]
    dW = DeviceArray.from_host(p["WEIGHT"])
    dDS = DeviceArray.from_host(np.asfortranarray(p["DS"]).T.copy())
    out = {}
    old = _lib.get_tuning(_lib.TUNE_ASSEMBLY_FORM)
    try:
        for form in (0, 1, 0):
            _lib.set_tuning(_lib.TUNE_ASSEMBLY_FORM, form)
            K = DeviceArray(P.nnzb * P.bs)
            P.assemble(p["n_q"], dD, dDS, dW, K)
            out.setdefault(form, []).append(K.get())
    finally:
        _lib.set_tuning(_lib.TUNE_ASSEMBLY_FORM, old)
    g, a = out[0][0], out[1][0]
    assert np.array_equal(out[0][0], out[0][1])
    assert np.abs(g - a).max() <= 1e-13 * np.abs(a).max()
    Vt = DeviceArray(P.nnz_free)
    P.extract_ref(DeviceArray.from_host(g), Vt)
    assert relK(Vt.get(), p["R"].global_tangent(p["DS"], p["WEIGHT"])) <= TOL_K


@pytest.mark.parametrize("dim", [2, 3])
def test_strain_and_force(dim):
    p = problem(dim, "medium")
    n_n = p["coord"].shape[1]
    P = A.Pattern(p["elem"], p["Q"], dim)
    rng = np.random.default_rng(9)
    U = rng.normal(size=(dim, n_n))
    dU = DeviceArray.from_host(U.T)  # node-major dim x n_n == (n_n, dim) C-order
    dD = [DeviceArray.from_host(x.ravel(order="F")) for x in p["DPhi"]]
    ns = 6 if dim == 3 else 3
    n_int = p["WEIGHT"].size
    dE = DeviceArray((n_int, ns))
    P.strain(p["n_q"], dD, dU, dE)
    E0 = O.strain(dim, p["elem"], p["DPhi"], U, p["n_q"])
    assert np.abs(dE.get().T - E0).max() <= 1e-12 * np.abs(E0).max()
    srows = 6 if dim == 3 else 4
    dS = DeviceArray.from_host(np.asfortranarray(p["S"]).T.copy())
    dWt = DeviceArray.from_host(p["WEIGHT"])
    F0 = O.force(dim, p["elem"], p["DPhi"], p["S"], p["WEIGHT"], p["n_q"], n_n)
    # the default form (element-node forces, then one ordered sum per node)
    # gives the same bits on every run; the FP64-atomic form agrees to rounding
    from slope_stability_octave_amd import _lib
    out = {}
    old = _lib.get_tuning(_lib.TUNE_ASSEMBLY_FORM)
    try:
        for form in (0, 1, 0):
            _lib.set_tuning(_lib.TUNE_ASSEMBLY_FORM, form)
            dF = DeviceArray((n_n, dim))
            dF.set(np.full((n_n, dim), np.nan))  # every entry must be written
            P.force(p["n_q"], dD, dS, srows, dWt, dF)
            out.setdefault(form, []).append(dF.get().T)
    finally:
        _lib.set_tuning(_lib.TUNE_ASSEMBLY_FORM, old)
    for F in (out[0][0], out[1][0]):
        assert np.abs(F - F0).max() <= 1e-12 * np.abs(F0).max()
    assert np.array_equal(out[0][0], out[0][1])


@pytest.mark.parametrize("dim,size", [(3, "small"), (2, "small"), (3, "medium"), (2, "medium")])
def test_fused_K_r_equals_two_pass_path(dim, size):
    """ssr_dev_assemble_K_r (element pair blocks, then one gather-and-
    regularise pass over the upper blocks writing each block and its mirror)
    against assemble + regularize: the same bits, for every element-kernel
    form; the element kernels agree to rounding, and K_r matches the
    oracle's regularised tangent at 1e-11."""
    from slope_stability_octave_amd import _lib
    p = problem(dim, size, seed=5)
    P = A.Pattern(p["elem"], p["Q"], dim)
    dD = [DeviceArray.from_host(x.ravel(order="F")) for x in p["DPhi"]]
    dW = DeviceArray.from_host(p["WEIGHT"])
    dDS = DeviceArray.from_host(np.asfortranarray(p["DS"]).T.copy())
    n_int = p["WEIGHT"].size
    ns = 6 if dim == 3 else 3
    dDSe = DeviceArray((n_int, ns * ns))
    A.ds_elastic_dev(dim, n_int, DeviceArray.from_host(p["shear"]), DeviceArray.from_host(p["bulk"]), dDSe)
    Kraw = DeviceArray(P.nnzb * P.bs)
    P.assemble(p["n_q"], dD, dDSe, dW, Kraw)
    Kel = DeviceArray(P.nnzb * P.bs)
    P.regularize(1.0, Kraw, Kraw, Kel)
    r = 0.21
    old = _lib.get_tuning(_lib.TUNE_ELEMENT_KERNEL)
    res = {}
    try:
        for form in (0, 1, 2):
            _lib.set_tuning(_lib.TUNE_ELEMENT_KERNEL, form)
            Kt = DeviceArray(P.nnzb * P.bs)
            P.assemble(p["n_q"], dD, dDS, dW, Kt)
            K2 = DeviceArray(P.nnzb * P.bs)
            P.regularize(r, Kel, Kt, K2)
            K1 = DeviceArray(P.nnzb * P.bs)
            K1.set(np.full(P.nnzb * P.bs, np.nan))  # every entry must be written
            P.assemble_K_r(p["n_q"], dD, dDS, dW, r, Kel, K1)
            a, b = K1.get(), K2.get()
            assert np.array_equal(a, b), (form, np.abs(a - b).max())
            res[form] = (Kt.get(), a)
    finally:
        _lib.set_tuning(_lib.TUNE_ELEMENT_KERNEL, old)
    t0, t1 = res[0][0], res[1][0]
    assert np.abs(t0 - t1).max() <= 1e-14 * np.abs(t1).max()
    Vr = DeviceArray(P.nnz_free)
    P.extract_ref(DeviceArray.from_host(res[0][1]), Vr)
    Vr0 = p["R"].build_K_r(r, p["R"].global_tangent(p["DS"], p["WEIGHT"]))
    assert relK(Vr.get(), Vr0) <= TOL_K


class _Offset:
    """A device address a few bytes into a DeviceArray (to present an input
    that is not 16-byte aligned)."""

    def __init__(self, base, nbytes):
        self.base, self.ptr = base, base.ptr + nbytes


@pytest.mark.parametrize("n_drop", [0, 1])
def test_element_kernel_forms_same_bits(n_drop):
    """The element kernels that give each lane two pairs of one column node
    (the persistent, prefetching form for aligned MEX-layout DS; the plain
    form otherwise) against the one-pair-per-lane kernel: the same terms in
    the same order, so K_t and K_r agree bit for bit, with an even and an odd
    element count (the last element pair then holds one element), 16-byte
    aligned and unaligned DS, MEX-layout and packed DS."""
    from slope_stability_octave_amd import _lib
    p = problem(3, "medium", seed=7)
    n_q = p["n_q"]
    keep = np.arange(p["elem"].shape[1])
    if n_drop:  # drop one element whose nodes all lie in other elements too
        cnt = np.bincount(p["elem"].ravel())
        k = next(k for k in range(keep.size) if (cnt[p["elem"][:, k]] > 1).all())
        keep = np.delete(keep, k)
    elem = p["elem"][:, keep]
    n_e = elem.shape[1]
    assert n_e % 2 == (1 if n_drop else 0)
    assert np.unique(elem).size == p["coord"].shape[1]  # every node still in an element
    pts = (keep[:, None] * n_q + np.arange(n_q)).ravel()
    n_int = pts.size
    DPhi = [x[:, pts] for x in p["DPhi"]]
    WEIGHT, DS = p["WEIGHT"][pts], p["DS"][:, pts]
    P = A.Pattern(elem, p["Q"], 3)
    dD = [DeviceArray.from_host(x.ravel(order="F")) for x in DPhi]
    dW = DeviceArray.from_host(WEIGHT)
    DSm = np.asfortranarray(DS).T.copy()  # (n_int, 36)
    dDS = DeviceArray.from_host(DSm)
    # the same DS one double into a larger buffer: 8-byte but not 16-byte aligned
    buf = DeviceArray.from_host(np.concatenate([[0.0], DSm.ravel()]))
    dDSu = _Offset(buf, 8)
    iu = np.triu_indices(6)
    DSp = np.ascontiguousarray(DSm.reshape(n_int, 6, 6).transpose(0, 2, 1)[:, iu[0], iu[1]])  # row-major upper
    dDSp = DeviceArray.from_host(DSp)
    Kel = DeviceArray(P.nnzb * P.bs)
    Kel.set(np.abs(np.random.default_rng(1).normal(size=P.nnzb * P.bs)))
    old = _lib.get_tuning(_lib.TUNE_ELEMENT_KERNEL)
    res = {}
    try:
        for form in (2, 0):
            _lib.set_tuning(_lib.TUNE_ELEMENT_KERNEL, form)
            for name, ds, lay in (("mex", dDS, 0), ("unaligned", dDSu, 0), ("packed", dDSp, 1)):
                Kt = DeviceArray(P.nnzb * P.bs)
                P.assemble(n_q, dD, ds, dW, Kt, ds_layout=lay)
                Kr = DeviceArray(P.nnzb * P.bs)
                P.assemble_K_r(n_q, dD, ds, dW, 0.3, Kel, Kr, ds_layout=lay)
                res[(form, name)] = (Kt.get(), Kr.get())
    finally:
        _lib.set_tuning(_lib.TUNE_ELEMENT_KERNEL, old)
    ref_t, ref_r = res[(2, "mex")]
    for key, (kt, kr) in res.items():
        assert np.array_equal(kt, ref_t), key
        assert np.array_equal(kr, ref_r), key
    # and the values themselves against the oracle's element assembly
    R = FO.RefPattern(elem, p["Q"], 3, DPhi, WEIGHT, p["shear"][pts], p["bulk"][pts])
    Vt = DeviceArray(P.nnz_free)
    P.extract_ref(DeviceArray.from_host(ref_t), Vt)
    assert relK(Vt.get(), R.global_tangent(DS, WEIGHT)) <= TOL_K
