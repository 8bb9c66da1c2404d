"""GPU SpMV, deflated FGMRES / CG and A-orthogonalisation vs scipy.

The operators are assembled on the GPU from a small slope problem; the
reference solutions come from scipy's sparse direct solver on K(Q,Q) in the
reference's (I, J, V) layout.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from slope_stability_octave_amd import _lib, problems
from slope_stability_octave_amd._lib import DeviceArray
from slope_stability_octave_amd.linear_solvers import DCG, DFGMRES, dot, spmv

pytestmark = pytest.mark.gpu


def operator(dim=2, r=1.0, h=2.0, counts=(1, 1, 1, 1, 1, 1)):
    if dim == 2:
        prob = problems.slope_2d(h=h)
    else:
        prob = problems.slope_3d(counts, 2.0, 2.0, 2.0, 2.0, 2.0, 2.0, problems.HOMO_2D)
    cmb = problems.make_constitutive(prob)
    cmb.reduction(1.0)
    U = DeviceArray((cmb.n_n, cmb.dim), zero=True)
    cmb.build_F_and_DS_all(1.0, U)
    K = cmb.build_K_r_QQ_vals(r)
    P = cmb.pattern
    Vref = DeviceArray(P.nnz_free)
    P.extract_ref(K, Vref)
    # reference-order triplets over the free DOFs
    rp = P.ref_pos
    rows = np.repeat(np.arange(P.n_n), np.diff(P.row_ptr))
    blk = rp // P.bs
    ri, cj = np.divmod(rp % P.bs, dim)
    I = dim * rows[blk] + ri
    J = dim * P.col[blk] + cj
    qf = np.flatnonzero(P.mask)
    free = np.full(P.n_dof, -1)
    free[qf] = np.arange(qf.size)
    Kq = sp.csr_matrix((Vref.get(), (free[I], free[J])), shape=(qf.size, qf.size))
    return cmb, K, Kq, qf


def _full(vq, qf, n_dof, dim):
    x = np.zeros(n_dof)
    x[qf] = vq
    return x.reshape(-1, dim)


@pytest.mark.parametrize("group", [0, 4, 8, 16, 32, 64])
@pytest.mark.parametrize("dim", [2, 3])
def test_spmv_matches_scipy(dim, group):
    """Every SpMV form (lanes per block row; 64 = one wavefront per row with
    lanes over the values) against scipy."""
    cmb, K, Kq, qf = operator(dim, counts=(2, 2, 2, 2, 2, 2))
    rng = np.random.default_rng(1)
    xq = rng.normal(size=qf.size)
    x = DeviceArray.from_host(_full(xq, qf, cmb.pattern.n_dof, dim))
    y = DeviceArray((cmb.n_n, dim))
    old = _lib.get_tuning(_lib.TUNE_SPMV_GROUP)
    try:
        _lib.set_tuning(_lib.TUNE_SPMV_GROUP, group)
        spmv(cmb.pattern, K, x, y)
    finally:
        _lib.set_tuning(_lib.TUNE_SPMV_GROUP, old)
    yq = y.get().ravel()[qf]
    ref = Kq @ xq
    assert np.abs(yq - ref).max() <= 1e-12 * np.abs(ref).max()
    # Dirichlet entries stay zero
    yfull = y.get().ravel()
    assert not np.any(yfull[np.setdiff1d(np.arange(yfull.size), qf)])


@pytest.mark.parametrize("dim", [2, 3])
def test_float_spmv_matches_scipy(dim):
    """The float operator copy (ssr_dev_convert_f64_f32) and the float SpMV
    against scipy in double: within float rounding of the row sums."""
    cmb, K, Kq, qf = operator(dim, counts=(2, 2, 2, 2, 2, 2))
    P = cmb.pattern
    nv = P.nnzb * P.bs
    K32 = DeviceArray(nv, dtype=np.float32)
    _lib.call("ssr_dev_convert_f64_f32", nv, _lib.ptr(K), _lib.ptr(K32))
    rng = np.random.default_rng(7)
    xq = rng.normal(size=qf.size)
    x = DeviceArray.from_host(_full(xq, qf, P.n_dof, dim).astype(np.float32))
    y = DeviceArray((cmb.n_n, dim), dtype=np.float32)
    _lib.call("ssr_dev_spmv_bsr_f32", dim, P.n_n, _lib.ptr(P.d_row_ptr), _lib.ptr(P.d_col), _lib.ptr(K32),
              _lib.ptr(x), _lib.ptr(y), 1.0, 0.0)
    ref = Kq @ xq.astype(np.float32).astype(np.float64)
    yq = y.get().astype(np.float64).ravel()[qf]
    assert np.abs(yq - ref).max() <= 1e-5 * np.abs(ref).max()


@pytest.mark.parametrize("dim", [2, 3])
def test_spmv_lds_staged_form_equals_direct(dim):
    """The SpMV whose block values go through LDS gives the same bits as the
    direct form (the same products in the same order), on a mesh with rows
    of every length."""
    cmb, K, Kq, qf = operator(dim, counts=(3, 2, 3, 2, 3, 2))
    rng = np.random.default_rng(4)
    x = DeviceArray.from_host(rng.normal(size=(cmb.n_n, dim)))
    out = {}
    old = _lib.get_tuning(_lib.TUNE_SPMV_STAGE)
    try:
        for stage in (0, 1):
            _lib.set_tuning(_lib.TUNE_SPMV_STAGE, stage)
            y = DeviceArray((cmb.n_n, dim))
            spmv(cmb.pattern, K, x, y)
            out[stage] = y.get()
    finally:
        _lib.set_tuning(_lib.TUNE_SPMV_STAGE, old)
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("cls", [DFGMRES, DCG])
def test_krylov_solves_to_tolerance(cls):
    cmb, K, Kq, qf = operator(2, r=1.0)
    rng = np.random.default_rng(2)
    bq = rng.normal(size=qf.size)
    xref = spla.spsolve(Kq.tocsc(), bq)
    s = cls(cmb.pattern, 1e-10, 2000 if cls is DCG else 800, 1e-3)
    s.setup_preconditioner(K)
    b = DeviceArray.from_host(_full(bq, qf, cmb.pattern.n_dof, 2))
    x, it = s.solve(K, b)
    xq = x.get().ravel()[qf]
    assert it > 0
    assert np.linalg.norm(Kq @ xq - bq) <= 1.1e-10 * np.linalg.norm(bq)
    assert np.linalg.norm(xq - xref) <= 1e-6 * np.linalg.norm(xref)


def test_A_orthogonalize_and_deflated_solve():
    cmb, K, Kq, qf = operator(2, r=1.0)
    rng = np.random.default_rng(3)
    n = cmb.pattern.n_dof
    s = DFGMRES(cmb.pattern, 1e-8, 800, 1e-3)
    s.setup_preconditioner(K)
    vecs = [rng.normal(size=qf.size) for _ in range(5)]
    vecs.append(vecs[0] * 2.0 + vecs[1])  # dependent: must be dropped
    for v in vecs:
        s.expand_deflation_basis(DeviceArray.from_host(_full(v, qf, n, 2)))
    s.A_orthogonalize(K)
    Wb = s.deflation_basis()[qf]
    assert Wb.shape[1] == 5
    G = Wb.T @ (Kq @ Wb)
    assert np.abs(G - np.eye(5)).max() <= 1e-9
    # same solution with deflation; x0 = W W'b already resolves the basis span
    bq = Kq @ vecs[2]
    b = DeviceArray.from_host(_full(bq, qf, n, 2))
    x, it = s.solve(K, b)
    assert it <= 1
    assert np.linalg.norm(x.get().ravel()[qf] - vecs[2]) <= 1e-6 * np.linalg.norm(vecs[2])


@pytest.mark.parametrize("spmm_group", [0, 32])
def test_A_orthogonalize_large_basis_matches_oracle(spmm_group):
    """72 columns (two dependent) on the 3D operator: two 64-column tiles of
    the Gram product, a partial 8-column group of the multi-vector SpMV (two
    lane shapes) and several row chunks; the basis must equal the oracle's
    column for column, and a second pass must leave an A-orthonormal basis
    unchanged."""
    from oracle import krylov_oracle as KO
    from slope_stability_octave_amd import _lib
    old = _lib.get_tuning(_lib.TUNE_SPMM_GROUP)
    _lib.set_tuning(_lib.TUNE_SPMM_GROUP, spmm_group)
    try:
        _check_A_orthogonalize_large_basis(KO)
    finally:
        _lib.set_tuning(_lib.TUNE_SPMM_GROUP, old)


@pytest.mark.parametrize("dim,spmm_group", [(3, 0), (3, 8), (3, 32), (2, 0), (2, 16)])
def test_A_orthogonalize_spmm_halving_sum_same_bits(dim, spmm_group):
    """Tuning key 12: the multi-vector SpMV's lane sums by halving exchanges
    give the basis of the full-butterfly form bit for bit (70 columns: full
    and partial 8-vector passes; lane shapes with odd and even final counts)."""
    from slope_stability_octave_amd import _lib
    cmb, K, Kq, qf = operator(dim, counts=(2, 2, 2, 2, 2, 2)) if dim == 3 else operator(dim)
    rng = np.random.default_rng(9)
    n = cmb.pattern.n_dof
    Wq = rng.normal(size=(qf.size, 70))
    out = {}
    old = (_lib.get_tuning(_lib.TUNE_SPMM_GROUP), _lib.get_tuning(_lib.TUNE_HALVING_SUMS))
    try:
        _lib.set_tuning(_lib.TUNE_SPMM_GROUP, spmm_group)
        for mode in (0, 1):
            _lib.set_tuning(_lib.TUNE_HALVING_SUMS, mode)
            s = DFGMRES(cmb.pattern, 1e-8, 100, 1e-3)
            s.setup_preconditioner(K)
            for c in range(Wq.shape[1]):
                s.expand_deflation_basis(DeviceArray.from_host(_full(Wq[:, c], qf, n, dim)))
            s.A_orthogonalize(K)
            out[mode] = s.deflation_basis()
    finally:
        _lib.set_tuning(_lib.TUNE_SPMM_GROUP, old[0])
        _lib.set_tuning(_lib.TUNE_HALVING_SUMS, old[1])
    assert out[0].shape == out[1].shape and out[0].shape[1] == 70
    assert np.array_equal(out[0], out[1])


def _check_A_orthogonalize_large_basis(KO):
    cmb, K, Kq, qf = operator(3, counts=(2, 2, 2, 2, 2, 2))
    rng = np.random.default_rng(5)
    n = cmb.pattern.n_dof
    Wq = rng.normal(size=(qf.size, 70))
    Wq = np.concatenate([Wq, (Wq[:, 3] - 0.5 * Wq[:, 40])[:, None], (Wq[:, 64] + Wq[:, 65])[:, None]], axis=1)
    s = DFGMRES(cmb.pattern, 1e-8, 100, 1e-3)
    s.setup_preconditioner(K)
    for c in range(Wq.shape[1]):
        s.expand_deflation_basis(DeviceArray.from_host(_full(Wq[:, c], qf, n, 3)))
    s.A_orthogonalize(K)
    Wb = s.deflation_basis()[qf]
    Wo, sg = KO.a_orthogonalize(Wq, Kq, 1e-3)
    assert Wo.shape == (qf.size, 70) and Wb.shape == Wo.shape
    assert np.abs(Wb - Wo).max() <= 1e-9 * np.abs(Wo).max()
    s.A_orthogonalize(K)
    Wb2 = s.deflation_basis()[qf]
    assert np.abs(np.abs(Wb2) - np.abs(Wb)).max() <= 1e-9 * np.abs(Wb).max()


@pytest.mark.parametrize("k", [1, 5, 8, 13, 128])
def test_multidot_and_gemv_update_match_numpy(k):
    """Tall-skinny V'u (8-column passes, shuffle reduction) and u = a u + b V h
    against numpy at a size with a partial last block."""
    import ctypes
    from slope_stability_octave_amd.linear_solvers import call, ptr  # noqa: F401  (binds signatures)
    n = 100_003
    rng = np.random.default_rng(k)
    V = rng.normal(size=(k, n))  # k columns of length n, column stride n
    u = rng.normal(size=n)
    dV, du = DeviceArray.from_host(V), DeviceArray.from_host(u)
    dh = DeviceArray(k)
    call("ssr_dev_multidot", n, k, ptr(dV), n, ptr(du), ptr(dh))
    ref = V @ u
    assert np.abs(dh.get() - ref).max() <= 1e-12 * np.abs(V).max() * np.abs(u).max() * np.sqrt(n)
    call("ssr_dev_gemv_update", n, k, ptr(dV), n, ptr(dh), ctypes.c_double(0.5), ctypes.c_double(-2.0), ptr(du))
    ref_u = 0.5 * u - 2.0 * (V.T @ dh.get())
    assert np.abs(du.get() - ref_u).max() <= 1e-12 * np.abs(ref_u).max() * k


def test_shared_amg_hierarchy_outlives_its_owner_and_rebuilds_on_request():
    """Copies of a solver share its AMG object: each copy solves to the
    tolerance with it, the object survives the original's destruction, and
    an invalidated hierarchy is rebuilt on the next setup (same solution)."""
    import gc
    cmb, K, Kq, qf = operator(3, counts=(2, 2, 2, 2, 2, 2))
    rng = np.random.default_rng(7)
    n = cmb.pattern.n_dof
    from slope_stability_octave_amd import mesh
    coord, _, _ = mesh.mesh_P2_3D(2, 2, 2, 2, 2, 2, 2.0, 2.0, 2.0, 2.0, 2.0, 2.0)
    orig = DFGMRES(cmb.pattern, 1e-8, 300, 1e-3, "amg", amg_options=dict(coord=coord))
    bq = rng.normal(size=qf.size)
    b = DeviceArray.from_host(_full(bq, qf, n, 3))
    c1 = orig.copy()
    c1.setup_preconditioner(K)
    x1, it1 = c1.solve(K, b)
    c2 = orig.copy()
    del orig, c1
    gc.collect()
    c2.setup_preconditioner(K)  # shared hierarchy: kept
    x2, it2 = c2.solve(K, b)
    assert it2 == it1 and np.array_equal(x1.get(), x2.get())
    c2.invalidate_preconditioner()
    c2.setup_preconditioner(K)  # rebuilt from the same operator
    x3, it3 = c2.solve(K, b)
    assert it3 == it1 and np.abs(x3.get() - x1.get()).max() <= 1e-12 * np.abs(x1.get()).max()
    for x in (x1, x3):
        assert np.linalg.norm(Kq @ x.get().ravel()[qf] - bq) <= 1.1e-8 * np.linalg.norm(bq)


@pytest.mark.parametrize("prec", [64, 32])
def test_symmetric_projector_matches_spmv_projector(prec):
    """With a symmetric operator the deflation projector takes W'A x as
    (A W)' x from the A-orthogonalisation; the solves must match the form that
    applies A every iteration, in double and in float."""
    cmb, K, Kq, qf = operator(3, r=0.3, counts=(2, 2, 2, 2, 2, 2))
    rng = np.random.default_rng(9)
    n = cmb.pattern.n_dof
    vecs = [rng.normal(size=qf.size) for _ in range(12)]
    bq = rng.normal(size=qf.size)
    b = DeviceArray.from_host(_full(bq, qf, n, 3))
    out = {}
    for sym in (True, False):
        s = DFGMRES(cmb.pattern, 1e-9, 400, 1e-3, precision=prec, symmetric=sym)
        s.setup_preconditioner(K)
        for v in vecs:
            s.expand_deflation_basis(DeviceArray.from_host(_full(v, qf, n, 3)))
        s.A_orthogonalize(K)
        x, it = s.solve(K, b)
        out[sym] = (x.get().ravel()[qf], it)
        assert np.linalg.norm(Kq @ out[sym][0] - bq) <= 1.1e-9 * np.linalg.norm(bq)
    assert abs(out[True][1] - out[False][1]) <= 2
    assert np.linalg.norm(out[True][0] - out[False][0]) <= 1e-7 * np.linalg.norm(out[False][0])


def test_fgmres_relative_tolerance_stop():
    """With tol 1e-1 the solver stops at the first iterate whose relative
    residual estimate is below 0.1 and reports it."""
    cmb, K, Kq, qf = operator(2, r=1e-4)
    rng = np.random.default_rng(4)
    bq = rng.normal(size=qf.size)
    s = DFGMRES(cmb.pattern, 1e-1, 100, 1e-3)
    s.setup_preconditioner(K)
    x, it = s.solve(K, DeviceArray.from_host(_full(bq, qf, cmb.pattern.n_dof, 2)))
    rel = np.linalg.norm(Kq @ x.get().ravel()[qf] - bq) / np.linalg.norm(bq)
    assert 1 <= it <= 100
    assert s.last_relres < 1e-1 and abs(rel - s.last_relres) <= 1e-6 + 1e-6 * rel


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("block", [0, 1])
def test_amg_preconditioned_fgmres(dim, block):
    """SA-AMG (finest-level Chebyshev with scalar or node-block diagonal
    scaling): converges to tolerance and needs far fewer iterations than
    node-block Jacobi on the same operator."""
    from slope_stability_octave_amd import _lib
    old = _lib.get_tuning(_lib.TUNE_AMG_BLOCK_SMOOTH)
    _lib.set_tuning(_lib.TUNE_AMG_BLOCK_SMOOTH, block)
    try:
        _amg_vs_jacobi(dim)
    finally:
        _lib.set_tuning(_lib.TUNE_AMG_BLOCK_SMOOTH, old)


def _amg_vs_jacobi(dim):
    from slope_stability_octave_amd import amg
    if dim == 2:
        prob = problems.slope_2d(h=1.0)
    else:
        prob = problems.slope_3d((2, 2, 2, 2, 2, 2), 2.0, 2.0, 2.0, 2.0, 2.0, 2.0, problems.HOMO_2D)
    cmb = problems.make_constitutive(prob)
    cmb.reduction(1.0)
    U = DeviceArray((cmb.n_n, cmb.dim), zero=True)
    cmb.build_F_and_DS_all(1.0, U)
    K = cmb.build_K_r_QQ_vals(1e-4)
    rng = np.random.default_rng(7)
    b = DeviceArray.from_host(rng.normal(size=(cmb.n_n, dim)) * cmb.Q.T)
    its = {}
    for name in ("block_jacobi", "amg"):
        s = DFGMRES(cmb.pattern, 1e-8, 1500, 1e-3, name, amg_options=dict(coord=prob["coord"]))
        s.setup_preconditioner(K)
        x, it = s.solve(K, b)
        its[name] = it
        r = DeviceArray((cmb.n_n, dim))
        spmv(cmb.pattern, K, x, r)
        res = np.linalg.norm(r.get() - b.get()) / np.linalg.norm(b.get())
        assert res <= 2e-8, (name, res)
    print("iterations", its)
    assert its["amg"] * 3 <= its["block_jacobi"]


@pytest.mark.parametrize("precond", ["block_jacobi", "amg"])
@pytest.mark.parametrize("tol", [1e-2, 1e-8])
def test_mixed_precision_refinement_meets_fp64_tolerance(precond, tol):
    """Float Krylov + float preconditioner with double refinement: the TRUE
    double residual meets the tolerance (also below float resolution, 1e-8,
    through repeated float solves), and the solution agrees with the FP64
    solver's to the accuracy both guarantee."""
    prob = problems.slope_3d((2, 2, 2, 2, 2, 2), 2.0, 2.0, 2.0, 2.0, 2.0, 2.0, problems.HOMO_2D)
    cmb = problems.make_constitutive(prob)
    cmb.reduction(1.0)
    U = DeviceArray((cmb.n_n, cmb.dim), zero=True)
    cmb.build_F_and_DS_all(1.0, U)
    K = cmb.build_K_r_QQ_vals(1e-4)
    rng = np.random.default_rng(11)
    b = DeviceArray.from_host(rng.normal(size=(cmb.n_n, 3)) * cmb.Q.T)
    xs = {}
    for prec in (64, 32):
        s = DFGMRES(cmb.pattern, tol, 1500, 1e-3, precond, amg_options=dict(coord=prob["coord"]), precision=prec)
        s.setup_preconditioner(K)
        # a deflation basis exercises the float projector as well
        for c in range(3):
            s.expand_deflation_basis(DeviceArray.from_host(rng.normal(size=(cmb.n_n, 3)) * cmb.Q.T))
        s.A_orthogonalize(K)
        x, it = s.solve(K, b)
        r = DeviceArray((cmb.n_n, 3))
        spmv(cmb.pattern, K, x, r)
        res = np.linalg.norm(r.get() - b.get()) / np.linalg.norm(b.get())
        print(prec, precond, tol, "iterations", it, "relres", res, s.last_relres, flush=True)
        assert res < tol * 1.0001, (prec, res)
        if prec == 32:
            assert abs(s.last_relres - res) <= 1e-3 * res + 1e-14  # reported residual is the true one
        xs[prec] = x.get()
    # both solutions solve K x = b to relative residual tol; the difference is bounded by
    # cond(K)*2*tol relative -- check a loose, tolerance-scaled agreement
    d = np.linalg.norm(xs[32] - xs[64]) / np.linalg.norm(xs[64])
    assert d < max(1e3 * tol, 1e-4), d


@pytest.mark.skip(reason="V-cycle graph replay is disabled: a profiled FP32 continuation ended in a host "
                         "SIGSEGV inside Amg::apply_graph<float> whose cause is not known")
@pytest.mark.parametrize("prec", [64, 32])
def test_amg_graph_replays_match_eager(prec):
    """The V-cycle replayed from its captured HIP graph equals the eager
    launch sequence bit for bit, on every replay, with other device work and
    a re-capture (new operator values through a new setup) in between."""
    from slope_stability_octave_amd import _lib
    prob = problems.slope_3d((2, 2, 2, 2, 2, 2), 2.0, 2.0, 2.0, 2.0, 2.0, 2.0, problems.HOMO_2D)
    cmb = problems.make_constitutive(prob)
    cmb.reduction(1.0)
    U = DeviceArray((cmb.n_n, cmb.dim), zero=True)
    cmb.build_F_and_DS_all(1.0, U)
    K = cmb.build_K_r_QQ_vals(1e-4)
    rng = np.random.default_rng(3)
    s = DFGMRES(cmb.pattern, 1e-8, 200, 1e-3, "amg", amg_options=dict(coord=prob["coord"]), precision=prec)
    s.setup_preconditioner(K)
    name = "ssr_solver_precond_apply"
    assert prec in (32, 64)
    xs = [rng.normal(size=(cmb.n_n, 3)) * cmb.Q.T for _ in range(4)]
    old = _lib.get_tuning(_lib.TUNE_AMG_GRAPH)
    try:
        for rep in range(2):
            outs = {}
            for mode in (0, 1, 1, 0, 1):
                _lib.set_tuning(_lib.TUNE_AMG_GRAPH, mode)
                res = []
                for x in xs:
                    dx = DeviceArray.from_host(x)
                    dy = DeviceArray((cmb.n_n, 3))
                    _lib.call(name, s.handle, _lib.ptr(dx), _lib.ptr(dy))
                    res.append(dy.get())
                    spmv(cmb.pattern, K, dx, dy)  # unrelated eager work between replays
                outs.setdefault(mode, []).append(res)
            ref = outs[0][0]
            for mode in (0, 1):
                for res in outs[mode]:
                    for a, b in zip(res, ref):
                        assert np.array_equal(a, b)
            # whole solves (float V-cycle when prec == 32): identical iterates
            b = DeviceArray.from_host(xs[0])
            sols = []
            for mode in (0, 1, 1):
                _lib.set_tuning(_lib.TUNE_AMG_GRAPH, mode)
                x, it = s.solve(K, b)
                sols.append((x.get(), it))
            assert sols[0][1] == sols[1][1] == sols[2][1]
            assert np.array_equal(sols[0][0], sols[1][0]) and np.array_equal(sols[0][0], sols[2][0])
            # new operator values: a new setup must invalidate the graph
            K = cmb.build_K_r_QQ_vals(0.3)
            s.setup_preconditioner(K)
    finally:
        _lib.set_tuning(_lib.TUNE_AMG_GRAPH, old)


@pytest.mark.parametrize("prec,dim", [(64, 3), (32, 3), (64, 2)])
def test_amg_coarse_block_csr_matches_scalar_csr(prec, dim):
    """The V-cycle with its coarse operators and transfers in block CSR over
    whole aggregates (the default) against the same hierarchy in scalar CSR:
    the same entries summed in another lane order, so the results agree to
    rounding (1e-12 in double, 1e-5 in float)."""
    from slope_stability_octave_amd import _lib
    if dim == 3:
        prob = problems.slope_3d((3, 3, 3, 3, 3, 3), 3.0, 3.0, 3.0, 3.0, 3.0, 3.0, problems.HOMO_2D)
    else:
        prob = problems.slope_2d(h=0.5)
    cmb = problems.make_constitutive(prob)
    cmb.reduction(1.0)
    U = DeviceArray((cmb.n_n, cmb.dim), zero=True)
    cmb.build_F_and_DS_all(1.0, U)
    K = cmb.build_K_r_QQ_vals(1e-4)
    rng = np.random.default_rng(11)
    x = DeviceArray.from_host(rng.normal(size=(cmb.n_n, cmb.dim)) * cmb.Q.T)
    old = _lib.get_tuning(_lib.TUNE_AMG_COARSE_BLOCKS)
    out = {}
    try:
        for mode in (1, 0):
            _lib.set_tuning(_lib.TUNE_AMG_COARSE_BLOCKS, mode)
            s = DFGMRES(cmb.pattern, 1e-8, 200, 1e-3, "amg", amg_options=dict(coord=prob["coord"], max_coarse=50),
                        precision=prec)
            s.setup_preconditioner(K)
            y = DeviceArray((cmb.n_n, cmb.dim))
            if prec == 64:
                _lib.call("ssr_solver_precond_apply", s.handle, _lib.ptr(x), _lib.ptr(y))
                out[mode] = y.get()
            else:
                xs, it = s.solve(K, x)
                out[mode] = xs.get()
    finally:
        _lib.set_tuning(_lib.TUNE_AMG_COARSE_BLOCKS, old)
    a, b = out[1], out[0]
    rel = np.abs(a - b).max() / np.abs(b).max()
    print(f"prec {prec} dim {dim}: block vs scalar CSR, max rel {rel:.2e}", flush=True)
    assert rel <= (1e-12 if prec == 64 else 1e-5)


@pytest.mark.parametrize("dim,n_n", [(3, 1), (3, 200_003), (2, 77_777)])
def test_node_norm_max_matches_numpy(dim, n_n):
    """U_max of the continuation histories (max over nodes of the node
    vector's Euclidean norm) against numpy on the same field; the largest
    node placed at the end, in the middle and (n_n = 1) alone."""
    from slope_stability_octave_amd.continuation import _umax
    rng = np.random.default_rng(n_n)
    u = rng.standard_normal((n_n, dim))
    u[n_n // 2] *= 5.0
    want = float(np.sqrt((u * u).sum(axis=1)).max())
    got = _umax(DeviceArray.from_host(u))
    print(f"dim {dim} n_n {n_n}: {got!r} vs {want!r}", flush=True)
    assert abs(got - want) <= 1e-15 * want
    u[-1] *= 10.0
    want = float(np.sqrt((u * u).sum(axis=1)).max())
    assert abs(_umax(DeviceArray.from_host(u)) - want) <= 1e-15 * want


@pytest.mark.parametrize("k", [1, 7, 8, 20, 92])
def test_float_multidot_and_basis_update(k):
    """The float forms the precision-32 iterations run: h = V' u with float
    V, u summed in double, and u = a u + b V h, against numpy in double on
    the same float inputs."""
    n = 300_007
    rng = np.random.default_rng(k)
    V = rng.standard_normal((k, n)).astype(np.float32)
    u = rng.standard_normal(n).astype(np.float32)
    Vd, ud, h = DeviceArray.from_host(V), DeviceArray.from_host(u), DeviceArray(k)
    _lib.call("ssr_dev_multidot_f32", n, k, _lib.ptr(Vd), n, _lib.ptr(ud), _lib.ptr(h))
    want = V.astype(np.float64) @ u.astype(np.float64)
    got = h.get()
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    hh = rng.standard_normal(k)
    hd = DeviceArray.from_host(hh)
    _lib.call("ssr_dev_gemv_update_f32", n, k, _lib.ptr(Vd), n, _lib.ptr(hd), 0.5, -2.0, _lib.ptr(ud))
    want_u = (0.5 * u.astype(np.float64) - 2.0 * (V.astype(np.float64).T @ hh)).astype(np.float32)
    got_u = ud.get()
    assert np.abs(got_u.astype(np.float64) - want_u).max() <= 2e-6 * np.abs(want_u).max()



@pytest.mark.parametrize("key", ["TUNE_HALVING_SUMS", "TUNE_CHEB_PER_DOF", "TUNE_BRC_PREFETCH"])
@pytest.mark.parametrize("prec,dim", [(64, 3), (32, 3), (64, 2), (32, 2)])
def test_vcycle_halving_sums_same_bits(prec, dim, key):
    """Tuning keys 12, 13 and 18 on the V-cycle: the block-CSR SpMVs of the
    coarse levels, prolongators and restrictions with lane sums by halving
    exchanges (12) and with each lane's next block loaded with its current
    one (18), and the finest-level Chebyshev block step with one thread per
    unknown (13), give the other forms' results bit for bit (one V-cycle in
    double; a whole float solve, V-cycles in float, in precision 32)."""
    from slope_stability_octave_amd import _lib
    if dim == 3:
        prob = problems.slope_3d((3, 3, 3, 3, 3, 3), 3.0, 3.0, 3.0, 3.0, 3.0, 3.0, problems.HOMO_2D)
    else:
        prob = problems.slope_2d(h=0.5)
    cmb = problems.make_constitutive(prob)
    cmb.reduction(1.0)
    U = DeviceArray((cmb.n_n, cmb.dim), zero=True)
    cmb.build_F_and_DS_all(1.0, U)
    K = cmb.build_K_r_QQ_vals(1e-4)
    rng = np.random.default_rng(13)
    x = DeviceArray.from_host(rng.normal(size=(cmb.n_n, cmb.dim)) * cmb.Q.T)
    old = _lib.get_tuning(getattr(_lib, key))
    out = {}
    try:
        for mode in (0, 1):
            _lib.set_tuning(getattr(_lib, key), mode)
            s = DFGMRES(cmb.pattern, 1e-8, 200, 1e-3, "amg", amg_options=dict(coord=prob["coord"], max_coarse=50),
                        precision=prec)
            s.setup_preconditioner(K)
            if prec == 64:
                y = DeviceArray((cmb.n_n, cmb.dim))
                _lib.call("ssr_solver_precond_apply", s.handle, _lib.ptr(x), _lib.ptr(y))
                out[mode] = y.get()
            else:
                xs, it = s.solve(K, x)
                out[mode] = (xs.get(), it)
    finally:
        _lib.set_tuning(getattr(_lib, key), old)
    if prec == 64:
        assert np.abs(out[1]).max() > 0
        assert np.array_equal(out[0], out[1])
    else:
        assert out[0][1] == out[1][1]
        assert np.array_equal(out[0][0], out[1][0])


@pytest.mark.parametrize("k", [1, 5, 8, 13, 70])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_multidot_and_update_prefetch_same_bits(k, dtype):
    """Tuning key 17: the multi-dot with a pass's loads issued before its
    products gives the guarded form's results bit for bit (long and short
    columns, partial column groups, both precisions); the update after it
    as well."""
    import ctypes
    from slope_stability_octave_amd.linear_solvers import call, ptr  # noqa: F401  (binds signatures)
    f32 = dtype == np.float32
    out = {}
    for n in (300_007, 20_011):
        rng = np.random.default_rng(k + n)
        V = rng.normal(size=(k, n)).astype(dtype)
        u = rng.normal(size=n).astype(dtype)
        dV, du = DeviceArray.from_host(V), DeviceArray.from_host(u)
        old = _lib.get_tuning(_lib.TUNE_KRYLOV_PREFETCH)
        try:
            for mode in (0, 1):
                _lib.set_tuning(_lib.TUNE_KRYLOV_PREFETCH, mode)
                dh = DeviceArray(k)
                call("ssr_dev_multidot_f32" if f32 else "ssr_dev_multidot", n, k, ptr(dV), n, ptr(du), ptr(dh))
                du2 = DeviceArray.from_host(u)
                call("ssr_dev_gemv_update_f32" if f32 else "ssr_dev_gemv_update", n, k, ptr(dV), n, ptr(dh),
                     ctypes.c_double(0.5), ctypes.c_double(-2.0), ptr(du2))
                out[(n, mode)] = (dh.get(), du2.get())
        finally:
            _lib.set_tuning(_lib.TUNE_KRYLOV_PREFETCH, old)
        assert np.array_equal(out[(n, 0)][0], out[(n, 1)][0])
        assert np.array_equal(out[(n, 0)][1], out[(n, 1)][1])
        ref = V.astype(np.float64) @ u.astype(np.float64)
        assert np.abs(out[(n, 1)][0] - ref).max() <= (1e-4 if f32 else 1e-11) * np.abs(V).max() * np.abs(u).max() * np.sqrt(n)
