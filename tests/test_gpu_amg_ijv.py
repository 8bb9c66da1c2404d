"""The BoomerAMG binding's commands (hypre_boomeramg_mex.cpp:751-1270) over the
C ABI: setup_ijv from the reference-order COO triplets of K_r(Q,Q) (1-based
I, J as doubles, V) with the rigid-body near-null space, apply, update_values,
clear and info, judged against scipy:

* the preconditioner is symmetric (y1'r2 = r1'y2 to rounding, as a V-cycle
  with matching pre/post Chebyshev smoothing is) and positive;
* CG on K(Q,Q) preconditioned by it (scipy's cg with the apply as M)
  reaches 1e-8 in far fewer iterations than unpreconditioned CG, and its
  solution matches scipy's direct solve;
* update_values with scaled values scales the preconditioner inversely;
* the id-keyed errors of the reference (unknown id, wrong length).
The AMG itself is this library's, not BoomerAMG (parity unpinned by design).
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from slope_stability_octave_amd import amg, problems
from slope_stability_octave_amd._lib import DeviceArray, SSRError

pytestmark = pytest.mark.gpu


def reference_triplets():
    prob = problems.slope_3d((2, 2, 2, 2, 2, 2), 4.0, 4.0, 4.0, 4.0, 4.0, 4.0, problems.HOMO_2D)
    cmb = problems.make_constitutive(prob)
    cmb.reduction(1.0)
    U = DeviceArray((cmb.n_n, 3), zero=True)
    cmb.build_F_and_DS_all(1.0, U)
    K = cmb.build_K_r_QQ_vals(1e-4)
    P = cmb.pattern
    Vref = DeviceArray(P.nnz_free)
    P.extract_ref(K, Vref)
    rp = P.ref_pos
    rows = np.repeat(np.arange(P.n_n), np.diff(P.row_ptr))
    blk = rp // P.bs
    ri, cj = np.divmod(rp % P.bs, 3)
    I = 3 * rows[blk] + ri
    J = 3 * P.col[blk] + cj
    qf = np.flatnonzero(P.mask)
    free = np.full(P.n_dof, -1)
    free[qf] = np.arange(qf.size)
    Z = amg.rigid_body_modes(prob["coord"])[qf]
    return free[I] + 1.0, free[J] + 1.0, Vref.get(), qf.size, Z


def test_amg_ijv_commands_against_scipy():
    I, J, V, n, Z = reference_triplets()
    A = sp.csr_matrix((V, (I.astype(int) - 1, J.astype(int) - 1)), shape=(n, n))
    print(f"n {n}, nnz {V.size}, n mod 3 = {n % 3}", flush=True)
    amg.hypre_boomeramg_clear()
    info = amg.hypre_boomeramg_setup_ijv(I, J, V, n, Z, None, "Kr")
    assert info["n"] == n and info["nnz"] == V.size and info["num_functions"] == 3
    assert info["num_interp_vectors"] == 6
    assert amg.hypre_boomeramg_info()["instances"] == 1
    rng = np.random.default_rng(0)
    r1, r2 = rng.normal(size=n), rng.normal(size=n)
    y1, y2 = amg.hypre_boomeramg_apply(r1, "Kr"), amg.hypre_boomeramg_apply(r2, "Kr")
    sym = abs(y1 @ r2 - r1 @ y2) / (np.linalg.norm(y1) * np.linalg.norm(r2))
    print(f"symmetry defect {sym:.2e}, r'My {r1 @ y1:.3e}", flush=True)
    assert sym <= 1e-10
    assert r1 @ y1 > 0
    # preconditioned CG
    b = rng.normal(size=n)
    x_ref = spla.spsolve(A.tocsc(), b)
    counts = {}
    for name, M in (("none", None), ("amg", spla.LinearOperator((n, n), lambda v: amg.hypre_boomeramg_apply(v, "Kr")))):
        k = [0]
        x, flag = spla.cg(A, b, rtol=1e-8, maxiter=5000, M=M, callback=lambda _: k.__setitem__(0, k[0] + 1))
        counts[name] = k[0]
        assert flag == 0
        if name == "amg":
            assert np.linalg.norm(x - x_ref) <= 1e-6 * np.linalg.norm(x_ref)
    print("CG iterations", counts, flush=True)
    assert counts["amg"] * 5 <= counts["none"]
    # values-only update (setup_ijv on the same id takes the same path)
    t = amg.hypre_boomeramg_update_values(4.0 * V, "Kr")
    assert t >= 0
    y1b = amg.hypre_boomeramg_apply(r1, "Kr")
    assert np.linalg.norm(4.0 * y1b - y1) <= 1e-8 * np.linalg.norm(y1)
    info2 = amg.hypre_boomeramg_setup_ijv(I, J, V, n, Z, None, "Kr")
    assert info2["n"] == n
    assert np.linalg.norm(amg.hypre_boomeramg_apply(r1, "Kr") - y1) <= 1e-8 * np.linalg.norm(y1)
    # the reference's errors
    with pytest.raises(SSRError, match="does not match"):
        amg.hypre_boomeramg_apply(r1[:-1], "Kr")
    with pytest.raises(SSRError, match="unknown instance"):
        amg.hypre_boomeramg_apply(r1, "other")
    with pytest.raises(SSRError, match="values"):
        amg.hypre_boomeramg_update_values(V[:-1], "Kr")
    amg.hypre_boomeramg_clear("Kr")
    assert amg.hypre_boomeramg_info()["instances"] == 0
    with pytest.raises(SSRError, match="unknown instance"):
        amg.hypre_boomeramg_apply(r1, "Kr")
