"""CONSTITUTIVE on the GPU vs its CPU restatement, and the SSR factor of
safety end to end.

Tolerances: internal forces within 1e-12 of the largest entry; K_r values
within 1e-11 relative (normwise); factor of safety |dFoS| <= 1e-4 against the
CPU oracle path (FP64, direct linear solves) on the same mesh.
"""
import numpy as np
import pytest

from oracle import newton_path as NP
from slope_stability_octave_amd import problems
from slope_stability_octave_amd._lib import DeviceArray

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def coarse2d():
    prob = problems.slope_2d(h=2.0)
    cpu = NP.ConstitutiveCPU(prob["coord"], prob["elem"], prob["Q"], 2, prob["DPhi"], prob["WEIGHT"], prob["mats"])
    return prob, cpu


def test_build_F_and_K_r_match_cpu(coarse2d):
    prob, cpu = coarse2d
    cmb = problems.make_constitutive(prob)
    rng = np.random.default_rng(0)
    U = rng.normal(0, 5e-3, size=(2, prob["coord"].shape[1])) * prob["Q"]
    dU = DeviceArray.from_host(U.T)
    for lam in (0.8, 1.3):
        F = cmb.build_F_and_DS_all(lam, dU).get().T
        F0 = cpu.build_F_and_DS_all(lam, U)
        assert np.abs(F - F0).max() <= 1e-12 * np.abs(F0).max()
        Fs = cmb.build_F_all(lam + 1e-7, dU).get().T
        Fs0 = cpu.build_F_all(lam + 1e-7, U)
        assert np.abs(Fs - Fs0).max() <= 1e-12 * np.abs(Fs0).max()
        for r in (1e-4, 0.3):
            V = cmb.build_K_r_QQ_vals_ref(r)
            V0 = cpu.build_K_r(r).tocoo()
            Vref = cpu.R.build_K_r(r, __import__("oracle").oracle.assemble(
                2, prob["DPhi"], cpu.DS, prob["WEIGHT"], cpu.smap, prob["n_q"], cpu.R.nnz))
            assert np.abs(V - Vref).max() <= 1e-11 * np.abs(Vref).max()
            assert V0.nnz == V.size


def test_fos_matches_cpu_oracle_coarse_2d(coarse2d):
    """Same continuation parameters on both sides; the GPU Krylov tolerance is
    tightened so both paths solve the Newton systems to the same accuracy
    (with the reference's 1e-1 the continuation may take one more or one
    fewer step, which moves lambda by the last increment)."""
    prob, cpu = coarse2d
    p = dict(problems.PARAMS_2D)
    p["lambda_init"] = 1.0
    p["linear_tol"] = 1e-10
    p["linear_maxit"] = 1500
    res = problems.run_ssr(prob, p)
    U0, lh0, oh0 = NP.SSR_indirect_continuation(
        1.0, p["d_lambda_init"], p["d_lambda_min"], p["d_lambda_diff_scaled_min"], p["step_max"],
        p["omega_max_stop"], p["it_newt_max"], p["it_damp_max"], p["tol"], p["r_min"], prob["f"], cpu)
    print("GPU FoS", res["FoS"], "CPU FoS", lh0[-1], "steps", len(res["lambda_hist"]), len(lh0))
    assert abs(res["FoS"] - lh0[-1]) <= 1e-4
    # the lambda-omega curve is monotone (slope_stability_2D_homo_SSR.m:229-234)
    assert np.all(np.diff(res["lambda_hist"]) >= -1e-10)
    assert np.all(np.diff(res["omega_hist"]) >= -1e-10)


def test_fos_heterogeneous_3d_matches_cpu_oracle():
    """The four-layer 3D slope (per-point materials through reduction, return
    mapping and assembly) on a coarse mesh: GPU continuation, with one AMG
    hierarchy shared by the steps, against the CPU oracle path."""
    prob = problems.slope_3d((2, 2, 2, 1, 2, 2), 20.0, 20.0, 20.0, 10.0, 20.0, 20.0, problems.HETERO_3D_LAYERS,
                             layers=problems.HETERO_3D_LEVELS)
    assert len(np.unique(prob["mats"]["c0"])) == 4
    cpu = NP.ConstitutiveCPU(prob["coord"], prob["elem"], prob["Q"], 3, prob["DPhi"], prob["WEIGHT"], prob["mats"])
    p = dict(problems.PARAMS_3D)
    p["lambda_init"] = 0.7
    p["linear_tol"] = 1e-10
    p["linear_maxit"] = 1500
    res = problems.run_ssr(prob, p, preconditioner="amg")
    U0, lh0, oh0 = NP.SSR_indirect_continuation(
        0.7, p["d_lambda_init"], p["d_lambda_min"], p["d_lambda_diff_scaled_min"], p["step_max"],
        p["omega_max_stop"], p["it_newt_max"], p["it_damp_max"], p["tol"], p["r_min"], prob["f"], cpu)
    print("GPU FoS", res["FoS"], "CPU FoS", lh0[-1], "steps", len(res["lambda_hist"]), len(lh0), flush=True)
    assert abs(res["FoS"] - lh0[-1]) <= 1e-4


@pytest.mark.parametrize("solver,precond,precision", [("dfgmres", "block_jacobi", 64), ("dcg", "block_jacobi", 64),
                                                      ("dfgmres", "amg", 32)])
def test_native_newton_driver_matches_python_driver(coarse2d, solver, precond, precision):
    """ssr_newton_ind_SSR (the Newton loop inside the library) against the
    Python loop over the same C entry points: the same kernels run in the same
    order, so every accepted (omega, lambda), the Newton and Krylov iteration
    counts and the final field agree to rounding (in practice exactly)."""
    prob, _ = coarse2d
    p = dict(problems.PARAMS_2D)
    p["lambda_init"] = 1.0
    runs = {}
    for driver in ("python", "native"):
        print(f"driver={driver} solver={solver} precond={precond} precision={precision}", flush=True)
        runs[driver] = problems.run_ssr(prob, p, solver=solver, preconditioner=precond, precision=precision,
                                        driver=driver)
    a, b = runs["python"], runs["native"]
    print("FoS", a["FoS"], b["FoS"], "Newton", a["newton_iterations"], b["newton_iterations"], flush=True)
    assert len(a["lambda_hist"]) == len(b["lambda_hist"])
    assert np.abs(np.asarray(a["lambda_hist"]) - b["lambda_hist"]).max() <= 1e-12
    assert np.abs(np.asarray(a["omega_hist"]) - b["omega_hist"]).max() <= 1e-12 * np.abs(a["omega_hist"]).max()
    assert a["newton_iterations"] == b["newton_iterations"]
    assert a["stats"]["init_newton_iterations"] == b["stats"]["init_newton_iterations"]
    assert a["counts"]["DFGMRES.solve.n_iters"] == b["counts"]["DFGMRES.solve.n_iters"]
    assert a["counts"]["DFGMRES.deflation_size.max"] == b["counts"]["DFGMRES.deflation_size.max"]
    Ua, Ub = a["U"].get(), b["U"].get()
    assert np.abs(Ua - Ub).max() <= 1e-12 * np.abs(Ua).max()
    for ha, hb in zip(a["stats"]["newton_histories"], b["stats"]["newton_histories"]):
        assert ha["status"] == hb["status"]
        assert ha["linear_iters_W"] == hb["linear_iters_W"] and ha["linear_iters_V"] == hb["linear_iters_V"]
        assert ha["alpha"] == hb["alpha"] and ha["r"] == hb["r"]
        assert np.allclose(ha["residual"], hb["residual"], rtol=1e-12, atol=0)


def test_native_newton_driver_heterogeneous_3d():
    """The four-layer 3D slope (per-point materials) with the shared AMG
    hierarchy: both Newton drivers give the same factor of safety."""
    prob = problems.slope_3d((2, 2, 2, 1, 2, 2), 20.0, 20.0, 20.0, 10.0, 20.0, 20.0, problems.HETERO_3D_LAYERS,
                             layers=problems.HETERO_3D_LEVELS)
    p = dict(problems.PARAMS_3D)
    p["lambda_init"] = 0.7
    fos = {}
    for driver in ("python", "native"):
        print(f"3D heterogeneous, driver={driver}", flush=True)
        res = problems.run_ssr(prob, p, preconditioner="amg", driver=driver)
        fos[driver] = (res["FoS"], res["newton_iterations"], len(res["lambda_hist"]))
    print(fos, flush=True)
    assert abs(fos["python"][0] - fos["native"][0]) <= 1e-12
    assert fos["python"][1:] == fos["native"][1:]
