"""The AMG hierarchy built on the device against the host build.

Both builds run the same algorithm on the same operator with the same
aggregation function (strength from node-block norms, greedy passes), so the
hierarchies differ only by the order of floating-point sums.  Checks:
* one V-cycle from each hierarchy agrees to 1e-11 relative in double (1e-5
  in float) -- different aggregates or transfers would differ at O(1);
* a preconditioned solve takes the same number of iterations;
* two device builds give the same bits (every output block is summed by one
  lane in a fixed order);
on the 3D and 2D slopes, several levels deep (max_coarse 50).
"""
import numpy as np
import pytest

from slope_stability_octave_amd import _lib, problems
from slope_stability_octave_amd._lib import DeviceArray
from slope_stability_octave_amd.linear_solvers import DFGMRES

pytestmark = pytest.mark.gpu


def _operator(dim):
    if dim == 3:
        prob = problems.slope_3d((3, 3, 3, 3, 3, 3), 3.0, 3.0, 3.0, 3.0, 3.0, 3.0, problems.HOMO_2D)
    else:
        prob = problems.slope_2d(h=0.5)
    cmb = problems.make_constitutive(prob)
    cmb.reduction(1.0)
    U = DeviceArray((cmb.n_n, cmb.dim), zero=True)
    cmb.build_F_and_DS_all(1.0, U)
    return prob, cmb, cmb.build_K_r_QQ_vals(1e-4)


@pytest.mark.parametrize("prec,dim", [(64, 3), (32, 3), (64, 2)])
def test_device_hierarchy_matches_host(prec, dim):
    prob, cmb, K = _operator(dim)
    rng = np.random.default_rng(5)
    x = DeviceArray.from_host(rng.normal(size=(cmb.n_n, cmb.dim)) * cmb.Q.T)
    old = _lib.get_tuning(_lib.TUNE_AMG_DEVICE_SETUP)
    out, its = {}, {}
    try:
        for mode in (1, 0, 1):
            _lib.set_tuning(_lib.TUNE_AMG_DEVICE_SETUP, mode)
            s = DFGMRES(cmb.pattern, 1e-8, 200, 1e-3, "amg", amg_options=dict(coord=prob["coord"], max_coarse=50),
                        precision=prec)
            s.setup_preconditioner(K)
            y = DeviceArray((cmb.n_n, cmb.dim))
            if prec == 64:
                _lib.call("ssr_solver_precond_apply", s.handle, _lib.ptr(x), _lib.ptr(y))
                val = y.get()
            else:
                xs, _ = s.solve(K, x)
                val = xs.get()
            xs, it = s.solve(K, x)
            out.setdefault(mode, []).append(val)
            its.setdefault(mode, []).append(it)
    finally:
        _lib.set_tuning(_lib.TUNE_AMG_DEVICE_SETUP, old)
    a, b = out[1][0], out[0][0]
    rel = np.abs(a - b).max() / np.abs(b).max()
    print(f"prec {prec} dim {dim}: device vs host hierarchy, max rel {rel:.2e}, iterations {its}", flush=True)
    assert rel <= (1e-11 if prec == 64 else 1e-5)
    assert its[1][0] == its[0][0]
    assert np.array_equal(out[1][0], out[1][1]) and its[1][0] == its[1][1]


def test_device_hierarchy_in_a_continuation():
    """2D slope at h = 1 solved to its FoS with the device-built hierarchy
    and with the host-built one: the same continuation to 1e-8."""
    old = _lib.get_tuning(_lib.TUNE_AMG_DEVICE_SETUP)
    res = {}
    try:
        for mode in (1, 0):
            _lib.set_tuning(_lib.TUNE_AMG_DEVICE_SETUP, mode)
            res[mode] = problems.run_ssr(problems.slope_2d(1.0), params=dict(linear_tol=1e-10), preconditioner="amg")
    finally:
        _lib.set_tuning(_lib.TUNE_AMG_DEVICE_SETUP, old)
    assert abs(res[1]["FoS"] - res[0]["FoS"]) <= 1e-8
    n = min(len(res[1]["lambda_hist"]), len(res[0]["lambda_hist"]))
    assert np.abs(np.asarray(res[1]["lambda_hist"][:n]) - np.asarray(res[0]["lambda_hist"][:n])).max() <= 1e-8
