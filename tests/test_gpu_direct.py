"""Direct SSR continuation and the potential on the GPU.

* Potential per point vs the oracle restatement of potential_{2D,3D}.m
  (itself pinned by S = dPsi/dE against the reference-pinned return mapping):
  1e-12 of the sample's largest |Psi|.
* Direct continuation on the 2D homogeneous slope at h = 1/4 against the run
  recorded with the reference (slope_stability/slope_stability_2D_homo_SSR.ipynb:
  "lambda1 = 0.9, lambda2 = 0.95, omega1 = 0.58214, omega2 = 0.71547" and the
  accepted steps lambda = 1, 1.05, 1.1, 1.15 with omega = 0.879293, 1.08038,
  1.33543, 1.70916): omega at fixed lambda is a property of the converged
  solutions, compared at 1e-4 relative (printed precision allowing).
"""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import potential_oracle as PO
from slope_stability_octave_amd import constitutive_mex as CM
from slope_stability_octave_amd import problems
from slope_stability_octave_amd import workloads as W
from slope_stability_octave_amd._lib import DeviceArray

pytestmark = pytest.mark.gpu

REF_DIRECT = [(0.9, 0.58214), (0.95, 0.71547), (1.0, 0.879293), (1.05, 1.08038), (1.1, 1.33543), (1.15, 1.70916)]


@pytest.mark.parametrize("dim", [3, 2])
def test_potential_matches_oracle(dim):
    n = 50_001
    d = W.config2_points(n, dim=dim)
    cb, sp = O.reduction(d["c0"], d["phi"], d["psi"], d["lam"], "B")
    E = np.ascontiguousarray(d["E"])
    args = [DeviceArray.from_host(x) for x in (E, cb, sp, d["shear"], d["bulk"], d["lame"])]
    out = DeviceArray(n)
    CM.potential_dev(dim, n, *args, out)
    fn = PO.potential_3d if dim == 3 else PO.potential_2d
    ref, br = fn(E.T, cb, sp, d["shear"], d["bulk"], d["lame"])
    got = out.get()
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    assert len(set(br.tolist())) == 5


def test_direct_continuation_recorded_omegas():
    prob = problems.slope_2d(h=0.25)
    p = dict(problems.PARAMS_2D, linear_tol=1e-3, linear_maxit=300, step_max=6)
    res = problems.run_ssr_direct(prob, p, preconditioner="amg")
    lam, om = res["lambda_hist"], res["omega_hist"]
    print("lambda", list(lam), "omega", list(om), flush=True)
    assert len(lam) == len(REF_DIRECT)
    for (l_ref, o_ref), l, o in zip(REF_DIRECT, lam, om):
        assert abs(l - l_ref) < 1e-12
        assert abs(o - o_ref) <= 1e-4 * o_ref + 5e-6


def test_direct_continuation_native_driver_matches_python_driver():
    """The direct continuation's Newton solves (newton.m at each lambda and at
    lambda - eps) through ssr_newton against the Python loop: same omegas."""
    prob = problems.slope_2d(h=1.0)
    p = dict(problems.PARAMS_2D, linear_tol=1e-3, linear_maxit=300, step_max=5)
    runs = {}
    for driver in ("python", "native"):
        print(f"direct driver={driver}", flush=True)
        runs[driver] = problems.run_ssr_direct(prob, p, preconditioner="amg", driver=driver)
    a, b = runs["python"], runs["native"]
    print("lambda", list(a["lambda_hist"]), "omega", list(a["omega_hist"]), list(b["omega_hist"]), flush=True)
    assert len(a["lambda_hist"]) == len(b["lambda_hist"]) >= 3
    assert np.abs(np.asarray(a["lambda_hist"]) - b["lambda_hist"]).max() <= 1e-12
    assert np.abs(np.asarray(a["omega_hist"]) - b["omega_hist"]).max() <= 1e-12 * np.abs(a["omega_hist"]).max()
