"""2D homogeneous slope, h = 1/4 (83280 unknowns), indirect SSR continuation
on the GPU against the run recorded with the reference
(slope_stability/slope_stability_2D_homo_SSR.ipynb:434-472): omega_init
425.2321 at lambda 0.9, omega2 425.352 at lambda 0.95, the (omega, lambda)
of every accepted step and the final factor of safety 1.16411213 (:615).

The GPU path uses its smoothed-aggregation preconditioner instead of
BoomerAMG and a Krylov tolerance of 1e-2 instead of 1e-1; lambda at a given
omega is a property of the converged Newton state (tolerance 1e-4), so each
accepted lambda is compared at 1e-4 and each omega at 1e-5 relative (the
recorded omegas carry six significant digits).
"""
import numpy as np
import pytest

from slope_stability_octave_amd import problems

pytestmark = pytest.mark.gpu

REF_OMEGA12 = (425.2321, 425.352)
# "accepted: lambda=..., omega=..." of steps 3..13
REF_STEPS = [(425.472, 0.991505), (425.712, 1.05786), (425.951, 1.10866), (426.431, 1.15644), (426.91, 1.15846),
             (427.869, 1.15996), (429.787, 1.16124), (433.623, 1.16228), (441.295, 1.16307), (456.638, 1.16366),
             (487.325, 1.16411)]
REF_FOS = 1.16411213


@pytest.mark.parametrize("precision", [64, 32])
def test_fos2d_quarter_recorded_trajectory(precision):
    prob = problems.slope_2d(h=0.25)
    assert prob["n_dof_free"] == 83280
    p = dict(problems.PARAMS_2D)  # the notebook's parameters (lambda_init 0.9)
    p["linear_tol"] = 1e-2
    print(f"2D h=1/4 indirect SSR, precision {precision}", flush=True)
    res = problems.run_ssr(prob, p, preconditioner="amg", precision=precision)
    om, lam = np.asarray(res["omega_hist"]), np.asarray(res["lambda_hist"])
    print("omega", list(om), "lambda", list(lam), "FoS", res["FoS"], "time", res["time_to_fos_s"], flush=True)
    assert abs(om[0] - REF_OMEGA12[0]) <= 1e-5 * REF_OMEGA12[0]
    assert abs(om[1] - REF_OMEGA12[1]) <= 1e-5 * REF_OMEGA12[1]
    assert lam[0] == 0.9 and abs(lam[1] - 0.95) < 1e-12
    assert len(lam) == 2 + len(REF_STEPS)
    for (o_ref, l_ref), o, l in zip(REF_STEPS, om[2:], lam[2:]):
        assert abs(o - o_ref) <= 1e-5 * o_ref, (o, o_ref)
        assert abs(l - l_ref) <= 1e-4, (l, l_ref)
    assert abs(res["FoS"] - REF_FOS) <= 1e-4
