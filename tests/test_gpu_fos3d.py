"""3D homogeneous slope, indirect SSR continuation on the reference's own mesh
(SSR_homo_ada_L1: 25591 nodes, 16615 P2 tetrahedra), against the run recorded
with the reference (slope_stability/slope_stability_3D_homo_SSR.ipynb:366-410):
71947 unknowns, K(Q,Q) nnz 5536583, omega_init 2201.8807 at lambda 0.9,
omega2 2204.5328 at lambda 1, and lambda after every accepted step up to
omega = 3000 (lambda 1.24087).

The GPU path uses its smoothed-aggregation preconditioner instead of
BoomerAMG and a Krylov tolerance of 1e-2 instead of 1e-1; lambda at a given
omega is a property of the converged Newton solution (tolerance 1e-4), so the
comparison holds at 1e-4 per step.
"""
import os

import numpy as np
import pytest

from slope_stability_octave_amd import problems

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MESH = os.path.join(ROOT, "tests", "golden", "mesh_SSR_homo_ada_L1.npz")

# slope_stability_3D_homo_SSR.ipynb:375,378 (omega1, omega2) and :381-408 (accepted steps)
REF_OMEGA12 = (2201.8807, 2204.5328)
REF_STEPS = [(2207.18, 1.05277), (2212.49, 1.12594), (2217.79, 1.17816), (2228.4, 1.19948), (2249.62, 1.2106),
             (2292.05, 1.22001), (2376.92, 1.2281), (2546.65, 1.23479), (2886.12, 1.23989), (3000.0, 1.24087)]


def _mesh():
    with np.load(MESH) as z:
        return dict(z)


def test_mesh_loader_matches_recorded_sizes():
    prob = problems.slope_3d_mesh(_mesh())
    # slope_stability_3D_homo_SSR.ipynb: "number of nodes =25591 number of unknowns =71947
    # number of elements =16615 number of integration points =182765"
    assert prob["coord"].shape == (3, 25591)
    assert prob["elem"].shape == (10, 16615)
    assert prob["n_dof_free"] == 71947
    assert prob["n_int"] == 182765
    # the body is 3000 m^3 (sum of quadrature weights) under gamma = 20 kN/m^3
    assert abs(prob["WEIGHT"].sum() - 3000.0) < 1e-8
    assert abs(prob["f"][1].sum() + 20.0 * 3000.0) < 1e-6
    assert abs(prob["f"][0].sum()) < 1e-8 and abs(prob["f"][2].sum()) < 1e-8


@pytest.mark.gpu
def test_fos3d_recorded_trajectory():
    prob = problems.slope_3d_mesh(_mesh())
    p = dict(problems.PARAMS_3D_NOTEBOOK)
    p["linear_tol"] = 1e-2
    res = problems.run_ssr(prob, p, preconditioner="amg")
    om, lam = res["omega_hist"], res["lambda_hist"]
    print("omega", list(om), "lambda", list(lam), flush=True)
    assert abs(om[0] - REF_OMEGA12[0]) <= 1e-4 * REF_OMEGA12[0]
    assert abs(om[1] - REF_OMEGA12[1]) <= 1e-4 * REF_OMEGA12[1]
    assert len(lam) == 2 + len(REF_STEPS)
    for (o_ref, l_ref), o, l in zip(REF_STEPS, om[2:], lam[2:]):
        assert abs(o - o_ref) <= 1e-4 * o_ref
        assert abs(l - l_ref) <= 1e-4
    assert abs(res["FoS"] - 1.24087) <= 1e-4
