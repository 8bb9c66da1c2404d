"""Write the factor-of-safety fixtures of the CPU oracle path
(oracle/newton_path.py: FP64, direct sparse solves) for the end-to-end
configurations whose oracle run takes minutes, so the GPU tests compare with
them without rerunning the oracle on the GPU machine.

    python tests/golden/make_fos_fixtures.py [name ...]

Cases (parameters of slope_stability_2D_homo_SSR.m:114-137 and
slope_stability_3D_homo_SSR.m:118-140 unless stated):
  c1_2d_h05       2D homogeneous slope, P2 h = 0.5 (20,840 unknowns),
                  lambda_init = 1.0 (the plain-method configuration);
  c3_3d_coarse    3D homogeneous slope, 60x30x20 m box with a 20 m 1:1 slope,
                  cells (4,4,4,2,4,4), lambda_init = 0.7;
  c4_3d_layers    the four-layer 3D slope, cells (2,2,2,1,2,2), lambda_init = 0.7.
Each fixture holds lambda_hist, omega_hist (accepted steps) and the run time.
"""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import newton_path as NP  # noqa: E402
from slope_stability_octave_amd import problems  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
GEOM_3D = (20.0, 20.0, 20.0, 10.0, 20.0, 20.0)


def case(name):
    if name == "c1_2d_h05":
        return problems.slope_2d(h=0.5), dict(problems.PARAMS_2D, lambda_init=1.0)
    if name == "c3_3d_coarse":
        prob = problems.slope_3d((4, 4, 4, 2, 4, 4), *GEOM_3D, problems.HOMO_2D)
        return prob, dict(problems.PARAMS_3D, lambda_init=0.7)
    if name == "c4_3d_layers":
        prob = problems.slope_3d((2, 2, 2, 1, 2, 2), *GEOM_3D, problems.HETERO_3D_LAYERS,
                                 layers=problems.HETERO_3D_LEVELS)
        return prob, dict(problems.PARAMS_3D, lambda_init=0.7)
    raise KeyError(name)


NAMES = ("c1_2d_h05", "c3_3d_coarse", "c4_3d_layers")


def run(name):
    prob, p = case(name)
    cpu = NP.ConstitutiveCPU(prob["coord"], prob["elem"], prob["Q"], prob["dim"], prob["DPhi"], prob["WEIGHT"],
                             prob["mats"])
    t0 = time.perf_counter()
    _, lam, om = NP.SSR_indirect_continuation(
        p["lambda_init"], p["d_lambda_init"], p["d_lambda_min"], p["d_lambda_diff_scaled_min"], p["step_max"],
        p["omega_max_stop"], p["it_newt_max"], p["it_damp_max"], p["tol"], p["r_min"], prob["f"], cpu,
        log=lambda s: print(f"[{name}] {s}", flush=True))
    el = time.perf_counter() - t0
    out = dict(name=name, free_dof=int(prob["n_dof_free"]), params=p, lambda_hist=[float(x) for x in lam],
               omega_hist=[float(x) for x in om], FoS=float(lam[-1]), oracle_seconds=el)
    with open(os.path.join(OUT, f"fos_{name}.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    print(f"[{name}] FoS {lam[-1]:.8f}, {len(lam)} steps, {el:.1f} s", flush=True)


if __name__ == "__main__":
    for n in sys.argv[1:] or NAMES:
        run(n)
