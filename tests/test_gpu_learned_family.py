"""The learned + mixed-precision form on held-out members of the seeded 2D
slope family (slope angle ~ U[26, 63] deg, c0 ~ U[5, 30] kPa, phi ~ U[20, 40]
deg) at 20k-40k free DOF, against the plain GPU form (secant predictor
SSR_indirect_continuation.m:119, FP64 Krylov).

The network (3 layers, 128 units, bf16 weights, trained on the GPU by the
library's kernels) learns from plain-solver trajectories of other seeds on
coarse meshes.  What it must show on the held-out members:
* its proposed starts have a smaller Newton residual than the secant starts
  on most continuation steps (median ratio below 1: the gain over the method
  without it; where it loses, the residual judge keeps the secant start);
* the chosen form traces the plain form's lambda(omega) curve within 1e-4;
* the factor of safety within 1e-4 wherever both runs stop on the same step
  (the stop rule, SSR_indirect_continuation.m:203-206, fires on a scaled
  lambda increment that either form's Newton tolerance can move across the
  threshold; a run that stops one step earlier reports lambda one increment
  lower, and the test then requires the difference to be that increment).
Newton iterations and times are reported.
"""
import numpy as np
import pytest

from slope_stability_octave_amd import learned, problems

pytestmark = pytest.mark.gpu

HELD_OUT = [(2000, 0.5), (2001, 0.45), (2002, 0.4), (2003, 0.35)]


def curve_gap(a, b):
    oa, la = np.asarray(a["omega_hist"]), np.asarray(a["lambda_hist"])
    ob, lb = np.asarray(b["omega_hist"]), np.asarray(b["lambda_hist"])
    sel = (ob >= oa[0]) & (ob <= oa[-1])
    return float(np.abs(np.interp(ob[sel], oa, la) - lb[sel]).max())


def test_learned_fp32_on_held_out_family_members():
    log = lambda s: print(s, flush=True)  # noqa: E731
    train = learned.generate(range(8), h=1.0, linear_tol=1e-2)
    val = learned.generate(range(1000, 1002), h=1.0, linear_tol=1e-2)
    net, norm, rep = learned.fit(train, val, epochs=1500)
    log(f"trained on {rep['samples']} samples: val L_U {rep['val']['L_U']:.3f} (secant 1), "
        f"L_lambda {rep['val']['L_lambda']:.3e} (secant {rep['val']['secant_L_lambda']:.3e})")
    r_pred, r_sec, its = [], [], {"plain": 0, "chosen": 0}
    times = {"plain": 0.0, "chosen": 0.0}
    for seed, h in HELD_OUT:
        prob, mat = problems.slope_2d_family(seed, h=h)
        assert 20000 <= prob["n_dof_free"] <= 200000
        p = dict(problems.PARAMS_FAMILY, linear_tol=1e-2)
        plain = problems.run_ssr(prob, p, preconditioner="amg", precision=64)
        chosen = problems.run_ssr(prob, p, preconditioner="amg", precision=32,
                                  predictor=learned.predictor_factory(net, norm, mat))
        pr = chosen["predictor"]
        for a, b in pr.start_residuals:
            r_pred.append(a)
            r_sec.append(b)
        its["plain"] += plain["newton_iterations"]
        its["chosen"] += chosen["newton_iterations"]
        times["plain"] += plain["time_to_fos_s"]
        times["chosen"] += chosen["time_to_fos_s"]
        gap = curve_gap(plain, chosen)
        lp, lc = plain["lambda_hist"], chosen["lambda_hist"]
        log(f"seed {seed} h {h}: {prob['n_dof_free']} DOF, FoS plain {plain['FoS']:.7f} ({len(lp)} steps, "
            f"{plain['newton_iterations']} Newton, {plain['time_to_fos_s']:.2f} s), chosen {chosen['FoS']:.7f} "
            f"({len(lc)} steps, {chosen['newton_iterations']} Newton, {chosen['time_to_fos_s']:.2f} s), "
            f"predictor used {pr.used} rejected {pr.rejected}, curve gap {gap:.1e}")
        assert gap <= 1e-4
        if len(lp) == len(lc):
            assert abs(plain["FoS"] - chosen["FoS"]) <= 1e-4
        else:
            # one step apart: the longer run's last increment separates them
            longer, shorter = (lp, lc) if len(lp) > len(lc) else (lc, lp)
            assert len(longer) == len(shorter) + 1
            assert abs(longer[-2] - shorter[-1]) <= 1e-4
    r_pred, r_sec = np.asarray(r_pred), np.asarray(r_sec)
    ratio = r_pred / r_sec
    judged = np.minimum(r_pred, r_sec)
    log(f"start residual network/secant over {ratio.size} steps: median {np.median(ratio):.3f}, geometric mean "
        f"{np.exp(np.mean(np.log(ratio))):.3f}, network better on {np.mean(ratio < 1):.0%}; mean start residual "
        f"with the judge {judged.mean():.3e} vs secant {r_sec.mean():.3e}; Newton iterations plain {its['plain']}, "
        f"chosen {its['chosen']}; time plain {times['plain']:.2f} s, chosen {times['chosen']:.2f} s")
    assert ratio.size >= 8
    assert np.median(ratio) < 1.0
