"""The indirect SSR continuation in the library (ssr_SSR_indirect_continuation,
driver "library") against the host loop over the same native Newton solves
(driver "native"): the same reductions, solver copies, basis expansions and
step control in the same order, so the same accepted (lambda, omega) history,
the same final field and the same Newton and Krylov iteration counts, bit for
bit -- on the 2D slope (AMG and node-block Jacobi) and the 3D slope (FP64 and
FP32 Krylov) -- and the reference's error when the initial lambda is too
large."""
import numpy as np
import pytest

from slope_stability_octave_amd import problems
from slope_stability_octave_amd.continuation import ContinuationError

pytestmark = pytest.mark.gpu

GEOM_3D = (20.0, 20.0, 20.0, 10.0, 20.0, 20.0)


@pytest.mark.parametrize("case", ["2d_amg", "2d_jacobi", "3d_64", "3d_32"])
def test_library_continuation_equals_host_loop(case):
    if case.startswith("2d"):
        make = lambda: problems.slope_2d(1.0)  # noqa: E731
        prec, pc = 64, ("amg" if case == "2d_amg" else "block_jacobi")
    else:
        make = lambda: problems.slope_3d((4, 4, 4, 2, 4, 4), *GEOM_3D, problems.HOMO_2D, device=True)  # noqa: E731
        prec, pc = (32 if case == "3d_32" else 64), "amg"
    res = {d: problems.run_ssr(make(), params=dict(linear_tol=1e-2), preconditioner=pc, precision=prec, driver=d)
           for d in ("native", "library")}
    a, b = res["native"], res["library"]
    print(f"{case}: FoS {a['FoS']!r} / {b['FoS']!r}, steps {len(a['lambda_hist'])}, Newton {a['newton_iterations']} "
          f"/ {b['newton_iterations']}, Krylov {a['counts'].get('DFGMRES.solve.n_iters')} / "
          f"{b['counts'].get('DFGMRES.solve.n_iters')}", flush=True)
    assert a["FoS"] == b["FoS"]
    assert np.array_equal(np.asarray(a["lambda_hist"]), np.asarray(b["lambda_hist"]))
    assert np.array_equal(np.asarray(a["omega_hist"]), np.asarray(b["omega_hist"]))
    assert np.array_equal(np.asarray(a["Umax_hist"]), np.asarray(b["Umax_hist"]))
    assert a["newton_iterations"] == b["newton_iterations"]
    assert a["stats"]["attempt_newton_iterations"] == b["stats"]["attempt_newton_iterations"]
    assert a["stats"]["init_newton_iterations"] == b["stats"]["init_newton_iterations"]
    assert a["counts"].get("DFGMRES.solve.n_iters") == b["counts"].get("DFGMRES.solve.n_iters")
    assert np.array_equal(a["U"].get(), b["U"].get())


def test_library_continuation_reference_error():
    for d in ("native", "library"):
        with pytest.raises(ContinuationError, match="Initial choice of lambda seems to be too large"):
            problems.run_ssr(problems.slope_2d(1.0), params=dict(linear_tol=1e-2, lambda_init=4.0, it_newt_max=30),
                             preconditioner="amg", driver=d)
