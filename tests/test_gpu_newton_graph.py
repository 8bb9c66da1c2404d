"""Element phases of the Newton loops replayed from captured HIP graphs
(tuning key 16).

Each build_F (strain, stress and tangent, internal force) and each K_r
assembly of a Newton solve runs eagerly once, is captured on its second use
and replayed afterwards; the kernels and their arguments are those of the
eager launches, so whole continuations must give the eager form's histories
bit for bit: the indirect SSR continuation in 2D and 3D (device setup with
coordinate gradients, and host setup with stored gradients; FP64 and FP32),
the direct continuation (plain Newton) and the LL continuation.
"""
import numpy as np
import pytest

from slope_stability_octave_amd import _lib, problems

pytestmark = pytest.mark.gpu

GEOM_3D = (20.0, 20.0, 20.0, 10.0, 20.0, 20.0)


def _both(run):
    old = _lib.get_tuning(_lib.TUNE_NEWTON_GRAPH)
    out = {}
    try:
        for mode in (0, 1):
            _lib.set_tuning(_lib.TUNE_NEWTON_GRAPH, mode)
            out[mode] = run()
    finally:
        _lib.set_tuning(_lib.TUNE_NEWTON_GRAPH, old)
    return out


@pytest.mark.parametrize("case", ["2d", "3d_device_64", "3d_device_32", "3d_host_64"])
def test_indirect_continuation_graph_replay_same_bits(case):
    if case == "2d":
        make, prec = (lambda: problems.slope_2d(1.0)), 64
    else:
        dev = "device" in case
        make = lambda: problems.slope_3d((4, 4, 4, 2, 4, 4), *GEOM_3D, problems.HOMO_2D, device=dev)  # noqa: E731
        prec = 32 if case.endswith("32") else 64
    out = _both(lambda: problems.run_ssr(make(), params=dict(linear_tol=1e-2), preconditioner="amg",
                                         precision=prec))
    a, b = out[0], out[1]
    print(f"{case}: FoS {a['FoS']!r} / {b['FoS']!r}, steps {len(a['lambda_hist'])}, "
          f"Newton iterations {a['newton_iterations']}", flush=True)
    assert a["FoS"] == b["FoS"] and a["newton_iterations"] == b["newton_iterations"]
    assert np.array_equal(np.asarray(a["lambda_hist"]), np.asarray(b["lambda_hist"]))
    assert np.array_equal(np.asarray(a["omega_hist"]), np.asarray(b["omega_hist"]))
    assert np.array_equal(a["U"].get(), b["U"].get())


def test_direct_continuation_graph_replay_same_bits():
    out = _both(lambda: problems.run_ssr_direct(problems.slope_2d(1.0), params=dict(linear_tol=1e-2)))
    a, b = out[0], out[1]
    assert np.array_equal(np.asarray(a["lambda_hist"]), np.asarray(b["lambda_hist"]))
    assert np.array_equal(np.asarray(a["omega_hist"]), np.asarray(b["omega_hist"]))


def test_ll_continuation_graph_replay_same_bits():
    out = _both(lambda: problems.run_ll(problems.slope_2d(1.0), dict(linear_tol=1e-2), preconditioner="amg"))
    a, b = out[0], out[1]
    assert a["limit_load"] == b["limit_load"]
    assert np.array_equal(np.asarray(a["t_hist"]), np.asarray(b["t_hist"]))
    assert np.array_equal(np.asarray(a["omega_hist"]), np.asarray(b["omega_hist"]))
