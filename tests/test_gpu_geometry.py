"""Strain and internal force with each point's gradients computed from the
node coordinates (ssr_dev_strain_geo, ssr_dev_force_geo) against the same
operators reading the stored DPhi of the device setup.

The computed gradients go through the arithmetic that produced the stored
ones, so the bar is bit-equality: E = B U and F = B'(w S) on random fields,
both element orders of the force (ordered per-node sums, atomics), 2D and 3D,
and a continuation whose Newton path reads no DPhi reaching the same lambda
history, bit for bit, as the one that reads it.  The stored-DPhi operators are
themselves pinned against the CPU oracle (test_gpu_assembly, test_gpu_newton)."""
import numpy as np
import pytest

from slope_stability_octave_amd import _lib, problems
from slope_stability_octave_amd._lib import DeviceArray

pytestmark = pytest.mark.gpu


def _problems():
    return {
        "2d": problems.slope_2d(1.0, device=True),
        "3d": problems.slope_3d((3, 2, 3, 2, 2, 2), 20.0, 20.0, 20.0, 10.0, 20.0, 20.0, problems.HETERO_3D_LAYERS,
                                layers=problems.HETERO_3D_LEVELS, device=True),
    }


@pytest.fixture(scope="module")
def probs():
    return _problems()


@pytest.mark.parametrize("case", ["2d", "3d"])
def test_strain_from_coordinates_equals_stored_gradients(probs, case):
    prob = probs[case]
    cmb = problems.make_constitutive(dict(prob, stored_gradients=True))
    P, dim, n_int, ns = cmb.pattern, cmb.dim, cmb.n_int, cmb.n_strain
    print(f"{case}: {cmb.n_n} nodes, {n_int} points", flush=True)
    rng = np.random.default_rng(7)
    U = DeviceArray.from_host(rng.standard_normal((cmb.n_n, dim)))
    E0, E1 = DeviceArray((n_int, ns)), DeviceArray((n_int, ns))
    P.strain(cmb.n_q, cmb.DPhi, U, E0)
    P.strain_geo(cmb.n_q, prob["coord_dev"], U, E1)
    a, b = E0.get(), E1.get()
    assert np.isfinite(a).all() and np.abs(a).max() > 0
    assert np.array_equal(a, b)


@pytest.mark.parametrize("form", [0, 1])
@pytest.mark.parametrize("case", ["2d", "3d"])
def test_force_from_coordinates_equals_stored_gradients(probs, case, form):
    """form 0: element-node forces and ordered per-node sums (bit for bit);
    form 1: atomics (the order of the additions varies: 1e-13)."""
    prob = probs[case]
    cmb = problems.make_constitutive(dict(prob, stored_gradients=True))
    P, dim, n_int = cmb.pattern, cmb.dim, cmb.n_int
    rng = np.random.default_rng(11)
    S = DeviceArray.from_host(rng.standard_normal((n_int, cmb.s_rows)))
    F0, F1 = DeviceArray((cmb.n_n, dim)), DeviceArray((cmb.n_n, dim))
    old = _lib.get_tuning(_lib.TUNE_ASSEMBLY_FORM)
    try:
        _lib.set_tuning(_lib.TUNE_ASSEMBLY_FORM, form)
        P.force(cmb.n_q, cmb.DPhi, S, cmb.s_rows, cmb.WEIGHT, F0)
        P.force_geo(cmb.n_q, prob["coord_dev"], S, cmb.s_rows, cmb.WEIGHT, F1)
    finally:
        _lib.set_tuning(_lib.TUNE_ASSEMBLY_FORM, old)
    a, b = F0.get(), F1.get()
    assert np.abs(a).max() > 0
    if form == 0:
        assert np.array_equal(a, b)
    else:
        assert np.abs(a - b).max() <= 1e-13 * np.abs(a).max()


def test_continuation_without_stored_gradients_is_bitwise(probs):
    """The native Newton path with coordinates (no DPhi read) and with the
    stored DPhi: the same lambda history, Newton and Krylov counts."""
    prob = probs["3d"]
    p = dict(problems.PARAMS_3D, lambda_init=0.7)
    res = {}
    for stored in (True, False):
        print(f"3D layered slope, stored_gradients={stored}", flush=True)
        res[stored] = problems.run_ssr(dict(prob, stored_gradients=stored), p, preconditioner="amg")
    a, b = res[True], res[False]
    print(a["FoS"], b["FoS"], flush=True)
    assert a["FoS"] == b["FoS"]
    assert np.array_equal(np.asarray(a["lambda_hist"]), np.asarray(b["lambda_hist"]))
    assert a["newton_iterations"] == b["newton_iterations"]
    assert a["counts"]["DFGMRES.solve.n_iters"] == b["counts"]["DFGMRES.solve.n_iters"]
