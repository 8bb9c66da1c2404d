"""The autotuner times every 3D kernel form and SpMV shape on the caller's
data, installs the fastest, caches the choice, and every form it may pick
gives the same S and DS bit for bit (they share the point code)."""
import json

import numpy as np
import pytest

from oracle import oracle as O
from slope_stability_octave_amd import _lib, problems
from slope_stability_octave_amd import constitutive_mex as CM
from slope_stability_octave_amd import workloads as W
from slope_stability_octave_amd._lib import DeviceArray
from slope_stability_octave_amd.autotune import CONSTITUTIVE_CANDIDATES, Autotuner

pytestmark = pytest.mark.gpu


def test_constitutive_forms_identical_and_tuner_picks_one(tmp_path):
    n = 200_003
    d = W.config2_points(n)
    cb, sp = O.reduction(d["c0"], d["phi"], d["psi"], d["lam"], "B")
    args = [DeviceArray.from_host(x) for x in (np.ascontiguousarray(d["E"]), cb, sp, d["shear"], d["bulk"],
                                                d["lame"])]
    S, DS = DeviceArray((n, 6)), DeviceArray((n, 36))
    old = _lib.get_tuning(_lib.TUNE_CONSTITUTIVE_VARIANT)
    outs = []
    try:
        for v in CONSTITUTIVE_CANDIDATES:
            _lib.set_tuning(_lib.TUNE_CONSTITUTIVE_VARIANT, v)
            CM.constitutive_dev(3, n, *args, S, DS)
            outs.append((S.get(), DS.get()))
        for s_, d_ in outs[1:]:
            assert np.array_equal(s_, outs[0][0]) and np.array_equal(d_, outs[0][1])
        cache = tmp_path / "tune.json"
        rec = Autotuner(str(cache)).constitutive(n, *args, S, DS, reps=3)
        assert rec["best"] in CONSTITUTIVE_CANDIDATES
        assert _lib.get_tuning(_lib.TUNE_CONSTITUTIVE_VARIANT) == rec["best"]
        assert json.loads(cache.read_text())[f"constitutive3d:{n}"]["best"] == rec["best"]
    finally:
        _lib.set_tuning(_lib.TUNE_CONSTITUTIVE_VARIANT, old)


def test_spmv_tuner():
    prob = problems.slope_3d((2, 2, 2, 2, 2, 2), 2.0, 2.0, 2.0, 2.0, 2.0, 2.0, problems.HOMO_2D)
    cmb = problems.make_constitutive(prob)
    old = _lib.get_tuning(_lib.TUNE_SPMV_GROUP)
    try:
        rec = Autotuner().spmv(cmb.pattern, cmb.K_elast, reps=5)
        assert rec["best"] in (4, 8, 16, 32, 64)
        assert _lib.get_tuning(_lib.TUNE_SPMV_GROUP) == rec["best"]
    finally:
        _lib.set_tuning(_lib.TUNE_SPMV_GROUP, old)
