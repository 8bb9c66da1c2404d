"""Autotuner for the launch-shape knobs of the hot kernels.

Times each candidate on the caller's own data shape with HIP events on the
library stream and installs the fastest through ssr_set_tuning:
  * the 3D return-mapping kernel form (TUNE_CONSTITUTIVE_VARIANT):
      2  packed 21-entry LDS staging, non-temporal stores
      9  36-wide LDS staging of DS, 32-point phases
      10 as 9 with the inputs read straight into registers
      11 as 10 with 16-point phases (smallest LDS footprint)
      14 as 11 with each wavefront's 16-point phases spread over the array
         (neighbouring wavefronts write neighbouring spans), non-temporal
      15 as 14 with ordinary stores
  * the lanes per block row of the node-block SpMV (TUNE_SPMV_GROUP: 4, 8, 16, 32, or 64 for one
    wavefront per row with lanes over the values).
Results are kept per (kind, shape) key in a JSON cache file when a path is
given, so a production run tunes once per problem size.
"""
from __future__ import annotations

import json
import os

import numpy as np

from . import _lib
from ._lib import DeviceArray, EventTimer

CONSTITUTIVE_CANDIDATES = (2, 9, 10, 11, 14, 15)
SPMV_CANDIDATES = (4, 8, 16, 32, 64)


def _time(fn, reps):
    fn()
    _lib.synchronize()
    t = EventTimer()
    t.start()
    for _ in range(reps):
        fn()
    t.stop()
    return t.elapsed_ms() / reps


class Autotuner:
    def __init__(self, cache_path=None):
        self.cache_path = cache_path
        self.cache = {}
        if cache_path and os.path.exists(cache_path):
            with open(cache_path) as fh:
                self.cache = json.load(fh)

    def _store(self, key, value):
        self.cache[key] = value
        if self.cache_path:
            tmp = self.cache_path + ".tmp"
            with open(tmp, "w") as fh:
                json.dump(self.cache, fh, indent=1)
            os.replace(tmp, self.cache_path)

    def constitutive(self, n, E, c_bar, sin_phi, shear, bulk, lame, S, DS, reps=10,
                     candidates=CONSTITUTIVE_CANDIDATES):
        """Pick the 3D S+DS kernel form for n points (device arrays)."""
        from .constitutive_mex import constitutive_dev
        key = f"constitutive3d:{n}"
        if key not in self.cache:
            old = _lib.get_tuning(_lib.TUNE_CONSTITUTIVE_VARIANT)
            times = {}
            try:
                for v in candidates:
                    _lib.set_tuning(_lib.TUNE_CONSTITUTIVE_VARIANT, v)
                    times[str(v)] = _time(lambda: constitutive_dev(3, n, E, c_bar, sin_phi, shear, bulk, lame, S, DS),
                                          reps)
            finally:
                _lib.set_tuning(_lib.TUNE_CONSTITUTIVE_VARIANT, old)
            best = min(times, key=times.get)
            self._store(key, dict(best=int(best), ms=times))
        rec = self.cache[key]
        _lib.set_tuning(_lib.TUNE_CONSTITUTIVE_VARIANT, rec["best"])
        return rec

    def spmv(self, pattern, K, reps=20, candidates=SPMV_CANDIDATES):
        """Pick the SpMV lanes-per-row for this operator."""
        from .linear_solvers import spmv
        key = f"spmv{pattern.dim}d:{pattern.n_n}:{pattern.nnzb}"
        if key not in self.cache:
            x = DeviceArray.from_host(np.random.default_rng(0).normal(size=(pattern.n_n, pattern.dim)))
            y = DeviceArray((pattern.n_n, pattern.dim))
            old = _lib.get_tuning(_lib.TUNE_SPMV_GROUP)
            times = {}
            try:
                for g in candidates:
                    _lib.set_tuning(_lib.TUNE_SPMV_GROUP, g)
                    times[str(g)] = _time(lambda: spmv(pattern, K, x, y), reps)
            finally:
                _lib.set_tuning(_lib.TUNE_SPMV_GROUP, old)
            best = min(times, key=times.get)
            self._store(key, dict(best=int(best), ms=times))
        rec = self.cache[key]
        _lib.set_tuning(_lib.TUNE_SPMV_GROUP, rec["best"])
        return rec
