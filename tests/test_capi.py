"""The C-ABI library loads (no GPU needed) and exports every symbol that
include/*.h declares; the product path refuses to run without a GPU."""
import ctypes
import glob
import os
import re

import pytest

from slope_stability_octave_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(ssr_\w+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "ssr_constitutive_problem_3D_SDS" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_signatures_cover_header():
    from slope_stability_octave_amd import amg, assembly, device_setup, learned, linear_solvers, newton  # noqa: F401
    missing = [n for n in declared_symbols() if n not in _lib.SIGNATURES]
    assert not missing, missing


def test_error_reporting_without_compute():
    lib = _lib.load()
    # argument checks run before any device work
    rc = lib.ssr_constitutive_problem_2D_S(ctypes.c_int64(4), ctypes.c_int64(6), None, None, None,
                                           None, None, None, None)
    assert rc == 1
    assert b"3 rows" in lib.ssr_last_error()


def test_newton_driver_argument_checks():
    """ssr_newton_ind_SSR rejects an incomplete problem description before
    touching the device."""
    from slope_stability_octave_amd import newton
    lib = _lib.load()
    d = newton._Constitutive()
    out = [ctypes.c_double(0.0), ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)]
    fake = ctypes.c_void_p(16)  # never dereferenced: the checks fail first
    rc = lib.ssr_newton_ind_SSR(ctypes.byref(d), fake, 0, fake, ctypes.c_double(1.0), ctypes.c_double(1.0), 10, 10,
                                ctypes.c_double(1e-4), ctypes.c_double(1e-4), fake, ctypes.c_int64(0), fake,
                                *[ctypes.byref(o) for o in out], None, None)
    assert rc == 1
    assert b"incomplete constitutive description" in lib.ssr_last_error()
    rc = lib.ssr_newton_ind_SSR(None, fake, 0, fake, ctypes.c_double(1.0), ctypes.c_double(1.0), 10, 10,
                                ctypes.c_double(1e-4), ctypes.c_double(1e-4), fake, ctypes.c_int64(0), fake,
                                *[ctypes.byref(o) for o in out], None, None)
    assert rc == 1
    assert b"null argument" in lib.ssr_last_error()
    n = ctypes.c_int64(0)
    assert lib.ssr_newton_work_len(ctypes.byref(d), ctypes.byref(n)) == 1
    rc = lib.ssr_newton(ctypes.byref(d), fake, 0, fake, 10, 10, ctypes.c_double(1e-4), ctypes.c_double(1e-4), fake,
                        ctypes.c_int64(0), fake, *[ctypes.byref(o) for o in out[1:]], None, None)
    assert rc == 1
    assert b"incomplete constitutive description" in lib.ssr_last_error()


def test_no_cpu_fallback_when_no_gpu():
    lib = _lib.load()
    c = ctypes.c_int(0)
    lib.ssr_device_count(ctypes.byref(c))
    if c.value > 0:
        pytest.skip("a GPU is visible")
    from slope_stability_octave_amd import constitutive_mex as cm
    import numpy as np
    with pytest.raises(_lib.SSRLibraryError):
        cm.constitutive_problem_3D_S_mex(np.zeros((6, 2)), *[np.ones(2)] * 5)
