"""Host side of the native Newton drivers (no GPU): the ctypes image of
ssr_constitutive matches the C declaration, and the description refuses
arrays too small for the mesh before anything reaches the device."""
import ctypes
import os
import re
import types

import numpy as np
import pytest

from slope_stability_octave_amd import newton

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_struct_fields_follow_header_order():
    text = open(os.path.join(ROOT, "include", "ssr_hip.h")).read()
    body = re.search(r"typedef struct ssr_constitutive \{(.*?)\} ssr_constitutive;", text, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        for part in decl.split(","):
            names.append(re.findall(r"\w+", part)[-1])
    fields = [f[0].rstrip("_") for f in newton._Constitutive._fields_]
    assert fields == names
    # 4 ints, one int64, then pointers: no padding surprises
    assert newton._Constitutive.pattern.offset == 24
    assert ctypes.sizeof(newton._Constitutive) == 24 + 8 * (len(names) - 5)


class _Dev:
    """Stands in for a DeviceArray: an address and a byte count."""

    def __init__(self, n, addr=0x1000):
        self.nbytes = 8 * n
        self.ptr = addr


def _fake_cmb(dim=2, n_n=10, n_e=3, n_q=7, n_p=6, nnzb=40, short=None):
    n_int = n_e * n_q
    ns = 3 if dim == 2 else 6
    s_rows = 4 if dim == 2 else 6
    bs = dim * dim
    pat = types.SimpleNamespace(nnzb=nnzb, bs=bs, n_dof=dim * n_n, handle=ctypes.c_void_p(0x2000))
    sizes = dict(WEIGHT=n_int, c0=n_int, phi=n_int, psi=n_int, shear=n_int, bulk=n_int, lame=n_int, c_bar=n_int,
                 sin_phi=n_int, E=n_int * ns, S=n_int * s_rows, DS=n_int * ns * ns, K_elast=nnzb * bs,
                 K_tang=nnzb * bs, K_r=nnzb * bs)
    if short:
        sizes[short] -= 1
    cmb = types.SimpleNamespace(dim=dim, n_q=n_q, n_int=n_int, n_p=n_p, n_strain=ns, s_rows=s_rows, Davis_type="B",
                                pattern=pat, pattern_initialized=True, mask=_Dev(dim * n_n),
                                DPhi=[_Dev(n_p * n_int) for _ in range(dim)], _lambda=_Dev(1),
                                _phi_u=np.array([0.5]), _psi_u=np.zeros(1))
    for k, v in sizes.items():
        setattr(cmb, k, _Dev(v))
    return cmb


def test_description_of_a_consistent_problem():
    cmb = _fake_cmb()
    d = newton.native_description(cmb, _Dev(cmb.pattern.n_dof, 0x3000))
    assert (d.dim, d.n_q, d.davis, d.s_rows, d.n_int) == (2, 7, ord("B"), 4, 21)
    assert d.pattern == 0x2000 and d.f == 0x3000 and not d.DPhi3
    assert d.phi_host == cmb._native_host[0].ctypes.data and d.n_friction == 1


@pytest.mark.parametrize("short", ["E", "DS", "K_r", "c_bar", "WEIGHT"])
def test_description_refuses_short_arrays(short):
    cmb = _fake_cmb(short=short)
    with pytest.raises(ValueError, match=short):
        newton.native_description(cmb, _Dev(cmb.pattern.n_dof))


def test_description_refuses_short_load_and_unpaired_materials():
    cmb = _fake_cmb()
    with pytest.raises(ValueError, match="DPhi or f"):
        newton.native_description(cmb, _Dev(cmb.pattern.n_dof - 1))
    cmb = _fake_cmb()
    cmb._phi_u = np.array([0.5, 0.6])
    with pytest.raises(ValueError, match="must pair up"):
        newton.native_description(cmb, _Dev(cmb.pattern.n_dof))


def test_unknown_driver_is_rejected():
    cmb = _fake_cmb()
    with pytest.raises(ValueError, match="unknown Newton driver"):
        newton.newton_ind_SSR(None, 1.0, 1.0, 5, 5, 1e-4, 1e-4, None, None, None, cmb, None, driver="gpu")
    with pytest.raises(ValueError, match="unknown Newton driver"):
        newton.newton(None, 1e-4, 5, 5, 1e-4, None, None, None, cmb, None, driver="gpu")
