"""The node-block SpMV that reads a symmetric operator's blocks on and above
the block diagonal only (tuning key 14; the blocks left of the diagonal are
read as the transposes of their mirrors).

K_r is symmetrised by construction (r K_el + (1 - r) (K_t + K_t') / 2 under
a symmetric mask), so its mirrored blocks are transposes bit for bit; the
mirror form then forms the same products in the same order as the plain
form and must give the same bits:
* the operator's bit symmetry on a plastic state, 2D and 3D;
* one product in double and in float, plain against mirror, bit for bit;
* whole continuations (2D, and 3D in both precisions) with the mirror form
  in the solver and the AMG finest level: the same lambda and omega
  histories, bit for bit.
"""
import numpy as np
import pytest

from slope_stability_octave_amd import _lib, linear_solvers, problems  # noqa: F401  (C signatures)
from slope_stability_octave_amd._lib import DeviceArray

pytestmark = pytest.mark.gpu


def _plastic_operator(dim):
    if dim == 3:
        prob = problems.slope_3d((2, 2, 2, 2, 2, 2), 2.0, 2.0, 2.0, 2.0, 2.0, 2.0, problems.HOMO_2D)
    else:
        prob = problems.slope_2d(h=1.0)
    cmb = problems.make_constitutive(prob)
    cmb.reduction(1.2)
    rng = np.random.default_rng(3)
    # displacements of a few mm: strains far past the yield strain, so every
    # return-mapping branch contributes to the tangent
    U = DeviceArray.from_host(rng.normal(size=(cmb.n_n, cmb.dim)) * 2e-3 * cmb.Q.T)
    cmb.build_F_and_DS_all(1.2, U)
    return cmb, cmb.build_K_r_QQ_vals(0.3)


def _tables(P):
    return P.row_ptr, P.col, P.table(4), P.table(3)


@pytest.mark.parametrize("dim", [2, 3])
def test_K_r_mirrored_blocks_are_transposes_bit_for_bit(dim):
    cmb, K = _plastic_operator(dim)
    P = cmb.pattern
    rp, col, diag, tblk = _tables(P)
    v = K.get().reshape(P.nnzb, dim, dim)
    assert np.array_equal(v[tblk], np.transpose(v, (0, 2, 1)))
    rows = np.repeat(np.arange(P.n_n), np.diff(rp))
    assert np.array_equal(col[diag], np.arange(P.n_n))
    assert np.array_equal(rows[tblk], col) and np.array_equal(col[tblk], rows)
    # a plastic state: the tangent is not the elastic one
    assert np.abs(v).max() > 0


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("bits", [64, 32])
def test_mirror_spmv_same_bits(dim, bits):
    cmb, K = _plastic_operator(dim)
    P = cmb.pattern
    rng = np.random.default_rng(11)
    dt = np.float64 if bits == 64 else np.float32
    x = DeviceArray.from_host((rng.normal(size=(cmb.n_n, dim)) * cmb.Q.T).astype(dt))
    z = DeviceArray.from_host(rng.normal(size=(cmb.n_n, dim)).astype(dt))
    if bits == 32:
        nv = P.nnzb * P.bs
        Kb = DeviceArray(nv, dtype=np.float32)
        _lib.call("ssr_dev_convert_f64_f32", nv, _lib.ptr(K), _lib.ptr(Kb))
        plain, sym = "ssr_dev_spmv_bsr_f32", "ssr_dev_spmv_bsr_sym_f32"
    else:
        Kb = K
        plain, sym = "ssr_dev_spmv_bsr", "ssr_dev_spmv_bsr_sym"
    ptrs = (_lib.ptr(P.d_row_ptr), _lib.ptr(P.d_col))
    d_diag, d_tblk = P._dptr(4), P._dptr(3)
    out = {}
    old = _lib.get_tuning(_lib.TUNE_SPMV_MIRROR)
    try:
        for mode in (0, 1):
            _lib.set_tuning(_lib.TUNE_SPMV_MIRROR, mode)
            for beta in (0.0, -0.5):
                y0 = DeviceArray.from_host(z.get())
                y1 = DeviceArray.from_host(z.get())
                _lib.call(plain, dim, P.n_n, *ptrs, _lib.ptr(Kb), _lib.ptr(x), _lib.ptr(y0), 1.5, beta)
                _lib.call(sym, dim, P.n_n, *ptrs, _lib.ptr(d_diag), _lib.ptr(d_tblk), _lib.ptr(Kb), _lib.ptr(x),
                          _lib.ptr(y1), 1.5, beta)
                out[(mode, beta)] = (y0.get(), y1.get())
    finally:
        _lib.set_tuning(_lib.TUNE_SPMV_MIRROR, old)
    for beta in (0.0, -0.5):
        ref = out[(0, beta)][0]
        assert np.array_equal(out[(0, beta)][1], ref)  # key off: the plain form
        assert np.array_equal(out[(1, beta)][0], ref)  # no tables: the plain form
        assert np.array_equal(out[(1, beta)][1], ref)  # the mirror form
        assert np.abs(ref).max() > 0


@pytest.mark.parametrize("case", ["2d", "3d_64", "3d_32"])
def test_mirror_spmv_in_a_continuation_same_bits(case):
    if case == "2d":
        prob, prec = problems.slope_2d(1.0), 64
    else:
        # the 60 x 30 x 20 m slope geometry on a coarse mesh
        prob = problems.slope_3d((4, 4, 4, 2, 4, 4), 20.0, 20.0, 20.0, 10.0, 20.0, 20.0, problems.HOMO_2D)
        prec = 64 if case == "3d_64" else 32
    old = _lib.get_tuning(_lib.TUNE_SPMV_MIRROR)
    res = {}
    try:
        for mode in (0, 1):
            _lib.set_tuning(_lib.TUNE_SPMV_MIRROR, mode)
            res[mode] = problems.run_ssr(prob, params=dict(linear_tol=1e-2), preconditioner="amg", precision=prec)
    finally:
        _lib.set_tuning(_lib.TUNE_SPMV_MIRROR, old)
    print(f"{case}: FoS {res[0]['FoS']!r} / {res[1]['FoS']!r}, steps {len(res[0]['lambda_hist'])}", flush=True)
    assert res[0]["FoS"] == res[1]["FoS"]
    assert np.array_equal(np.asarray(res[0]["lambda_hist"]), np.asarray(res[1]["lambda_hist"]))
    assert np.array_equal(np.asarray(res[0]["omega_hist"]), np.asarray(res[1]["omega_hist"]))
