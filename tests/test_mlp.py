"""Learned-predictor network: the CPU oracle is pinned by finite differences
and exact bf16 rounding cases; the HIP forward, backward, loss and training
kernels are checked against it.

Tolerances: forward/backward relative 1e-5 of the largest entry (float
accumulation over at most 128 terms against float64); loss gradient 1e-5.
"""
import numpy as np
import pytest

from oracle import mlp_oracle as MO

WIDTHS = [12, 128, 128, 5]


def _params(seed=0, widths=WIDTHS):
    rng = np.random.default_rng(seed)
    out = []
    for nin, nout in zip(widths[:-1], widths[1:]):
        out.append(rng.uniform(-1, 1, nout * nin) * np.sqrt(6.0 / (nin + nout)))
        out.append(rng.normal(0, 0.1, nout))
    return np.concatenate(out).astype(np.float32)


def test_bf16_round_known_values():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 1.0 + 2 ** -7 + 2 ** -8, 3.0e-3],
                    dtype=np.float32)
    got = MO.bf16_round(vals)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even mantissa (1.0) wins;
    # 1 + 3*2^-9 rounds up; 1 + 2^-7 + 2^-8 ties to the even 1 + 2^-6
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == 1.0 + 2 ** -7 and got[3] == -2.5
    assert got[4] == 1.0 + 2 ** -6
    assert abs(got[5] - 3.0e-3) <= 2 ** -8 * 3.0e-3


def test_oracle_backward_matches_finite_differences():
    widths = [4, 7, 6, 3]
    p = _params(1, widths).astype(np.float64)
    rng = np.random.default_rng(2)
    X = rng.normal(size=(5, 4))
    dY = rng.normal(size=(5, 3))
    # differentiate the float64 network without bf16 rounding: patch rounding out
    orig = MO.bf16_round
    try:
        MO.bf16_round = lambda w: np.asarray(w, dtype=np.float64)
        g = MO.backward(p, widths, X, dY)
        num = np.zeros_like(p)
        h = 1e-6
        for k in range(p.size):
            e = np.zeros_like(p)
            e[k] = h
            # forward casts parameters through float32; use float64 directly
            yp = _fwd64(p + e, widths, X)
            ym = _fwd64(p - e, widths, X)
            num[k] = ((yp - ym) * dY).sum() / (2 * h)
    finally:
        MO.bf16_round = orig
    assert np.abs(g - num).max() <= 1e-6 * max(1.0, np.abs(num).max())


def _fwd64(p, widths, X):
    A = X
    off = 0
    L = len(widths) - 1
    for l, (nin, nout) in enumerate(zip(widths[:-1], widths[1:])):
        W = p[off:off + nout * nin].reshape(nout, nin)
        off += nout * nin
        b = p[off:off + nout]
        off += nout
        Z = A @ W.T + b
        A = np.tanh(Z) if l + 1 < L else Z
    return A


def test_oracle_predictor_loss_gradient():
    rng = np.random.default_rng(3)
    B, nc = 6, 4
    Y = rng.normal(size=(B, nc + 1))
    M = rng.normal(size=(B, nc, nc))
    G = np.einsum("sij,skj->sik", M, M)
    g = rng.normal(size=(B, nc))
    h = np.ones(B)
    t = rng.normal(size=B)
    L, dY = MO.predictor_loss(Y, G.reshape(B, -1), g, h, t, 0.7)
    num = np.zeros_like(Y)
    for s in range(B):
        for k in range(nc + 1):
            e = np.zeros_like(Y)
            e[s, k] = 1e-6
            num[s, k] = (MO.predictor_loss(Y + e, G.reshape(B, -1), g, h, t, 0.7)[0].mean()
                         - MO.predictor_loss(Y - e, G.reshape(B, -1), g, h, t, 0.7)[0].mean()) / 2e-6
    assert np.abs(dY - num).max() <= 1e-6
    # at delta = 0 the loss is h + w t^2 (the baseline)
    L0, _ = MO.predictor_loss(np.c_[np.zeros((B, nc)), np.zeros(B)], G.reshape(B, -1), g, h, t, 0.7)
    assert np.allclose(L0, 1.0 + 0.7 * t ** 2)


@pytest.mark.gpu
def test_gpu_forward_backward_match_oracle():
    from slope_stability_octave_amd import learned
    net = learned.MLP(12, 128, 5, 3, seed=5, out_scale=1.0)
    p = _params(4)
    net.set_params(p)
    assert np.array_equal(net.get_params(), p)
    rng = np.random.default_rng(6)
    X = rng.normal(size=(37, 12)).astype(np.float32)
    Y = net.forward(X)
    Y0, _ = MO.forward(p, WIDTHS, X)
    assert np.abs(Y - Y0).max() <= 1e-5 * np.abs(Y0).max()
    dY = rng.normal(size=Y.shape).astype(np.float32)
    g = net.backward(X, dY)
    g0 = MO.backward(p, WIDTHS, X, dY)
    assert np.abs(g - g0).max() <= 1e-5 * np.abs(g0).max()
    # per-layer check as well (the output layer's gradients are much larger)
    off = 0
    for nin, nout in zip(WIDTHS[:-1], WIDTHS[1:]):
        n = nout * nin + nout
        seg, seg0 = g[off:off + n], g0[off:off + n]
        assert np.abs(seg - seg0).max() <= 2e-5 * np.abs(seg0).max()
        off += n


@pytest.mark.gpu
def test_gpu_predictor_loss_and_training_reduce_loss():
    from slope_stability_octave_amd import learned
    from slope_stability_octave_amd._lib import DeviceArray, call, ptr
    rng = np.random.default_rng(8)
    B, nc = 64, 4
    X = rng.normal(size=(B, 12)).astype(np.float32)
    M = rng.normal(size=(B, nc, nc))
    G = (np.einsum("sij,skj->sik", M, M) / nc + np.eye(nc)).astype(np.float32)
    # targets: an affine function of the features (learnable)
    A = rng.normal(size=(12, nc)) * 0.3
    dstar = X @ A
    g = np.einsum("sij,sj->si", G, dstar).astype(np.float32)
    h = np.einsum("si,si->s", dstar, g).astype(np.float32)  # L = (d - d*)'G(d - d*) + w (y - t)^2 >= 0
    t = (X[:, :3].sum(1) * 0.2).astype(np.float32)
    net = learned.MLP(12, 128, 5, 3, seed=1)
    Y = net.forward(X)
    dY = DeviceArray((B, 5), dtype=np.float32)
    loss = DeviceArray(B, dtype=np.float32)
    dd = [DeviceArray.from_host(a) for a in (Y, G.reshape(B, -1), g, h, t)]
    call("ssr_mlp_predictor_loss", B, nc, *[ptr(a) for a in dd], 0.5, ptr(dY), ptr(loss))
    L0, dY0 = MO.predictor_loss(Y, G.reshape(B, -1), g, h, t, 0.5)
    assert np.abs(loss.get() - L0).max() <= 1e-4 * np.abs(L0).max()
    assert np.abs(dY.get() - dY0).max() <= 1e-5 * np.abs(dY0).max()
    hist, _ = learned.train(net, (X, G, g, t, h), epochs=400, lr=3e-3, wd=0.0, w_lam=0.5)
    print("loss", hist[0], hist[-1])
    assert hist[-1] < 0.05 * hist[0]
    # the reported loss is the oracle's loss of the parameters before each step
    p = net.get_params()
    Lend, _ = MO.predictor_loss(MO.forward(p, WIDTHS, X)[0], G.reshape(B, -1), g, h, t, 0.5)
    assert abs(Lend.mean() - hist[-1]) <= 0.5 * hist[-1] + 1e-3
