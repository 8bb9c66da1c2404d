"""Comparison helpers shared by the parity tests."""
import numpy as np


def pointwise_rel(a, b):
    """Per-point relative error max_k |a-b| / max_k |b| (columns = points,
    rows = components).  Points whose reference block is exactly zero must
    match exactly (returns inf otherwise)."""
    a = np.asarray(a)
    b = np.asarray(b)
    num = np.max(np.abs(a - b), axis=0)
    den = np.max(np.abs(b), axis=0)
    out = np.where(den > 0, num / np.where(den > 0, den, 1.0), np.where(num == 0, 0.0, np.inf))
    return out


def yield_value_3d(S, sin_phi, c_bar):
    """Mohr-Coulomb yield function (1+s) s_max - (1-s) s_min - c_bar of the
    stress S (6 x n, Voigt 11,22,33,12,23,13, tensor shear components)."""
    n = S.shape[1]
    T = np.zeros((n, 3, 3))
    T[:, 0, 0], T[:, 1, 1], T[:, 2, 2] = S[0], S[1], S[2]
    T[:, 0, 1] = T[:, 1, 0] = S[3]
    T[:, 1, 2] = T[:, 2, 1] = S[4]
    T[:, 0, 2] = T[:, 2, 0] = S[5]
    ev = np.linalg.eigvalsh(T)
    return (1 + sin_phi) * ev[:, 2] - (1 - sin_phi) * ev[:, 0] - c_bar


def yield_value_2d(S, sin_phi, c_bar):
    """Same for plane strain S = (s11, s22, s12, s33)."""
    n = S.shape[1]
    T = np.zeros((n, 3, 3))
    T[:, 0, 0], T[:, 1, 1], T[:, 2, 2] = S[0], S[1], S[3]
    T[:, 0, 1] = T[:, 1, 0] = S[2]
    ev = np.linalg.eigvalsh(T)
    return (1 + sin_phi) * ev[:, 2] - (1 - sin_phi) * ev[:, 0] - c_bar
