"""The A-orthonormalisation oracle (A_orthogonalize.m restated) against what
fixes it: closed forms for a diagonal indefinite operator and a single
column, the A-orthonormality W'AW = diag(sign) and span preservation for a
symmetric operator, the drop of a dependent column, and the order rules
(reverse processing, forward output)."""
import numpy as np

from oracle import krylov_oracle as KO


def _spd(n, rng):
    B = rng.normal(size=(n, n))
    return B @ B.T + n * np.eye(n)


def test_single_column_closed_form():
    A = np.diag([4.0, 9.0, 1.0])
    w = np.array([1.0, 1.0, 0.0])
    Wo, s = KO.a_orthogonalize(w[:, None], A, 1e-3)
    assert Wo.shape == (3, 1) and s.tolist() == [1.0]
    np.testing.assert_allclose(Wo[:, 0], w / np.sqrt(13.0), rtol=0, atol=1e-15)


def test_indefinite_diagonal_signs():
    A = np.diag([2.0, -3.0, 5.0])
    W = np.eye(3)[:, :2]
    Wo, s = KO.a_orthogonalize(W, A, 1e-3)
    np.testing.assert_allclose(Wo, np.diag([1 / np.sqrt(2.0), 1 / np.sqrt(3.0), 0.0])[:, :2], atol=1e-15)
    assert s.tolist() == [1.0, -1.0]


def test_small_norm_dropped():
    A = np.diag([1.0, 1e-4])
    Wo, s = KO.a_orthogonalize(np.eye(2), A, 1e-3)
    # the second column (processed first) has w'Aw = 1e-4 <= 1e-3
    assert Wo.shape == (2, 1)
    np.testing.assert_allclose(Wo[:, 0], [1.0, 0.0], atol=1e-15)


def test_symmetric_invariants_and_dependent_drop():
    rng = np.random.default_rng(0)
    n, k = 60, 12
    A = _spd(n, rng)
    W = rng.normal(size=(n, k))
    W = np.concatenate([W, (2 * W[:, 0] + W[:, 1])[:, None]], axis=1)
    Wo, s = KO.a_orthogonalize(W, A, 1e-3)
    # one column is dependent on the rest
    assert Wo.shape == (n, k) and np.all(s == 1.0)
    np.testing.assert_allclose(Wo.T @ A @ Wo, np.eye(k), atol=1e-10)
    # span(Wo) = span(W)
    P = Wo @ np.linalg.lstsq(Wo, W, rcond=None)[0]
    assert np.abs(P - W).max() <= 1e-10 * np.abs(W).max()
    # reverse processing: the last column is the first accepted, so the last
    # output column is parallel to it
    last = W[:, -1] / np.sqrt(W[:, -1] @ A @ W[:, -1])
    np.testing.assert_allclose(Wo[:, -1], last, atol=1e-12)
    # and the dependent combination makes the FIRST input column redundant
    Wo2, _ = KO.a_orthogonalize(W[:, 1:], A, 1e-3)
    np.testing.assert_allclose(np.abs(Wo2), np.abs(Wo), atol=1e-10)


def test_indefinite_symmetric_signs():
    rng = np.random.default_rng(1)
    n, k = 40, 8
    Qm, _ = np.linalg.qr(rng.normal(size=(n, n)))
    d = np.concatenate([np.linspace(1, 3, n // 2), -np.linspace(1, 3, n - n // 2)])
    A = Qm @ np.diag(d) @ Qm.T
    W = rng.normal(size=(n, k))
    Wo, s = KO.a_orthogonalize(W, A, 1e-3)
    np.testing.assert_allclose(Wo.T @ A @ Wo, np.diag(s), atol=1e-9)
    assert set(s.tolist()) <= {1.0, -1.0}
