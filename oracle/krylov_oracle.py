"""CPU restatement of the deflation-basis A-orthonormalisation (test
infrastructure only: the product path never imports this file).

``a_orthogonalize`` follows +LINEAR_SOLVERS/A_orthogonalize.m:23-57 column by
column with dense numpy: the columns of W are taken in reverse order; each is
reduced by the classical Gram-Schmidt product against the columns accepted so
far (weighted by their signs), kept when |t'Aw| > eps_add, scaled by
1/sqrt|t'Aw| with the sign recorded, and the accepted columns are returned in
forward order.  Pinned in tests/test_oracle_krylov.py by closed forms
(diagonal indefinite A, a single column), the A-orthonormality invariant for
symmetric A, span preservation and the drop of a dependent column.
"""
from __future__ import annotations

import numpy as np


def a_orthogonalize(W, A, eps_add):
    """Return (W_orth, norms) for W (m x n) and an operator A (matrix or
    anything with ``@``), A_orthogonalize.m:23-57."""
    W = np.asarray(W, dtype=np.float64)
    m, n = W.shape
    W_orth = np.zeros((m, n))
    norms = np.ones(n)
    keep = np.zeros(n, dtype=bool)
    for i in range(n):                      # A_orthogonalize.m:30
        temp = W[:, n - 1 - i].copy()       # :32
        A_temp = A @ temp                   # :33
        if i > 0:                           # :35-37
            Wp = W_orth[:, :i]
            temp = temp - (Wp * norms[:i]) @ (A_temp @ Wp)
        nt = float(temp @ A_temp)           # :39
        if abs(nt) > eps_add:               # :40-48
            if nt > 0:
                nt = np.sqrt(nt)
            else:
                nt = np.sqrt(abs(nt))
                norms[i] = -1.0
            W_orth[:, i] = temp / nt
            keep[i] = True
    W_orth = W_orth[:, keep][:, ::-1]       # :52-56
    norms = norms[keep][::-1]
    return W_orth, norms
