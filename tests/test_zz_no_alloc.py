"""ssr_dev_multidot_noalloc (partial sums in a static device array, no
allocation) against ssr_dev_multidot (library-allocated partial sums): the
same kernels in the same order, so the results agree bit for bit.  Two calls
in a row, the second on a smaller input, then the size limit."""
import numpy as np
import pytest

from slope_stability_octave_amd import _lib
from slope_stability_octave_amd._lib import DeviceArray, SSRError

pytestmark = pytest.mark.gpu


def _multidot(name, n, k, V, ld, u):
    h = DeviceArray(max(k, 1))
    _lib.call(name, n, k, _lib.ptr(V), ld, _lib.ptr(u), _lib.ptr(h))
    _lib.synchronize()
    return h.get()[:k]


@pytest.mark.parametrize("n,k", [(1_200_000, 100), (70_001, 9)])
def test_multidot_noalloc_matches_allocating_form(n, k):
    _lib.require_gpu()
    rng = np.random.default_rng(n + k)
    Vh = rng.normal(size=(k, n))  # k columns of length n, column-major
    uh = rng.normal(size=n)
    V, u = DeviceArray.from_host(Vh), DeviceArray.from_host(uh)
    print(f"multidot n={n} k={k}", flush=True)
    a = _multidot("ssr_dev_multidot", n, k, V, n, u)
    b = _multidot("ssr_dev_multidot_noalloc", n, k, V, n, u)
    assert np.array_equal(a, b)
    ref = Vh @ uh
    assert np.abs(b - ref).max() <= 1e-12 * np.abs(Vh).max() * np.abs(uh).max() * n
    # second call on a smaller input: no stale partial sums of the first
    m, kk = n // 3 + 17, max(1, k // 2)
    a2 = _multidot("ssr_dev_multidot", m, kk, V, n, u)
    b2 = _multidot("ssr_dev_multidot_noalloc", m, kk, V, n, u)
    assert np.array_equal(a2, b2)
    assert np.abs(b2 - Vh[:kk, :m] @ uh[:m]).max() <= 1e-12 * m * np.abs(Vh).max() * np.abs(uh).max()


def test_multidot_noalloc_rejects_too_many_columns():
    _lib.require_gpu()
    V, u, h = DeviceArray(129 * 10), DeviceArray(10), DeviceArray(129)
    with pytest.raises(SSRError, match="exceeds"):
        _lib.call("ssr_dev_multidot_noalloc", 10, 129, _lib.ptr(V), 10, _lib.ptr(u), _lib.ptr(h))
