"""The moment form of the polynomial surrogate kernels (tests/poly_moments.py)
equals the oracle's direct sum over all pairs for each of the 16 kernels."""
import numpy as np
import pytest

from poly_moments import KERNELS, poly_kernel_sum


@pytest.mark.parametrize("kernel", KERNELS)
def test_moments_match_direct_sum(oracle, kernel):
    rng = np.random.default_rng(3)
    pts = rng.uniform(-7, 9, (300, 3))
    q = rng.normal(size=300)
    d = oracle.direct(kernel, pts, pts, q, mode=1)
    m = poly_kernel_sum(kernel, pts, q)
    assert np.abs(m - d).max() < 1e-11 * np.abs(d).max()
