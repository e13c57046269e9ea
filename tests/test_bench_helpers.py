"""CPU checks of bench.py's measurement helpers (no GPU)."""
import numpy as np
import pytest

from oracle import ls as ol


@pytest.mark.parametrize("shape", [(8, 6, 4), (5, 8, 6), (12, 10)])
def test_eq_residual_matches_oracle(shape):
    import torch
    import bench
    rng = np.random.default_rng(3)
    d = len(shape)
    comps = ((0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)) if d == 3 else ((0, 0), (1, 1), (0, 1))
    soa = rng.standard_normal((len(comps),) + shape)
    aos = np.zeros(shape + (d, d))
    for c, (i, j) in enumerate(comps):
        aos[..., i, j] = soa[c]
        aos[..., j, i] = soa[c]
    got = bench.eq_residual(torch.from_numpy(soa))
    ref = ol.equilibrium_residual(aos)
    assert abs(got - ref) <= 1e-12 * ref


def test_eq_residual_zero_for_equilibrated_field():
    # sigma = constant + a divergence-free field sig_xy(z) = sin(2 pi z / n)
    import torch
    import bench
    n = 8
    z = np.arange(n)
    soa = np.zeros((6, n, n, n))
    soa[0] = 1.0
    soa[5] = np.sin(2 * np.pi * z / n)[None, None, :]
    assert bench.eq_residual(torch.from_numpy(soa)) < 1e-14
