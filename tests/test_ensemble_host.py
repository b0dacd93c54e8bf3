"""Host side of the batched ensemble: the workload generator is bit-identical
to the same calls made against the reference library (fixtures written by
tests/golden/gen_golden.c)."""
import numpy as np
import pytest

from conftest import golden


@pytest.mark.parametrize("tag,seed,kmin,kmax,n", [("ensemble200_seed99", 99, 1, 12, 200),
                                                  ("ensemble8_big_seed5", 5, 30, 64, 8)])
def test_planetary_systems_bitwise(amd, tag, seed, kmin, kmax, n):
    nb, st, dt = amd.planetary_systems(n, seed, kmin, kmax)
    assert np.array_equal(nb, golden(tag + "_nbodies.npy"))
    assert np.array_equal(dt, golden(tag + "_dt.npy"))
    assert np.array_equal(st, golden(tag + "_init.npy"))


def test_planetary_systems_distribution(amd):
    nb, st, dt = amd.planetary_systems(2000, 3, 2, 8)
    K = nb - 1
    assert K.min() == 2 and K.max() == 8
    counts = np.bincount(K, minlength=9)[2:]
    assert counts.min() > 200          # uniform over 7 values: ~286 each
    first = np.concatenate([[0], np.cumsum(nb)[:-1]])
    # centre of mass at rest at the origin
    m = st[:, 9]
    for s in range(0, 2000, 97):
        sl = slice(first[s], first[s] + nb[s])
        com = (st[sl, :6] * m[sl, None]).sum(0) / m[sl].sum()
        assert np.abs(com).max() < 1e-15
    # dt = inner period / 20
    assert np.allclose(dt, 2 * np.pi * np.sqrt(1 / (1 + m[first + 1])) / 20, rtol=0, atol=0)
