"""The OpenMP generator library agrees with the numpy restatement."""
import numpy as np
import pytest

from visondf_amd import datagen as d


@pytest.mark.skipif(d._gen() is None, reason="libvdf_datagen.so not built")
def test_library_matches_numpy():
    assert np.array_equal(d.splitmix64(5, 3, 70_001, start=11), d.splitmix64(5, 3, 70_001, start=11, use_lib=False))
    assert np.array_equal(d.uniform01(5, 3, 70_001), d.uniform01(5, 3, 70_001, use_lib=False))
    assert np.array_equal(d.zipf_ranks(5, 0, 100_000, 1.1, 1 << 16), d.zipf_ranks(5, 0, 100_000, 1.1, 1 << 16,
                                                                                     use_lib=False))
    a, b = d.normal01(5, 1, 50_000), d.normal01(5, 1, 50_000, use_lib=False)
    assert np.all(np.abs(a - b) <= 4 * np.spacing(np.abs(b)))


def test_zipf_shape():
    r = d.zipf_ranks(1, 0, 200_000, 1.1, 1 << 24)
    assert r.min() >= 0 and r.max() < (1 << 24)
    # P(rank 0) = 1 / H_{2^24, 1.1} ~ 0.1151
    assert abs(np.mean(r == 0) - 0.1151) < 0.005
    s = d.scramble24(np.arange(1 << 24))
    assert np.unique(s).size == 1 << 24
