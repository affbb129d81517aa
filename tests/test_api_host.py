"""Host-side pieces of the user API (no GPU): array resampling as
engine/resample.go:22-42 does it (nearest cell by integer division)."""
import numpy as np

from mmx.api import resample


def test_resample_nearest_floor():
    a = np.arange(2 * 3 * 1, dtype=np.float64).reshape(1, 2, 3, 1)
    out = resample(a, (4, 6, 2))
    assert out.shape == (1, 4, 6, 2)
    for i in range(4):
        for j in range(6):
            for k in range(2):
                assert out[0, i, j, k] == a[0, i * 2 // 4, j * 3 // 6, k * 1 // 2]
    # the reference's temperature test: a 1x2x1 mask on a 32x32x1 grid
    m = resample([[[[0], [1]]]], (32, 32, 1))
    assert m.shape == (1, 32, 32, 1)
    assert np.all(m[0, :, :16] == 0) and np.all(m[0, :, 16:] == 1)
    # shrinking takes every n1/n2-th cell; same size is unchanged
    b = np.random.default_rng(0).normal(size=(3, 8, 4, 2))
    assert np.array_equal(resample(b, (4, 2, 1)), b[:, ::2, ::2, :1])
    assert resample(b, (8, 4, 2)) is not None and np.array_equal(resample(b, (8, 4, 2)), b)
