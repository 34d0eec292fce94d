"""Pins oracle/preprocess_ref.depth_normalize with closed forms: a view
whose valid depths span [a, b] has mid (a+b)/2 and distance b-a; the
normalised value of a is -1/2 (-1 with divided2), of b +1/2 (+1); depths
outside the valid range are masked (multi_mask) and do not move the range;
a view with no valid depth keeps the 99999 / -99999 sentinels."""
import numpy as np

from oracle.preprocess_ref import depth_normalize


def test_range_closed_form():
    d = np.full((1, 2, 3, 4), 2.0, np.float32)
    d[0, 0, 0, 0], d[0, 0, 2, 3], d[0, 0, 1, 1] = 1.0, 3.0, 7.0   # 7.0 is invalid
    d[0, 1] = 0.05                                                   # no valid depth
    r, mask, mid, dist = depth_normalize(d, 0.2, 5.0, -1, True, False)
    assert mid[0, 0, 0, 0] == 2.0 and abs(dist[0, 0, 0, 0] - 2.0) < 1e-6
    assert r[0, 0, 0, 0] == -0.5 and r[0, 0, 2, 3] == 0.5 and r[0, 0, 0, 1] == 0.0
    assert mask[0, 0, 1, 1] == 0.0 and r[0, 0, 1, 1] == 0.0
    assert mid[0, 1, 0, 0] == 0.0 and dist[0, 1, 0, 0] == np.float32(-199998.0)
    r2, *_ = depth_normalize(d, 0.2, 5.0, -1, False, True)
    assert r2[0, 0, 0, 0] == -1.0 and r2[0, 0, 2, 3] == 1.0 and r2[0, 0, 1, 1] == 1.0  # clamped
    r3, *_ = depth_normalize(d, 0.2, 5.0, 4.0, True, False)
    assert r3[0, 0, 2, 3] == 0.25


def test_rgb_normalizer_closed_form():
    from oracle.preprocess_ref import rgb_normalizer
    rgb = np.array([0.0, 0.25, 0.5, 1.0], np.float32).reshape(1, 1, 2, 2).repeat(3, 1)
    mask = np.array([1, 1, 0, 1], np.float32).reshape(1, 1, 2, 2)
    out = rgb_normalizer(rgb, mask)
    np.testing.assert_array_equal(out[0, :, 0, 0], -1.0)
    np.testing.assert_array_equal(out[0, :, 0, 1], -0.5)
    np.testing.assert_array_equal(out[0, :, 1, 0], 0.0)      # masked
    np.testing.assert_array_equal(out[0, :, 1, 1], 1.0)
    out2 = rgb_normalizer(rgb, mask, mid=0.0, dist=1.0)
    np.testing.assert_array_equal(out2, rgb * mask)
