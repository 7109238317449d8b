"""dort's statistical image comparison (TEST INFRASTRUCTURE ONLY).

Restates `test_convergence` (src/dort/convergence_test.cpp:10-101),
`bias_variance` (src/dort/lua_image.cpp:179-211), the acceptance rule of
test/test.lua:88-110 and `normal_cdf_inverse` (src/dort/math.cpp:4-19,
Abramowitz & Stegun 26.2.23).
"""
import math

import numpy as np


def _rational_approx(t):
    c0, c1, c2 = 2.515517, 0.802853, 0.010328
    d1, d2, d3 = 1.432788, 0.189269, 0.001308
    return t - ((c2 * t + c1) * t + c0) / (((d3 * t + d2) * t + d1) * t + 1.0)


def normal_cdf_inverse(p):
    if p < 0.5:
        return -_rational_approx(math.sqrt(-2.0 * math.log(p)))
    return _rational_approx(math.sqrt(-2.0 * math.log(1.0 - p)))


def test_convergence(reference, tested, min_tile_size=32, variation=8.0, bias=0.0, p_value=0.01):
    """Multi-scale tile z-test; returns "" on success or the first failing tile."""
    ref = np.asarray(reference, dtype=np.float64)
    tst = np.asarray(tested, dtype=np.float64)
    assert ref.shape == tst.shape
    h, w = ref.shape[:2]
    max_depth = 0
    while (min(w, h) >> (max_depth + 1)) >= min_tile_size:
        max_depth += 1
    test_count = float(((1 << (2 * max_depth + 2)) - 1) // 3)
    test_p = 1.0 - (1.0 - p_value) ** (1.0 / test_count)
    inv_phi = normal_cdf_inverse(test_p * 0.5)
    ref_mean = ref.reshape(-1, 3).mean(0)
    for depth in range(max_depth + 1):
        ts = min(w, h) >> depth
        for ty in range((h + ts - 1) // ts):
            for tx in range((w + ts - 1) // ts):
                y0, y1, x0, x1 = ty * ts, min((ty + 1) * ts, h), tx * ts, min((tx + 1) * ts, w)
                tr = ref[y0:y1, x0:x1].reshape(-1, 3).mean(0)
                tt = tst[y0:y1, x0:x1].reshape(-1, 3).mean(0)
                mean_diff = (tt - tr).max()
                sd = max((ref_mean * variation).max(), (tr * variation).max())
                b = min((ref_mean * bias).max(), (tr * bias).max())
                count = float((x1 - x0) * (y1 - y0))
                max_diff = -sd / math.sqrt(count) * inv_phi + b
                if abs(mean_diff) > max_diff:
                    return f"tile ({x0},{y0}):({x1},{y1}) has diff {mean_diff}, but threshold is {max_diff}"
    return ""


def bias_variance(reference, tested):
    d = np.asarray(reference, np.float64) - np.asarray(tested, np.float64)
    return float(d.sum()), float((d * d).sum())


def compare_with_ref(reference, tested, min_tile_size=32, variation=8.0, bias=0.0):
    """test.lua:88-110: ("ok" | "warn" | "err", message)."""
    h, w = reference.shape[:2]
    _, var = bias_variance(reference, tested)
    norm_sd = math.sqrt(var / (3 * w * h))
    err = test_convergence(reference, tested, min_tile_size, variation, bias, 0.01)
    if err:
        return "err", err
    return ("warn" if norm_sd > 1.5 else "ok"), ""


def rel_mse(x, r):
    """relMSE = mean over pixels and channels of (x - r)^2 / (r^2 + 1e-2)."""
    x = np.asarray(x, np.float64)
    r = np.asarray(r, np.float64)
    return float(np.mean((x - r) ** 2 / (r * r + 1e-2)))
