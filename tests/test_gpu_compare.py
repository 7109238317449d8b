"""dgpu_test_convergence / dgpu_bias_variance (the GPU comparator) against dort's own
test_convergence (oracle/_ref, when built) and the oracle's restatement of it and of
bias_variance: same verdicts, same first failing tile, thresholds and differences to
float precision (dort sums tile pixels in float one by one, the GPU in double)."""
import os
import re

import numpy as np
import pytest

from dort_amd import gpu
from oracle import convergence, ref, rgbe

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _golden():
    return rgbe.read_hdr(os.path.join(HERE, "golden", "box_diff_area_pin.hdr"))


def _noisy(img, sd, seed):
    rng = np.random.default_rng(seed)
    return np.maximum(img * (1.0 + sd * rng.normal(size=img.shape)), 0).astype(np.float32)


def _scaled(img, box, k):
    y0, y1, x0, x1 = box
    out = img.copy()
    out[y0:y1, x0:x1] *= k
    return out


def _tile(msg):
    m = re.match(r"tile \((\d+),(\d+)\):\((\d+),(\d+)\) has diff (\S+), but threshold is (\S+)", msg)
    assert m, msg
    return tuple(int(x) for x in m.groups()[:4]), float(m.group(5)), float(m.group(6))


def _check_same(got, want):
    if want == "":
        assert got == "", got
        return
    tg, dg, thg = _tile(got)
    tw, dw, thw = _tile(want)
    assert tg == tw, (got, want)
    assert abs(dg - dw) <= 1e-3 * abs(dw) + 1e-6 and abs(thg - thw) <= 1e-3 * abs(thw) + 1e-6, (got, want)


CASES = [
    ("identical", lambda g: g.copy(), {}),
    ("noise", lambda g: _noisy(g, 0.5, 1), {}),
    ("bright tile", lambda g: _scaled(g, (280, 320, 120, 160), 4.0), {}),
    # dark regions pass at dort's default variation (8): |diff| stays under the threshold
    ("dark corner", lambda g: _scaled(g, (0, 512, 400, 512), 0.1), {}),
    ("within bias", lambda g: g * 1.01, {"bias": 0.02}),
    ("biased", lambda g: g * 1.3, {}),
    ("crop 300x217 tile 16", lambda g: _noisy(g[:217, :300], 0.2, 2), {"min_tile_size": 16}),
]


@pytest.mark.parametrize("name,make,kw", CASES, ids=[c[0] for c in CASES])
def test_convergence_matches_dort(gpu_ctx, name, make, kw):
    g = _golden()
    t = make(g)
    r = g[: t.shape[0], : t.shape[1]]
    got = gpu_ctx.test_convergence(r, t, **kw)
    _check_same(got, convergence.test_convergence(r, t, **kw))
    if ref.available():
        _check_same(got, ref.test_convergence(r, t, **kw))
    if name == "identical":
        assert got == ""
    if name in ("bright tile", "biased"):
        assert got != ""


def test_convergence_device_buffers(gpu_ctx):
    g = _golden()
    t = _scaled(_noisy(g, 0.5, 3), (280, 320, 120, 160), 4.0)
    assert convergence.test_convergence(g, t) != ""
    da, db = gpu_ctx.device_array(g), gpu_ctx.device_array(t)
    assert gpu_ctx.test_convergence(da, db) == gpu_ctx.test_convergence(g, t) != ""


def test_bias_variance(gpu_ctx):
    g = _golden()
    t = _noisy(g, 0.3, 4)
    b, v = gpu_ctx.bias_variance(g, t)
    rb, rv = convergence.bias_variance(g, t)
    scale = np.abs(g.astype(np.float64) - t).sum()
    assert abs(b - rb) <= 1e-9 * scale and abs(v - rv) <= 1e-9 * rv
    b2, v2 = gpu_ctx.bias_variance(g, t, rect=(10, 20, 300, 100))
    d = g[20:100, 10:300].astype(np.float64) - t[20:100, 10:300]
    assert abs(b2 - d.sum()) <= 1e-9 * np.abs(d).sum() and abs(v2 - (d * d).sum()) <= 1e-9 * (d * d).sum()
    with pytest.raises(gpu.DgpuError):
        gpu_ctx.bias_variance(g, t, rect=(0, 0, 600, 10))
