"""Pins the oracle's geometry-aware blur (gi/gi_blur.frag, CPU only).

gi_blur.frag weights a sample at distance r (1..3) along the pass direction
by max(0, (3 - r + 1)/4 - dz^2 - (5 (1 - max(0, n_c . n_s)))^2) and the centre
by 1; the result is the weighted mean, x pass then y pass. Checked against:
  * identity: a constant image over a flat, equally oriented surface;
  * the closed form of an impulse on such a surface: w_i w_j / 16 with
    w_0 = 1, w_1..3 = 0.75, 0.5, 0.25 (each pass divides by 1 + 2 * 1.5 = 4);
  * no bleeding across a 90-degree crease (normal factor 5 -> weight 0) or a
    view-space depth step of 1 or more (0.75 - 1 < 0);
  * FP32 within 1e-6 of FP64.
A dropped +1 in the radius term, a missing factor 5 or a swapped pass order
changes at least one of these.
"""
import numpy as np

import mlgi
import oracle


def flat(w, h, depth=0.9, normal=(0.0, 0.0, 1.0)):
    d = np.full((h, w), depth, np.float32)
    n = np.empty((h, w, 3), np.float32)
    n[...] = np.asarray(normal, np.float32) * 0.5 + 0.5
    return d, n


def proj():
    return mlgi.Camera.default(1.0).matrices()[1]


def test_constant_image_is_unchanged():
    d, n = flat(24, 20)
    gi = np.full((20, 24, 3), 0.37, np.float32)
    out = oracle.gi_blur(gi, n, d, proj())
    assert np.allclose(out, float(np.float32(0.37)), rtol=1e-15, atol=0)


def test_impulse_closed_form():
    w, h = 31, 29
    d, n = flat(w, h)
    gi = np.zeros((h, w, 3), np.float32)
    x0, y0 = 15, 14
    gi[y0, x0] = (1.0, 2.0, 4.0)
    out = oracle.gi_blur(gi, n, d, proj())
    wt = {0: 1.0, 1: 0.75, 2: 0.5, 3: 0.25}
    ref = np.zeros((h, w))
    for i in range(-3, 4):
        for j in range(-3, 4):
            ref[y0 + j, x0 + i] = wt[abs(i)] * wt[abs(j)] / 16.0
    for c, s in enumerate((1.0, 2.0, 4.0)):
        assert np.allclose(out[..., c], ref * s, rtol=1e-12, atol=1e-15)


def test_no_bleeding_across_crease_or_depth_step():
    w, h = 24, 12
    d, n = flat(w, h)
    n[:, 12:] = np.asarray((1.0, 0.0, 0.0), np.float32) * 0.5 + 0.5  # 90 degrees
    gi = np.zeros((h, w, 3), np.float32)
    gi[:, :12] = 1.0
    gi[:, 12:] = 3.0
    out = oracle.gi_blur(gi, n, d, proj())
    assert np.allclose(out[:, 3:12], 1.0) and np.allclose(out[:, 12:21], 3.0)
    # a depth step of more than 1 view-space unit
    d2, n2 = flat(w, h)
    p = proj()
    A, B = float(p[10]), float(p[14])
    lin = lambda z: -B / (A + (2 * z - 1))
    far = 0.9
    while lin(far) - lin(0.9) > -1.5:  # move the right half >= 1.5 units farther
        far = (far + 1.0) / 2
    d2[:, 12:] = far
    out2 = oracle.gi_blur(gi, n2, d2, p)
    assert np.allclose(out2[:, 3:12], 1.0) and np.allclose(out2[:, 12:21], 3.0)


def test_fp32_close_to_fp64():
    rng = np.random.default_rng(5)
    w, h = 40, 30
    d = rng.uniform(0.95, 0.99, (h, w)).astype(np.float32)
    nv = rng.normal(size=(h, w, 3))
    nv /= np.linalg.norm(nv, axis=-1, keepdims=True)
    n = (nv * 0.5 + 0.5).astype(np.float32)
    gi = rng.uniform(0, 2, (h, w, 3)).astype(np.float32)
    a = oracle.gi_blur(gi, n, d, proj(), np.float64)
    b = oracle.gi_blur(gi, n, d, proj(), np.float32)
    assert np.abs(a - b).max() <= 1e-6 * np.abs(a).max()
