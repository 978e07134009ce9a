"""Pins the oracle's helper functions (CPU only).

The reference ships no tests or golden vectors, so every helper is pinned
against what the mathematics or a specification fixes:
  * pack3SN/unpack3SN (floatpacking.glsl:26-66): hand-derived codes and every
    value of each 10-bit field (the fields are independent, so 1024 cases per
    field cover all 2^30 codes);
  * packUnorm4x8 (GLSL 4.30 8.4): round(clamp(c, 0, 1) * 255), c / 255 back;
  * packHalf2x16: numpy's IEEE binary16 conversion, an independent
    implementation, on special values, subnormals, ties and random floats;
  * the one-at-a-time hash (random.glsl:34-66): an independent Python
    transcription on random keys plus fixed points;
  * lookAtRH / paraboloid_project (ism_utils.glsl): orthonormality, the
    normal direction at the tile centre, 90 degrees on the rim, v.z =
    dist / zFar, the tile offset (idx mod n, idx div n) / n;
  * ism grid side: int(pow(2, ceil(log2(n) / 2))) known answers.
A planted mistake in any of these (a dropped +0.5, a swapped field, a wrong
sign) fails at least one case.
"""
import math

import numpy as np
import pytest

import oracle


# ------------------------------------------------------------- pack3SN
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_pack3sn_known_codes(dt):
    # x, y: uint(c*0.5+0.5)*1023+0.5); z: uint((c*0.5+0.5)*1021+1.5)
    assert oracle.pack3sn((0.0, 0.0, 0.0), dt) == 512 | (512 << 10) | (512 << 20)
    assert oracle.pack3sn((1.0, 1.0, 1.0), dt) == 1023 | (1023 << 10) | (1022 << 20)
    assert oracle.pack3sn((-1.0, -1.0, -1.0), dt) == 0 | (0 << 10) | (1 << 20)
    assert oracle.pack3sn((1.0, -1.0, 0.0), dt) == 1023 | (0 << 10) | (512 << 20)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_unpack3sn_known_values(dt):
    v = oracle.unpack3sn(512 | (512 << 10) | (512 << 20), dt)
    assert np.allclose(v, [1024 / 1023 - 1, 1024 / 1023 - 1, 2 * (512 - 1.5) / 1021 - 1], atol=1e-6)
    v = oracle.unpack3sn(1023 | (0 << 10) | (1 << 20), dt)
    assert np.allclose(v, [1.0, -1.0, 2 * (1 - 1.5) / 1021 - 1], atol=1e-6)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_pack3sn_every_field_value(dt):
    """x/y codes unpack to their cell centres k/1023 and re-pack exactly. The z
    code unpacks to (k - 1.5)/1021, the lower edge of its cell
    [(k-1.5)/1021, (k-0.5)/1021) -- floatpacking.glsl:35 vs :57 -- so the path
    never re-packs it; the cell mid-point packs back to k."""
    for code in range(1024):
        z = min(max(code, 1), 1022)
        bits = code | (code << 10) | (z << 20)
        v = oracle.unpack3sn(bits, dt)
        assert (oracle.pack3sn(v, dt) & 0xFFFFF) == (bits & 0xFFFFF), code
        assert abs(float(v[2]) - (2 * (z - 1.5) / 1021 - 1)) < 1e-6
        zmid = 2 * (z - 1.0) / 1021 - 1  # centre of the z cell, as a signed normal component
        assert (oracle.pack3sn((v[0], v[1], zmid), dt) >> 20) == z, code


def test_pack3sn_quantisation_error_bound():
    rng = np.random.default_rng(0)
    n = rng.normal(size=(2000, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    for v in n:
        q = oracle.unpack3sn(oracle.pack3sn(v, np.float64), np.float64)
        # half a step of 2/1023 for x, y (cell centres); up to a full step of
        # 2/1021 for z (unpack returns the lower cell edge)
        assert np.all(np.abs(q - v) <= np.array([1 / 1023, 1 / 1023, 2 / 1021]) + 1e-12)


# -------------------------------------------------------- packUnorm4x8
def test_pack_unorm4x8_spec():
    assert oracle.pack_unorm4x8((0.0, 1.0, 0.5, 0.25)) == 0 | (255 << 8) | (128 << 16) | (64 << 24)
    assert oracle.pack_unorm4x8((-3.0, 7.0, 1 / 255, 254.5 / 255)) == 0 | (255 << 8) | (1 << 16) | (254 << 24)
    for b in (0, 1, 127, 128, 254, 255):
        got = oracle.unpack_unorm4x8(b | (b << 8) | (b << 16) | (b << 24))
        assert np.all(got == np.float32(b) / np.float32(255))


# -------------------------------------------------------- packHalf2x16
def test_f32_to_f16_matches_numpy():
    rng = np.random.default_rng(1)
    vals = np.concatenate([
        rng.normal(scale=1e3, size=4000).astype(np.float32),
        rng.uniform(-2, 2, size=4000).astype(np.float32),
        (rng.uniform(-1, 1, size=2000) * 1e-5).astype(np.float32),  # subnormal halves
        np.array([0.0, -0.0, 1.0, -1.0, 65504.0, 65520.0, 1e9, -1e9, 6.1035156e-05,
                  5.9604645e-08, 2.9802322e-08, 2.98023259e-08, np.inf, -np.inf], np.float32),
        # exact ties between two halves
        (np.arange(1, 2000, dtype=np.float32) * np.float32(2 ** -11) + np.float32(1.0)),
    ])
    for v in vals:
        want = int(np.float16(v).view(np.uint16))
        assert oracle.f32_to_f16(float(v)) == want, v
        assert oracle.f16_to_f32(want) == np.float32(np.float16(v)), v


# --------------------------------------------------------------- hash
def _py_hash(x):
    m = 0xFFFFFFFF
    x = (x + (x << 10)) & m
    x ^= x >> 6
    x = (x + (x << 3)) & m
    x ^= x >> 11
    x = (x + (x << 15)) & m
    return x


def test_hash_independent_transcription():
    rng = np.random.default_rng(2)
    assert oracle.hash1(0) == 0
    for x in list(rng.integers(0, 2 ** 32, size=500)) + [1, 2, 0xFFFFFFFF, 0x80000000]:
        x = int(x)
        assert oracle.hash1(x) == _py_hash(x)
    for _ in range(200):
        a, b, c = (int(v) for v in rng.integers(0, 2 ** 32, size=3))
        assert oracle.hash3(a, b, c) == _py_hash(a ^ _py_hash(b) ^ _py_hash(c))


def test_random_float_construct_range():
    rng = np.random.default_rng(3)
    xs = rng.uniform(0, 10, size=(3000, 3)).astype(np.float32)
    r = np.array([oracle.random3(*map(float, v)) for v in xs])
    assert r.min() >= 0.0 and r.max() < 1.0
    # the low 23 bits of the hash become the mantissa of 1.m - 1
    x = xs[0]
    bits = x.view(np.uint32)
    h = _py_hash(int(bits[0]) ^ _py_hash(int(bits[1])) ^ _py_hash(int(bits[2])))
    assert oracle.random3(*map(float, x)) == np.float32((h & 0x7FFFFF) / 2 ** 23)
    # roughly uniform
    assert abs(r.mean() - 0.5) < 0.02


# ------------------------------------------------------- ISM geometry
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_look_at_rh_orthonormal(dt):
    rng = np.random.default_rng(4)
    ns = rng.normal(size=(300, 3))
    ns /= np.linalg.norm(ns, axis=1, keepdims=True)
    ns = np.concatenate([ns, [[0, 1, 0], [0, -1, 0], [1, 0, 0], [0, 0, 1], [0, 0, -1]]])
    tol = 1e-12 if dt == np.float64 else 2e-6
    for n in ns:
        s, u = oracle.look_at_rh(n, dt)
        m = np.array([s, u, -np.asarray(n, dt)], np.float64)
        assert np.allclose(m @ m.T, np.eye(3), atol=5 * tol)
        assert abs(np.linalg.det(m) - 1.0) < 10 * tol
        assert np.allclose(m @ np.asarray(n, np.float64), [0, 0, -1], atol=5 * tol)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_paraboloid_project_invariants(dt):
    rng = np.random.default_rng(5)
    n1d, z_far = 8, 50.0
    tol = 1e-9 if dt == np.float64 else 1e-5
    for _ in range(200):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        s, u = oracle.look_at_rh(n, np.float64)
        idx = int(rng.integers(0, n1d * n1d))
        tile = np.array([idx % n1d, idx // n1d], np.float64)
        d = float(rng.uniform(0.5, 20.0))
        # along the normal: centre of the ISM tile, depth d / zFar
        v = oracle.paraboloid_project(n * d, d, n, z_far, idx, n1d, True, dt)
        assert np.allclose(v[:2], (tile + 0.5) / n1d, atol=tol)
        assert abs(v[2] - d / z_far) < tol * 10
        # 90 degrees (along s): the rim, half a tile from the centre
        v = oracle.paraboloid_project(s * d, d, n, z_far, idx, n1d, False, dt)
        assert np.allclose(v[:2], (tile + [1.0, 0.5]) / n1d, atol=tol * 10)
        # 90 degrees along -u: opposite rim in y
        v = oracle.paraboloid_project(-u * d, d, n, z_far, idx, n1d, False, dt)
        assert np.allclose(v[:2], (tile + [0.5, 0.0]) / n1d, atol=tol * 10)
        # 60 degrees off the normal towards s: x offset tan(30 deg)/2 of a tile
        r = (math.cos(math.pi / 3) * n + math.sin(math.pi / 3) * s) * d
        v = oracle.paraboloid_project(r, d, n, z_far, idx, n1d, False, dt)
        # paraboloid: x_local = sin(t) / (1 + cos(t)) = tan(t/2)
        want = (tile + [0.5 + 0.5 * math.tan(math.pi / 6), 0.5]) / n1d
        assert np.allclose(v[:2], want, atol=tol * 10)
        # behind the VPL (preserveSign): depth sign flips
        v = oracle.paraboloid_project(-n * d, d, n, z_far, idx, n1d, True, dt)
        assert v[2] < 0


def test_ism_grid_side_known_answers():
    cases = {1: 1, 2: 2, 4: 2, 5: 4, 16: 4, 17: 8, 256: 16, 1000: 32, 1024: 32, 1025: 64, 2048: 64,
             4096: 64, 16384: 128}
    for n, want in cases.items():
        assert oracle.ism_indices_1d(n) == want, n
        assert 2 ** math.ceil(math.log2(n) / 2) == want
