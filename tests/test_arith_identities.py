"""Pins the exact arithmetic rewrites of the pull-push and splat kernels (CPU only).

The kernel replaces two float operations of the reference's shaders by
cheaper forms that give the same bits:
  * x / w for the push and pull weight sums (w = 1..16) and x / 3.14 in the
    radius decode become RN(q + r y) with y = RN(1/w), q = RN(x y) and the
    exact FMA remainder r (Markstein's theorem);
  * sqrt(x) > R in the radius tests becomes x > RD(m^2), m the midpoint
    between R and the next float, with RD(m^2) = RD(R next(R)) formed in
    FP32 from the rounded product and its FMA error (no FP64 on the GPU).
The splat's unpackUnorm4x8 (c / 255 for a byte c) uses the first form with
divisor 255, checked for all 256 bytes, and its hemisphere tests decide
RN(s + t) < 0 as s < -t, checked on random pairs with exact cancellations,
signed zeros and subnormals; its keep mask (the first four of the 16
candidates that pass, ism.comp:80-103) is built from the mask of all passing
candidates by clearing the lowest set bit four times, checked for all 2^16
masks. The float rewrites are checked here against IEEE
float division and sqrtf: the division
over every significand of two binades per divisor and sign (scaling by a
power of two leaves all three roundings unchanged), the comparison on random
radii at the 65 floats around R^2 where the decision flips.
"""
import ctypes as C
import subprocess
from pathlib import Path

import numpy as np
import pytest

SRC = Path(__file__).with_name("arith_identities.c")


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    so = tmp_path_factory.mktemp("arith") / "libarith.so"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-shared", str(SRC), "-o", str(so), "-lm"],
                   check=True)
    L = C.CDLL(str(so))
    L.check_div_by_recip.argtypes = [C.c_float, C.c_int, C.c_int, C.c_int]
    L.check_div_by_recip.restype = C.c_long
    L.check_sqrt_bound.argtypes = [C.c_long, C.c_uint64, C.c_int]
    L.check_sqrt_bound.restype = C.c_long
    L.check_rd_bound.argtypes = [C.c_int, C.c_int]
    L.check_rd_bound.restype = C.c_long
    L.check_rd_bound_sqrt.argtypes = [C.c_long, C.c_uint64]
    L.check_rd_bound_sqrt.restype = C.c_long
    L.check_unorm8.argtypes = [C.c_int]
    L.check_unorm8.restype = C.c_long
    L.check_sign_of_sum.argtypes = [C.c_long, C.c_uint64, C.c_int]
    L.check_sign_of_sum.restype = C.c_long
    return L


@pytest.mark.parametrize("w", list(range(1, 17)) + [3.14])
def test_division_by_reciprocal_is_exact(lib, w):
    assert lib.check_div_by_recip(float(np.float32(w)), 0, -3, -1) == 0


def test_sqrt_comparison_without_root(lib):
    assert lib.check_sqrt_bound(200_000, 12345, 1) == 0


def test_fp32_bound_equals_the_fp64_midpoint_bound(lib):
    # the kernel's radii are at most 15 px and shrink with the level scale
    # and with small ISMs; checked over every float of [2^-16, 2^5), 0 and
    # denormals (the identity itself holds for every normal R)
    assert lib.check_rd_bound(-16, 5) == 0


def test_sqrt_comparison_with_the_fp32_bound(lib):
    assert lib.check_rd_bound_sqrt(200_000, 54321) == 0


def test_planted_mistakes_are_caught(lib):
    # a reciprocal far enough off that one correction step cannot repair it
    # (one ulp low already fails for 3.14), or the naive bound R^2, is caught
    assert lib.check_div_by_recip(3.0, 4096, -1, 0) > 0
    assert lib.check_div_by_recip(3.14, -1, -1, 0) > 0
    assert lib.check_sqrt_bound(20_000, 12345, 0) > 0


def test_unorm8_from_reciprocal_is_exact(lib):
    # the splat's unpackUnorm4x8 (point normals and radius byte): c / 255 for
    # all 256 bytes; the reciprocal product alone is wrong for about half
    assert lib.check_unorm8(0) == 0
    assert lib.check_unorm8(1) > 100


def test_sign_of_sum_without_the_addition(lib):
    # the splat's hemisphere tests: dot < 0 as s < -t without the last
    # addition (exact cancellations, signed zeros, subnormals included)
    assert lib.check_sign_of_sum(2_000_000, 99, 0) == 0
    assert lib.check_sign_of_sum(200_000, 99, 1) > 1000


def first_collect_counted(passing, n_test=16, n_collect=4):
    # ism.comp:80-103 as written: candidates in order, the first n_collect
    # that pass are kept
    keep, found = 0, 0
    for i in range(n_test):
        if (passing >> i) & 1 and found < n_collect:
            keep |= 1 << i
            found += 1
    return keep


def first_collect_bits(passing, n_collect=4):
    # the splat kernel's form: clear the lowest set bit n_collect times and
    # keep what was cleared
    rest = passing
    for _ in range(n_collect):
        rest &= rest - 1
    return passing ^ rest


def test_first_four_passing_candidates_by_bit_clearing():
    # the splat's keep mask: every 16-bit set of passing candidates
    masks = np.arange(1 << 16, dtype=np.uint32)
    rest = masks.copy()
    for _ in range(4):
        rest &= rest - np.uint32(1)
    got = masks ^ rest
    want = np.array([first_collect_counted(int(m)) for m in masks], dtype=np.uint32)
    assert np.array_equal(got, want)
    # set bits 0, 9, 10, 12, 13, 15: the first four are 0, 9, 10, 12
    assert first_collect_bits(0b1011_0110_0000_0001) == 0b0001_0110_0000_0001
    # a planted mistake (three clears) is caught
    assert any(first_collect_bits(int(m), 3) != int(w) for m, w in zip(masks[:64], want[:64]))
