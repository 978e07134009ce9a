"""The final gather's exact NEAREST texel decision (CPU only).

Plain FP32 rounds a lit pair's ISM texel coordinate by up to ~1e-5 texel, and
at 1080p that alone moves the irradiance by ~1e-4 rel-L2 (the plain form's
bar): each flipped texel changes the pair's shadow term wholesale. The gather
therefore re-decides every pair whose FP32 coordinate lies within 2.5e-4
texel of an edge, in float-float arithmetic (about 44 bits, FP32 operations
only; no FP64 on the GPU), from the FP32 estimate of the coordinate: the
estimate names the nearest edge and the exact arithmetic decides on which
side of it the coordinate lies. Here the host build of that code
(many-lights-gi_amd/csrc/float_float.hpp) is compared with an FP64
evaluation of the same formula on random lit pairs seen by a 1080p camera in
the room, with estimates off by up to 0.45 texel: every decision farther
than 1e-10 texel from an edge must agree.
"""
import ctypes as C
import subprocess
from pathlib import Path

import numpy as np
import pytest

SRC = Path(__file__).with_name("ff_texel.cpp")


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    so = tmp_path_factory.mktemp("ff") / "libff.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", str(SRC), "-o", str(so)],
                   check=True)
    L = C.CDLL(str(so))
    L.check_ff_texel.argtypes = [C.c_long, C.c_uint64, C.c_int, C.c_double, C.POINTER(C.c_long)]
    L.check_ff_texel.restype = C.c_long
    return L


@pytest.mark.parametrize("S", [16, 32, 64])
def test_float_float_texel_equals_fp64(lib, S):
    near = C.c_long(0)
    bad = lib.check_ff_texel(2_000_000, 1234 + S, S, float(np.float32(0.01)), C.byref(near))
    assert near.value > 500  # decisions the kernel re-decides are covered
    assert bad == 0


def test_planted_mistake_is_caught(lib):
    # an FP64 reference with a slightly different lookAtRH up vector
    near = C.c_long(0)
    assert lib.check_ff_texel(200_000, 99, 64, 0.0101, C.byref(near)) > 0
