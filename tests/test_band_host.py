"""CPU checks of the banded kernel's separable-PSF passes (band_passes.h, host
build): forward and adjoint against a direct correlation and its transpose, on
buffers of the kernel's exact sizes with NaN guard zones, over image sizes,
supersampling, PSF widths, ranks and band heights (bands that do and do not
divide the image, bands wider than the image).  No GPU needed."""
import ctypes
import os
import subprocess
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "native", "band_harness.hip")


@pytest.fixture(scope="module")
def bh():
    out = os.path.join(tempfile.mkdtemp(), "libbh.so")
    cmd = ["/opt/rocm/bin/hipcc", "-O2", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-o", out, SRC]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail(r.stderr)
    L = ctypes.CDLL(out)
    L.h_band_check.restype = ctypes.c_int
    L.h_band_check.argtypes = [ctypes.c_int] * 5 + [ctypes.c_uint, ctypes.POINTER(ctypes.c_double),
                                                    ctypes.POINTER(ctypes.c_double)]
    L.h_band_geom.restype = None
    L.h_band_geom.argtypes = [ctypes.c_int] * 5 + [ctypes.POINTER(ctypes.c_int)]
    L.h_band_check_m.restype = ctypes.c_int
    L.h_band_check_m.argtypes = [ctypes.c_int] * 5 + [ctypes.c_uint, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                                      ctypes.POINTER(ctypes.c_double)]
    return L


def run(bh, N, ss, k, R, BY, seed=0):
    ef, ea = ctypes.c_double(), ctypes.c_double()
    rc = bh.h_band_check(N, ss, k, R, BY, seed, ctypes.byref(ef), ctypes.byref(ea))
    return rc, ef.value, ea.value


CASES = [
    # (N, ss, k, R, BY): the uneven-batch geometries (11x11 Gaussian, its sub-grid kernels 21 / 33 / 43)
    (48, 2, 21, 1, 16), (64, 3, 33, 5, 9), (48, 4, 43, 1, 32), (128, 4, 43, 1, 64), (100, 3, 33, 5, 23),
    # the 83x83 Moffat at ss 4 (rank 7), the demo PSF at ss 1
    (40, 4, 83, 7, 8), (30, 1, 25, 3, 5),
    # small and degenerate: one pixel, PSF wider than the image, k = 1, even k, band of one row
    (1, 1, 1, 1, 1), (1, 3, 5, 2, 4), (2, 2, 21, 1, 3), (5, 4, 1, 1, 64), (7, 2, 4, 2, 1), (9, 3, 6, 8, 2),
    (17, 1, 11, 1, 64), (13, 4, 11, 3, 7),
]


@pytest.mark.parametrize("N,ss,k,R,BY", CASES)
@pytest.mark.parametrize("generic", [False, True])
def test_band_passes_match_direct(bh, N, ss, k, R, BY, generic):
    """Compile-time passes where (ss, kf, R) has them (the uneven workload's PSFs),
    and the generic passes on every geometry."""
    ef, ea = ctypes.c_double(), ctypes.c_double()
    rc = bh.h_band_check_m(N, ss, k, R, BY, N * 7 + ss, 5 if generic else 0, ctypes.byref(ef), ctypes.byref(ea))
    assert rc == 0, {1: "write outside a buffer", 2: "bad geometry", 3: "output not written exactly once"}.get(rc, rc)
    # FP32 accumulation over at most R * kf^2 terms against the FP64 direct sum
    assert ef.value < 2e-5 and ea.value < 2e-5, (ef.value, ea.value)


def test_fixed_pass_sets_are_exercised(bh):
    """The compile-time pass sets are the ones the cases above reach."""
    out = (ctypes.c_int * 18)()
    fixed = set()
    for N, ss, k, R, BY in CASES:
        bh.h_band_geom(N, ss, k, R, BY, out)
        fixed.add(out[15])
    assert {1, 2, 3} <= fixed and 0 in fixed


@pytest.mark.parametrize("mutate", [1, 2, 3, 4])
@pytest.mark.parametrize("N,ss,k,R,BY", [(48, 4, 43, 1, 32), (64, 3, 33, 5, 9), (13, 2, 11, 2, 5), (30, 2, 21, 1, 7)])
def test_band_checks_catch_planted_mistakes(bh, mutate, N, ss, k, R, BY):
    """A geometry mistake planted in the harness (V rows without margins, band rows
    without halo, one tap per phase too few, the PSF centre one sub-pixel off) fails the
    check: the passing cases above are not vacuous."""
    ef, ea = ctypes.c_double(), ctypes.c_double()
    rc = bh.h_band_check_m(N, ss, k, R, BY, 3, mutate, ctypes.byref(ef), ctypes.byref(ea))
    assert rc != 0 or ef.value > 1e-3 or ea.value > 1e-3, (rc, ef.value, ea.value)
