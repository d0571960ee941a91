"""The constant-divisor division of the step kernels, checked on the CPU.

The kernels form d2p/dx/dx (and /dz/dz, and the stored (.)/dt/dt) as
(a/b)/b with the Markstein sequence q0 = a*y, r = fma(-q0, b, a),
q = fma(r, y, q0), y = RN(1/b), for first dividends a == 0 or
2^-80 <= |a| < 2^80 and divisors 2^-16 <= |b| <= 2^16 (fwi2d_step.h).  The
same sequence is compiled here for the host with C's fmaf and compared with
IEEE division for extreme and ordinary
divisors, over every significand of the binades at both ends of the
dividend range (first division) -- which puts the second division's
dividend at both ends of its range [2^-96, 2^96] -- and, in between, a
sample.  Below the range the identity is allowed to fail, which is why the
kernels test the range: a wave with any other value divides with the
fallback RN_f(RN_d(a * RN_d(1/b))), compared with IEEE division bit for bit
at the end of this file.
"""
import ctypes
import subprocess

import numpy as np
import pytest

SRC = r"""
#include <math.h>
#include <stdint.h>
#include <string.h>

static float fdiv(float a, float b, float y)
{
    const float q0 = a * y;
    const float r = fmaf(-q0, b, a);
    return fmaf(r, y, q0);
}

/* Mismatches of the one- and two-division sequences against IEEE over the
 * significands m = 0, stride, 2*stride, ... of the binade [2^e, 2^(e+1)),
 * both signs. */
long check(float b, int e, unsigned stride, long *bad2)
{
    volatile float vb = b, one = 1.0f;
    const float y = one / vb;
    long bad1 = 0;
    *bad2 = 0;
    for (uint32_t m = 0; m < (1u << 23); m += stride) {
        for (int sgn = 0; sgn < 2; ++sgn) {
            uint32_t bits = ((uint32_t)sgn << 31) | ((uint32_t)(127 + e) << 23) | m;
            float a;
            memcpy(&a, &bits, 4);
            volatile float va = a;
            const float q = va / vb;
            const float qq = q / vb;
            const float f = fdiv(a, b, y);
            if (f != q) ++bad1;
            if (fdiv(f, b, y) != qq) ++*bad2;
        }
    }
    return bad1;
}
"""

DIVISORS = [2.0 ** 16, 2.0 ** -16, 12.345, 7e-4, 10.0, 20.0, 0.001, 0.002, 3.0, 6.25]


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    d = tmp_path_factory.mktemp("fastdiv")
    (d / "fastdiv.c").write_text(SRC)
    lib = d / "libfastdiv.so"
    subprocess.run(["gcc", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-o", str(lib), str(d / "fastdiv.c"),
                    "-lm"], check=True)
    L = ctypes.CDLL(str(lib))
    L.check.argtypes = [ctypes.c_float, ctypes.c_int, ctypes.c_uint, ctypes.POINTER(ctypes.c_long)]
    L.check.restype = ctypes.c_long

    def run(b, e, stride=1):
        bad2 = ctypes.c_long()
        bad1 = L.check(float(np.float32(b)), e, stride, ctypes.byref(bad2))
        return bad1, bad2.value
    return run


@pytest.mark.parametrize("b", DIVISORS)
def test_identity_at_the_ends_of_the_range(checker, b):
    for e in (-80, 79):                       # every significand of the extreme binades
        assert checker(b, e) == (0, 0), (b, e)
    for e in (-79, -40, -1, 0, 1, 40, 78):    # a sample in between
        assert checker(b, e, 97) == (0, 0), (b, e)


def test_identity_can_fail_below_the_range(checker):
    # with subnormal intermediates the sequence is no longer exact for every
    # divisor: the kernels must test the range (here: some divisor fails)
    fails = sum(sum(checker(b, -126, 31)) for b in DIVISORS)
    assert fails > 0


# The kernels' fallback division: RN_f(RN_d(a * RN_d(1/b))) must equal the
# IEEE quotient RN_f(a/b) bit for bit (signs of zero included) for every float
# a, so that a wave with any out-of-range lane can use it for all its lanes.

def _ddiv(a32, b):
    y = 1.0 / np.float64(np.float32(b))
    return (a32.astype(np.float64) * y).astype(np.float32)


def _binade(e, stride=1):
    m = np.arange(0, 1 << 23, stride, dtype=np.uint32)
    bits = ((np.uint32(127 + e) << np.uint32(23)) | m) if e > -127 else m   # e = -127: the subnormals
    return np.concatenate([bits, bits | np.uint32(1 << 31)]).view(np.float32)


@pytest.mark.parametrize("b", DIVISORS)
def test_double_product_fallback_is_exact(b):
    b32 = np.float32(b)
    with np.errstate(over="ignore", under="ignore"):
        for e, stride in ((-127, 1), (-126, 3), (-100, 7), (-80, 7), (0, 1), (79, 7), (127, 3)):
            a = _binade(e, stride)
            ref = a / b32
            got = _ddiv(a, b)
            assert np.array_equal(ref.view(np.uint32), got.view(np.uint32)), (b, e)
            ref2 = ref / b32     # the second division of (a/b)/b
            got2 = _ddiv(got, b)
            assert np.array_equal(ref2.view(np.uint32), got2.view(np.uint32)), (b, e)
