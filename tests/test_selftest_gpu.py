"""Device arithmetic the traversal relies on.

The traversal forms ray and triangle reciprocals as the hardware estimate
refined by one fused Newton step (traverse.h: recip).  Bit-exact closest hits
need that value to equal the IEEE division 1.0f / x the reference computes;
wf_selftest_reciprocal checks it on all 2^32 float inputs."""
import ctypes as C

import pytest

from wavefrontpathtracing_amd import _lib as L

pytestmark = pytest.mark.gpu


def test_reciprocal_equals_division_on_every_float():
    mism = C.c_uint64(0)
    first = C.c_uint32(0)
    L.check(L.lib().wf_selftest_reciprocal(C.byref(mism), C.byref(first)), what="selftest_reciprocal")
    assert mism.value == 0, f"{mism.value} inputs differ from 1.0f / x; smallest bit pattern {first.value:#010x}"
    assert first.value == 0xFFFFFFFF
