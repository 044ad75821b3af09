"""parcel_hip_dot_noalloc through the C ABI: the allocation-free dot gives the
bits of parcel_hip_dot on two calls in a row (the second one smaller, so a
stale partial left in the library's fixed scratch array would show), with the
result delivered to a host and to a device pointer, and rejects host inputs
and negative sizes with an error instead of staging them."""
import ctypes
import sys

import numpy as np
import pytest

import parcel_hip as ph

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _init():
    ph.init(0)


def _err():
    return (ph.load().parcel_hip_last_error() or b"").decode()


def _ref(n, xd, yd):
    r = ctypes.c_double(0.0)
    assert ph.load().parcel_hip_dot(n, xd.ptr, yd.ptr, ctypes.byref(r)) == 0, _err()
    return r.value


def _noalloc_host(n, xd, yd):
    r = ctypes.c_double(0.0)
    assert ph.load().parcel_hip_dot_noalloc(n, xd.ptr, yd.ptr, ctypes.byref(r)) == 0, _err()
    return r.value


def _noalloc_dev(n, xd, yd):
    out = ph.DeviceArray(n=1)
    assert ph.load().parcel_hip_dot_noalloc(n, xd.ptr, yd.ptr, out.ptr) == 0, _err()
    v = float(out.numpy()[0])
    out.free()
    return v


@pytest.mark.parametrize("n1,n2", [(3_000_001, 1000), (524_288, 257), (70_000, 1)])
def test_noalloc_dot_matches_dot(n1, n2):
    say("dot_noalloc", n1, n2)
    x = ph.uniform(n1, 11)
    y = ph.uniform(n1, 12)
    xd, yd = ph.DeviceArray(x), ph.DeviceArray(y)
    for n in (n1, n2):
        want = _ref(n, xd, yd)
        got_h = _noalloc_host(n, xd, yd)
        got_d = _noalloc_dev(n, xd, yd)
        assert np.float64(got_h).tobytes() == np.float64(want).tobytes(), (n, got_h, want)
        assert np.float64(got_d).tobytes() == np.float64(want).tobytes(), (n, got_d, want)
        # and the value itself: the fixed-order sum is within rounding of an exact one
        assert abs(got_h - float(np.dot(x[:n], y[:n]))) <= 1e-12 * max(1.0, float(np.abs(x[:n] * y[:n]).sum()))
    xd.free()
    yd.free()


def test_noalloc_dot_empty_and_errors():
    say("dot_noalloc errors")
    xd = ph.DeviceArray(np.ones(8))
    assert _noalloc_host(0, xd, xd) == 0.0
    r = ctypes.c_double(0.0)
    lib = ph.load()
    assert lib.parcel_hip_dot_noalloc(-1, xd.ptr, xd.ptr, ctypes.byref(r)) < 0
    h = np.ones(8)
    hp = ctypes.c_void_p(h.ctypes.data)
    assert lib.parcel_hip_dot_noalloc(8, hp, xd.ptr, ctypes.byref(r)) < 0
    assert "device pointers" in _err()
    # the library still works after the rejected calls
    assert _noalloc_host(8, xd, xd) == 8.0
    xd.free()
