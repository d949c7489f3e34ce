"""The allocation-free dot product (ugx_dot_f64_noalloc) against ugx_dot_f64
through the C ABI: the same kernels and grid, so the same value bit for bit,
on two calls in a row (the second with a smaller input, so stale partial sums
of the first call would show), on the library's stream and on an explicit
one, plus the argument checks. The entry point has no counterpart in the
reference."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_noalloc_dot_equals_dot_on_consecutive_calls():
    import ugcore_amd as U
    U.init()                       # the set-up takes the scratch array's address
    L = U.lib()
    rng = np.random.default_rng(42)
    # sizes: many partial sums (about 8000 workgroups), then a few, then one
    # workgroup with a ragged tail
    for n in (4_000_037, 100_003, 333):
        say(f"case noalloc dot n={n}")
        a = U.DeviceArray(rng.normal(size=n))
        b = U.DeviceArray(rng.normal(size=n))
        out = U.DeviceArray(np.full(1, np.nan))
        ref = C.c_double()
        assert L.ugx_dot_f64(n, a.ptr, b.ptr, C.byref(ref), None) == 0
        for stream in (None, L.ugx_stream()):
            assert L.ugx_dot_f64_noalloc(n, a.ptr, b.ptr, out.ptr, stream) == 0, L.ugx_last_error()
            assert L.ugx_synchronize() == 0
            got = out.numpy()[0]
            assert got == ref.value, (n, got, ref.value)
        exact = float(np.dot(a.host, b.host))
        assert abs(ref.value - exact) <= 1e-12 * np.linalg.norm(a.host) * np.linalg.norm(b.host)


def test_noalloc_dot_empty_and_invalid_arguments():
    import ugcore_amd as U
    U.init()
    L = U.lib()
    say("case noalloc dot n=0 and invalid arguments")
    out = U.DeviceArray(np.full(1, np.nan))
    assert L.ugx_dot_f64_noalloc(0, None, None, out.ptr, None) == 0
    assert L.ugx_synchronize() == 0
    assert out.numpy()[0] == 0.0
    a = U.DeviceArray(np.ones(8))
    assert L.ugx_dot_f64_noalloc(-1, a.ptr, a.ptr, out.ptr, None) < 0
    assert b"negative" in L.ugx_last_error()
    assert L.ugx_dot_f64_noalloc(8, a.ptr, a.ptr, None, None) < 0
    assert L.ugx_dot_f64_noalloc(8, None, a.ptr, out.ptr, None) < 0
    assert b"null" in L.ugx_last_error()
