"""gvm_vec_dot_dev_noalloc (the FP64 dot product with no context and no
allocation: block partials in the library's static device array) against
gvm_vec_dot_dev, through the C ABI: bit-identical results on two calls in a
row, the second on a smaller input, and sizes beyond the maximum refused
before anything is launched."""
import ctypes as C

import numpy as np
import pytest

from gpuvmem_amd import native
from gpuvmem_amd.device import Context, DeviceBuffer
from gpuvmem_amd.problem import small_problem

pytestmark = pytest.mark.gpu


def _dot_ctx(ctx, a, b, n):
    out = DeviceBuffer(8)
    native.check(ctx.lib.gvm_vec_dot_dev(ctx.ctx, a.ptr, b.ptr, n, out.ptr), "gvm_vec_dot_dev")
    ctx.sync()
    v = out.download(np.float64)[0]
    out.free()
    return v


def _dot_noalloc(lib, a, b, n):
    out = DeviceBuffer(8)
    native.check(lib.gvm_vec_dot_dev_noalloc(a.ptr, b.ptr, n, out.ptr, None),
                 "gvm_vec_dot_dev_noalloc")
    # enqueued on the null stream: the blocking copy below waits for it
    v = out.download(np.float64)[0]
    out.free()
    return v


def test_noalloc_dot_matches_context_dot_on_two_calls():
    p = small_problem(npix=64, n_antennas=4, seed=5)
    ctx = Context(p, add_data=False)
    lib = ctx.lib
    rng = np.random.default_rng(77)
    n1, n2 = 3_000_017, 1_001
    print(f"case: noalloc dot n={n1} then n={n2}", flush=True)
    a = DeviceBuffer.from_array(rng.standard_normal(n1).astype(np.float32))
    b = DeviceBuffer.from_array(rng.standard_normal(n1).astype(np.float32))
    ref1 = _dot_ctx(ctx, a, b, n1)
    got1 = _dot_noalloc(lib, a, b, n1)
    ref2 = _dot_ctx(ctx, a, b, n2)
    got2 = _dot_noalloc(lib, a, b, n2)
    ha = a.download(np.float32).astype(np.float64)
    hb = b.download(np.float32).astype(np.float64)
    print(f"n={n1}: {got1!r} vs {ref1!r}; n={n2}: {got2!r} vs {ref2!r}", flush=True)
    assert got1 == ref1 and got2 == ref2
    # and both agree with an FP64 sum on the host
    assert abs(got1 - np.dot(ha, hb)) <= 1e-12 * np.sum(np.abs(ha * hb))
    assert abs(got2 - np.dot(ha[:n2], hb[:n2])) <= 1e-12 * np.sum(np.abs(ha[:n2] * hb[:n2]))
    # out-of-range sizes are refused on the host, with nothing launched
    big = C.c_long(native.GVM_NOALLOC_DOT_MAX_N + 1)
    for n in (0, -5, big.value):
        rc = lib.gvm_vec_dot_dev_noalloc(a.ptr, b.ptr, n, a.ptr, None)
        assert rc == native.GVM_ERR_ARG
        assert b"GVM_NOALLOC_DOT_MAX_N" in lib.gvm_last_error()
    a.free()
    b.free()
    ctx.close()
