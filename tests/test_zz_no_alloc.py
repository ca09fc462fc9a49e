"""lqt_mlp_forward_noalloc, the surrogate's forward pass that allocates
nothing, against lqt_mlp_forward on the same network and inputs: the same
bits on two calls in a row (the second smaller, so rows of the first call
left in the library's scratch must not leak into it), at the row limit,
and a clear refusal beyond it.  The test harness allocates the input and
output device buffers itself through the HIP runtime the library uses."""
import ctypes as C

import numpy as np
import pytest

import lqt1_amd as A
from lqt1_amd import _lib as L
from lqt1_amd import surrogate as S

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost
ROW_LIMIT_H128 = 11904


def _ok(rc):
    assert rc == 0, f"HIP error {rc}"


def _run_noalloc(hip, net, x, stream):
    """x (host) -> device, the no-allocation forward, device -> host."""
    n = x.shape[0]
    dx, dy = C.c_void_p(), C.c_void_p()
    _ok(hip.hipMalloc(C.byref(dx), x.nbytes))
    _ok(hip.hipMalloc(C.byref(dy), n * net.n_out * 4))
    try:
        _ok(hip.hipMemcpy(dx, L.ptr(x), x.nbytes, H2D))
        junk = np.full((n, net.n_out), np.nan, np.float32)  # outputs must be written
        _ok(hip.hipMemcpy(dy, L.ptr(junk), junk.nbytes, H2D))
        L.check(L.lib().lqt_mlp_forward_noalloc(net.h, n, dx, dy, stream))
        _ok(hip.hipStreamSynchronize(stream))
        y = np.zeros((n, net.n_out), np.float32)
        _ok(hip.hipMemcpy(L.ptr(y), dy, y.nbytes, D2H))
        return y
    finally:
        hip.hipFree(dx)
        hip.hipFree(dy)


def test_noalloc_forward_equals_forward(lqt, hip):
    stream = C.c_void_p()
    _ok(hip.hipStreamCreate(C.byref(stream)))
    net = S.MLP(seed=11)
    rng = np.random.default_rng(5)
    try:
        # two calls in a row, the second with a smaller input, then the limit
        for n in (5000, 777, ROW_LIMIT_H128):
            x = rng.uniform(-1.5, 1.5, (n, S.N_IN)).astype(np.float32)
            ref = net.forward(x)
            got = _run_noalloc(hip, net, x, stream)
            assert np.isfinite(got).all()
            assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), f"n={n}"
        # the null stream works too
        x = rng.uniform(-1.0, 1.0, (64, S.N_IN)).astype(np.float32)
        assert np.array_equal(_run_noalloc(hip, net, x, None).view(np.uint32), net.forward(x).view(np.uint32))
    finally:
        hip.hipStreamDestroy(stream)
        net.close()


def test_noalloc_forward_refuses_beyond_scratch(lqt):
    net = S.MLP(seed=1)
    try:
        st = L.lib().lqt_mlp_forward_noalloc(net.h, ROW_LIMIT_H128 + 1, C.c_void_p(16), C.c_void_p(16), None)
        assert st == L.LQT_ERR_ARG
        msg = L.lib().lqt_last_error().decode()
        assert "scratch limit of 11904 rows" in msg, msg
        assert L.lib().lqt_mlp_forward_noalloc(net.h, 0, C.c_void_p(16), C.c_void_p(16), None) == L.LQT_ERR_ARG
    finally:
        net.close()
