"""Entry points on a caller's HIP stream (lqt_cells_step_async,
lqt_strands_run_async, lqt_tissue_step_async) give the same bits as their
synchronous forms (lqt_cells_step, lqt_strands_run, lqt_tissue_step) on the
same inputs."""
import ctypes as C

import numpy as np
import pytest

from lqt1_amd import _lib as L

pytestmark = pytest.mark.gpu

H2D = 1


def _ok(rc):
    assert rc == 0, f"HIP error {rc}"


@pytest.fixture()
def stream(hip):
    s = C.c_void_p()
    _ok(hip.hipStreamCreate(C.byref(s)))
    yield s
    hip.hipStreamDestroy(s)


@pytest.mark.parametrize("prec", [0, 1])
def test_cells_step_async_equals_step(lqt, hip, stream, prec):
    ct = np.array([lqt.ENDO, lqt.MCELL, lqt.EPI] * 40, np.uint8)
    n = ct.size
    rng = np.random.default_rng(3)
    ist = np.where(np.arange(n) % 3 == 0, -52.0, 0.0)
    gap = rng.uniform(-0.5, 0.5, n)
    a = lqt.Cells(ct, lqt.preset("cpu_default"), prec, 0.02)
    b = lqt.Cells(ct, lqt.preset("cpu_default"), prec, 0.02)
    di, dg = C.c_void_p(), C.c_void_p()
    _ok(hip.hipMalloc(C.byref(di), ist.nbytes))
    _ok(hip.hipMalloc(C.byref(dg), gap.nbytes))
    try:
        _ok(hip.hipMemcpy(di, L.ptr(ist), ist.nbytes, H2D))
        _ok(hip.hipMemcpy(dg, L.ptr(gap), gap.nbytes, H2D))
        for _ in range(3):
            L.check(L.lib().lqt_cells_step(a.h, 50, L.ptr(ist), L.ptr(gap)))
            L.check(L.lib().lqt_cells_step_async(b.h, 50, di, dg, stream))
        _ok(hip.hipStreamSynchronize(stream))
        sa, sb = a.state(), b.state()
        assert np.array_equal(sa, sb)
        assert np.ptp(sa[:, 0]) > 1.0  # the stimulated cells moved
    finally:
        hip.hipFree(di)
        hip.hipFree(dg)
        a.close()
        b.close()


@pytest.mark.parametrize("prec", [0, 1])
def test_strands_run_async_equals_run(lqt, stream, prec):
    p = lqt.preset("cpu_default", iks_scale=0.4)
    specs = [lqt.StrandSpec(lqt.transmural(n), p, bcl=400.0, nbeats=1) for n in (40, 70, 130)]
    a = lqt.Strands(specs, prec, 0.02)
    b = lqt.Strands(specs, prec, 0.02)
    try:
        a.run()
        L.check(L.lib().lqt_strands_run_async(b.h, stream))
        # a second asynchronous run before the results are read is refused
        assert L.lib().lqt_strands_run_async(b.h, stream) == L.LQT_ERR_STATE
        ra, rb = a.results(), b.results()  # the first result call waits for the stream
        assert ra == rb
        for i in range(len(specs)):
            for x, y in zip(a.cell_results(i), b.cell_results(i)):
                assert np.array_equal(x, y)
            assert np.array_equal(a.ecg(i)[1], b.ecg(i)[1])
            assert np.array_equal(a.final_state(i), b.final_state(i))
        assert b.last_ms() > 0
        # the synchronous run after an asynchronous one on the same handle
        b.run()
        assert b.results() == ra
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("prec", [0, 1])
def test_tissue_step_async_equals_step(lqt, hip, stream, prec):
    lab = np.full((6, 7, 9), lqt.ENDO, np.uint8)
    lab[:, :, 5:] = lqt.EPI
    lab[0, 0, :] = lqt.NONE
    fib = np.zeros(lab.shape + (3,), np.float32)
    fib[..., 0] = 1.0
    ts = []
    for _ in range(2):
        T = lqt.Tissue(lab, fib, 0.2, 0.154, 0.0385, lqt.preset("cpu_default"), prec, 0.02)
        sites = np.flatnonzero(np.indices(lab.shape)[2].ravel() < 2)
        sites = sites[lab.ravel()[sites] != lqt.NONE]
        T.set_stimulus(sites, None, duration=1.0, strength=-80.0)
        ts.append(T)
    try:
        ts[0].step(300)
        L.check(L.lib().lqt_tissue_step_async(ts[1].h, 300, stream))
        _ok(hip.hipStreamSynchronize(stream))
        va, vb = ts[0].v(), ts[1].v()
        assert np.array_equal(va, vb)
        assert va.max() > 0.0  # the stimulus fired
        us, _ = ts[0].kernel_us()
        assert us > 0
    finally:
        for T in ts:
            T.close()
