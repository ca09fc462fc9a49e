"""Storage of the five slow concentrations (nai, ki, cansr, cajsr, CaMKt)
of the FP32 tissue, checked on the FP64 oracle.

The FP32 tissue stores them after every step as FP32 offsets from reference
values (the initial values) and updates them in double.  Applying that
rounding to the oracle's own FP64 state after every step of paced single
cells (every cell class, 3 beats at BCL 1000 ms) must leave APD90 of the
last beat where unrounded FP64 puts it (to within one step) and the
concentrations within 1e-4 mM; plain FP32 accumulators, the failure the
offsets avoid, must move APD90 by more than that (ki alone, about 1e-6 mM
per step at 145 mM, is below one FP32 ULP of its value), so the check can
tell the two apart.
"""
import numpy as np
import pytest

SLOW = [1, 3, 7, 8, 39]   # nai, ki, cansr, cajsr, CaMKt in the 40-line order


def _pace(O, mode, nbeats=3, bcl=1000.0, dt=0.02):
    ct = np.array([O.ENDO, O.EPI, O.MCELL], np.uint8)
    init = O.cell_init()
    x0 = init[SLOW].copy()
    s = np.tile(init, (ct.size, 1))
    p = O.params()
    nsteps = int(round(nbeats * bcl / dt))
    last0 = int(round((nbeats - 1) * bcl / dt))
    tr = O.MapTracker(ct.size)
    times = []
    for k in range(nsteps):
        t = k * dt
        tc = t - np.floor(t / bcl) * bcl
        ist = np.full(ct.size, -53.0 if 50.0 <= tc < 51.0 else 0.0)
        O.cells_update(s, ct, p, dt, ist)
        if mode == "offset":
            s[:, SLOW] = x0 + (s[:, SLOW] - x0).astype(np.float32).astype(np.float64)
        elif mode == "fp32":
            s[:, SLOW] = s[:, SLOW].astype(np.float32).astype(np.float64)
        if k >= last0:
            tr.update(k - last0, s[:, 0])
            times.append(t)
    _, apd = tr.results(np.array(times))
    return apd, s[:, SLOW].copy()


@pytest.fixture(scope="module")
def runs(oracle):
    return {m: _pace(oracle, m) for m in ("fp64", "offset", "fp32")}


def test_offsets_keep_apd90_and_concentrations(runs):
    a64, c64 = runs["fp64"]
    a_off, c_off = runs["offset"]
    assert np.all(a64 > 200.0)
    d_apd = np.max(np.abs(a_off - a64))
    d_c = np.max(np.abs(c_off - c64))
    print(f"FP32 offsets: max |d APD90| {d_apd:.3f} ms, max |d concentration| {d_c:.2e} mM")
    assert d_apd <= 0.02 + 1e-9
    assert d_c < 1e-4


def test_plain_fp32_accumulators_drift(runs):
    a64, c64 = runs["fp64"]
    a32, c32 = runs["fp32"]
    d_apd = np.max(np.abs(a32 - a64))
    d_ki = np.max(np.abs(c32[:, 1] - c64[:, 1]))
    print(f"plain FP32: max |d APD90| {d_apd:.3f} ms, max |d ki| {d_ki:.2e} mM")
    assert d_apd > 0.05
    assert d_ki > 1e-3
