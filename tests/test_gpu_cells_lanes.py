"""Cells on two or four lanes (lqt_cells_set_lanes(2 | 4): each cell's
ionic currents in two groups, or in four parts, on as many wavefronts that
swap them every step) against one lane per cell (lanes 1) on the same
inputs: states, V, the SingleCell.cc statistics and the trace must be equal
bit for bit, since every lane of a cell finishes every step with the same
arithmetic as the single lane.  Cell counts that leave part of the last
64-lane group idle, per-cell parameter sets, FP32 and FP64; the oracle
comparison of the other cell tests runs the automatic choice, four lanes
for a few cells.
"""
import numpy as np
import pytest

import lqt1_amd as lqt

pytestmark = pytest.mark.gpu


def _batch(n, prec, seed, lanes):
    rng = np.random.default_rng(seed)
    ct = rng.choice(np.array([3, 4, 2], np.uint8), n)
    params = [lqt.preset("config3", iks_scale=float(rng.uniform(0.0, 1.0)), iso=int(rng.integers(0, 2)),
                         suprep=1, dose_fmol=float(rng.uniform(0.0, 300.0))) for _ in range(n)]
    c = lqt.Cells(ct, params, prec, 0.02)
    c.set_lanes(lanes)
    assert c.lanes == lanes
    return c


@pytest.mark.parametrize("prec", [lqt.FP32, lqt.FP64], ids=["fp32", "fp64"])
@pytest.mark.parametrize("n", [3, 100])
def test_step_on_more_lanes_equals_single_lane(lqt, prec, n):
    rng = np.random.default_rng(7)
    ist = np.where(rng.uniform(size=n) < 0.7, -53.0, 0.0)
    dvg = rng.uniform(-0.5, 0.5, n)
    out = []
    for lanes in (1, 2, 4):
        c = _batch(n, prec, 1, lanes)
        c.step(50, ist, dvg)  # 1 ms of stimulus on most cells, a gap current on all
        c.step(7000)          # upstroke, plateau and part of repolarisation
        out.append((c.state(), c.v()))
        c.close()
    s0, v0 = out[0]
    for lanes, (s1, v1) in zip((2, 4), out[1:]):
        assert np.array_equal(v0, v1), lanes
        assert np.array_equal(s0, s1), (lanes, np.argwhere(s0 != s1)[:5])
    assert np.ptp(v0) > 20.0, "the batch must hold cells at different phases"


@pytest.mark.parametrize("prec", [lqt.FP32, lqt.FP64], ids=["fp32", "fp64"])
def test_pace_on_more_lanes_equals_single_lane(lqt, prec):
    n = 70  # a full group of 64 cells and a group of 6
    out = []
    for lanes in (1, 2, 4):
        c = _batch(n, prec, 2, lanes)
        st, tr = c.pace(bcl=800.0, nbeats=2, trace_cell=66, trace_every=25)
        out.append((st, tr, c.state()))
        c.close()
    st0, tr0, s0 = out[0]
    for lanes, (st1, tr1, s1) in zip((2, 4), out[1:]):
        assert np.array_equal(st0, st1), lanes
        assert np.array_equal(tr0, tr1), lanes
        assert np.array_equal(s0, s1), lanes
    assert tr0.shape[0] >= 3200 and tr0.shape[1] == 3
    assert np.median(st0[:, 9]) > 100.0, "the cells must beat"


def test_automatic_layout(lqt):
    one = lqt.Cells(np.array([4], np.uint8), lqt.preset("cpu_default"), lqt.FP64, 0.02)
    assert one.lanes == 4, "one cell runs on four lanes"
    one.close()
    many = lqt.Cells(np.full(1 << 18, 3, np.uint8), lqt.preset("cpu_default"), lqt.FP32, 0.02)
    assert many.lanes == 1, "2^18 cells fill every SIMD on one lane each"
    many.set_lanes(2)
    assert many.lanes == 2
    many.set_lanes(0)
    assert many.lanes == 1
    for bad in (3, 8, -1):
        with pytest.raises(lqt.LqtError):
            many.set_lanes(bad)
    many.close()
