"""Beat measurements the reference does not define (QT, APD90 per cell, EAD),
pinned on synthetic waveforms with known answers, for both the library's
QT (lqt_ecg_qt) and the oracle's restatements."""
import numpy as np
import pytest

import lqt1_amd as A
from lqt1_amd import _lib as L
import oracle as O


def _ap(t, t_act, apd90, rest=-87.0, peak=40.0, ead_at=None, ead_mv=0.0):
    """Piecewise-linear AP: upstroke over 1 ms, plateau sloping to the 90 %
    level at t_act + apd90, then back to rest; optional bump for an EAD."""
    v = np.full_like(t, rest)
    lvl = peak - 0.9 * (peak - rest)
    up = (t >= t_act) & (t < t_act + 1.0)
    v[up] = rest + (peak - rest) * (t[up] - t_act)
    pl = (t >= t_act + 1.0) & (t <= t_act + apd90)
    v[pl] = peak + (lvl - peak) * (t[pl] - t_act - 1.0) / (apd90 - 1.0)
    dn = (t > t_act + apd90) & (t < t_act + apd90 + 20.0)
    v[dn] = lvl + (rest - lvl) * (t[dn] - t_act - apd90) / 20.0
    if ead_at is not None:
        bump = np.exp(-0.5 * ((t - ead_at) / 5.0) ** 2) * ead_mv
        v = v + np.where(pl, bump, 0.0)
    return v


def test_tracker_apd_and_activation():
    t = np.arange(0, 1000, 0.02)
    v = _ap(t, 60.0, 300.0)
    act, apd, ead = O.track_trace(t, v)
    # -30 mV is crossed at 60 + 57/127 ms on the upstroke
    assert abs(act - (60.0 + 57.0 / 127.0)) <= 0.02 + 1e-9
    assert abs(apd - (300.0 - 57.0 / 127.0)) <= 0.04
    assert ead == 0


def test_tracker_ead_and_failure():
    t = np.arange(0, 1000, 0.02)
    v = _ap(t, 60.0, 400.0, ead_at=300.0, ead_mv=8.0)
    assert O.track_trace(t, v)[2] & 1
    # a small bump (< 2 mV) is not an EAD, nor is the early dome
    assert O.track_trace(t, _ap(t, 60.0, 400.0, ead_at=300.0, ead_mv=1.0))[2] == 0
    assert O.track_trace(t, _ap(t, 60.0, 400.0, ead_at=100.0, ead_mv=8.0))[2] == 0
    # never repolarising within the beat
    w = _ap(t, 60.0, 2000.0)
    assert O.track_trace(t, w)[2] & 2
    # entering the beat still depolarised from the one before (its first V
    # above -60 mV), even if it then falls to rest within the beat
    late = np.where(t < 150.0, -20.0 - 0.4 * t, -87.0)
    assert O.track_trace(t, late)[2] & 2
    assert O.track_trace(t, np.where(t < 150.0, -61.0, -87.0))[2] == 0


def _ecg(t, t_q, qrs_ms, t_peak, t_width, amp=0.06):
    phi = np.zeros_like(t)
    q = (t >= t_q) & (t < t_q + qrs_ms)
    phi[q] = 0.28 * np.sin(np.pi * (t[q] - t_q) / qrs_ms)
    phi += amp * np.exp(-0.5 * ((t - t_peak) / t_width) ** 2)
    return phi


@pytest.mark.parametrize("t_peak,width", [(330.0, 25.0), (420.0, 40.0)])
def test_qt_known_answer(t_peak, width):
    """T end is where the Gaussian T wave falls to 5 % of its peak:
    t_peak + width*sqrt(2 ln 20)."""
    dts = 0.2
    t = np.arange(0, 1000, dts)
    t_q = 50.0
    phi = _ecg(t, t_q, 60.0, t_peak, width)
    expect = t_peak + width * np.sqrt(2 * np.log(20.0)) - t_q
    q_lib = L.ecg_qt(phi, 0.0, dts, t_q, t_q + 70.0, 1000.0)
    q_or = O.qt_from_phi(t, phi, t_q, t_q + 70.0, 1000.0)
    assert abs(q_lib - expect) <= dts + 1e-9
    assert abs(q_or - q_lib) < 1e-9
    # inverted T wave: same QT
    q_inv = L.ecg_qt(np.where(t > t_q + 70, -phi, phi), 0.0, dts, t_q, t_q + 70.0, 1000.0)
    assert abs(q_inv - q_lib) < 1e-9


def test_qt_missing_t_end():
    t = np.arange(0, 500, 0.2)
    phi = np.where(t > 200, 0.05, 0.0)  # never returns
    assert L.ecg_qt(phi, 0.0, 0.2, 50.0, 120.0, 500.0) == -1.0
    assert O.qt_from_phi(t, phi, 50.0, 120.0, 500.0) == -1.0


def test_map_tracker_is_the_beat_tracker():
    """oracle.MapTracker (the tissue maps' rule, many voxels fed one step at a
    time) gives the activation and APD90 of track_trace on every trace of a
    set that covers the cases: normal beats of varied timing, an EAD bump,
    no repolarisation within the window, no activation, noise."""
    t = np.arange(0, 800, 0.02)
    rng = np.random.default_rng(4)
    traces = [_ap(t, float(a), float(d)) for a, d in zip(rng.uniform(5, 200, 12), rng.uniform(150, 450, 12))]
    traces += [_ap(t, 60.0, 400.0, ead_at=300.0, ead_mv=8.0), _ap(t, 30.0, 2000.0), np.full_like(t, -85.0),
               _ap(t, 100.0, 250.0) + rng.normal(0.0, 0.3, t.size)]
    V = np.array(traces)
    tr = O.MapTracker(V.shape[0])
    for k in range(t.size):
        tr.update(k, V[:, k])
    act, apd = tr.results(t)
    a_step, r_step = tr.results()
    for i, v in enumerate(traces):
        a, d, _ = O.track_trace(t, v)
        assert (act[i], apd[i]) == (a, d), i
    assert a_step[-2] < 0 and act[-2] == -1.0 and apd[-3] == -1.0 and r_step[-3] < 0
