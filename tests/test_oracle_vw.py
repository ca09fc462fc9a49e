"""The oracle's S2 candidate run (OneD_VW.cc:239-398) against the reference
program's own output, plus the properties the vulnerable-window search
relies on: no S2 is the plain strand, S2 is the `else` branch of the S1
test, and a paused run resumes (or forks) exactly."""
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "ref_oned_vw.npz")


def _vw_strand():
    ct = np.array([O.ENDO] * 25 + [O.MCELL] * 35 + [O.EPI] * 40, np.uint8)
    co = np.full(100, 0.154)
    co[63] = 0.2 * 0.154
    return ct, co


def test_vw_candidate_matches_reference_program():
    """OneD_VW.cc as shipped runs s2 = 260 ms first; the oracle reproduces its
    printed peaks (6 significant digits) and its output file (10 decimals)."""
    z = np.load(GOLD)
    assert z["s2"][0] == 260.0
    ct, co = _vw_strand()
    r = O.StrandVW(ct, co, nbeats=5, s2_start=260.0).run()
    assert abs(r.vmax[0] - z["max_endo"][0]) <= 5e-6 * abs(z["max_endo"][0]) + 1e-5
    assert abs(r.vmax[1] - z["max_epi"][0]) <= 5e-6 * abs(z["max_epi"][0]) + 1e-5
    t, v = r.v_rows()
    ref = z["v0"]
    assert v.shape[0] == ref.shape[0] == int(z["shape0"][0])
    assert np.max(np.abs(v[:, z["cells"]] - ref)) < 2e-9
    # the reference stopped at its first candidate: the peaks straddle -10 mV
    vul = (z["max_endo"][0] + 10) * (z["max_epi"][0] + 10) <= 0
    assert bool(z["found"]) and vul


def _short(ct, co, **kw):
    a = dict(nbeats=2, bcl=300.0, stim=-52.0, stim_dur=3.0, stim_start=0.0, s2_first=3, s2_count=3,
             rec_every=25)
    a.update(kw)
    return O.StrandVW(ct, co, **a)


def test_no_s2_is_the_plain_strand():
    ct = np.array([O.ENDO] * 8 + [O.MCELL] * 8 + [O.EPI] * 8, np.uint8)
    co = np.full(ct.size, 0.154)
    r = _short(ct, co, s2_duration=0.0).run()
    p = O.strand_run(ct, co, dx=0.15, bcl=300.0, stop_time=600.0, stim=-52.0, stim_dur=3.0,
                     stim_start=0.0, rec_every=0, record=False)
    assert r.step.value == p["steps"]
    assert np.array_equal(r.state, p["state"])
    r0 = _short(ct, co, s2_strength=0.0, s2_start=100.0).run()
    assert np.array_equal(r0.state, p["state"])


def test_s2_is_the_else_branch_of_s1():
    """An S2 window that overlaps the S1 window is applied only where the S1
    window is closed (OneD_VW.cc:291)."""
    ct = np.array([O.ENDO] * 12, np.uint8)
    co = np.full(ct.size, 0.154)
    # S1 window of beat 1 is [300, 303); S2 asked for [298, 306) relative to 0
    a = _short(ct, co, s2_start=298.0, s2_duration=8.0).run()
    b = _short(ct, co, s2_start=298.0, s2_duration=8.0)
    b.s2 = O.S2(0.0, 298.0, 8.0, -104.0, 3, 3)
    b.run()
    # with base (nbeats-1)*bcl = 300 the window is [598, 606) instead: the two differ
    assert not np.array_equal(a.state, b.state)
    # the same S2 split around the S1 window gives the same run
    c = _short(ct, co)
    c.s2 = O.S2(0.0, 298.0, 2.0, -104.0, 3, 3)
    c.run(pause_at=303.0 - 1e-9)
    c.s2 = O.S2(0.0, 303.0, 3.0, -104.0, 3, 3)
    c.run()
    assert np.array_equal(b.state, c.state)


def test_pause_resume_and_fork_are_exact():
    ct = np.array([O.ENDO] * 6 + [O.MCELL] * 6 + [O.EPI] * 6, np.uint8)
    co = np.full(ct.size, 0.154)
    co[9] *= 0.2
    whole = _short(ct, co, s2_start=150.0).run()
    part = _short(ct, co, s2_start=150.0).run(pause_at=420.0)
    assert part.step.value < whole.step.value
    part.run()
    assert part.step.value == whole.step.value
    assert np.array_equal(part.state, whole.state)
    assert np.array_equal(part.vmax, whole.vmax)
    t1, v1 = part.v_rows()
    t2, v2 = whole.v_rows()
    assert np.array_equal(v1, v2) and np.array_equal(t1, t2)
    # candidates sharing the S1 prefix: fork before the earliest S2
    prefix = _short(ct, co, s2_start=140.0).run(pause_at=300.0 + 140.0)
    for s2 in (140.0, 160.0, 185.5):
        f = prefix.fork(s2).run()
        fresh = _short(ct, co, s2_start=s2).run()
        assert np.array_equal(f.state, fresh.state)
        assert np.array_equal(f.vmax, fresh.vmax)


def test_candidate_sequence_accumulates_like_the_reference():
    """MODE2 moves s2 by repeated `left += 0.1` (OneD_VW.cc:418-420)."""
    s = O.vw_candidates(260.0, 0.1, 40)
    ref = [260.0]
    x = 260.0
    for _ in range(39):
        x += 0.1
        ref.append(x)
    assert np.array_equal(s, np.array(ref))
    # which is not 260 + k*0.1 in floating point
    assert not np.array_equal(s, 260.0 + 0.1 * np.arange(40))
    down = O.vw_candidates(260.0, 0.1, 5, direction=-1)
    assert down[1] == 260.0 - 0.1 and down[4] < down[3]
