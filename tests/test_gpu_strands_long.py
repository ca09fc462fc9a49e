"""Strands in the long half of the sweep's length range (U{50..400} cells)
against the FP64 oracle of OneD.cc: 200, 257, 320 and 400 cells (4 to 7
wavefronts) at BCL 1500 ms, with a 100- and a 64-cell strand that the
packed layout puts beside the 257- and 400-cell ones (bundles of 7 and 8
wavefronts: the 7-wave strand in a full 512-lane workgroup), while layout 0
runs the 5- and 7-wave strands in its 320- and 448-lane classes and layout 3
runs the 200-, 100- and 64-cell strands split over two lanes per cell.

FP64: activation and APD90 of every cell to 1e-6 ms, V of every cell at
the end to 1e-6 mV, the pseudo-ECG to 1e-5 of its peak, QT to one sample
(0.2 ms).  FP32: activation within one step (0.02 ms), APD90 and QT within
1 ms.  Reference loop: OneD.cc:315-436.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LENGTHS = (200, 257, 320, 400, 100, 64)


def _specs(lqt):
    S, tm = lqt.StrandSpec, lqt.transmural
    pars = [lqt.preset("cpu_default", iks_scale=0.6, suprep=1, dose_fmol=80.0),
            lqt.preset("config1"),
            lqt.preset("cpu_default", iks_scale=0.2),
            lqt.preset("cpu_default", iks_scale=1.0, iso=0, suprep=1, dose_fmol=250.0),
            lqt.preset("cpu_default", iks_scale=0.4, iso=0),
            lqt.preset("cpu_default")]
    mf = (0.35, 0.25, 0.45, 0.3, 0.4, 0.2)
    return [S(tm(n, 0.5 * (1 - m), m), p, bcl=1500.0, nbeats=1) for n, p, m in zip(LENGTHS, pars, mf)]


def _oracle(O, spec):
    op = O.params(iks_scale=spec.params.iks_scale, iso=spec.params.iso, pca_iso=spec.params.pca_iso,
                  suprep=spec.params.suprep, dose_fmol=spec.params.dose_fmol)
    r = O.strand_run(spec.celltype, spec.coeff, op, dx=spec.dx, bcl=spec.bcl, stop_time=spec.nbeats * spec.bcl,
                     stim=spec.stim_strength, stim_dur=spec.stim_duration, stim_start=spec.stim_start,
                     nstim=spec.nstim, rec_every=spec.rec_every, act_beat=spec.nbeats - 1)
    L = spec.celltype.size
    phi = O.pseudo_ecg(r["v"], dx=spec.dx, x0=L * spec.dx + 20.0)
    keep = r["t"] >= (spec.nbeats - 1) * spec.bcl
    r["phi_t"], r["phi"] = r["t"][keep], phi[keep]
    t_q = (spec.nbeats - 1) * spec.bcl + spec.stim_start
    r["qt"] = O.qt_from_phi(r["phi_t"], r["phi"], t_q, r["act"].max() + 10.0, spec.nbeats * spec.bcl)
    r["v"] = None
    return r


@pytest.fixture(scope="module")
def oracle_runs(oracle, lqt):
    """One oracle strand per host thread (the oracle call releases the GIL)."""
    specs = _specs(lqt)
    n0 = oracle.num_threads()
    oracle.set_num_threads(1)
    try:
        with ThreadPoolExecutor(len(specs)) as ex:
            return list(ex.map(lambda s: _oracle(oracle, s), specs))
    finally:
        oracle.set_num_threads(n0)


def _layout_waves(lqt):
    specs = _specs(lqt)
    return [(s.celltype.size + 63) // 64 for s in specs]


@pytest.mark.parametrize("mode", [0, 2, 3])
def test_long_strands_fp64_match_oracle(lqt, oracle_runs, mode):
    specs = _specs(lqt)
    assert _layout_waves(lqt) == [4, 5, 5, 7, 2, 1]
    st = lqt.Strands(specs, lqt.FP64, 0.02)
    st.set_mode(mode)
    st.run()
    res = st.results()
    for i, (s, r) in enumerate(zip(specs, oracle_runs)):
        act, apd, ead = st.cell_results(i)
        assert res[i]["steps"] == r["steps"] == 75001, (i, res[i]["steps"], r["steps"])
        assert np.all(act > 0), i
        assert np.max(np.abs(act - r["act"])) < 1e-6, (i, np.max(np.abs(act - r["act"])))
        assert np.array_equal(apd >= 0, r["apd"] >= 0), i
        ok = apd >= 0
        assert np.max(np.abs(apd[ok] - r["apd"][ok]), initial=0) < 1e-6, i
        assert np.array_equal(ead, r["ead"]), i
        fin = st.final_state(i)
        assert np.max(np.abs(fin[:, 0] - r["state"][:, 0])) < 1e-6, i
        assert np.max(np.abs(fin - r["state"]) / np.maximum(1.0, np.abs(r["state"]))) < 1e-6, i
        t, phi = st.ecg(i)
        n = min(phi.size, r["phi"].size)
        assert n >= 7000
        assert np.max(np.abs(t[:n] - r["phi_t"][:n])) < 1e-6
        assert np.max(np.abs(phi[:n] - r["phi"][:n])) <= 1e-5 * np.max(np.abs(r["phi"])), i
        assert abs(res[i]["qt"] - r["qt"]) <= 0.2 + 1e-9, (i, res[i]["qt"], r["qt"])
    print("FP64 layout %d: QT %s ms" % (mode, [round(r["qt"], 1) for r in res]))


@pytest.mark.parametrize("mode", [0, 2, 3])
def test_long_strands_fp32_within_gates(lqt, oracle_runs, mode):
    specs = _specs(lqt)
    st = lqt.Strands(specs, lqt.FP32, 0.02)
    st.set_mode(mode)
    st.run()
    res = st.results()
    worst = [0.0, 0.0, 0.0]
    for i, (s, r) in enumerate(zip(specs, oracle_runs)):
        act, apd, _ = st.cell_results(i)
        d_act = np.max(np.abs(act - r["act"]))
        assert d_act <= 0.02 + 1e-9, (i, d_act)
        assert np.array_equal(apd >= 0, r["apd"] >= 0), i
        ok = apd >= 0
        d_apd = np.max(np.abs(apd[ok] - r["apd"][ok]), initial=0)
        assert d_apd <= 1.0, (i, d_apd)
        d_qt = abs(res[i]["qt"] - r["qt"])
        assert d_qt <= 1.0, (i, res[i]["qt"], r["qt"])
        worst = [max(worst[0], d_act), max(worst[1], d_apd), max(worst[2], d_qt)]
    print("FP32 layout %d: max |d act| %.3f ms, max |d APD90| %.3f ms, max |d QT| %.2f ms" % (mode, *worst))
