"""Batched strands (lqt_strands_*) against the FP64 oracle of OneD.cc.

Each strand of a mixed batch (lengths, cell types, parameters, BCL) is
compared with its own oracle run: FP64 activation times to 1e-6 ms, V of
every cell at the end to 1e-6 mV relative scale, pseudo-ECG to 1e-6 of its
peak; FP32 activation within one step (0.02 ms), APD90 within 1 ms, QT
within 1 ms.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _specs(lqt):
    S = lqt.StrandSpec
    tm = lqt.transmural
    return [
        S(tm(100), lqt.preset("config1"), bcl=1000.0, nbeats=1),
        S(tm(57, 0.3, 0.3), lqt.preset("cpu_default"), bcl=600.0, nbeats=2),
        S(tm(130, 0.2, 0.5), lqt.preset("gpu3d", v0=-87.5), bcl=800.0, nbeats=1),
        S(tm(80), lqt.preset("cpu_default", iks_scale=0.3, suprep=1, dose_fmol=60.0), bcl=1000.0,
          nbeats=1),
    ]


def _oracle(O, spec):
    op = O.params(iks_scale=spec.params.iks_scale, iso=spec.params.iso, pca_iso=spec.params.pca_iso,
                  suprep=spec.params.suprep, dose_fmol=spec.params.dose_fmol)
    stop = spec.nbeats * spec.bcl
    r = O.strand_run(spec.celltype, spec.coeff, op, dx=spec.dx, bcl=spec.bcl, stop_time=stop,
                     stim=spec.stim_strength, stim_dur=spec.stim_duration, stim_start=spec.stim_start,
                     nstim=spec.nstim, rec_every=spec.rec_every, act_beat=spec.nbeats - 1)
    L = spec.celltype.size
    phi = O.pseudo_ecg(r["v"], dx=spec.dx, x0=L * spec.dx + 20.0)
    keep = r["t"] >= (spec.nbeats - 1) * spec.bcl
    r["phi_t"], r["phi"] = r["t"][keep], phi[keep]
    t_q = (spec.nbeats - 1) * spec.bcl + spec.stim_start
    r["qt"] = O.qt_from_phi(r["phi_t"], r["phi"], t_q, r["act"].max() + 10.0, spec.nbeats * spec.bcl)
    return r


@pytest.fixture(scope="module")
def oracle_runs(oracle, lqt):
    return [_oracle(oracle, s) for s in _specs(lqt)]


@pytest.mark.parametrize("mode", [0, 2, 3])
def test_strands_fp64_match_oracle(lqt, oracle_runs, mode):
    """Layout 2 puts the 57-, 100- and 80-cell strands (60001, 50001 and
    50001 steps) in one workgroup and the 130-cell strand in another."""
    specs = _specs(lqt)
    st = lqt.Strands(specs, lqt.FP64, 0.02)
    st.set_mode(mode)
    st.run()
    res = st.results()
    for i, (s, r) in enumerate(zip(specs, oracle_runs)):
        act, apd, ead = st.cell_results(i)
        assert res[i]["steps"] == r["steps"], (i, res[i]["steps"], r["steps"])
        assert np.max(np.abs(act - r["act"])) < 1e-6, (i, np.max(np.abs(act - r["act"])))
        ok = (apd >= 0) & (r["apd"] >= 0)
        assert np.array_equal(apd >= 0, r["apd"] >= 0)
        assert np.max(np.abs(apd[ok] - r["apd"][ok]), initial=0) < 1e-6
        assert np.array_equal(ead, r["ead"])
        fin = st.final_state(i)
        assert np.max(np.abs(fin[:, 0] - r["state"][:, 0])) < 1e-6
        t, phi = st.ecg(i)
        n = min(phi.size, r["phi"].size)
        assert n > 0
        assert np.max(np.abs(t[:n] - r["phi_t"][:n])) < 1e-6
        # the device keeps samples in float
        assert np.max(np.abs(phi[:n] - r["phi"][:n])) <= 1e-5 * max(1.0, np.max(np.abs(r["phi"])))
        assert abs(res[i]["qt"] - r["qt"]) <= 0.2 + 1e-9, (i, res[i]["qt"], r["qt"])


def test_strands_fp32_within_gates(lqt, oracle_runs):
    specs = _specs(lqt)
    st = lqt.Strands(specs, lqt.FP32, 0.02)
    st.run()
    res = st.results()
    for i, (s, r) in enumerate(zip(specs, oracle_runs)):
        act, apd, ead = st.cell_results(i)
        dact = np.max(np.abs(act - r["act"]))
        ok = (apd >= 0) & (r["apd"] >= 0)
        dapd = np.max(np.abs(apd[ok] - r["apd"][ok]), initial=0)
        print(f"strand {i}: FP32 max |d act| = {dact:.4f} ms, max |d APD90| = {dapd:.4f} ms, "
              f"QT {res[i]['qt']:.2f} vs {r['qt']:.2f}")
        assert dact <= 0.02 + 1e-9, (i, dact)
        assert dapd <= 1.0, (i, dapd)
        # every value is either the -1 marker or a plausible duration
        assert np.all((apd == -1.0) | ((apd > 50.0) & (apd < 2000.0))), (i, apd)
        assert abs(res[i]["qt"] - r["qt"]) <= 1.0, (i, res[i]["qt"], r["qt"])


def test_launch_layouts_agree_and_autotune(lqt):
    """Every launch layout (per-size classes in parallel branches, one kernel
    at the largest block, strands packed several to a workgroup, per-size
    classes with each cell split over two lanes) gives identical per-strand
    results and final states; each layout runs on its own handle so that a
    strand a layout skipped would show up as missing output, and every
    strand reports its full step count."""
    specs = _specs(lqt)
    outs = []
    for mode in (0, 1, 2, 3):
        h = lqt.Strands(specs, lqt.FP32, 0.02)
        h.set_mode(mode)
        h.run()
        for i, r in enumerate(h.results()):
            n = 0
            t = 0.0
            while t <= specs[i].nbeats * specs[i].bcl:
                t += 0.02
                n += 1
            assert r["steps"] == n, (mode, i, r["steps"], n)
        outs.append([h.cell_results(i) for i in range(len(specs))] + [h.ecg(i)[1] for i in range(len(specs))] +
                    [h.final_state(i) for i in range(len(specs))])
        h.close()
    st = lqt.Strands(specs, lqt.FP32, 0.02)
    for other in outs[1:]:
        for k, (a, b) in enumerate(zip(outs[0], other)):
            if isinstance(a, tuple):
                for x, y in zip(a, b):
                    bad = np.flatnonzero(x != y)
                    assert bad.size == 0, (k, bad, x[bad], y[bad])
            else:
                assert np.array_equal(a, b)
    mode, ms = st.autotune(500)
    assert mode in (0, 1, 2, 3) and ms[mode] == min(ms)
    st.run()
    act, _, _ = st.cell_results(0)
    assert np.array_equal(act, outs[0][0][0])


@pytest.mark.parametrize("mode", [0, 2, 3])
def test_graph_replays_are_identical(lqt, mode):
    specs = _specs(lqt)[1:3]
    st = lqt.Strands(specs, lqt.FP32, 0.02)
    st.set_mode(mode)
    outs = []
    for _ in range(3):
        st.run()
        lqt.lib().lqt_device_synchronize()
        outs.append([st.cell_results(i) for i in range(len(specs))] + [st.ecg(i)[1] for i in range(len(specs))])
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            if isinstance(a, tuple):
                for x, y in zip(a, b):
                    assert np.array_equal(x, y)
            else:
                assert np.array_equal(a, b)
