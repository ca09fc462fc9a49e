"""Strands and tissue at time steps above the reference's 0.02 ms, up to the
FP32 limit LQT_FP32_MAX_DT (0.1 ms).

At dt = 0.1 ms the slow gates' series step works at z = dt/tau up to its
bound of 0.125, and the forward-Euler diffusion needs a coarser mesh to stay
stable (D dt / dx^2 <= 1/2 on a strand, 1/6 in 3D), so dt 0.05 runs at dx
0.15 mm and dt 0.1 at dx 0.25 mm (strands) or 0.4 mm (tissue).  FP64
against the oracle run at the same dt and dx (OneD.cc's loop, GPU3D.cu's
update): activation and APD90 to 1e-6 ms, V to 1e-6 mV; FP32 against FP64:
activation within one step, APD90 within 1 ms, QT within 1 ms.
"""
import numpy as np
import pytest

from test_gpu_tissue import _blob, _oracle_tissue

pytestmark = pytest.mark.gpu

CASES = [(0.05, 0.15), (0.1, 0.25)]


def _specs(lqt, dt, dx):
    S = lqt.StrandSpec
    rec = int(round(0.2 / dt))  # pseudo-ECG every 0.2 ms
    return [
        S(lqt.transmural(100), lqt.preset("config1"), bcl=1000.0, nbeats=2, dx=dx, rec_every=rec),
        S(lqt.transmural(80, 0.3, 0.35), lqt.preset("cpu_default", iks_scale=0.5, suprep=1, dose_fmol=60.0),
          bcl=1000.0, nbeats=2, dx=dx, rec_every=rec),
    ]


def _oracle(O, spec, dt):
    op = O.params(iks_scale=spec.params.iks_scale, iso=spec.params.iso, pca_iso=spec.params.pca_iso,
                  suprep=spec.params.suprep, dose_fmol=spec.params.dose_fmol)
    stop = spec.nbeats * spec.bcl
    r = O.strand_run(spec.celltype, spec.coeff, op, dx=spec.dx, dt=dt, bcl=spec.bcl, stop_time=stop,
                     stim=spec.stim_strength, stim_dur=spec.stim_duration, stim_start=spec.stim_start,
                     nstim=spec.nstim, rec_every=spec.rec_every, act_beat=spec.nbeats - 1)
    L = spec.celltype.size
    phi = O.pseudo_ecg(r["v"], dx=spec.dx, x0=L * spec.dx + 20.0)
    keep = r["t"] >= (spec.nbeats - 1) * spec.bcl
    t_q = (spec.nbeats - 1) * spec.bcl + spec.stim_start
    r["qt"] = O.qt_from_phi(r["t"][keep], phi[keep], t_q, r["act"].max() + 10.0, spec.nbeats * spec.bcl)
    return r


@pytest.mark.parametrize("dt,dx", CASES)
def test_strands_at_coarse_dt(lqt, oracle, dt, dx):
    specs = _specs(lqt, dt, dx)
    ref = [_oracle(oracle, s, dt) for s in specs]
    out = {}
    for prec in (lqt.FP64, lqt.FP32):
        st = lqt.Strands(specs, prec, dt)
        st.run()
        res = st.results()
        out[prec] = [(st.cell_results(i), res[i], st.final_state(i)) for i in range(len(specs))]
        st.close()
    for i, r in enumerate(ref):
        (act, apd, ead), res, fin = out[lqt.FP64][i]
        assert res["steps"] == r["steps"], (dt, i, res["steps"], r["steps"])
        assert np.all(r["act"] > 0) and np.all(r["apd"] > 100.0), (dt, i)  # every cell beats
        assert np.max(np.abs(act - r["act"])) < 1e-6, (dt, i)
        assert np.max(np.abs(apd - r["apd"])) < 1e-6, (dt, i)
        assert np.array_equal(ead, r["ead"]), (dt, i)
        assert np.max(np.abs(fin[:, 0] - r["state"][:, 0])) < 1e-6, (dt, i)
        assert abs(res["qt"] - r["qt"]) <= 0.2 + 1e-9, (dt, i, res["qt"], r["qt"])  # one ECG sample
        (a32, p32, e32), res32, _ = out[lqt.FP32][i]
        dact = np.max(np.abs(a32 - act))
        dapd = np.max(np.abs(p32 - apd))
        print(f"dt {dt} strand {i}: FP32 max |d act| {dact:.4f} ms, max |d APD90| {dapd:.4f} ms, "
              f"QT {res32['qt']:.2f} vs {res['qt']:.2f}")
        assert dact <= dt + 1e-9, (dt, i, dact)
        assert dapd <= 1.0, (dt, i, dapd)
        assert np.array_equal(e32, ead), (dt, i)
        assert abs(res32["qt"] - res["qt"]) <= 1.0, (dt, i)


def test_tissue_at_fp32_max_dt(lqt, oracle):
    """A 3D shell at dx 0.4 mm and dt 0.1 ms: FP64 V against the oracle after
    100 steps; FP32 activation and APD90 of every voxel against FP64 over the
    first beat (the tissue's own trackers)."""
    O = oracle
    lab, fib, sites = _blob(2)
    dx, dt = 0.4, 0.1
    vt, _ = _oracle_tissue(O, lab, fib, sites, dx, dt, 0.154, 0.0385, 100, O.params())
    maps = {}
    for prec in (lqt.FP64, lqt.FP32):
        T = lqt.Tissue(lab, fib, dx, 0.154, 0.0385, lqt.preset("cpu_default"), prec, dt)
        T.set_stimulus(sites, None, bcl=1000.0, duration=1.0, strength=-80.0)
        T.track(0.0, 1000.0)
        T.step(100)
        if prec == lqt.FP64:
            err = np.max(np.abs(T.v() - vt[-1]))
            assert err < 1e-7, err
            assert vt[-1].max() > 0, "the stimulus must have started a wave"
        T.step(9900)
        maps[prec] = T.track_results()
        T.close()
    (a64, p64), (a32, p32) = maps[lqt.FP64], maps[lqt.FP32]
    assert np.all(a64 >= 0) and np.all(p64 > 100.0), "every voxel activates and repolarises"
    assert np.array_equal(a32 >= 0, a64 >= 0) and np.array_equal(p32 >= 0, p64 >= 0)
    dact = np.max(np.abs(a32 - a64))
    dapd = np.max(np.abs(p32 - p64))
    print(f"tissue dt {dt}: {a64.size} voxels, FP32 max |d act| {dact:.4f} ms, max |d APD90| {dapd:.4f} ms")
    assert dact <= dt + 1e-9, dact
    assert dapd <= 1.0, dapd
