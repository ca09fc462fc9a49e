"""Edge cases of the three engines on the GPU.

* A tissue of one voxel (every stencil neighbour is a ghost mirroring it, so
  the diffusion term is exactly zero) steps exactly as one cell of the cell
  engine with the same stimulus: identical bits of all 40 states in FP64,
  isotropic and anisotropic.  In FP32 the tissue stores the five slow
  concentrations as FP32 offsets from reference values after every step
  (the cell engine keeps them in double), so there every state agrees to
  1e-5 relative after 1500 steps and V to 1e-3 mV.
* Strands at the wavefront edges (2, 63, 64, 65 and 128 cells) against the
  FP64 oracle of OneD.cc, in the per-size and the packed layouts.
* Empty inputs, strands outside the documented 2..512 cells and a
  pseudo-ECG electrode on a cell are refused with an error, not run.
* Strands whose BCL and stimulus times are not multiples of dt step as the
  FP64 oracle of OneD.cc's loop (the kernel's host-evaluated time tests).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _stim_steps(dur, dt, nsteps):
    """Steps whose time t (accumulated t += dt from 0) lies in [0, dur)."""
    t, on = 0.0, []
    for _ in range(nsteps):
        on.append(0.0 <= t < dur)
        t += dt
    return np.array(on)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("aniso", [False, True])
def test_one_voxel_tissue_is_one_cell(lqt, prec, aniso):
    P = lqt.FP64 if prec == "fp64" else lqt.FP32
    lab = np.full((3, 3, 3), 255, np.uint8)
    lab[1, 1, 1] = lqt.MCELL
    fib = None
    if aniso:
        fib = np.zeros((3, 3, 3, 3), np.float32)
        fib[1, 1, 1] = (0.6, 0.8, 0.0)
    p = lqt.preset("cpu_default")
    T = lqt.Tissue(lab, fib, 0.2, 0.154, 0.0385 if aniso else 0.154, p, P, 0.02)
    assert T.nvox == 1 and T.nghost == 26
    T.set_stimulus([13], None, bcl=1000.0, duration=1.0, strength=-52.0)
    c = lqt.Cells(np.array([lqt.MCELL], np.uint8), p, P, 0.02)
    nsteps = 1500
    on = _stim_steps(1.0, 0.02, nsteps)
    k = int(on.sum())
    assert on[:k].all() and not on[k:].any()
    T.step(nsteps)
    c.step(k, istim=-52.0)
    c.step(nsteps - k)
    a, b = T.state(0), c.state()[0]
    assert a[0] > -90.0
    if prec == "fp64":
        assert np.max(np.abs(a - b)) == 0.0, np.flatnonzero(a != b)
    else:
        rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-3)
        print(f"one-voxel FP32 tissue vs cell: max rel {rel.max():.2e}, |dV| {abs(a[0] - b[0]):.2e} mV")
        assert rel.max() < 1e-5, np.flatnonzero(rel >= 1e-5)
        assert abs(a[0] - b[0]) < 1e-3


def _oracle(O, spec):
    op = O.params(iks_scale=spec.params.iks_scale, iso=spec.params.iso, pca_iso=spec.params.pca_iso,
                  suprep=spec.params.suprep, dose_fmol=spec.params.dose_fmol)
    return O.strand_run(spec.celltype, spec.coeff, op, dx=spec.dx, bcl=spec.bcl, stop_time=spec.nbeats * spec.bcl,
                        stim=spec.stim_strength, stim_dur=spec.stim_duration, stim_start=spec.stim_start,
                        nstim=spec.nstim, rec_every=0, record=False, act_beat=spec.nbeats - 1)


def test_strands_at_wavefront_edges_match_oracle(lqt, oracle):
    O = oracle
    lengths = [2, 63, 64, 65, 128]
    specs = [lqt.StrandSpec(lqt.transmural(n), lqt.preset("cpu_default"), bcl=400.0, nbeats=1, stim_start=5.0,
                            nstim=min(3, n)) for n in lengths]
    ref = [_oracle(O, s) for s in specs]
    for mode in (0, 2, 3):
        st = lqt.Strands(specs, lqt.FP64, 0.02)
        st.set_mode(mode)
        st.run()
        res = st.results()
        for i, (n, r) in enumerate(zip(lengths, ref)):
            act, apd, ead = st.cell_results(i)
            assert act.size == n and res[i]["steps"] == r["steps"], (mode, n)
            assert np.all(act > 0), (mode, n, act)
            assert np.max(np.abs(act - r["act"])) < 1e-6, (mode, n)
            fin = st.final_state(i)
            assert np.max(np.abs(fin[:, 0] - r["state"][:, 0])) < 1e-6, (mode, n)
            rel = np.max(np.abs(fin - r["state"]) / np.maximum(np.abs(r["state"]), 1e-6))
            assert rel < 1e-6, (mode, n, rel)
        st.close()


def test_empty_inputs_are_refused(lqt):
    with pytest.raises(lqt.LqtError):
        lqt.Tissue(np.full((4, 4, 4), 255, np.uint8))
    with pytest.raises(lqt.LqtError):
        lqt.Strands([])
    for n in (1, 513):
        with pytest.raises(lqt.LqtError):
            lqt.Strands([lqt.StrandSpec(lqt.transmural(n), lqt.preset("cpu_default"), nbeats=1)])
    # the strand kernel carries step indices in 32 bits: 2^30 steps are refused
    with pytest.raises(lqt.LqtError, match="2\\^30"):
        lqt.Strands([lqt.StrandSpec(lqt.transmural(20), lqt.preset("cpu_default"), nbeats=1,
                                    stop_time=(1 << 30) * 0.02 + 1.0)])
    with pytest.raises(lqt.LqtError):
        lqt.Cells(np.zeros(0, np.uint8))
    # a pseudo-ECG electrode on a cell (here cell 0) would weight it by 1/0
    n = 20
    with pytest.raises(lqt.LqtError, match="electrode"):
        lqt.Strands([lqt.StrandSpec(lqt.transmural(n), lqt.preset("cpu_default"), nbeats=1,
                                    electrode_mm=-(n * 0.15))])


def test_longest_strand_matches_oracle(lqt, oracle):
    """The longest strand a handle takes, 512 cells: 8 wavefronts, a full
    512-lane workgroup with two wavefronts per SIMD, where the FP64 update
    spills part of its registers to scratch.  One beat of 600 ms in launch
    layouts 0, 1 and 2: FP64 activation and APD90 of every cell to 1e-6 ms
    and final V to 1e-6 mV against the oracle of OneD.cc's loop; FP32
    activation within one step and APD90 within 1 ms of the oracle."""
    O = oracle
    spec = lqt.StrandSpec(lqt.transmural(512, 0.3, 0.35), lqt.preset("cpu_default", iks_scale=0.5), bcl=600.0,
                          nbeats=1)
    pp = spec.params
    r = O.strand_run(spec.celltype, spec.coeff,
                     O.params(iks_scale=pp.iks_scale, iso=pp.iso, pca_iso=pp.pca_iso, suprep=pp.suprep,
                              dose_fmol=pp.dose_fmol),
                     dx=spec.dx, bcl=spec.bcl, stop_time=spec.bcl, stim=spec.stim_strength,
                     stim_dur=spec.stim_duration, stim_start=spec.stim_start, nstim=spec.nstim, record=False,
                     act_beat=0)
    assert np.all(r["act"] > 0), "the wave reaches the last cell"
    for prec in (lqt.FP64, lqt.FP32):
        for mode in (0, 1, 2):
            st = lqt.Strands([spec], prec, 0.02)
            st.set_mode(mode)
            st.run()
            act, apd, _ = st.cell_results(0)
            fin = st.final_state(0)
            st.close()
            assert np.array_equal(apd >= 0, r["apd"] >= 0), (prec, mode)
            ok = apd >= 0
            if prec == lqt.FP64:
                assert np.max(np.abs(act - r["act"])) < 1e-6, (mode, np.max(np.abs(act - r["act"])))
                assert np.max(np.abs(apd[ok] - r["apd"][ok]), initial=0) < 1e-6, mode
                assert np.max(np.abs(fin[:, 0] - r["state"][:, 0])) < 1e-6, mode
            else:
                assert np.max(np.abs(act - r["act"])) <= 0.02 + 1e-9, mode
                assert np.max(np.abs(apd[ok] - r["apd"][ok]), initial=0) <= 1.0, mode


def test_strand_time_tests_with_odd_protocol_match_oracle(lqt, oracle):
    """The loop's time tests arrive at the kernel as step indices evaluated on
    the host; with a BCL, stimulus start and duration that are not multiples
    of dt (and windows opening at the cycle's start, where floor(t/BCL)
    switches, one of them from a negative start) the steps they select are
    still those of the reference's loop: FP64 strands equal the oracle run
    of OneD.cc's loop."""
    O = oracle
    cases = [dict(bcl=333.3, stim_start=12.345, stim_duration=1.111),
             dict(bcl=400.0, stim_start=0.0, stim_duration=2.345),
             dict(bcl=400.0, stim_start=-1.0, stim_duration=3.0)]
    specs = [lqt.StrandSpec(lqt.transmural(24), lqt.preset("cpu_default"), nbeats=3, nstim=2, **c) for c in cases]
    st = lqt.Strands(specs, lqt.FP64, 0.02)
    st.run()
    res = st.results()
    for i, (spec, c) in enumerate(zip(specs, cases)):
        r = _oracle(O, spec)
        act, _, _ = st.cell_results(i)
        assert res[i]["steps"] == r["steps"], c
        assert np.all(act > 0) and np.max(np.abs(act - r["act"])) < 1e-6, c
        fin = st.final_state(i)
        assert np.max(np.abs(fin[:, 0] - r["state"][:, 0])) < 1e-6, c
    st.close()
