"""Benchmark of the TPORd monodomain hot path on one MI355X.

Workload (the metric's 3D configuration): synthetic biventricular shell on a
256x256x320 grid at 0.2 mm (about 6.4M tissue voxels), rotating fibres
(D_fibre 0.154, D_cross 0.0385 mm^2/ms), transmural endo/M/epi labels plus a
5 mm M-cell island, IKs x0.2 with ISO on, apical endocardial stimulus
(-52 pA/pF, 3 ms) at BCL 1000 ms, dt 0.02 ms.

A step is one fused tissue step (diffusion + TPORd update of every voxel,
plus the ghost refresh) with all state resident in HBM.  The headline value
is FP32 cell-steps/s; the plain FP64 path is timed on the same geometry and
steps, and the per-voxel activation and APD90 maps of the first beat in
both precisions give the accuracy gate ("gate").  Further legs: the
512-strand batched sweep (strands/s), the 2D annulus sheet, the
surrogate-steered dose sweep against the plain batched sweep of the same
grid ("surrogate"), and the same 3D step on a tissue the size of the
reference's own whole-ventricle grid ("reference_grid": 690x489x537 cells
at 0.2 mm, a 24.6M-voxel shell); the FP64 oracle on the host cores is timed
beside them.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-strands]
                    [--no-sheet] [--no-surrogate] [--no-accuracy3d] [--no-refgrid]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# FP32 mixed fused step: 2 x (V 4 + 28 gates x 4 + 4 fast concentrations x 4
# + 5 slow x 8) state + label 1 + fibre 12 + compact index 4
ALG_BYTES_PER_VOXEL_STEP = 361
# the same count for the stored format of the FP32 tissue, whose 5 slow
# concentrations are FP32 offsets (4 B each): 2 x (4 + 112 + 16 + 20) + 17
FORMAT_BYTES_PER_VOXEL_STEP = 321
# rocprofv3 --pmc summary of the FP32 tissue kernel on this workload
PROFILE_SUMMARY = os.path.join(ROOT, "profiles", "tissue_step_ventricle3d.md")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def tissue_run(A, geo, prec, warmup, steps, tune_steps=20):
    T = A.Tissue(geo.label, geo.fibre, geo.dx, 0.154, 0.0385, A.preset("config3"), prec, 0.02)
    T.set_stimulus(geo.stim_sites, geo.stim_offsets, bcl=1000.0, duration=3.0, strength=-52.0)
    # launch shape by the autotuner (state restored after each trial)
    best, trial_us = T.autotune(tune_steps) if tune_steps > 0 else (T.variant(), [])
    shapes = A.Tissue.variants()
    T.step(warmup)
    A.lib().lqt_device_synchronize()
    t0 = time.perf_counter()
    T.step(steps)  # synchronises at the end
    wall = time.perf_counter() - t0
    us_main, us_ghost = T.kernel_us()
    out = dict(nvox=T.nvox, nghost=T.nghost, wall=wall, us_main=us_main, us_ghost=us_ghost,
               bytes=T.bytes_per_step(), v=T.v(), guard=T.v_guard().tolist(),
               tune=dict(shapes=[list(x) for x in shapes], trial_us=[round(u, 1) for u in trial_us], best=best),
               shape=shapes[best])
    T.close()
    return out


def accuracy_3d(A, geo, t_end=1000.0):
    """Per-voxel activation time and APD90 of the ventricle's first beat in
    FP32 and in FP64 from identical inputs (the tissue's own trackers over
    [0, t_end) ms), compared voxel by voxel."""
    import numpy as np

    maps, vend = {}, {}
    for name, prec in (("fp32", A.FP32), ("fp64", A.FP64)):
        T = A.Tissue(geo.label, geo.fibre, geo.dx, 0.154, 0.0385, A.preset("config3"), prec, 0.02)
        T.set_stimulus(geo.stim_sites, geo.stim_offsets, bcl=1000.0, duration=3.0, strength=-52.0)
        T.track(0.0, t_end)
        t, n = 0.0, 0
        while t < t_end:  # the tissue clock's own accumulator
            t += 0.02
            n += 1
        T.step(n)
        maps[name] = T.track_results()
        vend[name] = T.v()
        T.close()
    (a32, p32), (a64, p64) = maps["fp32"], maps["fp64"]
    dv = np.abs(vend["fp32"] - vend["fp64"])
    both = (a32 >= 0) & (a64 >= 0)
    rep = both & (p32 >= 0) & (p64 >= 0)
    d_act = np.abs(a32[both] - a64[both])
    d_apd = np.abs(p32[rep] - p64[rep])
    return dict(window_ms=t_end, voxels=int(a64.size), activated=int(both.sum()),
                activated_one_only=int(((a32 >= 0) != (a64 >= 0)).sum()), repolarised=int(rep.sum()),
                repolarised_one_only=int((both & ((p32 >= 0) != (p64 >= 0))).sum()),
                act_fp64_ms=[float(a64[both].min()), float(a64[both].max())] if both.any() else None,
                apd90_fp64_ms_min_p1_p50_p99_max=[float(x) for x in np.percentile(p64[rep], [0, 1, 50, 99, 100])]
                if rep.any() else None,
                d_act_ms_max=float(d_act.max()) if d_act.size else None,
                d_act_steps_over_1=int((d_act > 0.02 + 1e-9).sum()),
                d_apd90_ms_max=float(d_apd.max()) if d_apd.size else None,
                d_apd90_ms_p99=float(np.percentile(d_apd, 99)) if d_apd.size else None,
                d_apd90_over_1ms=int((d_apd > 1.0).sum()),
                v_linf_mv_at_window_end=float(dv.max()),
                v_abs_diff_mv_p999_at_window_end=float(np.percentile(dv, 99.9)))


def gate_summary(vb):
    """The metric's accuracy gate (activation within one step, APD90 within
    1 ms) over the ventricle's first beat, from accuracy_3d's counts: met
    everywhere or not, and where not, how many voxels fail it."""
    over, one = vb["d_apd90_over_1ms"], vb["repolarised_one_only"]
    act_ok = vb["d_act_steps_over_1"] == 0 and vb["activated_one_only"] == 0
    return dict(criterion="FP32 vs FP64 per voxel over the first 1000 ms: activation <= 1 dt, dAPD90 <= 1 ms",
                met_everywhere=bool(act_ok and over == 0 and one == 0), activation_within_1dt_everywhere=bool(act_ok),
                apd90_over_1ms=over, repolarised_one_precision_only=one, repolarised_both=vb["repolarised"],
                exception_fraction_of_repolarised=(over + one) / max(1, vb["repolarised"] + one),
                voxels=vb["voxels"])


def cpu_baseline(geo, steps, threads):
    """The FP64 oracle's tissue step (OpenMP over voxels) on the same
    geometry for a bounded number of steps."""
    import numpy as np

    import oracle as O

    O.set_num_threads(threads)
    lab = geo.label
    mask = (lab != 255).astype(np.uint8)
    g, nvox = O.ghost_map(mask)
    cells = np.flatnonzero(mask.ravel())
    fib = np.ascontiguousarray(geo.fibre.reshape(-1, 3)[cells], np.float32)
    dten, drift = O.tissue_tensor(g, nvox, fib, 0.154, 0.0385, geo.dx)
    del fib
    ct = lab.ravel()[cells].astype(np.uint8)
    s = np.tile(O.cell_init(), (nvox, 1))
    ist = np.zeros(nvox)
    ist[g.ravel()[geo.stim_sites] - 1] = -52.0
    p = O.params(iks_scale=0.2, iso=1, pca_iso=4.0)
    O.tissue_step(g, nvox, dten, drift, geo.dx, 0.02, ct, p, s, ist)  # first touch
    t0 = time.perf_counter()
    for _ in range(steps):
        O.tissue_step(g, nvox, dten, drift, geo.dx, 0.02, ct, p, s, ist)
    wall = time.perf_counter() - t0
    return nvox * steps / wall, wall


def shape_key(shape):
    """Launch shape (block, XCD map, waves, LDS staging) as the four
    integers the profile summary lists."""
    return " ".join(str(int(x)) for x in shape)


def kernel_name(shape, real="float"):
    """Name rocprofv3 reports for the anisotropic 3D step kernel of a launch shape."""
    b, x, w, lds = shape
    return "tissue_step_kernel<%s, true, false, %d, %s, %d, %s>" % (real, b, str(x).lower(), w, str(lds).lower())


def measured_traffic(shape):
    """Bytes per launch the FP32 tissue kernel of launch shape (block, XCD
    map, waves, LDS staging) moves beyond L2 on this workload, as the
    committed rocprofv3 --pmc summary gives them (FETCH_SIZE and WRITE_SIZE
    per dispatch, each scaled by what the counter reads for a copy kernel of
    known size); None if the summary has no line for the shape."""
    key = "traffic_bytes_per_launch %s:" % shape_key(shape)
    try:
        with open(PROFILE_SUMMARY) as f:
            for line in f:
                if line.startswith(key):
                    return float(line.split(":", 1)[1].split()[0])
    except OSError:
        pass
    return None


def strand_accuracy(A):
    """FP32 vs FP64 on the transmural strand of OneD.cc (100 cells, ISO off,
    2 beats): worst activation-time and APD90 difference over all cells and
    the QT difference (a strand check, not the ventricle's gate)."""
    import numpy as np

    spec = [A.StrandSpec(A.transmural(100), A.preset("config1"), bcl=1000.0, nbeats=2)]
    res = {}
    for prec in (A.FP64, A.FP32):
        st = A.Strands(spec, prec, 0.02)
        st.run()
        res[prec] = st.cell_results(0) + (st.results()[0]["qt"],)
        st.close()
    a64, p64, _, q64 = res[A.FP64]
    a32, p32, _, q32 = res[A.FP32]
    return dict(workload="transmural strand, 100 cells, config-1 parameters (ISO off), 2 beats at BCL 1000 ms",
                d_act_ms=float(np.max(np.abs(a32 - a64))), d_apd90_ms=float(np.max(np.abs(p32 - p64))),
                d_qt_ms=float(abs(q32 - q64)))


def strand_specs(A, nstr, nbeats, seed=0):
    """The sweep's strands: length U{50..400}, IKs U[0,1], ISO {0,1}, dose
    U[0,1]*300 fmol, BCL {600,800,1000,1500}, M fraction U[0.2,0.5]."""
    import numpy as np

    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(nstr):
        L = int(rng.integers(50, 401))
        p = A.preset("cpu_default", iks_scale=float(rng.random()), iso=int(rng.integers(0, 2)), suprep=1,
                     dose_fmol=float(300.0 * rng.random()))
        mf = float(rng.uniform(0.2, 0.5))
        specs.append(A.StrandSpec(A.transmural(L, 0.5 * (1 - mf), mf), p,
                                  bcl=float(rng.choice([600, 800, 1000, 1500])), nbeats=nbeats))
    return specs


def sheet_run(A, sim_ms=2000.0, compare_ms=250.0, seed=0):
    """The 2D annulus sheet (600 x 600 @ 0.15 mm, isotropic D 0.154, about
    97.7k voxels): S1 from the endo rim every 1000 ms (3 ms), S2 over one
    quadrant at the seed's coupling interval (3 ms, -104 pA/pF), sim_ms of
    simulated time in FP32 and in FP64.  Reports cell-steps/s of each and
    V L-inf between them after the S1 wave (compare_ms)."""
    import numpy as np

    from lqt1_amd import geometry as G

    geo = G.annulus_2d(seed=seed)
    ci = float(geo.info["s2_ci"])
    dt = 0.02
    n_cmp = int(round(compare_ms / dt))
    n_all = int(round(sim_ms / dt))
    out, vcmp = {}, {}
    for name, prec in (("fp32", A.FP32), ("fp64", A.FP64)):
        T = A.Tissue(geo.label, None, geo.dx, 0.154, 0.154, A.preset("config1"), prec, dt)
        T.set_stimulus(geo.stim_sites, None, bcl=1000.0, duration=3.0, strength=-52.0, s2_sites=geo.s2_sites,
                       s2_time=ci, s2_duration=3.0, s2_strength=-104.0)
        T.step(n_cmp)
        vcmp[name] = T.v()
        A.lib().lqt_device_synchronize()
        t0 = time.perf_counter()
        T.step(n_all - n_cmp)
        wall = time.perf_counter() - t0
        # (no per-launch event time here: an event pair around a launch of
        # about 10 us adds about 3 us of its own; rocprofv3 gives the kernel)
        out[name] = dict(cell_steps_per_s=T.nvox * (n_all - n_cmp) / wall, ms_per_step=wall / (n_all - n_cmp) * 1e3,
                         guard=T.v_guard().tolist())
        nvox = T.nvox
        T.close()
    try:
        acc_s2 = sheet_accuracy(A, geo, ci, sim_ms, dt)
    except Exception as e:  # reported, not hidden
        acc_s2 = dict(error=str(e))
    return dict(workload=f"2D annulus 600x600 @0.15 mm, isotropic D 0.154, S1 BCL 1000 ms + S2 at {ci:.0f} ms "
                         f"over a quadrant, {sim_ms:.0f} ms", voxels=nvox, steps=n_all, **out,
                accuracy_after_s2=acc_s2,
                fp32_over_fp64=out["fp32"]["cell_steps_per_s"] / out["fp64"]["cell_steps_per_s"],
                v_linf_mv_fp32_vs_fp64_at_ms=[compare_ms, float(np.max(np.abs(vcmp["fp32"] - vcmp["fp64"])))])


def _map_diff(a32, p32, a64, p64):
    """FP32 against FP64 activation / APD90 maps of one window."""
    import numpy as np

    both = (a32 >= 0) & (a64 >= 0)
    rep = both & (p32 >= 0) & (p64 >= 0)
    da = np.abs(a32[both] - a64[both])
    dp = np.abs(p32[rep] - p64[rep])
    return dict(activated=int(both.sum()), activated_one_only=int(((a32 >= 0) != (a64 >= 0)).sum()),
                repolarised=int(rep.sum()), repolarised_one_only=int((both & ((p32 >= 0) != (p64 >= 0))).sum()),
                d_act_ms_max=float(da.max()) if da.size else None,
                d_act_over_1dt=int((da > 0.02 + 1e-9).sum()),
                d_apd90_ms_p99=float(np.percentile(dp, 99)) if dp.size else None,
                d_apd90_ms_max=float(dp.max()) if dp.size else None,
                d_apd90_over_1ms=int((dp > 1.0).sum()))


def sheet_accuracy(A, geo, ci, sim_ms=2000.0, dt=0.02):
    """FP32 against FP64 on the 2D annulus after the S2: activation and
    APD90 maps of the tissue trackers over [S2, 1000 ms) (the S2-induced
    wave and re-entry) and over [1000, 2000) ms (the second S1 beat), from
    identical inputs."""
    maps = {}
    for name, prec in (("fp32", A.FP32), ("fp64", A.FP64)):
        T = A.Tissue(geo.label, None, geo.dx, 0.154, 0.154, A.preset("config1"), prec, dt)
        T.set_stimulus(geo.stim_sites, None, bcl=1000.0, duration=3.0, strength=-52.0, s2_sites=geo.s2_sites,
                       s2_time=ci, s2_duration=3.0, s2_strength=-104.0)
        out = []
        t, n = 0.0, 0
        for lo, hi in ((ci, 1000.0), (1000.0, sim_ms)):
            k = 0
            while t < lo:  # the tissue clock's own accumulator
                t += dt
                k += 1
            T.step(k)
            T.track(lo, hi)
            k = 0
            while t < hi:
                t += dt
                k += 1
            T.step(k)
            out.append(T.track_results())
        maps[name] = out
        T.close()
    return {f"window_{lo:.0f}_{hi:.0f}_ms": _map_diff(*maps["fp32"][w], *maps["fp64"][w])
            for w, (lo, hi) in enumerate(((ci, 1000.0), (1000.0, sim_ms)))}


def strands_cpu(specs, nsample, threads):
    """The FP64 oracle's strand loop (OneD.cc) on the first nsample strands
    of the sweep for one beat each, one strand per host thread; returns
    strands/s scaled to the sweep's beats and the sample's cell-steps/s."""
    from concurrent.futures import ThreadPoolExecutor

    import oracle as O

    O.set_num_threads(1)

    def one(sp):
        p = sp.params
        op = O.params(iks_scale=p.iks_scale, iso=p.iso, pca_iso=p.pca_iso, suprep=p.suprep, dose_fmol=p.dose_fmol)
        r = O.strand_run(sp.celltype, sp.coeff, op, dx=sp.dx, bcl=sp.bcl, stop_time=sp.bcl, stim=sp.stim_strength,
                         stim_dur=sp.stim_duration, stim_start=sp.stim_start, nstim=sp.nstim, record=False)
        return r["steps"] * sp.celltype.size

    sample = specs[:nsample]
    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        cs = sum(ex.map(one, sample))
    wall = time.perf_counter() - t0
    beats = specs[0].nbeats
    return dict(value=nsample / beats / wall, unit="strands/s", cores=threads, kind="port",
                cell_steps_per_s=cs / wall,
                sample=f"FP64 oracle strand loop, first {nsample} strands of the sweep for 1 of their {beats} "
                       f"beats, one strand per thread, {wall:.1f} s")


def reference_grid_run(A, steps=20, warmup=5):
    """The fused 3D step (default launch shape) on an ellipsoidal shell that
    fills GPU3D.cu's 690 x 489 x 537 grid at 0.2 mm (24.6M voxels, the size
    of the reference's whole-ventricle model), with the ventricle's
    parameters (IKs x0.2, ISO on, anisotropic), FP32 and FP64."""
    from lqt1_amd import geometry as G

    t0 = time.perf_counter()
    geo = G.ellipsoid_shell_3d()
    out = dict(workload="ellipsoidal shell on the reference's 690x489x537 grid @0.2 mm, anisotropic D_f 0.154/"
                        "D_c 0.0385, IKs x0.2 ISO on, apical stimulus", grid=[690, 489, 537],
               geometry_s=round(time.perf_counter() - t0, 1))
    for name, prec in (("fp32", A.FP32), ("fp64", A.FP64)):
        T = A.Tissue(geo.label, geo.fibre, geo.dx, 0.154, 0.0385, A.preset("config3"), prec, 0.02)
        T.set_stimulus(geo.stim_sites, geo.stim_offsets, bcl=1000.0, duration=3.0, strength=-52.0)
        T.step(warmup)
        A.lib().lqt_device_synchronize()
        t0 = time.perf_counter()
        T.step(steps)
        wall = time.perf_counter() - t0
        out["voxels"] = T.nvox
        out[name] = dict(cell_steps_per_s=T.nvox * steps / wall, ms_per_step=wall / steps * 1e3,
                         kernel_us=round(T.kernel_us()[0], 1), bytes_per_step=T.bytes_per_step(),
                         guard=T.v_guard().tolist())
        T.close()
    out["fp32_over_fp64"] = out["fp32"]["cell_steps_per_s"] / out["fp64"]["cell_steps_per_s"]
    return out


def strands_throughput(A, nstr, nbeats, seed=0):
    """Batched strand sweep (strand_specs), one persistent launch."""
    import numpy as np

    specs = strand_specs(A, nstr, nbeats, seed)
    st = A.Strands(specs, A.FP32, 0.02)
    mode, trial = st.autotune(2000)  # also captures both graphs
    t0 = time.perf_counter()
    st.run()
    wall = time.perf_counter() - t0
    cs = sum(s.celltype.size * s.nbeats * s.bcl / 0.02 for s in specs)
    r = st.results()
    out = dict(strands=nstr, beats=nbeats, wall_s=wall, strands_per_s=nstr / wall, cell_steps_per_s=cs / wall,
               layout_mode=mode, trial_ms=trial, max_block=st.block(), ead=int(sum(x["ead"] for x in r)),
               qt_median_ms=float(np.median([x["qt"] for x in r if x["qt"] > 0])))
    st.close()
    out["_specs"] = specs
    return out


def surrogate_leg(args):
    """The surrogate-steered dose sweep (config 5): training strands from the
    batched solver, the MLP trained on the GPU, the plain batched sweep of
    the 100 x 100 x 2 grid once and one steered sweep per QT threshold (the
    50th, 75th, 90th and 97.5th percentiles of the training QT), each
    compared point by point with the plain sweep."""
    from lqt1_amd import sweep as SW

    a = SW.parser().parse_args(["--train", str(args.surrogate_train), "--grid", str(args.surrogate_grid),
                                "--beats", "10", "--plain", "--seed", str(args.seed)])
    t0 = time.perf_counter()
    out = SW.run(a, log=log)
    out["wall_s"] = time.perf_counter() - t0
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--fp64-steps", type=int, default=0, help="default: --steps")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=40, help="FP64 oracle steps of the ventricle timed on the host (about 10 s on 16 threads)")
    ap.add_argument("--tune-steps", type=int, default=20, help="autotuner trial steps per launch shape (0 = off)")
    ap.add_argument("--no-strands", action="store_true")
    ap.add_argument("--no-sheet", action="store_true", help="skip the 2D annulus sheet leg")
    ap.add_argument("--no-refgrid", action="store_true",
                    help="skip the 3D step on the reference's 690x489x537 grid (24.6M voxels)")
    ap.add_argument("--no-accuracy3d", action="store_true",
                    help="skip the per-voxel FP32 vs FP64 activation / APD90 comparison over the first beat")
    ap.add_argument("--strand-beats", type=int, default=10, help="beats per strand of the 512-strand sweep")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the surrogate-steered sweep leg")
    ap.add_argument("--surrogate-train", type=int, default=20000, help="training strands of the surrogate leg")
    ap.add_argument("--surrogate-grid", type=int, default=100, help="IKs and dose points of the swept grid")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--profile", choices=["fp32", "fp64", "copy", "cells", "cell-latency", "strands",
                                          "strand-latency", "sheet"],
                    default=None,
                    help="short run of one kernel for rocprofv3 (no JSON line); cells = the ionic update "
                         "alone, --steps updates per launch with the state in registers; cell-latency = "
                         "time per step of the single-cell pacing protocol (2 beats at BCL 1000 ms) by cell "
                         "count, one, two and four lanes per cell; strands = the "
                         "512-strand sweep of --strand-beats beats; strand-latency = time per step of one "
                         "strand alone on the GPU by length (one beat at BCL 1000 ms); sheet = the 2D annulus "
                         "leg alone")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (one GPU per box, "
              "no multi-GPU path in this library)", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import lqt1_amd as A
    from lqt1_amd import geometry as G

    if A.device_count() < 1:
        raise SystemExit("no HIP device")
    dev = A.device_name()
    log("device", dev)
    if args.profile == "copy":
        # 2 GiB read + 2 GiB written per launch: the counters' reference
        nbytes = 2 << 30
        us = A._lib.bench_copy(nbytes, max(1, args.steps))
        log(f"copy {nbytes} B: {us:.1f} us per copy, {2 * nbytes / us / 1e3:.0f} GB/s read+write")
        return
    if args.profile == "strands":
        r = strands_throughput(A, 512, args.strand_beats, args.seed)
        r.pop("_specs")
        log(f"strands: {r}")
        return
    if args.profile == "sheet":
        log(f"sheet2d: {sheet_run(A)}")
        return
    if args.profile == "strand-latency":
        # one strand alone: its workgroup is the only one on its CU, so the
        # step time is the latency of one cell update at 1 or 2 wavefronts
        # per SIMD (the critical path of a sweep)
        # (layout 0: one lane per cell; layout 3: strands up to 256 cells
        # split over two lanes per cell)
        p = A.preset("cpu_default", iks_scale=0.5, iso=1, suprep=1, dose_fmol=100.0)
        for n in (64, 100, 128, 192, 256, 320, 400, 512):
            for mode in (0, 3):
                st = A.Strands([A.StrandSpec(A.transmural(n, 0.3, 0.3), p, bcl=1000.0, nbeats=1)], A.FP32, 0.02)
                st.set_mode(mode)
                st.run()
                st.run()
                us = st.last_ms() * 1e3 / (1000.0 / 0.02)
                st.close()
                log(f"strand-latency: {n} cells, {(n + 63) // 64} wavefronts, layout {mode}: {us:.2f} us per step")
        # the steered sweep's re-check strand: 100 cells, 10 beats at BCL
        # 1000 ms, its QT and EAD flag measured in the last beat only
        from lqt1_amd import surrogate as S

        spec = S.spec_of(S.Point(0.3, 1, 0.5, 1000.0, 100, 0.35), 10)
        for mode in (0, 3):
            st = A.Strands([spec], A.FP32, 0.02)
            st.set_mode(mode)
            st.run()
            st.run()
            us = st.last_ms() * 1e3 / (10 * 1000.0 / 0.02)
            st.close()
            log(f"strand-latency: re-check strand (100 cells, 10 beats), layout {mode}: {us:.2f} us per step")
        return
    if args.profile == "cell-latency":
        # the SingleCell.cc protocol on few cells: a step costs one lane's
        # dependent chain, which two or four lanes per cell shorten; 1 and
        # 16 cells leave most lanes of their wavefront idle
        for prec, pname in ((A.FP32, "fp32"), (A.FP64, "fp64")):
            for n in (1, 16, 17, 64, 4096):
                for lanes in (1, 2, 4):
                    c = A.Cells(np.full(n, A.MCELL, np.uint8), A.preset("cpu_default"), prec, 0.02)
                    c.set_lanes(lanes)
                    t0 = time.perf_counter()
                    st, _ = c.pace(bcl=1000.0, nbeats=2)
                    us = (time.perf_counter() - t0) / (2 * 1000.0 / 0.02) * 1e6
                    c.close()
                    log(f"cell-latency: {pname}, {n} cells, {lanes} lane{'s' if lanes > 1 else ''} per cell: "
                        f"{us:.2f} us per step, APD90 {st[0, 9]:.2f} ms")
        return
    if args.profile == "cells":
        # the ionic update alone on the ventricle's voxel count and type mix
        n = 6356728
        ct = np.array([A.ENDO, A.MCELL, A.EPI], np.uint8)[np.arange(n) * 3 // n]
        c = A.Cells(ct, A.preset("config3"), A.FP32, 0.02)
        c.step(max(1, args.warmup))
        A.lib().lqt_device_synchronize()
        t0 = time.perf_counter()
        c.step(max(1, args.steps))
        A.lib().lqt_device_synchronize()
        dt_s = time.perf_counter() - t0
        log(f"cells: {n} cells x {args.steps} updates in one launch: {dt_s * 1e3:.1f} ms, "
            f"{n * args.steps / dt_s / 1e9:.2f} G cell-steps/s")
        return
    t0 = time.perf_counter()
    geo = G.ventricle_3d(seed=args.seed)
    log(f"geometry {time.perf_counter() - t0:.1f} s, {geo.nvox} voxels, {geo.stim_sites.size} stimulus sites")
    if args.profile in ("fp32", "fp64"):
        r = tissue_run(A, geo, A.FP32 if args.profile == "fp32" else A.FP64, args.warmup, args.steps,
                       args.tune_steps)
        log(f"{args.profile}: shape {r['shape']} (trials {r['tune']['trial_us']} us), "
            f"kernel {r['us_main']:.1f} us, ghost {r['us_ghost']:.1f} us, "
            f"{r['bytes'] / (r['us_main'] * 1e-6) / 1e9:.0f} GB/s algorithmic, {r['bytes']:.0f} B per launch")
        return

    r32 = tissue_run(A, geo, A.FP32, args.warmup, args.steps, args.tune_steps)
    log(f"FP32: shape {r32['shape']} (trials {r32['tune']['trial_us']} us), {r32['wall'] / args.steps * 1e3:.3f} "
        f"ms/step, kernel {r32['us_main']:.1f} us, ghost {r32['us_ghost']:.1f} us")
    k64 = args.fp64_steps or args.steps
    r64 = tissue_run(A, geo, A.FP64, args.warmup, k64, args.tune_steps)
    log(f"FP64: shape {r64['shape']} (trials {r64['tune']['trial_us']} us), {r64['wall'] / k64 * 1e3:.3f} ms/step, "
        f"kernel {r64['us_main']:.1f} us")

    nvox = r32["nvox"]
    v32 = nvox * args.steps / r32["wall"]
    v64 = nvox * k64 / r64["wall"]
    acc = {}
    try:
        acc["strand_config1"] = strand_accuracy(A)
    except Exception as e:  # reported, not hidden
        acc["strand_error"] = str(e)
    gate = None
    if not args.no_accuracy3d:
        try:
            t0 = time.perf_counter()
            acc["ventricle_first_beat"] = accuracy_3d(A, geo)
            gate = gate_summary(acc["ventricle_first_beat"])
            log(f"3D accuracy ({time.perf_counter() - t0:.1f} s): {acc['ventricle_first_beat']}")
        except Exception as e:
            acc["ventricle_error"] = str(e)
    # Algorithmic bytes per launch: 361 B per voxel-step (V, 28 gates and 4
    # fast concentrations in FP32 and 5 slow ones in FP64, read and written;
    # label, fibre and compact index read) times the voxels of one launch.
    # The FP32 tissue stores the 5 slow concentrations as FP32 offsets, so
    # its own format needs 321 B (achieved_format); the implementation's
    # count adds the stored divergence of D and the ghost-cell writes
    # (bytes_per_launch_impl, achieved_impl), and the copy comparison uses
    # those bytes.  The kernel time is the mean per launch of one pair of HIP
    # events around the timed steps on the tissue's stream.
    alg_bytes = ALG_BYTES_PER_VOXEL_STEP * nvox
    sec = r32["us_main"] * 1e-6
    achieved = alg_bytes / sec / 1e9
    achieved_impl = r32["bytes"] / sec / 1e9
    traffic = measured_traffic(r32["shape"])
    # the practical streaming ceiling of this GPU: a plain 2 GiB copy
    copy_us = A._lib.bench_copy(2 << 30, 20)
    copy_gbs = 2 * (2 << 30) / (copy_us * 1e-6) / 1e9
    roof = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic,
                traffic_over_algorithmic=round(traffic / alg_bytes, 4) if traffic else None,
                traffic_over_impl=round(traffic / r32["bytes"], 4) if traffic else None,
                copy_gbs=round(copy_gbs, 1), frac_of_copy=round(achieved_impl / copy_gbs, 4),
                kernel=kernel_name(r32["shape"]),
                kernel_us=round(r32["us_main"], 2), tune=r32["tune"],
                bytes_per_voxel_step=ALG_BYTES_PER_VOXEL_STEP, bytes_per_launch=alg_bytes,
                bytes_per_voxel_step_format=FORMAT_BYTES_PER_VOXEL_STEP,
                achieved_format=round(FORMAT_BYTES_PER_VOXEL_STEP * nvox / sec / 1e9, 1),
                bytes_per_launch_impl=r32["bytes"], achieved_impl=round(achieved_impl, 1))

    strands = None
    if not args.no_strands:
        try:
            strands = strands_throughput(A, 512, args.strand_beats)
            specs = strands.pop("_specs")
            log(f"strands: {strands}")
        except Exception as e:
            strands = dict(error=str(e))

    sheet = None
    if not args.no_sheet:
        try:
            sheet = sheet_run(A)
            log(f"sheet2d: {sheet}")
        except Exception as e:
            sheet = dict(error=str(e))

    refgrid = None
    if not args.no_refgrid:
        try:
            refgrid = reference_grid_run(A)
            log(f"reference grid: {refgrid}")
        except Exception as e:
            refgrid = dict(error=str(e))

    surrogate = None
    if not args.no_surrogate:
        try:
            surrogate = surrogate_leg(args)
            log(f"surrogate: {surrogate}")
        except Exception as e:
            surrogate = dict(error=str(e))

    cpu = None
    if not args.no_cpu:
        threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
        threads = max(1, min(threads, 16))
        try:
            cps, wall = cpu_baseline(geo, args.cpu_steps, threads)
            cpu = dict(value=cps, unit="cell-steps/s", cores=threads, kind="port",
                       sample=f"FP64 oracle tissue step (diffusion + TPORd update, OpenMP over voxels), "
                              f"{args.cpu_steps} steps of the same {nvox}-voxel ventricle, {wall:.1f} s")
            if strands is not None and "error" not in strands:
                strands["cpu_baseline"] = strands_cpu(specs, threads, threads)
                strands["gpu_over_cpu"] = strands["strands_per_s"] / strands["cpu_baseline"]["value"]
        except Exception as e:
            cpu = dict(error=str(e))

    line = {
        "metric": "3D tissue cell-steps/s FP32 vs FP64 @dAPD90<=1ms; strands/s batched+surrogate",
        "value": v32,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r32["wall"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (slow concentrations accumulated in f64, stored as f32 offsets)",
        "data": "synthetic",
        "config": {"workload": "3D synthetic ventricle 256x256x320 @0.2 mm, FP32 state + FP64 slow "
                               "concentrations, anisotropic D_f 0.154/D_c 0.0385, IKs x0.2 ISO on, dt 0.02 ms",
                   "voxels": nvox, "ghost_cells": r32["nghost"], "device": dev},
        "fp64": {"value": v64, "unit": "cell-steps/s", "steps": k64, "ms_per_step": r64["wall"] / k64 * 1e3,
                 "kernel_us": round(r64["us_main"], 2), "tune": r64["tune"]},
        "fp32_over_fp64": v32 / v64,
        "accuracy": acc,
        "gate": gate,
        "strands": strands,
        "sheet2d": sheet,
        "reference_grid": refgrid,
        "surrogate": surrogate,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
