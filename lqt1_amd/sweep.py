"""Surrogate-steered dose sweep, end to end on one GPU.

    python -m lqt1_amd.sweep [--train 20000] [--beats 10] [--grid 100]
                             [--epochs 200] [--plain] [--json out.json]

1. Training data: `--train` strands at seeded random parameters (IKs U[0,1],
   ISO {0,1} with PCa x 4 under ISO as in GPU3D.cu:767, dose U[0,1] x 300
   fmol, BCL {600,800,1000,1500}, length U{50..400}, M fraction
   U[0.2,0.5]) through the batched strand solver.
2. The 6->128->128->5 GELU MLP is trained on the GPU (BF16 MFMA, FP32
   accumulation, Adam).
3. Sweep: grid x grid IKs scale x dose at ISO off/on (BCL 1000 ms, 100 cells,
   M fraction 0.35); every point within 10 ms of the QT threshold or with an
   EAD probability in (0.2, 0.8) is re-run by the solver.  One sweep per QT
   threshold (`--qt-threshold`, or `--qt-percentiles` of the training QT
   and the clinical `--qt-thresholds-ms`).
4. With `--plain`, the whole grid is also run by the solver once, and the
   steered decisions of every threshold are compared with it point by point.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import surrogate as S


def parser():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=20000)
    ap.add_argument("--beats", type=int, default=10)
    ap.add_argument("--grid", type=int, default=100)
    ap.add_argument("--epochs", type=int, default=200)
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--qt-threshold", type=float, default=None, help="one QT threshold (ms); default: --qt-percentiles")
    ap.add_argument("--qt-percentiles", default="50,75,90,97.5",
                    help="QT thresholds as percentiles of the training QT at BCL 1000 ms, one sweep each")
    ap.add_argument("--qt-thresholds-ms", default="450",
                    help="further QT thresholds in ms (450: a clinical long-QT limit; at BCL 1000 ms QT is "
                         "its own rate-corrected value), one sweep each; '' for none")
    ap.add_argument("--pca-iso", type=float, default=S.PCA_ISO, help="PCa multiplier under ISO")
    ap.add_argument("--plain", action="store_true")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--json", default=None)
    return ap


def run(a, log=None):
    """The whole pipeline for parsed arguments `a`; returns the report dict."""
    if log is None:
        def log(*x):
            print(*x, file=sys.stderr, flush=True)

    def prog(done, total):
        log(f"  solver: {done}/{total} strands")

    out = dict(train_strands=a.train, beats=a.beats, grid=a.grid, pca_iso=a.pca_iso)
    pts = S.random_points(a.train, seed=a.seed)
    t0 = time.perf_counter()
    res, t_dev = S.simulate(pts, nbeats=a.beats, progress=prog, pca_iso=a.pca_iso)
    t_gen = time.perf_counter() - t0
    cs = sum(p.length * a.beats * p.bcl / 0.02 for p in pts)
    out.update(gen_s=t_gen, gen_strands_per_s=a.train / t_gen, gen_cell_steps_per_s=cs / t_gen)
    log(f"training data: {a.train} strands x {a.beats} beats in {t_gen:.1f} s "
        f"({a.train / t_gen:.1f} strands/s, {cs / t_gen / 1e9:.2f} G cell-steps/s)")
    x = S.points_features(pts)
    y, m = S.targets(res, pts)
    out["ead_fraction"] = float(y[:, S.N_REG].mean())
    out["qt_not_measurable_fraction"] = float(np.mean([r["qt"] < 0 for r in res]))
    net = S.MLP(seed=a.seed)
    t0 = time.perf_counter()
    losses = net.train(x, y, m, epochs=a.epochs, batch=a.batch, lr=a.lr, seed=a.seed)
    t_train = time.perf_counter() - t0
    nb = a.train // a.batch
    out.update(train_s=t_train, loss_first_epoch=float(losses[:nb].mean()), loss_last_epoch=float(losses[-nb:].mean()))
    log(f"training: {a.epochs} epochs in {t_train:.2f} s, loss {losses[:nb].mean():.4f} -> {losses[-nb:].mean():.4f}; "
        f"EAD in {out['ead_fraction']:.3f} of the training strands")
    # held-out error on fresh strands
    hold = S.random_points(512, seed=a.seed + 1)
    hres, _ = S.simulate(hold, nbeats=a.beats, pca_iso=a.pca_iso)
    hy, hm = S.targets(hres)
    qt_p, apd_p, p_p = S.decode(net.forward(S.points_features(hold)))
    ok = hm[:, 0] > 0
    out["heldout_qt_mae_ms"] = float(np.mean(np.abs(qt_p[ok] - (hy[ok, 0] * S.T_SCALE + S.T_SHIFT))))
    out["heldout_ead_acc"] = float(np.mean((p_p >= 0.5) == (hy[:, S.N_REG] > 0.5)))
    out["heldout_ead_fraction"] = float(hy[:, S.N_REG].mean())
    log(f"held-out: QT MAE {out['heldout_qt_mae_ms']:.2f} ms, EAD accuracy {out['heldout_ead_acc']:.3f} "
        f"(EAD in {out['heldout_ead_fraction']:.3f})")

    # QT thresholds: the given one, else percentiles of the training QT at
    # BCL 1000 ms of the strands without an EAD (the median puts the
    # threshold where QT values are densest, so its band holds the most grid
    # points)
    q = np.array([r["qt"] for r, p in zip(res, pts) if p.bcl == 1000.0 and r["qt"] > 0 and not r["ead"]])
    if a.qt_threshold is not None:
        thrs = [(None, a.qt_threshold)]
    else:
        pcts = [float(x) for x in a.qt_percentiles.split(",") if x.strip()]
        thrs = [(pc, float(np.percentile(q, pc)) if q.size else 450.0) for pc in pcts]
        thrs += [(None, float(x)) for x in a.qt_thresholds_ms.split(",") if x.strip()]
    g = np.linspace(0.0, 1.0, a.grid)
    I, D, SS = np.meshgrid(g, g, [0.0, 1.0], indexing="ij")
    grid = [S.Point(float(i), int(s), float(d), 1000.0, 100, 0.35) for i, d, s in zip(I.ravel(), D.ravel(), SS.ravel())]
    # the network's own EAD probability per grid point (which points the
    # EAD band sends to the solver)
    _, _, p_grid = S.decode(net.forward(S.points_features(grid)))
    ead_band = (p_grid > 0.2) & (p_grid < 0.8)
    out["grid_ead_band_points"] = int(ead_band.sum())
    plain = None
    if a.plain:
        t0 = time.perf_counter()
        pres, _ = S.simulate(grid, nbeats=a.beats, progress=prog, pca_iso=a.pca_iso)
        t_plain = time.perf_counter() - t0
        plain = (S.qt_of(pres), np.array([bool(r["ead"]) for r in pres]))
        out["t_plain_s"] = t_plain
        out["plain_ead_points"] = int(plain[1].sum())
        log(f"plain sweep of {len(grid)} points: {t_plain:.1f} s, EAD at {out['plain_ead_points']} points")
    runs = []
    for pc, thr in thrs:
        sw = S.sweep(net, g, g, (0, 1), qt_threshold=thr, band_ms=10.0, nbeats=a.beats, pca_iso=a.pca_iso)
        n_pts = sw.qt.size
        chk = sw.rechecked
        r = dict(qt_percentile=pc, qt_threshold=thr, grid_points=int(n_pts), rechecked=int(chk.sum()),
                 rechecked_ead_band=int((chk & ead_band).sum()), rechecked_ead_positive=int(sw.ead[chk].sum()),
                 t_surrogate_s=sw.t_surrogate, t_recheck_s=sw.t_recheck,
                 recheck_qt_mae_ms=float(np.nanmean(np.abs(sw.qt_err))) if np.isfinite(sw.qt_err).any() else 0.0,
                 recheck_ead_agree=sw.ead_agree)
        log(f"sweep at QT threshold {thr:.1f} ms: {n_pts} points, surrogate {sw.t_surrogate * 1e3:.1f} ms, "
            f"re-checked {chk.sum()} ({r['rechecked_ead_band']} by the EAD band) in {sw.t_recheck:.2f} s")
        if plain is not None:
            qt_p, ead_p = plain
            lq_p = S.long_qt(qt_p, thr, ead_p)
            r.update(speedup=out["t_plain_s"] / (sw.t_surrogate + sw.t_recheck),
                     rechecked_identical=bool(np.array_equal(sw.long_qt[chk], lq_p[chk]) and
                                              np.array_equal(sw.ead[chk], ead_p[chk])),
                     agree_long_qt=float(np.mean(sw.long_qt == lq_p)),
                     agree_ead=float(np.mean(sw.ead == ead_p)))
            log(f"  speed-up {r['speedup']:.1f}x, decisions agree {r['agree_long_qt']:.4f} (QT) "
                f"{r['agree_ead']:.4f} (EAD), re-checked identical {r['rechecked_identical']}")
        runs.append(r)
    out.update(runs[0])  # the first threshold, flat, as before
    out["thresholds"] = runs
    net.close()
    return out


def main(argv=None):
    a = parser().parse_args(argv)
    out = run(a)
    print(json.dumps(out), flush=True)
    if a.json:
        with open(a.json, "w") as f:
            json.dump(out, f, indent=1)
    return out


if __name__ == "__main__":
    main()
