"""Golden vectors of the reduced config-3 ventricle, written by the FP64
oracle (oracle/tpord_oracle.c, OpenMP over voxels).

    python tests/golden/make_config3_fixture.py      # about 10 min on 8 threads

The case is config 3 at a fifth of the size on the same grid spacing: the
biventricular shell of lqt1_amd.geometry.ventricle_3d with every length
scaled by 0.2 on a 52 x 52 x 64 grid at 0.2 mm (52,852 voxels), rotating
fibres (-60 deg epi to +60 deg endo, D_fibre 0.154, D_cross 0.0385 mm^2/ms),
transmural endo/M/epi labels with the seed's M-cell island, IKs x 0.2 with
ISO on and PCa x 4 (GPU3D.cu:767), the apical endocardial stimulus of
-52 pA/pF for 3 ms at BCL 1000 ms with per-site offsets 0, 0.5, .., 2 ms
(GPU3D.cu:397-409 reads them from the stimulation profile), dt 0.02 ms,
the first 1000 ms (the tissue clock's own double accumulator).

The fixture is data: the case parameters (so that the test rebuilds the
same inputs, and checksums of the label and fibre arrays it must find),
the per-voxel activation and 90 % repolarisation step indices of the beat
tracker (oracle.MapTracker), V of every voxel after the last step, V of
every 4th voxel after 100 and 300 ms, and the 40 states of 32 sampled
voxels after the last step.
"""
from __future__ import annotations

import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
import oracle as O  # noqa: E402
from lqt1_amd import geometry as G  # noqa: E402

OUT = os.path.join(HERE, "config3_reduced.npz")
SHAPE, SCALE, SEED = (52, 52, 64), 0.2, 0
D_FIBRE, D_CROSS, DT, BCL, DUR, STRENGTH, T_END = 0.154, 0.0385, 0.02, 1000.0, 3.0, -52.0, 1000.0
CHECK_STEPS = (5000, 15000)  # V of every 4th voxel after 100 and 300 ms


def case():
    """The inputs: geometry, stimulus sites and offsets, step start times."""
    geo = G.ventricle_3d(shape=SHAPE, scale=SCALE, seed=SEED)
    offs = (0.5 * (np.arange(geo.stim_sites.size) % 5)).astype(np.float32)
    times = []
    t = 0.0
    while t < T_END:
        times.append(t)
        t += DT
    return geo, offs, np.array(times)


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    geo, offs, times = case()
    lab = geo.label
    mask = (lab != 255).astype(np.uint8)
    g, nvox = O.ghost_map(mask)
    cells = np.flatnonzero(mask.ravel())
    dten, drift = O.tissue_tensor(g, nvox, np.ascontiguousarray(geo.fibre.reshape(-1, 3)[cells]), D_FIBRE,
                                  D_CROSS, geo.dx)
    ct = lab.ravel()[cells].astype(np.uint8)
    p = O.params(iks_scale=0.2, iso=1, pca_iso=4.0)
    s = np.tile(O.cell_init(), (nvox, 1))
    site_k = g.ravel()[geo.stim_sites] - 1
    assert np.all(site_k >= 0)
    off64 = offs.astype(np.float64)
    tr = O.MapTracker(nvox)
    sample = np.linspace(0, nvox - 1, 32).astype(np.int64)
    checks = {}
    t0 = time.perf_counter()
    for k, t in enumerate(times):
        tc = t - np.floor(t / BCL) * BCL
        ist = np.zeros(nvox)
        on = (tc >= off64) & (tc < off64 + DUR)
        ist[site_k[on]] = STRENGTH
        O.tissue_step(g, nvox, dten, drift, geo.dx, DT, ct, p, s, ist)
        tr.update(k, s[:, 0])
        if k + 1 in CHECK_STEPS:
            checks[k + 1] = s[::4, 0].copy()
        if (k + 1) % 5000 == 0:
            print(f"step {k + 1}/{times.size} ({time.perf_counter() - t0:.0f} s)", flush=True)
    act_step, rep_step = tr.results()
    np.savez_compressed(
        OUT, shape=np.array(SHAPE), scale=SCALE, seed=SEED, d_fibre=D_FIBRE, d_cross=D_CROSS, dt=DT, bcl=BCL,
        duration=DUR, strength=STRENGTH, t_end=T_END, iks_scale=0.2, iso=1, pca_iso=4.0,
        label_sha256=digest(lab), fibre_sha256=digest(geo.fibre), stim_sites=geo.stim_sites, stim_offsets=offs,
        nsteps=times.size, act_step=act_step.astype(np.int32), rep_step=rep_step.astype(np.int32),
        v_end=s[:, 0].copy(), v_step5000_every4=checks[5000], v_step15000_every4=checks[15000],
        sample=sample, state_end_sample=s[sample].copy())
    a = act_step >= 0
    r = rep_step >= 0
    print(f"{nvox} voxels, {times.size} steps in {time.perf_counter() - t0:.0f} s: {a.sum()} activate, "
          f"{r.sum()} repolarise within the window; wrote {OUT}")


if __name__ == "__main__":
    main()
