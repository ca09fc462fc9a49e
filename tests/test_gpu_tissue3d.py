"""The whole-ventricle 3D program (GPU3D.cu) through lqt_tissue3d.

A synthetic ellipsoidal shell in the reference's file format ("x y z fx fy fz
label", labels 0-20) with patches of every relabelled label, an M-island
sphere placed inside the grid, anisotropic conduction, and a stimulation
profile ("x y z f f f t_act s") whose activation times include sites that
start at 81 ms (cut to 1 ms by the 82 ms window of GPU3D.cu:1757) and at
90 ms (never applied):

* the label rules of GPU3D.cu:1447-1462 on every voxel;
* FP64 V of every voxel at 83 ms and 90 ms (ap_83.vtk, ap_90.vtk at
  stride 1) against the oracle stepping the same tissue with the same
  stimulus rule, to the 6 digits the file keeps; at 83 ms the same oracle
  without the 82 ms window differs by about 11 mV, so the comparison sees
  the window;
* the activation map the program writes with --maps against the oracle's;
* over a whole beat, both columns of the --maps file (activation, APD90)
  against the library's trackers, and the --maps-window cut-off.
"""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
X, Y, Z = 30, 28, 24
ISLAND = (24.0, 14.0, 12.0, 3.0)   # centre and radius, in the mid-wall of the shell


def _shell():
    x, y, z = np.meshgrid(np.arange(X), np.arange(Y), np.arange(Z), indexing="ij")
    cx, cy, cz = (X - 1) / 2, (Y - 1) / 2, (Z - 1) / 2
    r = np.sqrt(((x - cx) / 12.5) ** 2 + ((y - cy) / 11.5) ** 2 + ((z - cz) / 9.5) ** 2)
    tissue = (r <= 1.0) & (r >= 0.55)
    depth = (r - 0.55) / 0.45
    lab = np.full((X, Y, Z), -1, int)
    lab[tissue & (depth < 0.3)] = 2
    lab[tissue & (depth >= 0.3) & (depth < 0.65)] = 15
    lab[tissue & (depth >= 0.65)] = 8
    lab[tissue & (depth < 0.3) & (z < 8)] = 5                            # 4-6 -> 2
    lab[tissue & (depth < 0.3) & (z >= 15)] = 13                         # stays 13 (ENDO)
    lab[tissue & (depth >= 0.3) & (depth < 0.65) & (x < 10)] = 14        # 14 -> 12 outside the island
    lab[tissue & (depth >= 0.3) & (depth < 0.65) & (x > 20) & (y < 12)] = 16
    lab[tissue & (depth >= 0.65) & (y > 22)] = 19
    lab[tissue & (depth >= 0.65) & (y < 5) & (x < 15)] = 0
    th = np.arctan2(y - cy, x - cx)
    al = np.deg2rad(60 - 120 * np.clip(depth, 0, 1))
    f = np.stack([-np.sin(th) * np.cos(al), np.cos(th) * np.cos(al), np.sin(al)], -1)
    return tissue, lab, f, depth


def _rules(x, y, z, l):
    """GPU3D.cu:1447-1462 with the island of this test."""
    if 4 <= l <= 6:
        l = 2          # the reference tests the fibre's y component (<= 4): always true
    d = np.sqrt((x - ISLAND[0]) ** 2 + (y - ISLAND[1]) ** 2 + (z - ISLAND[2]) ** 2)
    if 14 <= l <= 16 and not d < ISLAND[3]:
        l = 12 if l <= 14 else 17
    return l


def _celltype(O, l):
    if 7 <= l <= 10 or 17 <= l <= 20:
        return O.EPI
    if 4 <= l <= 6 or 14 <= l <= 16:
        return O.MCELL
    return O.ENDO


def _files(tmp):
    tissue, lab, f, depth = _shell()
    pts = np.argwhere(tissue)
    with open(tmp / "geo.txt", "w") as fh:
        for px, py, pz in pts:
            fx, fy, fz = f[px, py, pz]
            fh.write(f"{px} {py} {pz} {fx:.6f} {fy:.6f} {fz:.6f} {lab[px, py, pz]}\n")
    rim = np.argwhere(tissue & (depth < 0.12))
    tact = np.zeros(len(rim))
    tact[::7] = 0.081
    tact[3::11] = 0.090
    with open(tmp / "stim.cnd", "w") as fh:
        for (px, py, pz), t in zip(rim, tact):
            fh.write(f"{px} {py} {pz} 0 0 0 {t:.3f} s\n")
        fh.write("0 0 0 0 0 0 0.000 s\n")          # outside the tissue: ignored
        fh.write(f"{X + 3} 1 1 0 0 0 0.000 s\n")   # outside the grid: ignored
    return tissue, lab, f, rim, tact


def _read_vtk(path, dims):
    with open(path) as fh:
        lines = fh.read().split("\n")
    assert lines[4] == "DIMENSIONS %d %d %d" % dims
    vals = np.array(" ".join(lines[10:]).split(), float)
    assert vals.size == dims[0] * dims[1] * dims[2]
    return vals.reshape(dims[2], dims[1], dims[0]).transpose(2, 1, 0)   # x fastest in the file


def _oracle(O, tissue, ct, fib, rim, tact, snaps, window):
    mask = tissue.astype(np.uint8)
    g, nvox = O.ghost_map(mask)
    cells = np.flatnonzero(mask.ravel())
    dten, drift = O.tissue_tensor(g, nvox, fib.reshape(-1, 3)[cells], 0.154, 0.0385, 0.2)
    site = g.ravel()[np.ravel_multi_index(rim.T, tissue.shape)] - 1
    off = (np.float32(1000.0) * tact.astype(np.float32)).astype(np.float64)  # float product (GPU3D.cu:403)
    p = O.params(iks_scale=0.05, iso=1, pca_iso=4.0, suprep=1, dose_fmol=100.0)
    s = np.tile(O.cell_init(-86.2), (nvox, 1))
    t = 0.0
    out = {}
    act = np.full(nvox, -1.0)  # first step whose new V >= -30 mV (OneD.cc:422), at its start time
    for n in range(1, max(snaps) + 1):
        tc = t - np.floor(t / 500.0) * 500.0
        ist = np.zeros(nvox)
        if window <= 0 or tc < window:
            on = (tc >= off) & (tc < off + 3.0)
            ist[site[on]] = -52.0
        O.tissue_step(g, nvox, dten, drift, 0.2, 0.02, ct, p, s, ist)
        up = (act < 0) & (s[:, 0] >= -30.0)
        act[up] = t
        t += 0.02
        if n in snaps:
            out[n] = s[:, 0].copy()
    out["act"] = act
    return out, cells


def test_ventricle_program_matches_oracle(lqt, oracle, tmp_path):
    O = oracle
    tissue, lab, f, rim, tact = _files(tmp_path)
    out = tmp_path / "out"
    out.mkdir()
    exe = os.path.join(ROOT, "lqt1_amd", "bin", "lqt_tissue3d")
    r = subprocess.run([exe, "--geometry", str(tmp_path / "geo.txt"), "--dims", str(X), str(Y), str(Z),
                        "--stim", str(tmp_path / "stim.cnd"), "--time", "90.5", "--fp64", "--dfibre", "0.154",
                        "--dcross", "0.0385", "--vtk-every", "1", "--vtk-stride", "1", "--outdir", str(out),
                        "--island", *[str(v) for v in ISLAND], "--labels-out", str(tmp_path / "lab.txt"),
                        "--maps", str(tmp_path / "maps" / "act.txt")],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert f"There are {rim.shape[0]} stimulus points." in r.stdout
    assert f"There are {int(tissue.sum())} ventricular points." in r.stdout
    # label rules on every voxel, with known answers for each relabelled kind
    rl = np.loadtxt(tmp_path / "lab.txt", dtype=int)
    assert rl.shape[0] == tissue.sum()
    got = {(a, b, c): (l, t) for a, b, c, l, t in rl}
    kinds = set()
    for (a, b, c), (l, t) in got.items():
        want = _rules(a, b, c, lab[a, b, c])
        assert l == want, ((a, b, c), lab[a, b, c], l, want)
        assert t == _celltype(O, want)
        kinds.add((lab[a, b, c], want))
    for k in [(5, 2), (15, 15), (15, 17), (14, 12), (16, 17), (13, 13), (8, 8), (19, 19), (0, 0), (2, 2)]:
        assert k in kinds, k
    # the oracle on the same tissue
    cells_xyz = np.argwhere(tissue)
    ct = np.array([got[tuple(p)][1] for p in cells_xyz], np.uint8)
    # the fibre as the file holds it, rounded to single precision as the driver does
    geo = np.loadtxt(tmp_path / "geo.txt")
    fib = np.zeros((X, Y, Z, 3), np.float32)
    gi = geo[:, :3].astype(int)
    fib[gi[:, 0], gi[:, 1], gi[:, 2]] = geo[:, 3:6].astype(np.float32)
    snaps = {4150: 83, 4500: 90}
    ref, cells = _oracle(O, tissue, ct, fib, rim, tact, snaps, 82.0)
    nowin, _ = _oracle(O, tissue, ct, fib, rim, tact, {4150: 83}, 0.0)
    for n, ms in snaps.items():
        v = _read_vtk(out / f"ap_{ms}.vtk", (X, Y, Z))
        err = np.max(np.abs(v.ravel()[cells] - ref[n]) / np.maximum(1.0, np.abs(ref[n])))
        print(f"3D ventricle program FP64 vs oracle at {ms} ms: max rel |dV| {err:.3g} over {cells.size} voxels")
        assert err < 1e-5, ms
        assert np.all(v[~tissue] == -100.0)
    # the activation map the program writes (--maps) against the oracle's
    # (times printed to 0.01 ms; the oracle ran 4500 of the program's 4526 steps)
    mp = np.loadtxt(tmp_path / "maps" / "act.txt")
    assert mp.shape == (cells.size, 5)
    assert np.array_equal(mp[:, :3].astype(int), np.argwhere(tissue))
    ao = ref["act"]
    hit = ao >= 0
    assert hit.sum() > 0.5 * cells.size
    assert np.max(np.abs(mp[hit, 3] - ao[hit])) < 0.006
    assert np.all((mp[~hit, 3] < 0) | (mp[~hit, 3] >= 89.99))
    print(f"activation map: {hit.sum()} of {cells.size} voxels by 90 ms, equal to the oracle's")
    d_win = np.max(np.abs(nowin[4150] - ref[4150]))
    print(f"the 82 ms window changes V at 83 ms by up to {d_win:.1f} mV")
    assert d_win > 5.0


def _driver_steps(tmax, dt=0.02):
    """Steps the program's time loop takes to pass tmax (time <= tmax, time += dt)."""
    tt, n = 0.0, 0
    while tt <= tmax:
        tt += dt
        n += 1
    return n


def test_ventricle_program_maps_over_a_beat(lqt, tmp_path):
    """--maps over a whole beat (BCL 500 ms, 480 ms run): both columns of the
    file equal the library's trackers on the same tissue (the trackers are
    pinned to the oracle in the tissue-track tests); a --maps-window that ends
    at the median activation time stops tracking there, so no voxel gets an
    APD90 and only the upstrokes before it an activation time."""
    tissue, lab, f, rim, tact = _files(tmp_path)
    exe = os.path.join(ROOT, "lqt1_amd", "bin", "lqt_tissue3d")
    common = [exe, "--geometry", str(tmp_path / "geo.txt"), "--dims", str(X), str(Y), str(Z),
              "--stim", str(tmp_path / "stim.cnd"), "--fp64", "--dfibre", "0.154", "--dcross", "0.0385",
              "--vtk-every", "0", "--island", *[str(v) for v in ISLAND], "--bcl", "500"]
    r = subprocess.run(common + ["--time", "480", "--labels-out", str(tmp_path / "lab.txt"),
                                 "--maps", str(tmp_path / "beat.txt")], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    # the same run through the library: cell types as the program assigned them,
    # the fibre as the file holds it in single precision, the profile's sites
    rl = np.loadtxt(tmp_path / "lab.txt", dtype=int)
    ct = np.full((X, Y, Z), lqt.NONE, np.uint8)
    ct[rl[:, 0], rl[:, 1], rl[:, 2]] = rl[:, 4]
    geo = np.loadtxt(tmp_path / "geo.txt")
    gi = geo[:, :3].astype(int)
    fib = np.zeros((X, Y, Z, 3), np.float32)
    fib[gi[:, 0], gi[:, 1], gi[:, 2]] = geo[:, 3:6].astype(np.float32)
    sites = np.ravel_multi_index(rim.T, (X, Y, Z))
    offs = np.float32(1000.0) * tact.astype(np.float32)
    T = lqt.Tissue(ct, fib, 0.2, 0.154, 0.0385, lqt.preset("gpu3d"), lqt.FP64, 0.02)
    T.set_stimulus(sites, offs, bcl=500.0, duration=3.0, strength=-52.0, window=82.0)
    T.track(0.0, 500.0)
    T.step(_driver_steps(480.0))
    act, apd = T.track_results()
    T.close()
    mp = np.loadtxt(tmp_path / "beat.txt")
    assert mp.shape == (tissue.sum(), 5)
    assert np.array_equal(mp[:, :3].astype(int), np.argwhere(tissue))
    assert np.all(act > 0) and np.all(apd > 100.0), "every voxel activates and repolarises within the beat"
    assert np.max(np.abs(mp[:, 3] - act)) < 0.006
    assert np.max(np.abs(mp[:, 4] - apd)) < 0.006
    print(f"maps over one beat: {act.size} voxels, activation {act.min():.2f}-{act.max():.2f} ms, "
          f"APD90 {apd.min():.2f}-{apd.max():.2f} ms, equal to the library's trackers")
    # a window that ends at the median activation time (between two step starts)
    w = float(np.median(act)) + 0.01
    r = subprocess.run(common + ["--time", f"{w + 10.0:.2f}", "--maps", str(tmp_path / "early.txt"),
                                 "--maps-window", "0", f"{w:.2f}"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    early = np.loadtxt(tmp_path / "early.txt")
    assert np.all(early[:, 4] == -1.0)
    hit = act < w
    assert 0 < hit.sum() < act.size
    assert np.max(np.abs(early[hit, 3] - act[hit])) < 0.006
    assert np.all(early[~hit, 3] == -1.0)
