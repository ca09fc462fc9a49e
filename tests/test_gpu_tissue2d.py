"""The real-slice 2D program (new_11.cu) through lqt_tissue2d.

A synthetic 690 x 489 slice in the reference's file format ("x y fx fy
label", labels 0-20, an annular wall with tangential fibres, a patch of
type-0 voxels, labels 3-6 and 13-16 inside and outside the M-island
rectangle) and a stimulation profile with rows on and off layer 157:

* the label rules of new_11.cu:1460-1489 on hand-picked points;
* FP64 V of the whole slice after 10 ms (the ap_10.vtk file) against the
  oracle stepping the same 2D sheet (9-point stencil, 2D ghost rule,
  type-0 voxels held at -100 mV) to the 6 digits the file keeps;
* the activation map the program writes with --maps against the oracle's.
"""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
X, Y = 690, 489


def _slice():
    cx, cy = 520.0, 300.0
    x, y = np.meshgrid(np.arange(X), np.arange(Y), indexing="ij")
    r = np.hypot(x - cx, y - cy)
    tissue = (r >= 40) & (r <= 75)
    depth = (r - 40) / 35.0
    lab = np.zeros((X, Y), int)
    lab[tissue & (depth < 0.3)] = 2            # LV endo
    lab[tissue & (depth >= 0.3) & (depth < 0.65)] = 15   # M (kept only inside the island)
    lab[tissue & (depth >= 0.65)] = 8          # epi
    # labels 3-6 become 2 or 7
    lab[tissue & (depth >= 0.3) & (depth < 0.65) & (y < 250)] = 5
    lab[tissue & (depth < 0.3) & (y > 330)] = 4
    # 13/14 outside the island -> 12, inside stays
    lab[tissue & (depth >= 0.3) & (depth < 0.65) & (x < 500)] = 13
    # a patch of type-0 voxels
    lab[tissue & (np.abs(x - cx) < 6) & (y - cy > 50)] = 0
    th = np.arctan2(y - cy, x - cx)
    fx, fy = -np.sin(th), np.cos(th)
    return tissue, lab, fx, fy, depth


def _write(tmp):
    tissue, lab, fx, fy, depth = _slice()
    pts = np.argwhere(tissue)
    with open(tmp / "slice.txt", "w") as f:
        for px, py in pts:
            f.write(f"{px} {py} {fx[px, py]:.6f} {fy[px, py]:.6f} {lab[px, py]}\n")
    rim = np.argwhere(tissue & (depth < 0.06))
    with open(tmp / "stim.cnd", "w") as f:
        for px, py in rim:
            f.write(f"{px} {py} 157 0 0 0 0.005 s\n")
        for px, py in rim[:50]:
            f.write(f"{px} {py} 156 0 0 0 0.0 s\n")   # another layer: ignored
        f.write("3 3 157 0 0 0 0.0 s\n")               # outside tissue: ignored
    return tissue, lab, fx, fy, rim


def _rules(px, py, l):
    """new_11.cu:1473-1489 (known answers are asserted below)."""
    t5 = l
    if 3 <= t5 <= 6:
        t5 = 2 if t5 <= 4 else 7
    if 13 <= t5 <= 16 and not (550 <= px < 689 and 150 <= py < 487):
        t5 = 12 if t5 <= 14 else 17
    return t5


def _read_vtk(path):
    with open(path) as f:
        lines = f.read().split("\n")
    assert lines[4] == f"DIMENSIONS {X} {Y} 1"
    vals = np.array(" ".join(lines[10:]).split(), float)
    assert vals.size == X * Y
    return vals.reshape(Y, X).T   # rows are y, x fastest


def test_slice_program_matches_oracle(lqt, oracle, tmp_path):
    O = oracle
    tissue, lab, fx, fy, rim = _write(tmp_path)
    out = tmp_path / "out"
    out.mkdir()
    exe = os.path.join(ROOT, "lqt1_amd", "bin", "lqt_tissue2d")
    r = subprocess.run([exe, "--geometry", str(tmp_path / "slice.txt"), "--stim", str(tmp_path / "stim.cnd"),
                        "--time", "10.5", "--fp64", "--outdir", str(out), "--labels-out", str(tmp_path / "lab.txt"),
                        "--maps", str(tmp_path / "act.txt")],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert f"There are {rim.shape[0]} stimulus points." in r.stdout
    rl = np.loadtxt(tmp_path / "lab.txt", dtype=int)
    got = {(a, b): (c, d) for a, b, c, d in rl}
    # known answers of the label rules: (label as read, region) -> label used
    isl = np.zeros((X, Y), bool)
    isl[550:689, 150:487] = True
    cases = [((lab == 13) & ~isl, 12), ((lab == 15) & isl, 15), ((lab == 15) & ~isl, 17), (lab == 5, 7),
             (lab == 4, 2), (lab == 8, 8), (lab == 2, 2), (lab == 0, 0)]
    for sel, want in cases:
        pts = np.argwhere(tissue & sel)
        assert pts.size, want
        for a, b in pts[:: max(1, len(pts) // 20)]:
            assert got[(a, b)][0] == want, ((a, b), lab[a, b], got[(a, b)], want)
    zero = np.argwhere(tissue & (lab == 0))
    assert zero.size and all(got[(a, b)][1] == (O.ENDO | 0x80) for a, b in zero)
    for (a, b), (t5, ct) in list(got.items())[::97]:
        assert t5 == _rules(a, b, lab[a, b])
    # the oracle on the same sheet
    mask = tissue[:, :, None].astype(np.uint8)
    g, nvox = O.ghost_map(mask)
    cells = np.flatnonzero(mask.ravel())
    fib = np.zeros((nvox, 3), np.float32)
    fib[:, 0] = fx.ravel()[cells]
    fib[:, 1] = fy.ravel()[cells]
    # the file keeps 6 decimals of the fibre
    fib[:, :2] = np.round(fib[:, :2].astype(np.float64), 6)
    dten, drift = O.tissue_tensor(g, nvox, fib, 0.154, 0.0385, 0.2)
    ctab = {}
    for (a, b), (t5, ct) in got.items():
        ctab[a * Y + b] = ct
    ct = np.array([ctab[c] for c in cells], np.uint8)
    s = np.tile(O.cell_init(-86.2), (nvox, 1))
    k = g.ravel()[rim[:, 0] * Y + rim[:, 1]] - 1
    p = O.params(iks_scale=0.05, iso=0)
    t = 0.0
    act = np.full(nvox, -1.0)  # first step whose new V >= -30 mV, at its start time
    for _ in range(500):
        ist = np.zeros(nvox)
        if t - np.floor(t / 1000.0) * 1000.0 < 3.0:
            ist[k] = -52.0
        O.tissue_step(g, nvox, dten, drift, 0.2, 0.02, ct, p, s, ist)
        up = (act < 0) & (s[:, 0] >= -30.0)
        act[up] = t
        t += 0.02
    v = _read_vtk(out / "ap_10.vtk")
    vg = v.ravel()[cells]
    err = np.max(np.abs(vg - s[:, 0]) / np.maximum(1.0, np.abs(s[:, 0])))
    print(f"2D slice FP64 vs oracle after 500 steps: max rel |dV| {err:.3g} over {nvox} voxels; "
          f"activated {(s[:, 0] > -30).sum()}")
    assert (s[:, 0] > -30).sum() > 100
    assert err < 1e-5
    assert np.all(v[~tissue] == -100.0)
    assert np.all(vg[ct >= 0x80] == -100.0)
    # the activation map the program writes (--maps), printed to 0.01 ms; the
    # program ran 526 steps to the oracle's 500
    mp = np.loadtxt(tmp_path / "act.txt")
    assert np.array_equal(mp[:, :2].astype(int), np.argwhere(tissue))
    hit = act >= 0
    assert hit.sum() > 100 and np.max(np.abs(mp[hit, 2] - act[hit])) < 0.006
    assert np.all((mp[~hit, 2] < 0) | (mp[~hit, 2] >= 9.99))
