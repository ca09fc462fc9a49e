"""The reference's own 3D grid (GPU3D.cu:28-42: 690 x 489 x 537 cells at
0.2 mm, 181.2M grid cells, about 22.6M tissue voxels in its ventricle), the
largest tissue the test suite runs.

A synthetic ellipsoidal shell of about 24.6M voxels fills that grid
(labels endo/M/epi by transmural depth, fibres rotating from +60 degrees at
the endocardium to -60 at the epicardium, anisotropic D 0.154/0.0385,
apical endocardial stimulus).  FP64 on the GPU against the FP64 oracle
(calc_du + Itot + new_u_v of GPU3D.cu, restated) after 6 steps with the
stimulus on: V of every voxel within 1e-7 mV.  This is the grid size at
which 32-bit index arithmetic of the padded grid and of the voxel list
would first matter, and at which the host-built ghost map is largest.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPE = (690, 489, 537)
DX = 0.2


def shell(shape=SHAPE, dx=DX):
    """Labels (uint8, 255 outside), fibres (float32, (nx, ny, nz, 3)) and
    stimulus sites (flat indices) of an ellipsoidal shell centred in the
    grid: outer semi-axes 55 x 42 x 48 mm, inner 47 x 34 x 40 mm."""
    nx, ny, nz = shape
    x = ((np.arange(nx, dtype=np.float32) - (nx - 1) / 2) * dx)[:, None, None]
    y = ((np.arange(ny, dtype=np.float32) - (ny - 1) / 2) * dx)[None, :, None]
    z = ((np.arange(nz, dtype=np.float32) - (nz - 1) / 2) * dx)[None, None, :]
    ro = (x / 55.0) ** 2 + (y / 42.0) ** 2 + (z / 48.0) ** 2
    ri = (x / 47.0) ** 2 + (y / 34.0) ** 2 + (z / 40.0) ** 2
    idx = np.flatnonzero((ro <= 1.0) & (ri > 1.0))
    del ro, ri
    ix, iy, iz = np.unravel_index(idx, shape)
    px = (ix.astype(np.float32) - (nx - 1) / 2) * dx
    py = (iy.astype(np.float32) - (ny - 1) / 2) * dx
    pz = (iz.astype(np.float32) - (nz - 1) / 2) * dx
    so = np.sqrt((px / 55.0) ** 2 + (py / 42.0) ** 2 + (pz / 48.0) ** 2)
    si = np.sqrt((px / 47.0) ** 2 + (py / 34.0) ** 2 + (pz / 40.0) ** 2)
    u = (si - 1.0) / np.maximum((si - 1.0) + (1.0 - so), 1e-6)  # 0 at the endocardium, 1 at the epicardium
    lab = np.full(shape, 255, np.uint8)
    lab.reshape(-1)[idx] = np.where(u < 0.30, 3, np.where(u < 0.65, 4, 2)).astype(np.uint8)
    ang = np.deg2rad(60.0 - 120.0 * u)
    rho = np.hypot(px, py) + 1e-6
    fib = np.zeros(shape + (3,), np.float32)
    f = fib.reshape(-1, 3)
    f[idx, 0] = -np.cos(ang) * py / rho
    f[idx, 1] = np.cos(ang) * px / rho
    f[idx, 2] = np.sin(ang)
    sites = idx[(u < 0.15) & (pz < -30.0)]
    return lab, fib, sites


def test_reference_grid_fp64_matches_oracle(lqt, oracle):
    O = oracle
    nsteps, dt = 6, 0.02
    t0 = time.perf_counter()
    lab, fib, sites = shell()
    print(f"shell: {np.count_nonzero(lab != 255)} voxels, {sites.size} stimulus sites, "
          f"{time.perf_counter() - t0:.1f} s", flush=True)
    t0 = time.perf_counter()
    T = lqt.Tissue(lab, fib, DX, 0.154, 0.0385, lqt.preset("cpu_default"), lqt.FP64, dt)
    T.set_stimulus(sites, None, bcl=1000.0, duration=1.0, strength=-80.0)
    T.step(nsteps)
    v = T.v()
    nvox = T.nvox
    T.close()
    print(f"GPU FP64 tissue: {nvox} voxels, create and {nsteps} steps {time.perf_counter() - t0:.1f} s",
          flush=True)
    t0 = time.perf_counter()
    mask = (lab != 255).astype(np.uint8)
    g, n_or = O.ghost_map(mask)
    del mask
    assert n_or == nvox
    cells = np.flatnonzero(g.reshape(-1) > 0)  # tissue cells in compaction order
    fibre = fib.reshape(-1, 3)[cells]
    del fib
    ct = lab.reshape(-1)[cells].copy()
    del cells
    dten, drift = O.tissue_tensor(g, nvox, fibre, 0.154, 0.0385, DX)
    del fibre
    s = np.tile(O.cell_init(), (nvox, 1))
    ist = np.zeros(nvox)
    ist[g.reshape(-1)[sites] - 1] = -80.0
    p = O.params()
    for _ in range(nsteps):  # t < 1 ms: the stimulus is on in every step
        O.tissue_step(g, nvox, dten, drift, DX, dt, ct, p, s, ist)
    print(f"oracle: {nsteps} steps {time.perf_counter() - t0:.1f} s", flush=True)
    err = float(np.max(np.abs(v - s[:, 0])))
    vs = s[g.reshape(-1)[sites] - 1, 0]
    print(f"max |dV| over {nvox} voxels: {err:.3g} mV; V over the stimulus sites {vs.min():.2f}..{vs.max():.2f} mV",
          flush=True)
    assert vs.min() > -85.0, "the stimulus must have raised V at every site"
    assert err < 1e-7, err
