"""The reference's own 3D grid (GPU3D.cu:28-42: 690 x 489 x 537 cells at
0.2 mm, 181.2M grid cells, about 22.6M tissue voxels in its ventricle), the
largest tissue the test suite runs.

A synthetic ellipsoidal shell of about 24.6M voxels fills that grid
(`geometry.ellipsoid_shell_3d`: labels endo/M/epi by transmural depth,
fibres rotating from +60 degrees at the endocardium to -60 at the
epicardium, anisotropic D 0.154/0.0385, apical endocardial stimulus).  FP64 on the GPU against the FP64 oracle
(calc_du + Itot + new_u_v of GPU3D.cu, restated) after 6 steps with the
stimulus on: V of every voxel within 1e-7 mV.  This is the grid size at
which 32-bit index arithmetic of the padded grid and of the voxel list
would first matter, and at which the host-built ghost map is largest.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DX = 0.2


def test_reference_grid_fp64_matches_oracle(lqt, oracle):
    O = oracle
    nsteps, dt = 6, 0.02
    from lqt1_amd import geometry as G

    t0 = time.perf_counter()
    geo = G.ellipsoid_shell_3d()
    lab, fib, sites = geo.label, geo.fibre, geo.stim_sites
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
