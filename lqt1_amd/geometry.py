"""Synthetic tissue geometries in the label/fibre layout of lqt_geometry.

The reference drivers read real geometries (GPU3D.cu:1276, new_11.cu:1296)
that are not shipped with it; these generators give seeded stand-ins of the
same kind: a 2D annular cross-section, a 3D biventricular shell with
rotating fibres, transmural endo/M/epi labels and an M-cell island, and an
ellipsoidal shell that fills the reference's own 3D grid.
Grid index is (x*ny + y)*nz + z; labels are enum CellType, NONE outside.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _lib as L


@dataclass
class Geometry:
    label: np.ndarray            # (nx, ny, nz) uint8
    fibre: np.ndarray | None     # (nx, ny, nz, 3) float32
    dx: float
    stim_sites: np.ndarray       # grid indices
    stim_offsets: np.ndarray     # ms
    s2_sites: np.ndarray | None = None
    depth: np.ndarray | None = None  # transmural depth per grid cell (0 endo .. 1 epi), -1 outside
    info: dict | None = None

    @property
    def nvox(self):
        return int((self.label != L.NONE).sum())


def label_by_depth(depth, split=(0.30, 0.35, 0.35)):
    lab = np.full(depth.shape, L.NONE, np.uint8)
    t = depth >= 0
    e1 = split[0]
    e2 = split[0] + split[1]
    lab[t & (depth < e1)] = L.ENDO
    lab[t & (depth >= e1) & (depth < e2)] = L.MCELL
    lab[t & (depth >= e2)] = L.EPI
    return lab


def strand_line(celltype, axis=2):
    """A 1x1xL tissue (the OneD.cc strand as a tissue)."""
    ct = np.asarray(celltype, np.uint8)
    shape = [1, 1, 1]
    shape[axis] = ct.size
    return ct.reshape(shape).copy()


def annulus_2d(n=600, dx=0.15, r_out=40.0, wall=10.0, split=(0.30, 0.35, 0.35), seed=0,
               ci_range=(280, 340)):
    """Annular cross-section on an n x n grid (nz = 1): endo = inner rim.

    S1 sites are the inner-rim voxels; S2 sites are the tissue of one quadrant.
    The seed picks the S2 coupling interval (returned in info['s2_ci'])."""
    rng = np.random.default_rng(seed)
    c = (n - 1) * dx / 2.0
    x = np.arange(n) * dx - c
    X, Y = np.meshgrid(x, x, indexing="ij")
    r = np.hypot(X, Y)
    r_in = r_out - wall
    tissue = (r <= r_out) & (r >= r_in)
    depth = np.where(tissue, (r - r_in) / wall, -1.0)
    lab = label_by_depth(depth, split)[:, :, None]
    rim = tissue & (r < r_in + 1.5 * dx)
    sites = np.flatnonzero(rim[:, :, None])
    quad = tissue & (X > 0) & (Y > 0)
    s2 = np.flatnonzero(quad[:, :, None])
    ci = int(rng.integers(ci_range[0], ci_range[1] + 1))
    return Geometry(lab, None, dx, sites.astype(np.int64), np.zeros(sites.size, np.float32),
                    s2.astype(np.int64), depth[:, :, None], dict(s2_ci=ci, r_in=r_in, r_out=r_out))


def _shell_depth(P, a_in, c_in, a_out, c_out, iters=16):
    """Depth s in [0,1] such that P lies on the ellipsoid with semi-axes
    (a(s), a(s), c(s)) interpolated between the inner and outer surfaces; -1
    outside the shell.  Bisection on the monotone implicit function, only
    for the points inside the shell."""
    x, y, z = P
    rho2 = x * x + y * y
    z2 = z * z
    inside = (rho2 / (a_in * a_in) + z2 / (c_in * c_in) >= 1.0) & \
             (rho2 / (a_out * a_out) + z2 / (c_out * c_out) <= 1.0)
    r2 = rho2[inside]
    zz = z2[inside]
    lo = np.zeros(r2.shape, np.float32)
    hi = np.ones(r2.shape, np.float32)
    for _ in range(iters):
        mid = 0.5 * (lo + hi)
        a = a_in + mid * (a_out - a_in)
        cc = c_in + mid * (c_out - c_in)
        up = r2 / (a * a) + zz / (cc * cc) > 1.0
        lo = np.where(up, mid, lo)
        hi = np.where(up, hi, mid)
    out = np.full(x.shape, -1.0, np.float32)
    out[inside] = 0.5 * (lo + hi)
    return out


def ventricle_3d(shape=(256, 256, 320), dx=0.2, seed=0, lv=(10.5, 44.0, 20.5, 54.0),
                 rv_wall=4.0, base_mm=58.0, island_mm=5.0, split=(0.30, 0.35, 0.35),
                 fibre_epi=-60.0, fibre_endo=60.0, apex_stim_mm=4.0, scale=1.0):
    """Biventricular shell: LV truncated ellipsoid (inner a, c; outer a, c in
    mm), RV crescent of wall `rv_wall` on the +x side, base plane at z =
    base_mm above the apex.  Fibre angle goes linearly from fibre_epi at the
    epicardium to fibre_endo at the endocardium.  A sphere of M cells of
    radius island_mm sits mid-wall at a seed-chosen site.  Stimulus: LV
    endocardial voxels within apex_stim_mm of the LV cavity apex.  `scale`
    multiplies every length (a smaller heart on the same dx)."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    if scale != 1.0:
        lv = tuple(scale * v for v in lv)
        rv_wall, base_mm, island_mm, apex_stim_mm = (scale * rv_wall, scale * base_mm, scale * island_mm,
                                                     scale * apex_stim_mm)
    a_in, c_in, a_out, c_out = lv
    # apex of the outer LV at z = 2 mm; long axis along +z
    cx = (nx - 1) * dx / 2.0 - 4.0 * scale
    cy = (ny - 1) * dx / 2.0
    cz = 2.0 * scale + c_out
    x = (np.arange(nx) * dx - cx).astype(np.float32)
    y = (np.arange(ny) * dx - cy).astype(np.float32)
    z = (np.arange(nz) * dx - cz).astype(np.float32)
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    below_base = Z <= (base_mm - c_out)
    dl = _shell_depth((X, Y, Z), a_in, c_in, a_out, c_out)
    dl = np.where(below_base, dl, -1.0)
    # RV: ellipsoid family centred at +x offset; wall rv_wall, outside the LV
    rv_off = a_out * 0.55
    ra_in, rc_in = a_out * 0.95, c_out * 0.80
    ra_out, rc_out = ra_in + rv_wall, rc_in + rv_wall
    Xr = X - rv_off
    dr = _shell_depth((Xr, Y, Z), ra_in, rc_in, ra_out, rc_out)
    lv_outer = (X * X + Y * Y) / (a_out * a_out) + Z * Z / (c_out * c_out) <= 1.0
    dr = np.where(below_base & (Xr > -1.0) & ~lv_outer, dr, -1.0)
    depth = np.where(dl >= 0, dl, dr)
    tissue = depth >= 0
    lab = label_by_depth(depth, split)
    # M-cell island (ellipsoidal ball) mid-wall in the LV free wall
    phi = rng.uniform(0.5 * np.pi, 1.5 * np.pi)       # away from the RV side
    hz = rng.uniform(-0.6, -0.2) * c_out
    s_mid = 0.5
    a_mid = a_in + s_mid * (a_out - a_in)
    c_mid = c_in + s_mid * (c_out - c_in)
    rad = a_mid * np.sqrt(max(0.0, 1.0 - (hz / c_mid) ** 2))
    ic = np.array([rad * np.cos(phi), rad * np.sin(phi), hz], np.float32)
    isl = ((X - ic[0]) ** 2 + (Y - ic[1]) ** 2 + (Z - ic[2]) ** 2 <= island_mm ** 2) & tissue
    lab[isl] = L.MCELL
    # fibres: circumferential / longitudinal frame of the ellipsoid family
    fib = np.zeros(shape + (3,), np.float32)
    e = np.clip(depth, 0.0, 1.0)
    alpha = np.deg2rad(fibre_endo + (fibre_epi - fibre_endo) * e).astype(np.float32)
    Xc = np.where(dl >= 0, X, Xr)
    rho = np.sqrt(Xc * Xc + Y * Y) + 1e-6
    cdir = np.stack([-Y / rho, Xc / rho, np.zeros_like(rho)], -1)
    # surface normal of the ellipsoid through the point
    a_s = np.where(dl >= 0, a_in + e * (a_out - a_in), ra_in + e * (ra_out - ra_in))
    c_s = np.where(dl >= 0, c_in + e * (c_out - c_in), rc_in + e * (rc_out - rc_in))
    nvec = np.stack([Xc / (a_s * a_s), Y / (a_s * a_s), Z / (c_s * c_s)], -1)
    nvec /= np.linalg.norm(nvec, axis=-1, keepdims=True) + 1e-12
    ldir = np.cross(nvec, cdir)
    ldir /= np.linalg.norm(ldir, axis=-1, keepdims=True) + 1e-12
    f = np.cos(alpha)[..., None] * cdir + np.sin(alpha)[..., None] * ldir
    f /= np.linalg.norm(f, axis=-1, keepdims=True) + 1e-12
    fib[tissue] = f[tissue]
    # apical LV endocardial stimulus
    apex_in = np.array([0.0, 0.0, -c_in], np.float32)
    d_apex = np.sqrt(X ** 2 + Y ** 2 + (Z - apex_in[2]) ** 2)
    stim = (dl >= 0) & (dl < 0.25) & (d_apex <= apex_stim_mm)
    sites = np.flatnonzero(stim).astype(np.int64)
    return Geometry(lab, fib, dx, sites, np.zeros(sites.size, np.float32), None,
                    depth.astype(np.float32), dict(island_centre_mm=ic.tolist(), seed=seed))


def ellipsoid_shell_3d(shape=(690, 489, 537), dx=0.2, outer=(55.0, 42.0, 48.0), inner=(47.0, 34.0, 40.0),
                       split=(0.30, 0.35, 0.35), fibre_endo=60.0, fibre_epi=-60.0, stim_depth=0.15,
                       stim_z_mm=-30.0):
    """An ellipsoidal shell centred in the grid (semi-axes `outer` and
    `inner` in mm).  On the default grid, GPU3D.cu's 690 x 489 x 537 cells at
    0.2 mm (GPU3D.cu:28-42), it holds 24.6M voxels, the size of the
    reference's whole-ventricle model (about 22.6M).  Labels by the
    transmural fraction u (0 endocardium, 1 epicardium) in the `split`
    proportions; fibres circumferential about z, rotated towards z by an
    angle going linearly from fibre_endo to fibre_epi; stimulus on the
    endocardial voxels (u < stim_depth) below z = stim_z_mm.  Computed on
    the tissue voxels only, so the full-size grid takes seconds."""
    nx, ny, nz = shape
    x = ((np.arange(nx, dtype=np.float32) - (nx - 1) / 2) * dx)[:, None, None]
    y = ((np.arange(ny, dtype=np.float32) - (ny - 1) / 2) * dx)[None, :, None]
    z = ((np.arange(nz, dtype=np.float32) - (nz - 1) / 2) * dx)[None, None, :]
    ao, bo, co = outer
    ai, bi, ci = inner
    ro = (x / ao) ** 2 + (y / bo) ** 2 + (z / co) ** 2
    ri = (x / ai) ** 2 + (y / bi) ** 2 + (z / ci) ** 2
    idx = np.flatnonzero((ro <= 1.0) & (ri > 1.0))
    del ro, ri
    ix, iy, iz = np.unravel_index(idx, shape)
    px = (ix.astype(np.float32) - (nx - 1) / 2) * dx
    py = (iy.astype(np.float32) - (ny - 1) / 2) * dx
    pz = (iz.astype(np.float32) - (nz - 1) / 2) * dx
    so = np.sqrt((px / ao) ** 2 + (py / bo) ** 2 + (pz / co) ** 2)
    si = np.sqrt((px / ai) ** 2 + (py / bi) ** 2 + (pz / ci) ** 2)
    u = (si - 1.0) / np.maximum((si - 1.0) + (1.0 - so), 1e-6)
    e1, e2 = split[0], split[0] + split[1]
    lab = np.full(shape, L.NONE, np.uint8)
    lab.reshape(-1)[idx] = np.where(u < e1, L.ENDO, np.where(u < e2, L.MCELL, L.EPI)).astype(np.uint8)
    ang = np.deg2rad(fibre_endo + (fibre_epi - fibre_endo) * u)
    rho = np.hypot(px, py) + 1e-6
    fib = np.zeros(shape + (3,), np.float32)
    f = fib.reshape(-1, 3)
    f[idx, 0] = -np.cos(ang) * py / rho
    f[idx, 1] = np.cos(ang) * px / rho
    f[idx, 2] = np.sin(ang)
    sites = idx[(u < stim_depth) & (pz < stim_z_mm)].astype(np.int64)
    return Geometry(lab, fib, dx, sites, np.zeros(sites.size, np.float32), None, None,
                    dict(voxels=int(idx.size), outer_mm=list(outer), inner_mm=list(inner)))
