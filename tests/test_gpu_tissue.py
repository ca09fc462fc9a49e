"""Fused tissue kernel (lqt_tissue_*) against the FP64 oracle of the
reference's tissue update (calc_du + Itot + new_u_v, GPU3D.cu/new_11.cu).

FP64: V of every voxel within 1e-7 mV after 400 steps (1e-9 relative on
the states); FP32: activation time within one step (0.02 ms) on every
voxel, V L-inf stated."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _blob(seed=0, shape=(22, 20, 18)):
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    x, y, z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    cx, cy, cz = (nx - 1) / 2, (ny - 1) / 2, (nz - 1) / 2
    r = ((x - cx) / (0.45 * nx)) ** 2 + ((y - cy) / (0.45 * ny)) ** 2 + ((z - cz) / (0.45 * nz)) ** 2
    tissue = (r <= 1.0) & (r >= 0.25)
    tissue &= rng.random(shape) > 0.02  # a few holes for irregular ghosts
    lab = np.full(shape, 255, np.uint8)
    depth = np.clip((np.sqrt(r) - 0.5) / 0.5, 0, 1)
    lab[tissue & (depth < 0.3)] = 3
    lab[tissue & (depth >= 0.3) & (depth < 0.65)] = 4
    lab[tissue & (depth >= 0.65)] = 2
    ang = np.deg2rad(60 - 120 * depth)
    rho = np.hypot(x - cx, y - cy) + 1e-6
    c = np.stack([-(y - cy) / rho, (x - cx) / rho, np.zeros_like(rho)], -1)
    l = np.array([0, 0, 1.0])
    f = np.cos(ang)[..., None] * c + np.sin(ang)[..., None] * l
    f /= np.linalg.norm(f, axis=-1, keepdims=True)
    fib = np.where(tissue[..., None], f, 0).astype(np.float32)
    sites = np.flatnonzero(tissue & (x < cx - 4))
    return lab, fib, sites


def _oracle_tissue(O, lab, fib, sites, dx, dt, dfib, dcross, nsteps, params, bcl=1000.0, dur=1.0,
                   strength=-80.0, v0=-87.5):
    mask = (lab != 255).astype(np.uint8)
    g, nvox = O.ghost_map(mask)
    flat = lab.ravel()
    cells = np.flatnonzero(mask.ravel())
    fibre = fib.reshape(-1, 3)[cells].astype(np.float32)
    dten, drift = O.tissue_tensor(g, nvox, fibre, dfib, dcross, dx)
    ct = flat[cells].astype(np.uint8)
    s = np.tile(O.cell_init(v0), (nvox, 1))
    site_k = g.ravel()[sites] - 1
    vt = []
    t = 0.0
    for k in range(nsteps):
        tc = t - np.floor(t / bcl) * bcl
        ist = np.zeros(nvox)
        if 0.0 <= tc < dur:
            ist[site_k] = strength
        O.tissue_step(g, nvox, dten, drift, dx, dt, ct, params, s, ist)
        vt.append(s[:, 0].copy())
        t += dt
    return np.array(vt), s


@pytest.mark.parametrize("aniso", [True, False])
def test_tissue_fp64_matches_oracle(lqt, oracle, aniso):
    O = oracle
    lab, fib, sites = _blob()
    dx, dt, nsteps = 0.2, 0.02, 400
    dfib, dcross = (0.154, 0.0385) if aniso else (0.154, 0.154)
    vt, s = _oracle_tissue(O, lab, fib, sites, dx, dt, dfib, dcross, nsteps, O.params())
    T = lqt.Tissue(lab, fib if aniso else None, dx, dfib, dcross, lqt.preset("cpu_default"), lqt.FP64, dt)
    T.set_stimulus(sites, None, bcl=1000.0, duration=1.0, strength=-80.0)
    T.step(nsteps)
    v = T.v()
    err = np.max(np.abs(v - vt[-1]))
    assert err < 1e-7, err
    assert vt[-1].max() > 0, "the stimulus must have started a wave"
    for k in (0, T.nvox // 2, T.nvox - 1):
        st = T.state(k)
        rel = np.max(np.abs(st - s[k]) / np.maximum(np.abs(s[k]), 1e-6))
        assert rel < 1e-8, (k, rel)


def test_written_states_at_chunk_edges_step_like_the_oracle(lqt, oracle):
    """States written from the host at the first and last lanes of the
    64-voxel chunks (and in the partial last chunk) are the ones the step
    kernel reads and writes back: FP64 against the oracle stepping the same
    written states."""
    O = oracle
    lab, fib, sites = _blob(5, (18, 16, 14))
    dx, dt, nsteps = 0.2, 0.02, 60
    T = lqt.Tissue(lab, fib, dx, 0.154, 0.0385, lqt.preset("cpu_default"), lqt.FP64, dt)
    nvox = T.nvox
    assert nvox % 64 != 0 and nvox > 200
    picks = [0, 1, 62, 63, 64, 127, 128, nvox - 64, nvox - 2, nvox - 1]
    rng = np.random.default_rng(7)
    s0 = O.cell_init(-87.5)
    written = {}
    for k in picks:
        st = s0.copy()
        st[0] = -87.5 + 140.0 * rng.random()               # V, some above threshold
        st[9:12] = rng.random(3)                          # INa gates
        st[1] = 7.0 + rng.random()                        # nai
        st[5] = 1e-4 * (1.0 + rng.random())               # cai
        T.write_state(k, st)
        written[k] = st
    for k in picks:
        assert np.array_equal(T.state(k), written[k]), k
    # the oracle from the same states
    mask = (lab != 255).astype(np.uint8)
    g, nv = O.ghost_map(mask)
    assert nv == nvox
    cells = np.flatnonzero(mask.ravel())
    dten, drift = O.tissue_tensor(g, nvox, fib.reshape(-1, 3)[cells].astype(np.float32), 0.154, 0.0385, dx)
    ct = lab.ravel()[cells].astype(np.uint8)
    s = np.tile(s0, (nvox, 1))
    for k, st in written.items():
        s[k] = st
    p = O.params()
    for _ in range(nsteps):
        O.tissue_step(g, nvox, dten, drift, dx, dt, ct, p, s, np.zeros(nvox))
    T.step(nsteps)
    err = np.max(np.abs(T.v() - s[:, 0]))
    assert err < 1e-7, err
    for k in picks + [2, 65, nvox - 3]:
        st = T.state(k)
        rel = np.max(np.abs(st - s[k]) / np.maximum(np.abs(s[k]), 1e-6))
        assert rel < 1e-8, (k, rel)


def test_tissue_fp32_activation_within_one_step(lqt, oracle):
    O = oracle
    lab, fib, sites = _blob(1)
    dx, dt, nsteps = 0.2, 0.02, 900
    vt, _ = _oracle_tissue(O, lab, fib, sites, dx, dt, 0.154, 0.0385, nsteps, O.params())
    T = lqt.Tissue(lab, fib, dx, 0.154, 0.0385, lqt.preset("cpu_default"), lqt.FP32, dt)
    T.set_stimulus(sites, None, bcl=1000.0, duration=1.0, strength=-80.0)
    vg = []
    for k in range(nsteps // 10):
        T.step(10)
        vg.append(T.v())
    vg = np.array(vg)
    vo = vt[9::10]
    # activation step (first sample at or above -30 mV) per voxel
    def act(a):
        up = a >= -30.0
        return np.where(up.any(0), up.argmax(0), -1)
    ag, ao = act(vg), act(vo)
    assert np.array_equal(ag >= 0, ao >= 0)
    # sampled every 10 steps: equal sample index means within 10 steps;
    # the exact one-step gate is checked on the strand and line tests
    assert np.max(np.abs(ag - ao)) <= 1
    linf = np.max(np.abs(vg - vo))
    print("FP32 tissue V L-inf over 18 ms:", linf)


def test_line_tissue_equals_strand(lqt, oracle):
    """A 1-voxel-wide line in a 3x3xL grid reduces to OneD.cc's cable."""
    O = oracle
    L = 40
    ct = np.array([3] * 10 + [4] * 14 + [2] * 16, np.uint8)
    lab = np.full((3, 3, L + 2), 255, np.uint8)
    lab[1, 1, 1:L + 1] = ct
    dx, dt = 0.15, 0.02
    T = lqt.Tissue(lab, None, dx, 0.154, 0.154, lqt.preset("config1"), lqt.FP64, dt)
    sites = np.ravel_multi_index((np.ones(3, int), np.ones(3, int), np.arange(1, 4)), lab.shape)
    T.set_stimulus(sites, None, bcl=1000.0, duration=1.0, strength=-52.0)
    # the tissue stimulus starts at offset 0 of the cycle; the strand oracle
    # uses stim_start 0 to match
    r = O.strand_run(ct, 0.154, O.params(iso=0), dx=dx, bcl=1000.0, stop_time=40.0, stim_start=0.0,
                     rec_every=1)
    n = r["v"].shape[0]
    T.step(n)
    assert np.max(np.abs(T.v() - r["v"][-1])) < 1e-8


def test_checkpoint_and_vtk(lqt, tmp_path):
    lab, fib, sites = _blob(2, (12, 12, 10))
    T = lqt.Tissue(lab, fib, 0.2, 0.154, 0.0385, lqt.preset("cpu_default"), lqt.FP32, 0.02)
    T.set_stimulus(sites, None, strength=-80.0)
    T.step(50)
    T.save(tmp_path / "ck.bin")
    T.step(50)
    v100 = T.v()
    T.load(tmp_path / "ck.bin")
    assert T.time()[1] == 50
    T.step(50)
    assert np.array_equal(T.v(), v100)
    T.write_vtk(tmp_path / "ap.vtk", 3)
    txt = (tmp_path / "ap.vtk").read_text().split("\n")
    assert txt[4].startswith("DIMENSIONS 4 4 4")
    assert T.v_guard()[2] == 0


def test_checkpoint_of_the_earlier_format(lqt, tmp_path):
    """A checkpoint whose FP32 slow concentrations are stored in double (the
    earlier format, magic LQTTIS2, no reference values) restores the same
    state: built here from a current checkpoint (LQTTIS3: header, time, 5
    reference values, the float slots, the 5 slow slots as float offsets,
    the V grid), it loads into a fresh tissue whose 40 states equal the
    saved ones to FP32 rounding of the offsets, and both continue alike."""
    lab, fib, sites = _blob(2, (12, 12, 10))
    def make():
        T = lqt.Tissue(lab, fib, 0.2, 0.154, 0.0385, lqt.preset("cpu_default"), lqt.FP32, 0.02)
        T.set_stimulus(sites, None, strength=-80.0)
        return T
    A = make()
    A.step(400)
    A.save(tmp_path / "v3.bin")
    raw = (tmp_path / "v3.bin").read_bytes()
    assert raw[:8] == b"LQTTIS3\0"
    nvox = A.nvox
    pitch = (nvox + 255) // 256 * 256
    head = 8 + 6 * 8 + 8
    dref = np.frombuffer(raw, np.float64, 5, head)
    o = head + 5 * 8
    nr = 4 * 32 * pitch
    r = raw[o:o + nr]
    d = np.frombuffer(raw, np.float32, 5 * pitch, o + nr)
    g = raw[o + nr + 4 * 5 * pitch:]
    # chunk-major slow slots: chunk c, slot j, lane l at (c * 5 + j) * 64 + l
    dd = d.astype(np.float64).reshape(-1, 5, 64) + dref[None, :, None]
    old = b"LQTTIS2\0" + raw[8:head] + r + dd.astype(np.float64).tobytes() + g
    (tmp_path / "v2.bin").write_bytes(old)
    B = make()
    B.load(tmp_path / "v2.bin")
    assert B.time() == A.time()
    for k in (0, 1, 63, 64, nvox - 1):
        sa, sb = A.state(k), B.state(k)
        rel = np.abs(sa - sb) / np.maximum(np.abs(sa), 1e-3)
        assert rel.max() < 1e-6, (k, rel.max())
    A.step(200)
    B.step(200)
    dv = np.max(np.abs(A.v() - B.v()))
    print(f"earlier-format checkpoint: {nvox} voxels, max |dV| after 200 more steps {dv:.2e} mV")
    assert dv < 1e-3


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_launch_shapes_agree_and_autotune(lqt, prec):
    """Every launch shape of the fused step (block size, XCD map, register
    budget, LDS staging) computes each voxel the same way: identical V and
    states, in 3D and on a 2D sheet (which runs shape 0 for every shape).
    The autotuner leaves the simulation where it was and keeps the
    fastest."""
    P = lqt.FP32 if prec == "fp32" else lqt.FP64
    lab, fib, sites = _blob(3, (16, 14, 12))
    shapes = lqt.Tissue.variants()
    assert len(shapes) >= 4 and shapes[0] == (256, False, 2, False)
    assert any(s[3] for s in shapes)
    # a slab of long z-runs (260 voxels per column), so that most
    # wavefronts are one run and the LDS-staged shapes take their LDS path
    slab = np.full((6, 6, 262), 255, np.uint8)
    slab[1:5, 1:5, 1:261] = 3
    slab[2, 2, 100] = 255  # a hole: ghosts inside the slab
    zz = np.arange(262)[None, None, :]
    sfib = np.zeros(slab.shape + (3,), np.float32)
    sfib[..., 0] = np.cos(zz / 40.0)
    sfib[..., 2] = np.sin(zz / 40.0)
    ssites = np.flatnonzero((slab != 255) & (zz < 4))
    # a shell of about 366k voxels (5,700 chunks of 64 voxels)
    big = _blob(4, (112, 104, 96))
    # a 2D ring sheet with holes and tangential fibres
    xx, yy = np.meshgrid(np.arange(90), np.arange(80), indexing="ij")
    rr = np.hypot(xx - 44.5, yy - 39.5)
    ring = (rr > 20) & (rr < 34) & (np.random.default_rng(2).random(rr.shape) > 0.02)
    sheet = np.where(ring, np.where(rr < 25, 3, np.where(rr < 29, 4, 2)), 255).astype(np.uint8)[:, :, None]
    th = np.arctan2(yy - 39.5, xx - 44.5)
    shf = np.zeros((90, 80, 1, 3), np.float32)
    shf[:, :, 0, 0], shf[:, :, 0, 1] = -np.sin(th), np.cos(th)
    shsites = np.flatnonzero(ring & (rr < 22))
    for geo in ((lab, fib, sites), (slab, sfib, ssites), big, (sheet, shf, shsites)):
        ref = None
        for i in range(len(shapes)):
            T = lqt.Tissue(geo[0], geo[1], 0.2, 0.154, 0.0385, lqt.preset("cpu_default"), P, 0.02)
            T.set_stimulus(geo[2], None, strength=-80.0)
            T.set_variant(i)
            T.step(57)
            T.step(263)
            out = (T.v(), T.state(T.nvox // 3))
            if ref is None:
                ref = out
                assert np.ptp(out[0]) > 1.0
            else:
                assert np.array_equal(out[0], ref[0]), shapes[i]
                assert np.array_equal(out[1], ref[1]), shapes[i]
            T.close()
    T = lqt.Tissue(lab, fib, 0.2, 0.154, 0.0385, lqt.preset("cpu_default"), P, 0.02)
    T.set_stimulus(sites, None, strength=-80.0)
    T.step(120)
    ref = (T.v(), T.state(T.nvox // 3))
    T.close()
    T = lqt.Tissue(lab, fib, 0.2, 0.154, 0.0385, lqt.preset("cpu_default"), P, 0.02)
    T.set_stimulus(sites, None, strength=-80.0)
    T.step(60)
    v60, clock = T.v(), T.time()
    best, us = T.autotune(8)
    print(f"{prec} shapes {shapes}: us {np.round(us, 1).tolist()}, best {best}")
    assert T.variant() == best and us[best] == min(us) and all(u > 0 for u in us)
    assert np.array_equal(T.v(), v60) and T.time() == clock
    T.step(60)
    assert np.array_equal(T.v(), ref[0])
