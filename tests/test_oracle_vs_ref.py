"""The oracle against the reference's own TPORd class (oracle/_ref, built
from /root/reference by oracle/Makefile), off the paced trajectory: random
states, membrane potentials and gap currents.  Skipped where the reference
tree (and so oracle/_ref) is absent."""
import ctypes as C

import numpy as np
import pytest

import oracle as O

VARIANTS = {
    "con_iso": dict(),
    "lqt1_iso": dict(iks_scale=0.05),
    "suprep_iso": dict(suprep=1, dose_fmol=150.0),
    "lqt1_suprep_iso": dict(iks_scale=0.05, suprep=1, dose_fmol=150.0),
}


def _random_state(rng):
    s = O.cell_init()
    s[0] = rng.uniform(-95.0, 55.0)
    s[1] = rng.uniform(5.0, 15.0)
    s[2] = s[1] * rng.uniform(0.95, 1.05)
    s[3] = rng.uniform(130.0, 150.0)
    s[4] = s[3] * rng.uniform(0.99, 1.01)
    s[5] = rng.uniform(5e-5, 1e-3)
    s[6] = s[5] * rng.uniform(0.5, 20.0)
    s[7] = rng.uniform(0.5, 3.0)
    s[8] = s[7] * rng.uniform(0.3, 1.2)
    for k in range(9, 39):
        s[k] = rng.uniform(0.0, 1.0)
    s[37] = rng.uniform(-5.0, 0.0)  # Jrelnp
    s[38] = rng.uniform(-5.0, 0.0)  # Jrelp
    s[39] = rng.uniform(0.0, 0.8)   # CaMKt
    return s


@pytest.mark.parametrize("variant", list(VARIANTS))
def test_random_states_bitwise(variant):
    if not O.ref_available(variant):
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(sum(map(ord, variant)))
    p = O.params(**VARIANTS[variant])
    for trial in range(200):
        ct = int(rng.choice([O.LVEPI, O.LVENDO, O.EPI, O.ENDO, O.MCELL]))
        s0 = _random_state(rng)
        ist = float(rng.choice([0.0, -52.0, rng.uniform(-80, 10)]))
        dvg = float(rng.uniform(-50, 50))
        dt = float(rng.choice([0.005, 0.02, 0.05]))
        ref = O.RefCell(ct, dt=dt, variant=variant)
        ref.set_state(s0)
        s = s0.copy()
        for _ in range(3):
            ref.step(ist, dvg)
            O.cell_update(s, ct, p, dt, ist, dvg)
        r = ref.state()
        same = (s == r) | (np.isnan(s) & np.isnan(r))
        assert same.all(), (variant, trial, np.flatnonzero(~same))


def test_strand_steps_bitwise():
    """Strand coupling of OneD.cc:364-391 with the reference class."""
    if not O.ref_available():
        pytest.skip("oracle/_ref not built (reference tree absent)")
    L = O.ref_lib()
    n = 30
    ct = np.array([O.ENDO] * 8 + [O.MCELL] * 10 + [O.EPI] * 12, np.uint8)
    cells = [O.RefCell(int(c), dt=0.02) for c in ct]
    arr = (C.c_void_p * n)(*[c.h for c in cells])
    tio = C.c_double(0.0)
    L.ref_strand_steps(arr, n, 0.154, 0.15, 0.02, 6000, 0.0, 1000.0, 50.0, 1.0, -52.0, 3, C.byref(tio))
    r = O.strand_run(ct, 0.154, O.params(), stop_time=119.99, rec_every=0, record=False)
    assert r["steps"] == 6000
    ref = np.array([c.state() for c in cells])
    assert np.array_equal(ref, r["state"])
