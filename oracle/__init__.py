"""ctypes front end of the FP64 CPU oracle (``oracle/liboracle.so``).

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` import this package, and only as the
reference that results are compared with.  The product package ``lqt1_amd``
never imports it.

Every function restates the reference CPU path (see ``oracle.h`` for the
file:line map); ``ref_*`` helpers load the reference's own TPORd class from
``oracle/_ref`` when that was built (only where the reference tree exists).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

NSTATE = 40
STATE_NAMES = (
    "v nai nass ki kss cai cass cansr cajsr sm sh sj hsp jp mL hL hLp a iF iS ap iFp iSp "
    "d ff fs fcaf fcas jca nca ffp fcafp xrf xrs xs1 xs2 xk1 Jrelnp Jrelp CaMKt"
).split()
LVEPI, LVENDO, EPI, ENDO, MCELL = 0, 1, 2, 3, 4


class Params(C.Structure):
    _fields_ = [
        ("iks_scale", C.c_double),
        ("iso", C.c_int),
        ("pca_iso", C.c_double),
        ("suprep", C.c_int),
        ("dose_fmol", C.c_double),
        ("sca", C.c_double),
        ("clamp", C.c_int),
    ]


def params(iks_scale=1.0, iso=1, pca_iso=1.6, suprep=0, dose_fmol=150.0, sca=1.0, clamp=0):
    """Runtime form of the reference's compile-time flags (TPORd.cc:30-36)."""
    return Params(iks_scale, iso, pca_iso, suprep, dose_fmol, sca, clamp)


class Track(C.Structure):
    _fields_ = [("act", C.c_double), ("rep", C.c_double), ("vrest", C.c_double),
                ("vmax", C.c_double), ("vmin", C.c_double), ("flags", C.c_int)]


class S2(C.Structure):
    _fields_ = [("base", C.c_double), ("start", C.c_double), ("duration", C.c_double),
                ("strength", C.c_double), ("first", C.c_int), ("count", C.c_int)]


class Diag(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("Iion", "dV_dt", "ICaL", "IKs", "IKr", "IK1", "INa")]


_lib = None


def build():
    """Compile liboracle.so (and oracle/_ref where the reference exists)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        u8 = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
        i32 = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        f32 = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
        L.or_cell_init.argtypes = [dp, C.c_double]
        L.or_cell_update.argtypes = [dp, C.c_int, C.POINTER(Params), C.c_double, C.c_double,
                                     C.c_double, C.POINTER(Diag)]
        L.or_cells_update.argtypes = [C.c_int, dp, u8, C.POINTER(Params), C.c_double,
                                      C.c_void_p, C.c_void_p]
        L.or_single_cell.restype = C.c_int64
        L.or_single_cell.argtypes = [C.c_int, C.POINTER(Params), C.c_double, C.c_int, C.c_double,
                                     C.c_double, C.c_double, C.c_double, dp, C.c_void_p,
                                     C.c_int, C.c_int64]
        L.or_strand_run.restype = C.c_int64
        L.or_strand_run.argtypes = [C.c_int, u8, dp, C.POINTER(Params), C.c_double, C.c_double,
                                    C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                    C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int64,
                                    C.POINTER(C.c_int64), C.c_int, C.c_void_p, C.c_void_p]
        L.or_strand_run2.restype = C.c_int64
        L.or_strand_run2.argtypes = [C.c_int, u8, dp, C.POINTER(Params), C.c_double, C.c_double,
                                     C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                     C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int64,
                                     C.POINTER(C.c_int64), C.c_int, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_void_p]
        L.or_strand_vw.restype = C.c_int64
        L.or_strand_vw.argtypes = [C.c_int, u8, dp, C.POINTER(Params), C.c_double, C.c_double,
                                   C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                   C.c_int, C.POINTER(S2), dp, C.POINTER(C.c_double),
                                   C.POINTER(C.c_int64), C.c_double, C.c_void_p, C.c_void_p,
                                   C.c_void_p, C.c_int, C.c_int, C.c_int64, C.POINTER(C.c_int64)]
        L.or_track_init.argtypes = [C.POINTER(Track)]
        L.or_track_step.argtypes = [C.POINTER(Track), C.c_double, C.c_double]
        L.or_track_result.restype = C.c_int
        L.or_track_result.argtypes = [C.POINTER(Track), C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.or_pseudo_ecg.argtypes = [C.c_int64, C.c_int, dp, C.c_double, C.c_double, dp]
        L.or_ghost_map.restype = C.c_int64
        L.or_ghost_map.argtypes = [C.c_int, C.c_int, C.c_int, u8, i32]
        L.or_tissue_tensor.argtypes = [C.c_int, C.c_int, C.c_int, i32, C.c_int64, f32, C.c_double,
                                       C.c_double, C.c_double, dp, dp]
        L.or_tissue_du.argtypes = [C.c_int, C.c_int, C.c_int, i32, C.c_int64, dp, dp, C.c_double,
                                   dp, dp]
        L.or_tissue_step.argtypes = [C.c_int, C.c_int, C.c_int, i32, C.c_int64, dp, dp, C.c_double,
                                     C.c_double, u8, C.POINTER(Params), dp, C.c_void_p, dp, dp]
        L.or_num_threads.restype = C.c_int
        L.or_set_num_threads.argtypes = [C.c_int]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def cell_init(v0=-87.5):
    s = np.zeros(NSTATE)
    lib().or_cell_init(s, v0)
    return s


def cell_update(s, celltype, p, dt, ist=0.0, dvgap=0.0):
    """One TPORd::update() (TPORd.cc:379-1021) in place; returns the diagnostics."""
    d = Diag()
    lib().or_cell_update(s, int(celltype), C.byref(p), dt, ist, dvgap, C.byref(d))
    return d


def cells_update(s, celltype, p, dt, ist=None, dvgap=None):
    """Independent update of n cells, s shaped (n, 40)."""
    n = s.shape[0]
    lib().or_cells_update(n, s, np.ascontiguousarray(celltype, np.uint8), C.byref(p), dt,
                          _ptr(ist), _ptr(dvgap))


def single_cell(celltype=MCELL, p=None, bcl=1000.0, nbeats=100, dt=0.02, stim=-53.0,
                stim_dur=1.0, stim_start=50.0, trace_every=0):
    """SingleCell.cc protocol; returns (stats dict, trace or None)."""
    p = p or params()
    stats = np.zeros(10)
    nsteps = int(round(nbeats * bcl / dt)) + 2
    trace = None
    if trace_every:
        trace = np.zeros(nsteps // trace_every + 2)
    lib().or_single_cell(int(celltype), C.byref(p), bcl, nbeats, dt, stim, stim_dur, stim_start,
                         stats, _ptr(trace), trace_every, 0 if trace is None else trace.size)
    keys = ("rest", "dvdt_max", "overshoot", "amplitude", "apd20", "apd25", "apd50", "apd75",
            "apd80", "apd90")
    return dict(zip(keys, stats.tolist())), trace


def strand_run(celltype, coeff, p=None, dx=0.15, dt=0.02, bcl=1000.0, stop_time=5000.0,
               stim=-52.0, stim_dur=1.0, stim_start=50.0, nstim=3, rec_every=10, act_beat=None,
               record=True):
    """OneD.cc strand; returns dict(v=(rows, n), t=(rows,), act=(n,), state=(n,40), steps)."""
    p = p or params()
    celltype = np.ascontiguousarray(celltype, np.uint8)
    n = celltype.size
    coeff = np.ascontiguousarray(np.broadcast_to(np.asarray(coeff, np.float64), (n,)))
    cap = int(stop_time / dt / max(rec_every, 1)) + 3 if record else 0
    vrec = np.zeros((cap, n)) if record else None
    rt = np.zeros(cap) if record else None
    act = np.zeros(n)
    apd = np.zeros(n)
    ead = np.zeros(n, np.int32)
    fin = np.zeros((n, NSTATE))
    rows = C.c_int64(0)
    ab = int(act_beat) if act_beat is not None else int(stop_time // bcl) - 1
    steps = lib().or_strand_run2(n, celltype, coeff, C.byref(p), dx, dt, bcl, stop_time, stim,
                                 stim_dur, stim_start, nstim, _ptr(vrec), _ptr(rt),
                                 rec_every if record else 0, cap, C.byref(rows), ab, _ptr(act),
                                 _ptr(apd), _ptr(ead), _ptr(fin))
    r = rows.value
    return dict(v=None if vrec is None else vrec[:r], t=None if rt is None else rt[:r],
                act=act, apd=apd, ead=ead, state=fin, steps=int(steps))


class StrandVW:
    """OneD_VW.cc:239-398 for one S2 candidate, resumable (oracle restatement).

    ``run(pause_at)`` advances to the first step with time >= pause_at (or
    past the stop time); ``fork(s2_start)`` copies the paused run for
    another candidate, which is exact because the S2 test is false for
    every step before its window.  ``vmax`` holds the peaks of cell 0 and
    cell n-1 (OneD_VW.cc:390-394), ``rows``/``row_t`` the recorded V."""

    def __init__(self, celltype, coeff, p=None, dx=0.15, dt=0.02, bcl=1000.0, nbeats=5, stim=-52.0,
                 stim_dur=3.0, stim_start=0.0, nstim=3, s2_start=260.0, s2_duration=3.0, s2_strength=-104.0,
                 s2_first=10, s2_count=3, rec_every=50, record=True):
        self.ct = np.ascontiguousarray(celltype, np.uint8)
        n = self.ct.size
        self.coeff = np.ascontiguousarray(np.broadcast_to(np.asarray(coeff, np.float64), (n,)))
        self.p = p or params()
        self.dx, self.dt, self.bcl, self.nbeats = dx, dt, bcl, nbeats
        self.stim = (stim, stim_dur, stim_start, nstim)
        self.s2 = S2((nbeats - 1) * bcl, s2_start, s2_duration, s2_strength, s2_first, s2_count)
        self.state = np.zeros((n, NSTATE))
        for i in range(n):
            self.state[i] = cell_init()
        self.time = C.c_double(0.0)
        self.step = C.c_int64(0)
        self.vmax = np.array([-100.0, -100.0])
        self.rec_every = rec_every if record else 0
        cap = int(bcl / dt / max(rec_every, 1)) + 2 if record else 0
        self.rows = np.zeros((cap, n)) if record else None
        self.row_t = np.zeros(cap) if record else None
        self.nrows = C.c_int64(0)

    def run(self, pause_at=np.inf):
        st, sd, ss, ns = self.stim
        lib().or_strand_vw(self.ct.size, self.ct, self.coeff, C.byref(self.p), self.dx, self.dt, self.bcl,
                           self.nbeats * self.bcl, st, sd, ss, ns, C.byref(self.s2), self.state,
                           C.byref(self.time), C.byref(self.step), float(pause_at), _ptr(self.vmax),
                           _ptr(self.rows), _ptr(self.row_t), self.rec_every, self.nbeats - 1,
                           0 if self.rows is None else self.rows.shape[0], C.byref(self.nrows))
        return self

    def fork(self, s2_start):
        import copy
        o = copy.copy(self)
        o.state = self.state.copy()
        o.time = C.c_double(self.time.value)
        o.step = C.c_int64(self.step.value)
        o.vmax = self.vmax.copy()
        o.s2 = S2(self.s2.base, s2_start, self.s2.duration, self.s2.strength, self.s2.first, self.s2.count)
        o.rows = None if self.rows is None else self.rows.copy()
        o.row_t = None if self.row_t is None else self.row_t.copy()
        o.nrows = C.c_int64(self.nrows.value)
        return o

    def v_rows(self):
        r = self.nrows.value
        return self.row_t[:r], self.rows[:r]


def vw_candidates(left=260.0, step=0.1, count=1, direction=1):
    """The S2 times OneD_VW.cc's MODE2 (direction +1, :405-423) or MODE3
    (-1, :427-443) visits: s2 = left, then left += step (or -=) each time,
    accumulated in double as the reference does."""
    out = []
    s = left
    for _ in range(count):
        out.append(s)
        s = s + step if direction > 0 else s - step
    return np.array(out)


def track_trace(t, v):
    """Beat measurements of one V trace (oracle restatement of the strand
    tracker): returns (act, apd90, ead_flags)."""
    tr = Track()
    lib().or_track_init(C.byref(tr))
    for ti, vi in zip(np.asarray(t, float), np.asarray(v, float)):
        lib().or_track_step(C.byref(tr), ti, vi)
    a, d = C.c_double(), C.c_double()
    e = lib().or_track_result(C.byref(tr), C.byref(a), C.byref(d))
    return a.value, d.value, e


class MapTracker:
    """The beat tracker of track_trace (activation at the first V >= -30 mV,
    APD90 at the first later V <= Vmax - 0.9 (Vmax - Vrest), Vrest the V of
    the first tracked step) for many voxels at once, fed one step at a time:
    the per-voxel maps of the tissue's trackers.  update(k, v) takes the
    step index k and every voxel's V after that step; results(times) gives
    the step indices (act_step, rep_step, -1 where none) and, with the step
    start times, the maps in ms (act, apd90; -1 where none)."""

    def __init__(self, n):
        self.act = np.full(n, -1, np.int64)
        self.rep = np.full(n, -1, np.int64)
        self.vrest = np.zeros(n)
        self.vmax = np.zeros(n)
        self.started = False

    def update(self, k, v):
        v = np.asarray(v, np.float64)
        if not self.started:
            self.vrest[:] = v
            self.started = True
        was = self.act >= 0
        new = ~was & (v >= -30.0)
        self.act[new] = k
        self.vmax[new] = v[new]
        live = was & (self.rep < 0)
        self.vmax[live] = np.maximum(self.vmax[live], v[live])
        down = live & (v <= self.vmax - 0.9 * (self.vmax - self.vrest))
        self.rep[down] = k

    def results(self, times=None):
        if times is None:
            return self.act.copy(), self.rep.copy()
        times = np.asarray(times, np.float64)
        act = np.where(self.act >= 0, times[np.maximum(self.act, 0)], -1.0)
        rep_t = times[np.maximum(self.rep, 0)]
        apd = np.where(self.rep >= 0, rep_t - act, -1.0)
        return act, apd


def qt_from_phi(t, phi, t_q, t_qrs_end, t_end):
    """QT of a pseudo-ECG (oracle restatement of the strand measurement):
    T peak = largest |phi| in [t_qrs_end, t_end); T end = first sample from
    the peak on with |phi| <= 5 % of it; QT = T end - t_q (-1 if none)."""
    t = np.asarray(t, float)
    a = np.abs(np.asarray(phi, float))
    win = (t >= t_qrs_end) & (t < t_end)
    if not win.any() or a[win].max() <= 0:
        return -1.0
    idx = np.flatnonzero(win)
    ipk = idx[np.argmax(a[win])]
    pk = a[ipk]
    for j in range(ipk, t.size):
        if t[j] >= t_end:
            break
        if a[j] <= 0.05 * pk:
            return float(t[j] - t_q)
    return -1.0


def pseudo_ecg(v, dx=0.15, x0=35.0):
    """plot1DECG_compare.py:89-105 on V rows (T, L)."""
    v = np.ascontiguousarray(v, np.float64)
    T, L = v.shape
    phi = np.zeros(T)
    lib().or_pseudo_ecg(T, L, v, dx, x0, phi)
    return phi


def ghost_map(mask):
    """GPU3D.cu ghost map on a (nx, ny, nz) mask; returns (g int32, nvox)."""
    mask = np.ascontiguousarray(mask, np.uint8)
    nx, ny, nz = mask.shape
    g = np.zeros(mask.shape, np.int32)
    nvox = lib().or_ghost_map(nx, ny, nz, mask, g)
    return g, int(nvox)


def tissue_tensor(g, nvox, fibre, d_fibre, d_cross, dx):
    nx, ny, nz = g.shape
    dten = np.zeros((nvox, 6))
    drift = np.zeros((nvox, 3))
    lib().or_tissue_tensor(nx, ny, nz, np.ascontiguousarray(g, np.int32), nvox,
                           np.ascontiguousarray(fibre, np.float32), d_fibre, d_cross, dx, dten, drift)
    return dten, drift


def tissue_du(g, nvox, dten, drift, dx, v):
    nx, ny, nz = g.shape
    du = np.zeros(nvox)
    lib().or_tissue_du(nx, ny, nz, np.ascontiguousarray(g, np.int32), nvox, dten, drift, dx,
                       np.ascontiguousarray(v, np.float64), du)
    return du


def tissue_step(g, nvox, dten, drift, dx, dt, celltype, p, s, ist=None):
    nx, ny, nz = g.shape
    vb = np.zeros(nvox)
    du = np.zeros(nvox)
    lib().or_tissue_step(nx, ny, nz, g, nvox, dten, drift, dx, dt,
                         np.ascontiguousarray(celltype, np.uint8), C.byref(p), s, _ptr(ist), vb, du)


def num_threads():
    return lib().or_num_threads()


def set_num_threads(n):
    lib().or_set_num_threads(int(n))


# ---- the reference's own class (oracle/_ref) -------------------------------

REF_VARIANTS = {
    "con_iso": "libtpord_ref_con_iso.so",
    "lqt1_iso": "libtpord_ref_lqt1_iso.so",
    "suprep_iso": "libtpord_ref_suprep_iso.so",
    "lqt1_suprep_iso": "libtpord_ref_lqt1_suprep_iso.so",
    "clamp": "libtpord_ref_clamp.so",
}
_ref_libs = {}


def ref_available(variant="con_iso"):
    return os.path.exists(os.path.join(REF_DIR, REF_VARIANTS[variant]))


def ref_lib(variant="con_iso"):
    if variant not in _ref_libs:
        L = C.CDLL(os.path.join(REF_DIR, REF_VARIANTS[variant]))
        dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        L.ref_cell_new.restype = C.c_void_p
        L.ref_cell_new.argtypes = [C.c_int]
        L.ref_cell_free.argtypes = [C.c_void_p]
        L.ref_cell_set_dt.argtypes = [C.c_void_p, C.c_double]
        L.ref_cell_step.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.ref_cell_get.argtypes = [C.c_void_p, dp]
        L.ref_cell_set.argtypes = [C.c_void_p, dp]
        for f in ("ref_cell_dvdt", "ref_cell_ical", "ref_cell_iks"):
            getattr(L, f).restype = C.c_double
            getattr(L, f).argtypes = [C.c_void_p]
        L.ref_strand_steps.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_double, C.c_double,
                                       C.c_double, C.c_long, C.c_double, C.c_double, C.c_double,
                                       C.c_double, C.c_double, C.c_int, C.POINTER(C.c_double)]
        _ref_libs[variant] = L
    return _ref_libs[variant]


class RefCell:
    """The reference's TPORd object (TPORd.cc:137) behind the harness."""

    def __init__(self, celltype, dt=0.02, variant="con_iso"):
        self.L = ref_lib(variant)
        self.h = self.L.ref_cell_new(int(celltype))
        self.L.ref_cell_set_dt(self.h, dt)

    def step(self, ist=0.0, dvgap=0.0):
        self.L.ref_cell_step(self.h, ist, dvgap)

    def state(self):
        o = np.zeros(NSTATE)
        self.L.ref_cell_get(self.h, o)
        return o

    def set_state(self, s):
        self.L.ref_cell_set(self.h, np.ascontiguousarray(s, np.float64))

    def dvdt(self):
        return self.L.ref_cell_dvdt(self.h)

    def __del__(self):
        try:
            self.L.ref_cell_free(self.h)
        except Exception:
            pass
