"""Batched 1D transmural strands (lqt_strands_*): many OneD.cc runs, each with
its own length, cell types, parameters and pacing, in one persistent launch."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L


@dataclass
class StrandSpec:
    celltype: np.ndarray            # per cell enum CellType
    params: L.Params
    bcl: float = 1000.0
    nbeats: int = 5
    dx: float = 0.15
    coeff: np.ndarray | float = 0.154
    stim_start: float = 50.0
    stim_duration: float = 1.0
    stim_strength: float = -52.0
    nstim: int = 3
    rec_every: int = 10             # pseudo-ECG every 0.2 ms at dt 0.02
    electrode_mm: float = 20.0      # beyond the last cell (plot1DECG_compare.py:86)
    stop_time: float | None = None  # default nbeats * bcl
    measure_beat: int | None = None  # default the last beat
    rec_all_beats: bool = False
    vrec_every: int = 0             # V of every cell every vrec_every steps (0 = off)
    rec_from: float | None = None   # default: start of the measured beat (0 with rec_all_beats)
    # S2 (OneD_VW.cc:291-301): time - s2_base in [s2_start, s2_start + s2_duration)
    s2_start: float = 0.0
    s2_duration: float = 0.0        # 0 = no S2
    s2_strength: float = -104.0
    s2_first: int = 10
    s2_count: int = 3
    s2_base: float | None = None    # default (nbeats - 1) * bcl
    extra: dict = field(default_factory=dict)


def transmural(n, endo=0.25, mid=0.35):
    """Cell types of a transmural strand: endo | M | epi (OneD.cc:61-63 uses 25/35/40 of 100)."""
    ne = int(round(endo * n))
    nm = int(round(mid * n))
    ct = np.full(n, L.EPI, np.uint8)
    ct[:ne] = L.ENDO
    ct[ne:ne + nm] = L.MCELL
    return ct


class Strands:
    def __init__(self, specs, precision=L.FP32, dt=0.02):
        self.specs = list(specs)
        n = len(self.specs)
        if n == 0:  # the library refuses an empty batch; let it say so
            L.check(L.lib().lqt_strands_create(0, None, None, None, None, 0, precision, dt, C.byref(C.c_void_p())))
        descs = (L.StrandDesc * n)()
        cts, coeffs, plist = [], [], []
        off = 0
        for i, s in enumerate(self.specs):
            ct = np.ascontiguousarray(s.celltype, np.uint8)
            L_ = ct.size
            stop = s.stop_time if s.stop_time is not None else s.nbeats * s.bcl
            mb = s.measure_beat if s.measure_beat is not None else int(stop // s.bcl) - 1
            rec_from = s.rec_from if s.rec_from is not None else (0.0 if s.rec_all_beats else mb * s.bcl)
            x0 = L_ * s.dx + s.electrode_mm
            s2b = s.s2_base if s.s2_base is not None else (s.nbeats - 1) * s.bcl
            descs[i] = L.StrandDesc(L_, off, i, s.dx, s.bcl, stop, s.stim_start, s.stim_duration,
                                    s.stim_strength, s.nstim, s.rec_every, rec_from, x0, mb, s.vrec_every,
                                    s2b, s.s2_start, s.s2_duration, s.s2_strength, s.s2_first, s.s2_count)
            cts.append(ct)
            coeffs.append(np.broadcast_to(np.asarray(s.coeff, np.float64), (L_,)))
            plist.append(s.params)
            off += L_
        self.celltype = np.ascontiguousarray(np.concatenate(cts), np.uint8)
        self.coeff = np.ascontiguousarray(np.concatenate(coeffs), np.float64)
        self._descs = descs
        self._params = L.params_array(plist)
        h = C.c_void_p()
        L.check(L.lib().lqt_strands_create(n, descs, L.ptr(self.celltype), L.ptr(self.coeff), self._params, n,
                                           precision, dt, C.byref(h)))
        self.h = h
        self.n = n
        self.dt = dt

    def run(self):
        L.check(L.lib().lqt_strands_run(self.h))

    def last_ms(self):
        v = C.c_double()
        L.check(L.lib().lqt_strands_last_ms(self.h, C.byref(v)))
        return v.value

    def autotune(self, trial_steps=2000):
        """Time every launch layout (0: per-size classes, 1: one strand per
        workgroup at the largest block, 2: packed bundles, 3: per-size
        classes with each cell of strands up to 256 cells split over two
        lanes) on the batch with its stop times scaled so that the longest
        strand runs trial_steps steps, and keep the fastest; returns (mode,
        [ms per layout])."""
        mode = C.c_int()
        ms = np.zeros(4)
        L.check(L.lib().lqt_strands_autotune(self.h, int(trial_steps), C.byref(mode), L.ptr(ms)))
        return mode.value, ms.tolist()

    def set_mode(self, mode):
        L.check(L.lib().lqt_strands_set_mode(self.h, int(mode)))

    def block(self):
        v = C.c_int()
        L.check(L.lib().lqt_strands_block(self.h, C.byref(v)))
        return v.value

    def results(self):
        out = (L.StrandResult * self.n)()
        L.check(L.lib().lqt_strands_results(self.h, out))
        return [dict(qt=r.qt, apd90=tuple(r.apd90), ead=r.ead, act_first=r.act_first,
                     act_last=r.act_last, cv=r.cv, steps=r.steps, vmax_first=r.vmax_first,
                     vmax_last=r.vmax_last) for r in out]

    def cell_results(self, i):
        n = self._descs[i].ncell
        act, apd, ead = np.zeros(n), np.zeros(n), np.zeros(n, np.int32)
        L.check(L.lib().lqt_strands_cell_results(self.h, i, L.ptr(act), L.ptr(apd), L.ptr(ead)))
        return act, apd, ead

    def ecg(self, i):
        n, t0 = C.c_int64(), C.c_double()
        L.check(L.lib().lqt_strands_ecg(self.h, i, None, C.byref(t0), C.byref(n)))
        phi = np.zeros(n.value)
        L.check(L.lib().lqt_strands_ecg(self.h, i, L.ptr(phi), C.byref(t0), C.byref(n)))
        d = self._descs[i]
        t = t0.value + np.arange(n.value) * d.rec_every * self.dt
        return t, phi

    def vtrace(self, i):
        n, t0 = C.c_int64(), C.c_double()
        L.check(L.lib().lqt_strands_vtrace(self.h, i, None, C.byref(t0), C.byref(n)))
        nc = self._descs[i].ncell
        v = np.zeros((n.value, nc))
        L.check(L.lib().lqt_strands_vtrace(self.h, i, L.ptr(v), C.byref(t0), C.byref(n)))
        t = t0.value + np.arange(n.value) * self._descs[i].vrec_every * self.dt
        return t, v

    def final_state(self, i):
        out = np.zeros((self._descs[i].ncell, 40))
        L.check(L.lib().lqt_strands_final_state(self.h, i, L.ptr(out)))
        return out

    def close(self):
        if getattr(self, "h", None):
            L.lib().lqt_strands_destroy(self.h)
            self.h = None

    __del__ = close
