"""Independent TPORd cells on the GPU (lqt_cells_*).

``Cells`` batches n cells; ``TPORd`` mirrors the reference's per-cell object
(SingleCell/TPORd.cc:137-376: setDt, setIstim, setDVgap_dt, update, getV,
outputAllStates, readinAllStates) on top of a one-cell batch.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L

NSTATE = 40
STATS_KEYS = ("rest", "dvdt_max", "overshoot", "amplitude", "apd20", "apd25", "apd50", "apd75",
              "apd80", "apd90")


class Cells:
    def __init__(self, celltype, params=None, precision=L.FP64, dt=0.02):
        celltype = np.ascontiguousarray(np.atleast_1d(celltype), np.uint8)
        self.n = celltype.size
        if params is None:
            params = [L.preset("cpu_default")]
        elif isinstance(params, L.Params):
            params = [params]
        arr = L.params_array(params)
        self._arr = arr
        h = C.c_void_p()
        L.check(L.lib().lqt_cells_create(self.n, L.ptr(celltype), arr, len(params), precision, dt,
                                         C.byref(h)))
        self.h = h
        self.dt = dt

    def step(self, nsteps=1, istim=None, dvgap=None):
        ist = None if istim is None else np.ascontiguousarray(np.broadcast_to(istim, (self.n,)), np.float64)
        dvg = None if dvgap is None else np.ascontiguousarray(np.broadcast_to(dvgap, (self.n,)), np.float64)
        L.check(L.lib().lqt_cells_step(self.h, int(nsteps), L.ptr(ist), L.ptr(dvg)))

    @property
    def lanes(self):
        """Lanes per cell (1, 2 or 4) of the next step or pace call."""
        v = C.c_int(0)
        L.check(L.lib().lqt_cells_get_lanes(self.h, C.byref(v)))
        return v.value

    def set_lanes(self, lanes):
        """Lanes per cell: 1, 2 or 4, or 0 for the automatic choice."""
        L.check(L.lib().lqt_cells_set_lanes(self.h, int(lanes)))

    def state(self):
        out = np.zeros((self.n, NSTATE))
        L.check(L.lib().lqt_cells_get_state(self.h, L.ptr(out)))
        return out

    def set_state(self, s):
        s = np.ascontiguousarray(s, np.float64).reshape(self.n, NSTATE)
        L.check(L.lib().lqt_cells_set_state(self.h, L.ptr(s)))

    def v(self):
        out = np.zeros(self.n)
        L.check(L.lib().lqt_cells_get_v(self.h, L.ptr(out)))
        return out

    def pace(self, bcl=1000.0, nbeats=100, stim_start=50.0, stim_duration=1.0, stim_strength=-53.0,
             trace_cell=-1, trace_every=0):
        """SingleCell.cc protocol for every cell; returns (stats (n,10), trace or None)."""
        p = L.Pacing(bcl, stim_start, stim_duration, stim_strength, nbeats)
        stats = np.zeros((self.n, 10))
        trace = None
        nt = C.c_int64(0)
        if trace_cell >= 0 and trace_every > 0:
            cap = int(nbeats * bcl / self.dt / trace_every) + 4
            trace = np.zeros((cap, 3))
        L.check(L.lib().lqt_cells_pace(self.h, C.byref(p), L.ptr(stats), int(trace_cell), int(trace_every),
                                       L.ptr(trace), 0 if trace is None else trace.shape[0], C.byref(nt)))
        if trace is not None:
            trace = trace[: nt.value]  # columns V, ICaL, IKs
        return stats, trace

    def close(self):
        if getattr(self, "h", None):
            L.lib().lqt_cells_destroy(self.h)
            self.h = None

    __del__ = close


class TPORd:
    """One cell with the method names of the reference class (TPORd.cc:137-376)."""

    def __init__(self, celltype, params=None, precision=L.FP64, dt=0.005):
        self._ct = celltype
        self._params = params
        self._prec = precision
        self._cells = Cells([celltype], params, precision, dt)
        self._ist = 0.0
        self._dvgap = 0.0

    def setDt(self, dt):
        s = self._cells.state()
        self._cells.close()
        self._cells = Cells([self._ct], self._params, self._prec, dt)
        self._cells.set_state(s)

    def setIstim(self, v):
        self._ist = float(v)

    def setDVgap_dt(self, v):
        self._dvgap = float(v)

    def update(self):
        self._cells.step(1, self._ist, self._dvgap)

    def getV(self):
        return float(self._cells.v()[0])

    def outputAllStates(self, path):
        L.check(L.lib().lqt_cells_write_state_file(self._cells.h, 0, str(path).encode()))

    def readinAllStates(self, path):
        L.check(L.lib().lqt_cells_read_state_file(self._cells.h, 0, str(path).encode()))

    def states(self):
        return self._cells.state()[0]
