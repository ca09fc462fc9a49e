"""Monodomain tissue on the GPU (lqt_tissue_*): the reference's 2D/3D drivers
(GPU3D.cu, new_11.cu) as a handle over a label grid with fibres."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L


class Tissue:
    def __init__(self, label, fibre=None, dx=0.2, d_fibre=0.154, d_cross=0.154, params=None,
                 precision=L.FP32, dt=0.02):
        label = np.ascontiguousarray(label, np.uint8)
        if label.ndim == 2:
            label = label[:, :, None]
        self.shape = label.shape
        self._label = label
        self._fibre = None if fibre is None else np.ascontiguousarray(fibre, np.float32).reshape(
            label.shape + (3,))
        g = L.Geometry(label.shape[0], label.shape[1], label.shape[2], dx, label.ctypes.data,
                       None if self._fibre is None else self._fibre.ctypes.data, d_fibre, d_cross)
        self.params = params if params is not None else L.preset("cpu_default")
        h = C.c_void_p()
        L.check(L.lib().lqt_tissue_create(C.byref(g), C.byref(self.params), precision, dt, C.byref(h)))
        self.h = h
        self.dt = dt
        self.precision = precision
        nv, ng = C.c_int64(), C.c_int64()
        L.check(L.lib().lqt_tissue_info(h, C.byref(nv), C.byref(ng)))
        self.nvox, self.nghost = nv.value, ng.value

    def set_stimulus(self, sites, offsets_ms=None, bcl=1000.0, duration=1.0, strength=-52.0,
                     s2_sites=None, s2_time=0.0, s2_duration=0.0, s2_strength=-104.0, window=0.0):
        """sites: grid indices (x*ny + y)*nz + z; offsets_ms: per-site start within the cycle;
        window: S1 only while the in-cycle time is below it (0 = no limit)."""
        sites = np.ascontiguousarray(sites, np.int64)
        offs = None if offsets_ms is None else np.ascontiguousarray(offsets_ms, np.float32)
        s2 = None if s2_sites is None else np.ascontiguousarray(s2_sites, np.int64)
        p = L.StimProtocol(bcl, duration, strength, s2_time, s2_duration, s2_strength, window)
        L.check(L.lib().lqt_tissue_set_stimulus(self.h, C.byref(p), sites.size, L.ptr(sites), L.ptr(offs),
                                                0 if s2 is None else s2.size, L.ptr(s2)))

    def step(self, nsteps=1):
        L.check(L.lib().lqt_tissue_step(self.h, int(nsteps)))

    def time(self):
        t, s = C.c_double(), C.c_int64()
        L.check(L.lib().lqt_tissue_time(self.h, C.byref(t), C.byref(s)))
        return t.value, s.value

    def v(self):
        out = np.zeros(self.nvox)
        L.check(L.lib().lqt_tissue_read_v(self.h, L.ptr(out)))
        return out

    def state(self, k):
        out = np.zeros(40)
        L.check(L.lib().lqt_tissue_read_state(self.h, int(k), L.ptr(out)))
        return out

    def write_state(self, k, s):
        s = np.ascontiguousarray(s, np.float64)
        L.check(L.lib().lqt_tissue_write_state(self.h, int(k), L.ptr(s)))

    def v_guard(self):
        out = np.zeros(3, np.int64)
        L.check(L.lib().lqt_tissue_v_guard(self.h, L.ptr(out)))
        return out

    def track(self, t_from=0.0, t_to=1e300):
        """Start (reset) the activation / APD90 trackers for steps starting
        in [t_from, t_to) ms of the tissue clock; t_to <= t_from stops them."""
        L.check(L.lib().lqt_tissue_track(self.h, float(t_from), float(t_to)))

    def track_results(self):
        """(activation time, APD90) per voxel in ms, -1 where none."""
        act, apd = np.zeros(self.nvox), np.zeros(self.nvox)
        L.check(L.lib().lqt_tissue_track_results(self.h, L.ptr(act), L.ptr(apd)))
        return act, apd

    def kernel_us(self):
        a, b = C.c_double(), C.c_double()
        L.check(L.lib().lqt_tissue_last_kernel_us(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def bytes_per_step(self):
        b = C.c_double()
        L.check(L.lib().lqt_tissue_bytes_per_step(self.h, C.byref(b)))
        return b.value

    @staticmethod
    def variants():
        """Launch shapes of the fused step kernel: [(block, xcd_map, waves, lds_stage), ...]."""
        n = C.c_int()
        L.check(L.lib().lqt_tissue_variants(C.byref(n)))
        out = []
        for i in range(n.value):
            b, x, w, s = C.c_int(), C.c_int(), C.c_int(), C.c_int()
            L.check(L.lib().lqt_tissue_variant_info(i, C.byref(b), C.byref(x), C.byref(w), C.byref(s)))
            out.append((b.value, bool(x.value), w.value, bool(s.value)))
        return out

    def autotune(self, trial_steps=20):
        """Time every launch shape from the current state (which is restored)
        and keep the fastest; returns (best index, [us per shape])."""
        n = len(self.variants())
        us = np.zeros(n)
        best = C.c_int()
        L.check(L.lib().lqt_tissue_autotune(self.h, int(trial_steps), C.byref(best), L.ptr(us)))
        return best.value, us.tolist()

    def set_variant(self, i):
        L.check(L.lib().lqt_tissue_set_variant(self.h, int(i)))

    def variant(self):
        i = C.c_int()
        L.check(L.lib().lqt_tissue_get_variant(self.h, C.byref(i)))
        return i.value

    def write_vtk(self, path, stride=3):
        L.check(L.lib().lqt_tissue_write_vtk(self.h, str(path).encode(), int(stride)))

    def save(self, path):
        L.check(L.lib().lqt_tissue_save(self.h, str(path).encode()))

    def load(self, path):
        L.check(L.lib().lqt_tissue_load(self.h, str(path).encode()))

    def close(self):
        if getattr(self, "h", None):
            L.lib().lqt_tissue_destroy(self.h)
            self.h = None

    __del__ = close
