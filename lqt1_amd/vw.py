"""Vulnerable-window search (lqt_vw_search): OneD_VW.cc's S1-S2 protocol on a
transmural strand, every S2 timing a strand of one batched launch.

The strand of OneD_VW.cc:42-131: 25 ENDO, 35 MCELL, 40 EPI cells at
dx 0.15 mm, D 0.154 mm^2/ms with a 0.2x coupling defect at cell 63 and the
one-sided derivative of D at the ends; S1 -52 pA/pF for 3 ms at t mod BCL in
[0, 3) on cells 0-2, five beats at BCL 1000 ms; S2 -104 pA/pF for 3 ms on
cells 10-12 at s2 ms into the last beat.  A candidate is vulnerable when one
end of the strand goes above -10 mV after the S2 and the other does not
(OneD_VW.cc:405-443).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .strands import transmural


def vw_strand(params=None, ncell=100, defect_cell=63, defect=0.2, d=0.154, numS1=5, bcl=1000.0,
              s2_duration=3.0, s2_strength=-104.0, s2_first=10, s2_count=3):
    """(desc, celltype, coeff) of OneD_VW.cc's strand for lqt_vw_search."""
    ct = transmural(ncell, 0.25, 0.35)
    coeff = np.full(ncell, d)
    if 0 <= defect_cell < ncell:
        coeff[defect_cell] = defect * d
    desc = L.StrandDesc()
    desc.ncell = ncell
    desc.cell_offset = 0
    desc.param_index = 0
    desc.dx = 0.15
    desc.bcl = bcl
    desc.stop_time = numS1 * bcl
    desc.stim_start = 0.0
    desc.stim_duration = 3.0
    desc.stim_strength = -52.0
    desc.nstim = 3
    desc.rec_every = 0
    desc.rec_from = 0.0
    desc.x0 = 0.0
    desc.measure_beat = numS1 - 1
    desc.vrec_every = 0
    desc.s2_base = (numS1 - 1) * bcl
    desc.s2_start = 0.0
    desc.s2_duration = s2_duration
    desc.s2_strength = s2_strength
    desc.s2_first = s2_first
    desc.s2_count = s2_count
    return desc, ct, coeff


@dataclass
class VWResult:
    found: bool
    s2: float
    vmax_first: float
    vmax_last: float
    visited_s2: np.ndarray
    visited_vmax: np.ndarray   # (n, 2): maxEndo, maxEpi
    simulated: int
    launches: int
    device_ms: float


def search(mode=2, left=260.0, right=260.0, step=0.1, boundary=10.0, batch=512, max_candidates=4096,
           params=None, precision=L.FP64, dt=0.02, strand=None):
    """lqt_vw_search on OneD_VW.cc's strand (or `strand` = (desc, celltype, coeff))."""
    p = params if params is not None else L.preset("cpu_default")
    desc, ct, coeff = strand if strand is not None else vw_strand()
    ct = np.ascontiguousarray(ct, np.uint8)
    coeff = np.ascontiguousarray(coeff, np.float64)
    vw = L.VWDesc(int(mode), float(left), float(right), float(step), float(boundary), int(batch),
                  int(max_candidates))
    out = L.VWResult()
    cap = int(max_candidates)
    s2 = np.zeros(cap)
    vm = np.zeros((cap, 2))
    L.check(L.lib().lqt_vw_search(C.byref(desc), L.ptr(ct), L.ptr(coeff), C.byref(p), precision, dt, C.byref(vw),
                                  C.byref(out), L.ptr(s2), L.ptr(vm), cap))
    n = min(out.visited, cap)
    return VWResult(bool(out.found), out.s2, out.vmax_first, out.vmax_last, s2[:n].copy(), vm[:n].copy(),
                    out.simulated, out.launches, out.device_ms)


def scan(s2_times, params=None, precision=L.FP32, dt=0.02, strand=None, boundary=10.0):
    """Every S2 time of `s2_times` as one strand of one launch: returns the
    (n, 2) peaks (maxEndo, maxEpi) and the vulnerable mask."""
    from .strands import Strands, StrandSpec

    p = params if params is not None else L.preset("cpu_default")
    desc, ct, coeff = strand if strand is not None else vw_strand()
    specs = []
    for s2 in np.asarray(s2_times, float):
        specs.append(StrandSpec(ct, p, bcl=desc.bcl, nbeats=int(round(desc.stop_time / desc.bcl)), dx=desc.dx,
                                coeff=coeff, stim_start=desc.stim_start, stim_duration=desc.stim_duration,
                                stim_strength=desc.stim_strength, nstim=desc.nstim, rec_every=0,
                                s2_start=float(s2), s2_duration=desc.s2_duration, s2_strength=desc.s2_strength,
                                s2_first=desc.s2_first, s2_count=desc.s2_count, s2_base=desc.s2_base))
    st = Strands(specs, precision, dt)
    st.run()
    r = st.results()
    ms = st.last_ms()
    st.close()
    vm = np.array([[x["vmax_first"], x["vmax_last"]] for x in r])
    vul = (vm[:, 0] + boundary) * (vm[:, 1] + boundary) <= 0
    return vm, vul, ms
