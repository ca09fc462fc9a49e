"""ctypes bindings of include/lqt.h (lqt1_amd/liblqt.so).

The library is the product: every compute call goes to its HIP kernels.
Loading fails loudly when liblqt.so has not been built, and every call
that needs a device raises LqtError when no HIP device is present; there is
no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblqt.so")

LQT_OK, LQT_ERR_ARG, LQT_ERR_HIP, LQT_ERR_IO, LQT_ERR_RANGE, LQT_ERR_STATE = range(6)
FP64, FP32 = 0, 1
LVEPI, LVENDO, EPI, ENDO, MCELL, NONE = 0, 1, 2, 3, 4, 255


class LqtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"lqt error {code}: {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [
        ("iks_scale", C.c_double),
        ("iso", C.c_int),
        ("pca_iso", C.c_double),
        ("suprep", C.c_int),
        ("dose_fmol", C.c_double),
        ("sca", C.c_double),
        ("clamp", C.c_int),
        ("v0", C.c_double),
    ]


class Pacing(C.Structure):
    _fields_ = [("bcl", C.c_double), ("stim_start", C.c_double), ("stim_duration", C.c_double),
                ("stim_strength", C.c_double), ("nbeats", C.c_int)]


class Geometry(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("dx", C.c_double),
                ("label", C.c_void_p), ("fibre", C.c_void_p), ("d_fibre", C.c_double),
                ("d_cross", C.c_double)]


class StimProtocol(C.Structure):
    _fields_ = [("bcl", C.c_double), ("duration", C.c_double), ("strength", C.c_double),
                ("s2_time", C.c_double), ("s2_duration", C.c_double), ("s2_strength", C.c_double),
                ("window", C.c_double)]


class StrandDesc(C.Structure):
    _fields_ = [
        ("ncell", C.c_int), ("cell_offset", C.c_int), ("param_index", C.c_int),
        ("dx", C.c_double), ("bcl", C.c_double), ("stop_time", C.c_double),
        ("stim_start", C.c_double), ("stim_duration", C.c_double), ("stim_strength", C.c_double),
        ("nstim", C.c_int), ("rec_every", C.c_int), ("rec_from", C.c_double), ("x0", C.c_double),
        ("measure_beat", C.c_int), ("vrec_every", C.c_int),
        ("s2_base", C.c_double), ("s2_start", C.c_double), ("s2_duration", C.c_double),
        ("s2_strength", C.c_double), ("s2_first", C.c_int), ("s2_count", C.c_int),
    ]


class SweepDesc(C.Structure):
    _fields_ = [("bcl", C.c_double), ("ncell", C.c_int), ("mfrac", C.c_double), ("nbeats", C.c_int),
                ("qt_threshold", C.c_double), ("band_ms", C.c_double), ("p_lo", C.c_double), ("p_hi", C.c_double),
                ("dose_max_fmol", C.c_double), ("chunk", C.c_int), ("autotune_steps", C.c_int),
                ("pca_iso", C.c_double)]


class SweepStats(C.Structure):
    _fields_ = [("points", C.c_int64), ("rechecked", C.c_int64), ("ead_agree", C.c_double),
                ("t_surrogate_s", C.c_double), ("t_recheck_s", C.c_double)]


class VWDesc(C.Structure):
    _fields_ = [("mode", C.c_int), ("left", C.c_double), ("right", C.c_double), ("step", C.c_double),
                ("boundary", C.c_double), ("batch", C.c_int), ("max_candidates", C.c_int64)]


class VWResult(C.Structure):
    _fields_ = [("found", C.c_int), ("s2", C.c_double), ("vmax_first", C.c_double), ("vmax_last", C.c_double),
                ("visited", C.c_int64), ("simulated", C.c_int64), ("launches", C.c_int),
                ("device_ms", C.c_double)]


class StrandResult(C.Structure):
    _fields_ = [("qt", C.c_double), ("apd90", C.c_double * 3), ("ead", C.c_int),
                ("act_first", C.c_double), ("act_last", C.c_double), ("cv", C.c_double),
                ("steps", C.c_int64), ("vmax_first", C.c_double), ("vmax_last", C.c_double)]


_lib = None
_VP = C.c_void_p
_S = C.c_int


def _sig(L, name, args, res=_S):
    f = getattr(L, name)
    f.argtypes = args
    f.restype = res


def lib():
    """Load liblqt.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run `python -m lqt1_amd.build` (hipcc, gfx950)")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    _sig(L, "lqt_last_error", [], C.c_char_p)
    _sig(L, "lqt_params_preset", [C.c_char_p, P(Params)])
    _sig(L, "lqt_suprep_fractions", [C.c_double, P(C.c_double), P(C.c_double)])
    _sig(L, "lqt_device_count", [], C.c_int)
    _sig(L, "lqt_device_synchronize", [])
    _sig(L, "lqt_device_name", [C.c_char_p, C.c_int])
    _sig(L, "lqt_bench_copy", [C.c_int64, C.c_int, P(C.c_double)])
    # cells
    _sig(L, "lqt_cells_create", [C.c_int, _VP, P(Params), C.c_int, C.c_int, C.c_double, P(_VP)])
    _sig(L, "lqt_cells_destroy", [_VP])
    _sig(L, "lqt_cells_step", [_VP, C.c_int, _VP, _VP])
    _sig(L, "lqt_cells_step_async", [_VP, C.c_int, _VP, _VP, _VP])
    _sig(L, "lqt_cells_set_lanes", [_VP, C.c_int])
    _sig(L, "lqt_cells_get_lanes", [_VP, P(C.c_int)])
    _sig(L, "lqt_cells_get_state", [_VP, _VP])
    _sig(L, "lqt_cells_set_state", [_VP, _VP])
    _sig(L, "lqt_cells_get_v", [_VP, _VP])
    _sig(L, "lqt_cells_write_state_file", [_VP, C.c_int, C.c_char_p])
    _sig(L, "lqt_cells_read_state_file", [_VP, C.c_int, C.c_char_p])
    _sig(L, "lqt_cells_pace", [_VP, P(Pacing), _VP, C.c_int, C.c_int, _VP, C.c_int64, P(C.c_int64)])
    # tissue
    _sig(L, "lqt_tissue_create", [P(Geometry), P(Params), C.c_int, C.c_double, P(_VP)])
    _sig(L, "lqt_tissue_create_from_file", [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                            C.c_double, P(Params), C.c_int, C.c_double, P(_VP)])
    _sig(L, "lqt_tissue_destroy", [_VP])
    _sig(L, "lqt_tissue_info", [_VP, P(C.c_int64), P(C.c_int64)])
    _sig(L, "lqt_geometry_ghost_map", [C.c_int, C.c_int, C.c_int, _VP, _VP, P(C.c_int64), P(C.c_int64)])
    _sig(L, "lqt_tissue_set_stimulus", [_VP, P(StimProtocol), C.c_int64, _VP, _VP, C.c_int64, _VP])
    _sig(L, "lqt_tissue_set_stimulus_file", [_VP, P(StimProtocol), C.c_char_p])
    _sig(L, "lqt_tissue_step", [_VP, C.c_int64])
    _sig(L, "lqt_tissue_step_async", [_VP, C.c_int64, _VP])
    _sig(L, "lqt_tissue_time", [_VP, P(C.c_double), P(C.c_int64)])
    _sig(L, "lqt_tissue_read_v", [_VP, _VP])
    _sig(L, "lqt_tissue_read_state", [_VP, C.c_int64, _VP])
    _sig(L, "lqt_tissue_write_state", [_VP, C.c_int64, _VP])
    _sig(L, "lqt_tissue_v_guard", [_VP, _VP])
    _sig(L, "lqt_tissue_track", [_VP, C.c_double, C.c_double])
    _sig(L, "lqt_tissue_track_results", [_VP, _VP, _VP])
    _sig(L, "lqt_tissue_write_vtk", [_VP, C.c_char_p, C.c_int])
    _sig(L, "lqt_tissue_save", [_VP, C.c_char_p])
    _sig(L, "lqt_tissue_load", [_VP, C.c_char_p])
    _sig(L, "lqt_tissue_last_kernel_us", [_VP, P(C.c_double), P(C.c_double)])
    _sig(L, "lqt_tissue_bytes_per_step", [_VP, P(C.c_double)])
    # strands
    _sig(L, "lqt_strands_create", [C.c_int, _VP, _VP, _VP, P(Params), C.c_int, C.c_int, C.c_double, P(_VP)])
    _sig(L, "lqt_strands_destroy", [_VP])
    _sig(L, "lqt_strands_run", [_VP])
    _sig(L, "lqt_strands_run_async", [_VP, _VP])
    _sig(L, "lqt_strands_results", [_VP, _VP])
    _sig(L, "lqt_strands_cell_results", [_VP, C.c_int, _VP, _VP, _VP])
    _sig(L, "lqt_strands_ecg", [_VP, C.c_int, _VP, P(C.c_double), P(C.c_int64)])
    _sig(L, "lqt_strands_final_state", [_VP, C.c_int, _VP])
    _sig(L, "lqt_strands_vtrace", [_VP, C.c_int, _VP, P(C.c_double), P(C.c_int64)])
    _sig(L, "lqt_strands_last_ms", [_VP, P(C.c_double)])
    _sig(L, "lqt_strands_block", [_VP, P(C.c_int)])
    _sig(L, "lqt_strands_autotune", [_VP, C.c_int, P(C.c_int), _VP])
    _sig(L, "lqt_strands_set_mode", [_VP, C.c_int])
    _sig(L, "lqt_tissue_variants", [P(C.c_int)])
    _sig(L, "lqt_tissue_variant_info", [C.c_int, P(C.c_int), P(C.c_int), P(C.c_int), P(C.c_int)])
    _sig(L, "lqt_tissue_set_variant", [_VP, C.c_int])
    _sig(L, "lqt_tissue_get_variant", [_VP, P(C.c_int)])
    _sig(L, "lqt_tissue_autotune", [_VP, C.c_int, P(C.c_int), _VP])
    _sig(L, "lqt_vw_search", [P(StrandDesc), _VP, _VP, P(Params), C.c_int, C.c_double, P(VWDesc), P(VWResult),
                              _VP, _VP, C.c_int64])
    _sig(L, "lqt_geometry_read_file", [C.c_char_p, C.c_int, C.c_int, C.c_int, _VP, _VP, P(C.c_int64)])
    _sig(L, "lqt_stimulus_read_file", [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int64, _VP, _VP, P(C.c_int64)])
    _sig(L, "lqt_mlp_create", [C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64, P(_VP)])
    _sig(L, "lqt_mlp_destroy", [_VP])
    _sig(L, "lqt_mlp_get_params", [_VP] + [_VP] * 6)
    _sig(L, "lqt_mlp_set_params", [_VP] + [_VP] * 6)
    _sig(L, "lqt_mlp_forward", [_VP, C.c_int64, _VP, _VP])
    _sig(L, "lqt_mlp_forward_noalloc", [_VP, C.c_int64, _VP, _VP, _VP])
    _sig(L, "lqt_mlp_grads", [_VP, C.c_int, _VP, _VP, _VP, _VP] + [_VP] * 6)
    _sig(L, "lqt_mlp_train", [_VP, C.c_int64, _VP, _VP, _VP, C.c_int, C.c_int, C.c_double, C.c_uint64, _VP])
    _sig(L, "lqt_surrogate_sweep", [_VP, C.c_int, _VP, C.c_int, _VP, C.c_int, _VP, P(SweepDesc), C.c_int, _VP, _VP,
                                    _VP, _VP, _VP, P(SweepStats)])
    _sig(L, "lqt_ecg_qt", [_VP, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                           P(C.c_double)])
    _lib = L
    return L


def check(status):
    if status != LQT_OK:
        raise LqtError(status, lib().lqt_last_error().decode(errors="replace"))


def ptr(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def preset(name="cpu_default", **overrides) -> Params:
    """Parameter set of a reference build (lqt_params_preset), with overrides."""
    p = Params()
    check(lib().lqt_params_preset(name.encode(), C.byref(p)))
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


def params_array(plist):
    arr = (Params * len(plist))()
    for i, p in enumerate(plist):
        arr[i] = p
    return arr


def device_count() -> int:
    return lib().lqt_device_count()


def device_name() -> str:
    buf = C.create_string_buffer(128)
    check(lib().lqt_device_name(buf, 128))
    return buf.value.decode()


def bench_copy(nbytes, reps=20):
    """Mean microseconds of one plain device copy of nbytes (lqt_bench_copy)."""
    us = C.c_double()
    check(lib().lqt_bench_copy(int(nbytes), int(reps), C.byref(us)))
    return us.value


def suprep_fractions(dose):
    a, b = C.c_double(), C.c_double()
    check(lib().lqt_suprep_fractions(dose, C.byref(a), C.byref(b)))
    return a.value, b.value


def ecg_qt(phi, t0, dts, t_q, t_qrs_end, t_end):
    """lqt_ecg_qt: QT of a pseudo-ECG trace (host only)."""
    phi = np.ascontiguousarray(phi, np.float64)
    q = C.c_double()
    check(lib().lqt_ecg_qt(ptr(phi), phi.size, t0, dts, t_q, t_qrs_end, t_end, C.byref(q)))
    return q.value


def ghost_map(label):
    """lqt_geometry_ghost_map on a (nx, ny, nz) uint8 label grid (host only)."""
    label = np.ascontiguousarray(label, np.uint8)
    nx, ny, nz = label.shape
    g = np.zeros(label.shape, np.int32)
    nv, ng = C.c_int64(), C.c_int64()
    check(lib().lqt_geometry_ghost_map(nx, ny, nz, ptr(label), ptr(g), C.byref(nv), C.byref(ng)))
    return g, nv.value, ng.value
