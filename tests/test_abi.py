"""The C ABI library loads, exports every function include/lqt.h declares, and
fails loudly (an error code, no silent CPU path) where no GPU is present."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import lqt1_amd as A
from lqt1_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "lqt.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(lqt_[a-z0-9_]+)\s*\(", txt)))


def test_every_declared_symbol_is_exported():
    names = _declared()
    assert len(names) > 40
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (lqt_\w+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    lib = A.lib()
    for n in names:
        getattr(lib, n)


def test_library_is_built_for_gfx950():
    out = subprocess.run(["strings", L.LIB_PATH], capture_output=True, text=True).stdout
    assert "gfx950" in out


def test_presets_follow_reference_builds():
    p = A.preset("cpu_default")
    assert (p.iks_scale, p.iso, p.pca_iso, p.suprep, p.v0) == (1.0, 1, 1.6, 0, -87.5)
    g = A.preset("gpu3d")
    assert (g.iks_scale, g.iso, g.pca_iso, g.suprep, g.dose_fmol, g.v0) == (0.05, 1, 4.0, 1, 100.0, -86.2)
    t = A.preset("gpu2d")
    assert (t.iks_scale, t.iso, t.v0) == (0.05, 0, -86.2)
    assert A.preset("config1").iso == 0
    assert A.preset("config3").iks_scale == 0.2
    with pytest.raises(A.LqtError):
        A.preset("nope")


def test_errors_are_reported_not_hidden():
    lib = A.lib()
    h = C.c_void_p()
    p = A.preset()
    ct = np.array([3], np.uint8)
    st = lib.lqt_cells_create(0, L.ptr(ct), C.byref(p), 1, A.FP64, 0.02, C.byref(h))
    assert st == L.LQT_ERR_ARG
    assert b"bad argument" in lib.lqt_last_error()
    if A.device_count() == 0:
        st = lib.lqt_cells_create(1, L.ptr(ct), C.byref(p), 1, A.FP64, 0.02, C.byref(h))
        assert st == L.LQT_ERR_HIP
        assert b"no HIP device" in lib.lqt_last_error()
        with pytest.raises(A.LqtError):
            A.Tissue(np.full((3, 3, 3), 3, np.uint8))
        with pytest.raises(A.LqtError):
            A.Strands([A.StrandSpec(A.transmural(10), A.preset())])


def test_product_does_not_import_the_oracle():
    src = os.path.join(ROOT, "lqt1_amd")
    for dirpath, _, files in os.walk(src):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".hpp", ".h")):
                txt = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in txt and "from oracle" not in txt, f
                assert "oracle.h" not in txt and "liboracle" not in txt, f


def test_fp32_step_limit_is_enforced():
    """FP32 handles refuse dt above LQT_FP32_MAX_DT (the slow gates' series
    Rush-Larsen step is accurate to dt/tau = 0.125); FP64 takes it."""
    lib = A.lib()
    h = C.c_void_p()
    p = A.preset()
    ct = np.array([3], np.uint8)
    st = lib.lqt_cells_create(1, L.ptr(ct), C.byref(p), 1, A.FP32, 0.2, C.byref(h))
    assert st == L.LQT_ERR_ARG and b"FP32 needs dt" in lib.lqt_last_error()
    with pytest.raises(A.LqtError, match="FP32 needs dt"):
        A.Tissue(np.full((3, 3, 3), 3, np.uint8), None, 0.2, 0.154, 0.154, A.preset(), A.FP32, 0.2)
    with pytest.raises(A.LqtError, match="FP32 needs dt"):
        A.Strands([A.StrandSpec(A.transmural(10), A.preset())], A.FP32, 0.2)
    st = lib.lqt_cells_create(1, L.ptr(ct), C.byref(p), 1, A.FP64, 0.2, C.byref(h))
    assert st != L.LQT_ERR_ARG  # no device here: LQT_ERR_HIP; on a GPU: created
    if st == L.LQT_OK:
        lib.lqt_cells_destroy(h)


def test_sizes_beyond_32_bit_offsets_are_refused():
    """The kernels address the V grid and the state slots by 32-bit byte
    offsets; handles that would overflow them are refused, not run."""
    lib = A.lib()
    h = C.c_void_p()
    p = A.preset()
    ct = np.array([3], np.uint8)
    st = lib.lqt_cells_create(1 << 29, L.ptr(ct), C.byref(p), 1, A.FP64, 0.02, C.byref(h))
    assert st == L.LQT_ERR_ARG and b"2^29" in lib.lqt_last_error()
    big = np.zeros((1024, 1024, 520), np.uint8)  # 549M padded cells: over 2^29 in FP64
    with pytest.raises(A.LqtError, match="limit of 2"):
        A.Tissue(big, None, 0.2, 0.154, 0.154, A.preset(), A.FP64, 0.02)
