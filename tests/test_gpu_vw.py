"""Vulnerable-window search (lqt_vw_search, OneD_VW.cc:199-449) on the GPU.

* The reference program's own run (MODE2 from s2 = 260 ms stops at its
  first candidate): same verdict, peaks to its 6 printed digits, its output
  file to 1e-6 mV, through the search and through the C++ driver.
* A scan of S2 times as one launch: FP64 peaks against the oracle (forked
  from a shared S1 prefix) on one candidate of each kind (no capture,
  unidirectional block, bidirectional), FP32 against FP64.
* The batched search replays the reference's sequential rule exactly:
  modes 2 and 3 against the first hit of a scan of the same S2 sequence,
  mode 1 (bisection, tree per launch) against a one-candidate-per-launch
  loop of OneD_VW.cc:396-404.
"""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden", "ref_oned_vw.npz")


@pytest.fixture(scope="module")
def vw():
    from lqt1_amd import vw as V
    return V


def test_reference_search_fp64(lqt, vw):
    z = np.load(GOLD)
    r = vw.search(mode=2, left=260.0, right=260.0, step=0.1, batch=64, precision=lqt.FP64)
    print(f"MODE2 from 260: found {r.found} at {r.s2}, peaks {r.vmax_first:.6g} / {r.vmax_last:.6g}, "
          f"{r.simulated} strands, {r.device_ms:.1f} ms")
    assert r.found == bool(z["found"])
    assert len(r.visited_s2) == len(z["max_endo"]) == 1
    assert r.s2 == 260.0
    assert abs(r.vmax_first - z["max_endo"][0]) <= 5e-6 * abs(z["max_endo"][0]) + 1e-5
    assert abs(r.vmax_last - z["max_epi"][0]) <= 5e-6 * abs(z["max_epi"][0]) + 1e-5


def test_driver_writes_reference_file(lqt, tmp_path):
    z = np.load(GOLD)
    exe = os.path.join(ROOT, "lqt1_amd", "bin", "lqt_oned_vw")
    r = subprocess.run([exe, "--outdir", str(tmp_path), "--batch", "32"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    assert "VM found!" in r.stdout and "All done." in r.stdout
    endo = float(r.stdout.split("maxEndo = ")[1].split()[0])
    assert abs(endo - z["max_endo"][0]) <= 5e-6 * abs(z["max_endo"][0]) + 1e-5
    data = np.loadtxt(tmp_path / "Transmural1D_s2@260.00.dat")
    assert data.shape == tuple(z["shape0"])
    err = np.max(np.abs(data[:, z["cells"]] - z["v0"]))
    print("driver file vs reference file: max |dV|", err)
    assert err < 1e-6


S2_SCAN = np.arange(200.0, 420.0 + 1e-9, 4.0)


@pytest.fixture(scope="module")
def scans(lqt, vw):
    out = {}
    for prec in (lqt.FP64, lqt.FP32):
        vm, vul, ms = vw.scan(S2_SCAN, precision=prec)
        out[prec] = (vm, vul)
        print(f"scan of {S2_SCAN.size} S2 times, {'FP64' if prec == lqt.FP64 else 'FP32'}: {ms:.1f} ms; "
              f"vulnerable at {S2_SCAN[vul]}")
    return out


def test_scan_fp64_matches_oracle(lqt, oracle, scans):
    O = oracle
    vm, vul = scans[lqt.FP64]
    thr = -10.0
    kinds = {"none": np.flatnonzero((vm[:, 0] < thr) & (vm[:, 1] < thr)),
             "block": np.flatnonzero(vul),
             "both": np.flatnonzero((vm[:, 0] > thr) & (vm[:, 1] > thr))}
    picks = [int(v[len(v) // 2]) for v in kinds.values() if v.size]
    assert len(picks) >= 2
    ct = np.array([O.ENDO] * 25 + [O.MCELL] * 35 + [O.EPI] * 40, np.uint8)
    co = np.full(100, 0.154)
    co[63] *= 0.2
    first = min(S2_SCAN[i] for i in picks)
    base = O.StrandVW(ct, co, nbeats=5, s2_start=first, record=False).run(pause_at=4000.0 + first)
    for i in picks:
        r = base.fork(S2_SCAN[i]).run()
        err = np.max(np.abs(r.vmax - vm[i]))
        print(f"s2 {S2_SCAN[i]}: oracle {r.vmax}, GPU {vm[i]}, |d| {err:.3g}")
        assert err < 1e-6


def test_scan_fp32_agrees_with_fp64(lqt, scans):
    vm64, vul64 = scans[lqt.FP64]
    vm32, vul32 = scans[lqt.FP32]
    # the verdict may only move at the window's edges
    diff = np.flatnonzero(vul64 != vul32)
    edges = np.flatnonzero(np.diff(vul64.astype(int)) != 0)
    assert all(np.min(np.abs(edges - d)) <= 1 for d in diff) if diff.size else True
    assert np.max(np.abs(vm32 - vm64)[vul64 == vul32]) < 5.0


def _sequence(left, step, count, direction):
    s, out = left, []
    for _ in range(count):
        out.append(s)
        s = s + step if direction > 0 else s - step
    return np.array(out)


@pytest.mark.parametrize("mode", [2, 3])
def test_stepping_modes_equal_scan(lqt, vw, mode):
    left = 200.0 if mode == 2 else 420.0
    seq = _sequence(left, 2.0, 120, 1 if mode == 2 else -1)
    seq = seq[(seq >= 0) & (seq <= 800)]
    vm, vul, _ = vw.scan(seq, precision=lqt.FP32)
    r = vw.search(mode=mode, left=left, right=left, step=2.0, batch=16, max_candidates=seq.size,
                  precision=lqt.FP32)
    if vul.any():
        k = int(np.argmax(vul))
        assert r.found and r.s2 == seq[k] and len(r.visited_s2) == k + 1
        assert np.array_equal(r.visited_vmax, vm[:k + 1])
    else:
        assert not r.found


def test_bisection_equals_sequential_rule(lqt, vw):
    left, right, res = 200.0, 420.0, 0.5
    r = vw.search(mode=1, left=left, right=right, step=res, batch=7, max_candidates=64, precision=lqt.FP32)
    # OneD_VW.cc:396-404 one candidate per launch
    seq = []
    s2 = 0.5 * (left + right)
    found = False
    while (right - left >= res or right == left) and len(seq) < 64:
        vm, vul, _ = vw.scan([s2], precision=lqt.FP32)
        e, p = vm[0]
        seq.append((s2, e, p))
        if e + 10 < 0 and p + 10 < 0:
            left = s2
            s2 = 0.5 * (left + right)
        elif e + 10 > 0 and p + 10 > 0:
            right = s2
            s2 = 0.5 * (left + right)
        else:
            found = True
            break
    print(f"bisection: {len(seq)} candidates, found {found} at {seq[-1][0]}; batched: {r.launches} launches, "
          f"{r.simulated} strands")
    assert r.found == found
    assert np.array_equal(r.visited_s2, np.array([x[0] for x in seq]))
    assert np.array_equal(r.visited_vmax, np.array([[x[1], x[2]] for x in seq]))
