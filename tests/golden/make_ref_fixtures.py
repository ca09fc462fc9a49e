"""Write the golden fixtures of tests/golden/ from the reference's own CPU build.

Runs only where the reference tree exists (it needs oracle/_ref, which
``make -C oracle`` compiles from /root/reference).  The outputs are data:
states, statistics and traces that the reference's own programs printed or
wrote, kept small so they can be committed.

    python tests/golden/make_ref_fixtures.py

Fixtures
  ref_cell_states.npz   TPORd state (40 values) after 1, 10, 100, 1000 and
                        5000 steps from init for ENDO/MCELL/EPI under each
                        flag variant, stimulus -53 pA/pF x 1 ms at 50 ms.
  ref_single_cell.json  Outputs/VentriSingleCellStats.dat of SingleCell.cc
                        (MCELL, CON+ISO1, 100 beats) and the last beat's V
                        every 1 ms from its results file.
  ref_oned.npz          OneD.cc as shipped (CON+ISO1, 6 stimuli, 6000 ms):
                        V of 9 cells every 1 ms in the last beat, the time
                        column, per-cell -30 mV activation times of the last
                        beat from the 0.2 ms output, and the printed CV.
  ref_oned_vw.npz       OneD_VW.cc as shipped (MODE2 from s2 = 260 ms): the
                        printed maxEndo/maxEpi of each candidate, whether it
                        printed "VM found!", and 16 columns of each
                        candidate's Transmural1D_s2@<s2>.dat (V every 1 ms of
                        the last beat).
"""
from __future__ import annotations

import json
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
import oracle as O  # noqa: E402

VARIANTS = ("con_iso", "lqt1_iso", "suprep_iso", "lqt1_suprep_iso", "clamp")
CELLTYPES = (O.ENDO, O.MCELL, O.EPI)
CHECKPOINTS = (1, 10, 100, 1000, 5000)
ONED_CELLS = (0, 12, 24, 25, 42, 59, 60, 80, 99)
VW_CELLS = (0, 5, 10, 11, 12, 20, 24, 25, 40, 59, 60, 62, 63, 64, 80, 99)


def cell_states():
    out = {}
    for var in VARIANTS:
        for ct in CELLTYPES:
            c = O.RefCell(ct, dt=0.02, variant=var)
            t = 0.0
            rows = []
            for step in range(1, CHECKPOINTS[-1] + 1):
                tc = t - np.floor(t / 1000.0) * 1000.0
                ist = -53.0 if (tc >= 50.0 and tc < 51.0) else 0.0
                c.step(ist, 0.0)
                t += 0.02
                if step in CHECKPOINTS:
                    rows.append(c.state())
            out[f"{var}_{ct}"] = np.array(rows)
    out["checkpoints"] = np.array(CHECKPOINTS)
    np.savez_compressed(os.path.join(HERE, "ref_cell_states.npz"), **out)


def single_cell():
    exe = os.path.join(O.REF_DIR, "model_single_cell")
    with tempfile.TemporaryDirectory() as d:
        os.mkdir(os.path.join(d, "Outputs"))
        r = subprocess.run([exe], cwd=d, capture_output=True, text=True, check=True)
        stats = {}
        with open(os.path.join(d, "Outputs", "VentriSingleCellStats.dat")) as f:
            for line in f:
                m = re.match(r"\s*(.+?)\s*=\s*([-0-9.eE+]+)", line)
                if m:
                    stats[m.group(1)] = float(m.group(2))
        data = np.loadtxt(os.path.join(d, "Outputs", "VentriSingleCellResults_LQT1_ISO_MCELL.dat"))
    t = data[:, 0]
    last = t >= 99000.0
    sel = np.where(last)[0][::50]  # every 1 ms (file rows are every 0.02 ms)
    out = dict(stats=stats, t=t[sel].tolist(), v=data[sel, 1].tolist(),
               nrows=int(data.shape[0]), stdout=r.stdout[-2000:])
    with open(os.path.join(HERE, "ref_single_cell.json"), "w") as f:
        json.dump(out, f, indent=1)


def oned():
    exe = os.path.join(O.REF_DIR, "model_oned")
    with tempfile.TemporaryDirectory() as d:
        os.mkdir(os.path.join(d, "Outputs"))
        r = subprocess.run([exe], cwd=d, capture_output=True, text=True, check=True)
        data = np.loadtxt(os.path.join(d, "Outputs", "VentOneDResults_CON_ISO.dat"))
    cv = float(re.search(r"CV = ([0-9.eE+-]+)", r.stdout).group(1))
    start = float(re.search(r"start = ([0-9.eE+-]+)", r.stdout).group(1))
    end = float(re.search(r"end = ([0-9.eE+-]+)", r.stdout).group(1))
    t = data[:, 0]
    v = data[:, 1:]
    beat = t >= 5000.0
    act = np.full(v.shape[1], -1.0)
    for i in range(v.shape[1]):
        idx = np.where(beat & (v[:, i] >= -30.0))[0]
        if idx.size:
            act[i] = t[idx[0]]
    sel = np.where(beat)[0][::5]  # rows are every 0.2 ms -> every 1 ms
    np.savez_compressed(os.path.join(HERE, "ref_oned.npz"), t=t[sel], v=v[np.ix_(sel, ONED_CELLS)],
                        cells=np.array(ONED_CELLS), act=act, cv=cv, cv_start=start, cv_end=end,
                        nrows=data.shape[0], t_first=t[:3], t_last=t[-3:],
                        phi_rows_t=t[beat], phi=O.pseudo_ecg(v[beat]))


def oned_vw():
    """OneD_VW.cc as shipped (CON+ISO1, MODE2 from s2 = 260 ms): the printed
    peaks of every candidate it runs, its verdict, and the output file of
    each candidate at the VW_CELLS columns."""
    exe = os.path.join(O.REF_DIR, "model_oned_vw")
    with tempfile.TemporaryDirectory() as d:
        os.mkdir(os.path.join(d, "Outputs"))
        r = subprocess.run([exe], cwd=d, capture_output=True, text=True, check=True, timeout=3600)
        names = sorted(f for f in os.listdir(os.path.join(d, "Outputs")) if f.startswith("Transmural1D_s2@"))
        files = {n: np.loadtxt(os.path.join(d, "Outputs", n)) for n in names}
    endo = [float(x) for x in re.findall(r"maxEndo = ([-0-9.eE+]+)", r.stdout)]
    epi = [float(x) for x in re.findall(r"maxEpi = ([-0-9.eE+]+)", r.stdout)]
    s2 = [float(re.search(r"@([0-9.]+)\.dat", n).group(1)) for n in names]
    out = dict(max_endo=np.array(endo), max_epi=np.array(epi), found=int("VM found!" in r.stdout),
               s2=np.array(s2), cells=np.array(VW_CELLS))
    for k, n in enumerate(names):
        out[f"v{k}"] = files[n][:, VW_CELLS]
        out[f"shape{k}"] = np.array(files[n].shape)
    np.savez_compressed(os.path.join(HERE, "ref_oned_vw.npz"), **out)


if __name__ == "__main__":
    if not O.ref_available():
        sys.exit("oracle/_ref is not built (needs the reference tree): run make -C oracle")
    cell_states()
    single_cell()
    oned()
    oned_vw()
    print("fixtures written to", HERE)
