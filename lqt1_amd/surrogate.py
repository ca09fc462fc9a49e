"""Dose-sweep surrogate: an MLP trained on strands the solver generates, used
to steer a dense IKs-scale x gene-therapy-dose sweep, with every point near
a decision threshold re-run by the batched strand solver.

The reference has no learned component: its dose-response work is 16
separate 1D runs post-processed in Python (PlotScript/plot1DECG_multiple.py:
55-99).  Inputs (6): IKs scale, ISO, dose fraction, BCL, strand length,
M-cell fraction.  Outputs (5): QT, APD90 of ENDO / M / EPI (squared error,
masked where the solver gives no value) and the EAD logit (cross-entropy).
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .strands import StrandSpec, Strands, transmural

N_IN, N_REG, N_BCE = 6, 4, 1
DOSE_MAX_FMOL = 300.0  # the fitted dose range of fit_dose_dependent.py:122
# PCa multiplier under ISO: the beta-adrenergic setting of the 3D program
# (GPU3D.cu:767) and of the config-3 solver the training strands come from;
# with it, low IKs under ISO gives early afterdepolarisations (TPORd.cc:629
# uses 1.6, under which 10-beat strands show none)
PCA_ISO = 4.0


class MLP:
    """lqt_mlp_*: BF16-MFMA MLP with FP32 accumulation and Adam."""

    def __init__(self, n_in=N_IN, hidden=128, n_reg=N_REG, n_bce=N_BCE, seed=0):
        h = C.c_void_p()
        L.check(L.lib().lqt_mlp_create(n_in, hidden, n_reg, n_bce, seed, C.byref(h)))
        self.h, self.n_in, self.hidden, self.n_out = h, n_in, hidden, n_reg + n_bce

    def params(self):
        H, ni, no = self.hidden, self.n_in, self.n_out
        p = [np.zeros((H, ni), np.float32), np.zeros(H, np.float32), np.zeros((H, H), np.float32),
             np.zeros(H, np.float32), np.zeros((no, H), np.float32), np.zeros(no, np.float32)]
        L.check(L.lib().lqt_mlp_get_params(self.h, *[L.ptr(a) for a in p]))
        return p

    def set_params(self, p):
        p = [np.ascontiguousarray(a, np.float32) for a in p]
        L.check(L.lib().lqt_mlp_set_params(self.h, *[L.ptr(a) for a in p]))

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((x.shape[0], self.n_out), np.float32)
        L.check(L.lib().lqt_mlp_forward(self.h, x.shape[0], L.ptr(x), L.ptr(y)))
        return y

    def grads(self, x, y, mask=None):
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        mk = None if mask is None else np.ascontiguousarray(mask, np.float32)
        H, ni, no = self.hidden, self.n_in, self.n_out
        g = [np.zeros((H, ni), np.float32), np.zeros(H, np.float32), np.zeros((H, H), np.float32),
             np.zeros(H, np.float32), np.zeros((no, H), np.float32), np.zeros(no, np.float32)]
        loss = C.c_float()
        L.check(L.lib().lqt_mlp_grads(self.h, x.shape[0], L.ptr(x), L.ptr(y), L.ptr(mk), C.byref(loss),
                                      *[L.ptr(a) for a in g]))
        return loss.value, g

    def train(self, x, y, mask=None, epochs=100, batch=512, lr=1e-3, seed=0):
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        mk = None if mask is None else np.ascontiguousarray(mask, np.float32)
        nb = x.shape[0] // batch
        losses = np.zeros(epochs * nb, np.float32)
        L.check(L.lib().lqt_mlp_train(self.h, x.shape[0], L.ptr(x), L.ptr(y), L.ptr(mk), epochs, batch, lr, seed,
                                      L.ptr(losses)))
        return losses

    def close(self):
        if getattr(self, "h", None):
            L.lib().lqt_mlp_destroy(self.h)
            self.h = None

    __del__ = close


# ---- features and targets ---------------------------------------------------

def features(iks, iso, dose, bcl, length, mfrac):
    """Normalised network inputs (each roughly in [-1, 1])."""
    iks, iso, dose, bcl, length, mfrac = np.broadcast_arrays(iks, iso, dose, bcl, length, mfrac)
    return np.stack([2 * iks - 1, 2 * iso - 1, 2 * dose - 1, (bcl - 1000.0) / 500.0, (length - 225.0) / 175.0,
                     (mfrac - 0.35) / 0.15], 1).astype(np.float32)


T_SCALE, T_SHIFT = 100.0, 350.0


QT_CLIP = (150.0, 750.0)    # ms: the regression's QT target range
APD_CLIP = (100.0, 900.0)   # ms


def targets(results, points=None):
    """(QT, APD90 x3) normalised + EAD flag; mask 0 where a value is missing.

    The QT of a strand with an EAD is masked (its pseudo-ECG's T end is not a
    repolarisation end, and the sweep's decision for it follows the EAD
    call).  With the points given, a strand that activates without an EAD
    but has no T end within its beat (QT not measurable, the longest QTs of
    all) gets the beat length as a censored target.  Targets are clipped to
    QT_CLIP / APD_CLIP, so that a few extreme strands do not dominate the
    squared error."""
    y = np.zeros((len(results), N_REG + N_BCE), np.float32)
    m = np.ones_like(y)
    for i, r in enumerate(results):
        qt = r["qt"]
        if r["ead"]:
            qt = None
        elif qt < 0 and points is not None and r["act_first"] >= 0:
            qt = points[i].bcl
        vals = [None if qt is None or qt < 0 else float(np.clip(qt, *QT_CLIP))]
        vals += [None if v is None or v < 0 else float(np.clip(v, *APD_CLIP)) for v in r["apd90"]]
        for j, v in enumerate(vals):
            if v is None:
                m[i, j] = 0.0
            else:
                y[i, j] = (v - T_SHIFT) / T_SCALE
        y[i, N_REG] = 1.0 if r["ead"] else 0.0
    return y, m


def qt_of(results):
    """Strand QT per result in ms, NaN where it is not measurable (no T end
    within the beat)."""
    q = np.array([r["qt"] for r in results], np.float64)
    return np.where(q < 0, np.nan, q)


def long_qt(qt, threshold, ead=None):
    """The sweep's decision: QT above the threshold, not measurable (NaN),
    or an EAD."""
    out = ~(np.asarray(qt) <= threshold)
    return out if ead is None else out | np.asarray(ead, bool)


def decode(out):
    """Network outputs -> (QT ms, APD90 ms x3, EAD probability)."""
    qt = out[:, 0] * T_SCALE + T_SHIFT
    apd = out[:, 1:N_REG] * T_SCALE + T_SHIFT
    p = 1.0 / (1.0 + np.exp(-out[:, N_REG].astype(np.float64)))
    return qt, apd, p


# ---- solver runs --------------------------------------------------------------

@dataclass
class Point:
    iks: float
    iso: int
    dose: float      # fraction of DOSE_MAX_FMOL
    bcl: float
    length: int
    mfrac: float


def spec_of(p: Point, nbeats: int, pca_iso: float = PCA_ISO) -> StrandSpec:
    par = L.preset("cpu_default", iks_scale=float(p.iks), iso=int(p.iso), suprep=1,
                   dose_fmol=float(p.dose) * DOSE_MAX_FMOL, pca_iso=float(pca_iso))
    endo = 0.5 * (1.0 - p.mfrac)
    return StrandSpec(transmural(int(p.length), endo, float(p.mfrac)), par, bcl=float(p.bcl), nbeats=nbeats)


def simulate(points, nbeats=10, precision=L.FP32, chunk=32768, progress=None, pca_iso=PCA_ISO):
    """Run every point as a strand through the batched solver (launch layout
    chosen by the autotuner per chunk of up to `chunk` strands, one
    persistent launch each); returns the per-strand results and the device
    time.  progress(done, total) is called after each chunk.  (One launch
    for the 20k training strands took 85.7 s on MI355X against 91.2 s in
    chunks of 4096: each chunk ends with its longest strand running alone.)"""
    out, t = [], 0.0
    for i in range(0, len(points), chunk):
        st = Strands([spec_of(p, nbeats, pca_iso) for p in points[i:i + chunk]], precision, 0.02)
        t0 = time.perf_counter()
        st.autotune(2000)
        st.run()
        t += time.perf_counter() - t0
        out.extend(st.results())
        st.close()
        if progress:
            progress(len(out), len(points))
    return out, t


def random_points(n, seed=0):
    """The training distribution: IKs U[0,1], ISO {0,1}, dose U[0,1], BCL
    {600,800,1000,1500}, length U{50..400}, M fraction U[0.2,0.5]."""
    rng = np.random.default_rng(seed)
    return [Point(float(rng.random()), int(rng.integers(0, 2)), float(rng.random()),
                  float(rng.choice([600.0, 800.0, 1000.0, 1500.0])), int(rng.integers(50, 401)),
                  float(rng.uniform(0.2, 0.5))) for _ in range(n)]


def points_features(points):
    a = np.array([[p.iks, p.iso, p.dose, p.bcl, p.length, p.mfrac] for p in points], np.float64)
    return features(*a.T)


# ---- the steered sweep ---------------------------------------------------------

@dataclass
class SweepResult:
    qt: np.ndarray            # final QT per grid point (solver value where re-checked)
    ead: np.ndarray           # final EAD decision
    long_qt: np.ndarray       # QT > threshold
    rechecked: np.ndarray     # bool mask of points re-run by the solver
    qt_err: np.ndarray        # surrogate QT error on the re-checked points (ms)
    ead_agree: float          # fraction of re-checked points where the surrogate's EAD call matched
    t_surrogate: float
    t_recheck: float


def sweep(mlp: MLP, iks_grid, dose_grid, iso_values=(0, 1), bcl=1000.0, length=100, mfrac=0.35,
          qt_threshold=450.0, band_ms=10.0, p_band=(0.2, 0.8), nbeats=10, precision=L.FP32, pca_iso=PCA_ISO):
    """Surrogate over the grid, then the solver on every point within band_ms
    of the QT threshold or with an EAD probability inside p_band
    (lqt_surrogate_sweep).  Grid order: iks outer, dose, iso inner."""
    iks = np.ascontiguousarray(iks_grid, np.float64)
    dose = np.ascontiguousarray(dose_grid, np.float64)
    iso = np.ascontiguousarray(iso_values, np.int32)
    n = iks.size * dose.size * iso.size
    qt, err = np.zeros(n), np.zeros(n)
    ead, long_qt, near = np.zeros(n, np.uint8), np.zeros(n, np.uint8), np.zeros(n, np.uint8)
    d = L.SweepDesc(float(bcl), int(length), float(mfrac), int(nbeats), float(qt_threshold), float(band_ms),
                    float(p_band[0]), float(p_band[1]), DOSE_MAX_FMOL, 32768, 2000, float(pca_iso))
    stats = L.SweepStats()
    L.check(L.lib().lqt_surrogate_sweep(mlp.h, iks.size, L.ptr(iks), dose.size, L.ptr(dose), iso.size, L.ptr(iso),
                                        C.byref(d), precision, L.ptr(qt), L.ptr(ead), L.ptr(long_qt), L.ptr(near),
                                        L.ptr(err), C.byref(stats)))
    near = near.astype(bool)
    return SweepResult(qt, ead.astype(bool), long_qt.astype(bool), near, err[near], stats.ead_agree,
                       stats.t_surrogate_s, stats.t_recheck_s)
