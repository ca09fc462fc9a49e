// C ABI: the surrogate-steered dose sweep.  No counterpart in the
// reference, whose dose-response work is separate 1D runs post-processed in
// Python (PlotScript/plot1DECG_multiple.py:55-99).
//
// Over an IKs-scale x dose x ISO grid of one strand shape the trained
// network (lqt_mlp) predicts QT and the EAD logit of every point; every
// point whose predicted QT lies within band_ms of the threshold, or whose
// EAD probability lies inside (p_lo, p_hi), is re-run as a strand of the
// batched solver (lqt_strands), whose QT and EAD then replace the
// prediction.  The strand of a point is the transmural strand of
// OneD.cc:40-448 (S1 of -52 pA/pF for 1 ms at cells 0-2 from 50 ms, dx
// 0.15 mm, D 0.154, pseudo-ECG at 20 mm beyond the last cell, QT/APD/EAD of
// the last beat) with IKs scale, ISO and the SUPREP dose of the point.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <vector>

#include "common.hpp"

namespace {

// network inputs and output scaling (the training side uses the same)
constexpr int kIn = 6, kOut = 5;
constexpr double kTScale = 100.0, kTShift = 350.0;

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// cell types of a transmural strand: ENDO | MCELL | EPI, the first two
// counts rounded half to even
void transmural(int n, double endo, double mid, uint8_t* out) {
    const int ne = (int)nearbyint(endo * n), nm = (int)nearbyint(mid * n);
    for (int i = 0; i < n; ++i) out[i] = i < ne ? LQT_ENDO : (i < ne + nm ? LQT_MCELL : LQT_EPI);
}

}  // namespace

extern "C" {

lqt_status lqt_surrogate_sweep(lqt_mlp* net, int n_iks, const double* iks, int n_dose, const double* dose,
                               int n_iso, const int* iso, const lqt_sweep_desc* d, lqt_precision prec, double* qt,
                               uint8_t* ead, uint8_t* long_qt, uint8_t* rechecked, double* qt_err,
                               lqt_sweep_stats* stats) {
    LQT_CHECK(net && iks && dose && iso && d && qt && ead && long_qt && rechecked && qt_err && n_iks > 0 &&
                  n_dose > 0 && n_iso > 0,
              LQT_ERR_ARG, "lqt_surrogate_sweep: bad argument");
    LQT_CHECK(d->bcl > 0 && d->ncell >= 2 && d->ncell <= 512 && d->mfrac >= 0 && d->mfrac <= 1 && d->nbeats > 0 &&
                  d->band_ms >= 0 && d->dose_max_fmol >= 0 && d->chunk >= 0 && d->autotune_steps >= 0 &&
                  d->pca_iso >= 0,
              LQT_ERR_ARG, "lqt_surrogate_sweep: bad sweep description");
    const int64_t np = (int64_t)n_iks * n_dose * n_iso;

    // 1. the network over the whole grid (iks outer, dose, iso inner)
    const double t0 = now_s();
    std::vector<float> x((size_t)np * kIn), y((size_t)np * kOut);
    int64_t p = 0;
    for (int a = 0; a < n_iks; ++a)
        for (int b = 0; b < n_dose; ++b)
            for (int c = 0; c < n_iso; ++c, ++p) {
                float* f = &x[(size_t)p * kIn];
                f[0] = (float)(2.0 * iks[a] - 1.0);
                f[1] = (float)(2.0 * iso[c] - 1.0);
                f[2] = (float)(2.0 * dose[b] - 1.0);
                f[3] = (float)((d->bcl - 1000.0) / 500.0);
                f[4] = (float)((d->ncell - 225.0) / 175.0);
                f[5] = (float)((d->mfrac - 0.35) / 0.15);
            }
    lqt_status st = lqt_mlp_forward(net, np, x.data(), y.data());
    if (st) return st;
    std::vector<int64_t> near;
    std::vector<double> pe(np);
    for (int64_t i = 0; i < np; ++i) {
        qt[i] = (double)y[(size_t)i * kOut] * kTScale + kTShift;
        pe[i] = 1.0 / (1.0 + exp(-(double)y[(size_t)i * kOut + 4]));
        ead[i] = pe[i] >= 0.5;
        qt_err[i] = NAN;
        rechecked[i] = fabs(qt[i] - d->qt_threshold) < d->band_ms || (pe[i] > d->p_lo && pe[i] < d->p_hi);
        if (rechecked[i]) near.push_back(i);
    }
    const double t1 = now_s();

    // 2. the solver on the points near a decision threshold
    const int chunk = d->chunk > 0 ? d->chunk : 4096;
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        ncu = 0;
    const int trial = d->autotune_steps > 0 ? d->autotune_steps : 2000;
    const int L = d->ncell;
    int64_t agree = 0;
    lqt_params base;
    if ((st = lqt_params_preset("cpu_default", &base))) return st;
    for (size_t c0 = 0; c0 < near.size(); c0 += (size_t)chunk) {
        const int n = (int)std::min<size_t>((size_t)chunk, near.size() - c0);
        std::vector<lqt_strand_desc> desc(n);
        std::vector<lqt_params> par(n, base);
        std::vector<uint8_t> ct((size_t)n * L);
        for (int k = 0; k < n; ++k) {
            const int64_t i = near[c0 + k];
            const int a = (int)(i / ((int64_t)n_dose * n_iso)), b = (int)(i / n_iso % n_dose), c = (int)(i % n_iso);
            par[k].iks_scale = iks[a];
            par[k].iso = iso[c];
            par[k].suprep = 1;
            par[k].dose_fmol = dose[b] * d->dose_max_fmol;
            if (d->pca_iso > 0) par[k].pca_iso = d->pca_iso;
            transmural(L, 0.5 * (1.0 - d->mfrac), d->mfrac, &ct[(size_t)k * L]);
            lqt_strand_desc& s = desc[k];
            memset(&s, 0, sizeof(s));
            s.ncell = L;
            s.cell_offset = k * L;
            s.param_index = k;
            s.dx = 0.15;
            s.bcl = d->bcl;
            s.stop_time = d->nbeats * d->bcl;
            s.stim_start = 50.0;  // OneD.cc:81
            s.stim_duration = 1.0;
            s.stim_strength = -52.0;
            s.nstim = 3;
            s.rec_every = 10;  // pseudo-ECG every 0.2 ms
            s.measure_beat = d->nbeats - 1;
            s.rec_from = (d->nbeats - 1) * d->bcl;
            s.x0 = L * 0.15 + 20.0;  // plot1DECG_compare.py:86
            s.s2_base = (d->nbeats - 1) * d->bcl;
            s.s2_strength = -104.0;
            s.s2_first = 10;
            s.s2_count = 3;
        }
        lqt_strands* sh = nullptr;
        if ((st = lqt_strands_create(n, desc.data(), ct.data(), nullptr, par.data(), n, prec, 0.02, &sh))) return st;
        std::vector<lqt_strand_result> res(n);
        int mode = 0;
        // A batch that leaves CUs idle (at most one strand of up to 128 cells
        // per CU) runs fastest split over two lanes per cell, one strand per
        // CU at one wavefront per SIMD: no trial runs on the critical path.
        // Larger batches are timed in every layout.
        if (n <= ncu && L <= 128) st = lqt_strands_set_mode(sh, 3);
        else st = lqt_strands_autotune(sh, trial, &mode, nullptr);
        if (!st) st = lqt_strands_run(sh);
        if (!st) st = lqt_strands_results(sh, res.data());
        lqt_strands_destroy(sh);
        if (st) return st;
        for (int k = 0; k < n; ++k) {
            const int64_t i = near[c0 + k];
            agree += (int64_t)(ead[i] == (res[k].ead != 0));
            if (res[k].qt >= 0) {
                qt_err[i] = qt[i] - res[k].qt;
                qt[i] = res[k].qt;
            } else {
                qt[i] = NAN;  // no T end within the beat: QT not measurable, and long
            }
            ead[i] = res[k].ead != 0;
        }
    }
    const double t2 = now_s();
    // long QT: above the threshold, not measurable (NaN, a re-checked point
    // without a T end in its beat), or an EAD (whose T end is not a
    // repolarisation end)
    for (int64_t i = 0; i < np; ++i) long_qt[i] = ead[i] || !(qt[i] <= d->qt_threshold);
    if (stats) {
        stats->points = np;
        stats->rechecked = (int64_t)near.size();
        stats->ead_agree = near.empty() ? 1.0 : (double)agree / (double)near.size();
        stats->t_surrogate_s = t1 - t0;
        stats->t_recheck_s = t2 - t1;
    }
    return LQT_OK;
}

}  // extern "C"
