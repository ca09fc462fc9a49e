// lqt_oned_vw: the vulnerable-window program of OneD_VW.cc:36-449 through
// the C ABI.  Same strand (25 ENDO, 35 MCELL, 40 EPI cells, dx 0.15 mm,
// D 0.154 mm^2/ms with 0.2x at cell 63, dt 0.02 ms), same protocol (five S1
// of -52 pA/pF x 3 ms on cells 0-2 at BCL 1000 ms; S2 of -104 pA/pF x 3 ms
// on cells 10-12 at s2 ms into the last beat), same search (MODE2 by
// default: s2 = 260, 260.1, ... until one end of the strand passes -10 mV
// and the other does not), same prints ("maxEndo = ", "maxEpi = ",
// "VM found!" / "No VM!") and the same per-candidate output file
// Outputs/Transmural1D_s2@<s2>.dat (V of all cells every 1 ms of the last
// beat, %4.10f tab separated).  The candidates run as strands of batched
// launches (lqt_vw_search); the files come from one more batched launch of
// the visited candidates with V recording on.
//
//   lqt_oned_vw [--mode 1|2|3] [--left MS] [--right MS] [--step MS]
//               [--batch N] [--max N] [--preset NAME] [--fp32] [--no-files]
//               [--outdir DIR]
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/lqt.h"
#include "outdir.hpp"

static int die(const char* what) {
    fprintf(stderr, "%s: %s\n", what, lqt_last_error());
    return 1;
}

int main(int argc, char** argv) {
    std::string preset = "cpu_default", outdir = "Outputs";
    lqt_vw_desc vw;
    memset(&vw, 0, sizeof(vw));
    vw.mode = 2;
    vw.left = vw.right = 260.0;  // OneD_VW.cc:201-202
    vw.step = 0.1;
    vw.boundary = 10.0;          // OneD_VW.cc:204
    vw.batch = 512;
    vw.max_candidates = 8000;
    bool fp32 = false, files = true;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&]() -> const char* { return i + 1 < argc ? argv[++i] : "0"; };
        if (a == "--mode") vw.mode = atoi(next());
        else if (a == "--left") vw.left = atof(next());
        else if (a == "--right") vw.right = atof(next());
        else if (a == "--step") vw.step = atof(next());
        else if (a == "--batch") vw.batch = atoi(next());
        else if (a == "--max") vw.max_candidates = atol(next());
        else if (a == "--preset") preset = next();
        else if (a == "--fp32") fp32 = true;
        else if (a == "--no-files") files = false;
        else if (a == "--outdir") outdir = next();
        else {
            fprintf(stderr, "unknown argument %s\n", a.c_str());
            return 2;
        }
    }
    lqt_params p;
    if (lqt_params_preset(preset.c_str(), &p)) return die("preset");
    const int n = 100, numS1 = 5;
    const double bcl = 1000.0, dt = 0.02, dx = 0.15, vent = 0.154;
    std::vector<uint8_t> ct(n);
    std::vector<double> coeff(n, vent);
    for (int i = 0; i < n; ++i) ct[i] = i < 25 ? LQT_ENDO : (i < 60 ? LQT_MCELL : LQT_EPI);
    coeff[63] = 0.2 * vent;  // OneD_VW.cc:103-104
    lqt_strand_desc d;
    memset(&d, 0, sizeof(d));
    d.ncell = n;
    d.dx = dx;
    d.bcl = bcl;
    d.stop_time = numS1 * bcl;
    d.stim_start = 0.0;
    d.stim_duration = 3.0;
    d.stim_strength = -52.0;
    d.nstim = 3;
    d.measure_beat = numS1 - 1;
    d.s2_base = (numS1 - 1) * bcl;
    d.s2_duration = 3.0;
    d.s2_strength = -104.0;
    d.s2_first = 10;
    d.s2_count = 3;
    const lqt_precision prec = fp32 ? LQT_FP32 : LQT_FP64;

    std::vector<double> s2((size_t)vw.max_candidates), vmax((size_t)vw.max_candidates * 2);
    lqt_vw_result r;
    if (lqt_vw_search(&d, ct.data(), coeff.data(), &p, prec, dt, &vw, &r, s2.data(), vmax.data(),
                      vw.max_candidates))
        return die("lqt_vw_search");
    for (int64_t i = 0; i < r.visited; ++i) {
        printf("s2startTime = %g ms\n", s2[i]);
        printf("maxEndo = %g\nmaxEpi = %g\n\n", vmax[2 * i], vmax[2 * i + 1]);
    }
    printf(r.found ? "VM found!\n" : "No VM!\n");
    fprintf(stderr, "%lld candidates visited, %lld strands in %d launches, %.1f ms on the device\n",
            (long long)r.visited, (long long)r.simulated, r.launches, r.device_ms);

    if (files && r.visited > 0 && !make_dirs(outdir)) {
        fprintf(stderr, "cannot create %s\n", outdir.c_str());
        return 1;
    }
    if (files && r.visited > 0) {
        // rows of the last beat every 50 steps (OneD_VW.cc:370-385): replay
        // the double time accumulator to find the recorded steps
        int64_t first_step = -1, nrows = 0;
        {
            double time = 0.0;
            for (int64_t step = 0; time <= d.stop_time; time += dt, ++step)
                if (floor(time / bcl) == numS1 - 1 && step % 50 == 0) {
                    if (first_step < 0) first_step = step;
                    ++nrows;
                }
        }
        const int chunk = 1024;
        for (int64_t c0 = 0; c0 < r.visited; c0 += chunk) {
            const int m = (int)std::min<int64_t>(chunk, r.visited - c0);
            std::vector<lqt_strand_desc> ds((size_t)m, d);
            std::vector<uint8_t> cts((size_t)m * n);
            std::vector<double> cos((size_t)m * n);
            for (int k = 0; k < m; ++k) {
                ds[k].cell_offset = k * n;
                ds[k].s2_start = s2[c0 + k];
                ds[k].vrec_every = 50;
                ds[k].rec_from = (numS1 - 1) * bcl;
                memcpy(&cts[(size_t)k * n], ct.data(), n);
                memcpy(&cos[(size_t)k * n], coeff.data(), sizeof(double) * n);
            }
            lqt_strands* s = nullptr;
            if (lqt_strands_create(m, ds.data(), cts.data(), cos.data(), &p, 1, prec, dt, &s))
                return die("lqt_strands_create");
            if (lqt_strands_run(s)) return die("lqt_strands_run");
            for (int k = 0; k < m; ++k) {
                int64_t rows = 0;
                double t0 = 0;
                lqt_strands_vtrace(s, k, nullptr, &t0, &rows);
                std::vector<double> v((size_t)rows * n);
                if (lqt_strands_vtrace(s, k, v.data(), &t0, &rows)) return die("lqt_strands_vtrace");
                char name[512];
                snprintf(name, sizeof(name), "%s/Transmural1D_s2@%.2f.dat", outdir.c_str(), s2[c0 + k]);
                FILE* f = fopen(name, "w");
                if (!f) {
                    fprintf(stderr, "cannot open %s\n", name);
                    return 1;
                }
                for (int64_t row = 0; row < rows && row < nrows; ++row) {
                    for (int j = 0; j < n; ++j) fprintf(f, "%4.10f\t", v[(size_t)row * n + j]);
                    fprintf(f, "\n");
                }
                fclose(f);
            }
            lqt_strands_destroy(s);
        }
    }
    printf("All done.\n");
    return 0;
}
