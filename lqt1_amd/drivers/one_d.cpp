// lqt_oned: the transmural strand program of OneD.cc:40-448 on the GPU
// through the C ABI (one strand of the batched strand kernel).  Same strand
// (25 ENDO, 35 MCELL, 40 EPI, dx 0.15 mm, D 0.154 mm^2/ms, dt 0.02 ms), same
// pacing (BCL 1000 ms, numS1 = 5, stop at numS1*BCL + BCL, -52 pA/pF for
// 1 ms at 50 ms on cells 0-2), same output file (time and 100 V every 10
// steps, %4.10f, tab separated) and the same CV report between cells 10
// and 90 in beat numS1.
//
//   lqt_oned [--preset NAME] [--numS1 N] [--bcl MS] [--stop MS] [--fp32]
//            [--out FILE]
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/lqt.h"
#include "outdir.hpp"

static int die(const char* what) {
    fprintf(stderr, "%s: %s\n", what, lqt_last_error());
    return 1;
}

int main(int argc, char** argv) {
    std::string preset = "cpu_default", out;
    int numS1 = 5;
    double bcl = 1000.0, dt = 0.02, stop = -1.0;
    bool fp32 = false;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&]() -> const char* { return i + 1 < argc ? argv[++i] : ""; };
        if (a == "--preset") preset = next();
        else if (a == "--numS1") numS1 = atoi(next());
        else if (a == "--bcl") bcl = atof(next());
        else if (a == "--stop") stop = atof(next());
        else if (a == "--fp32") fp32 = true;
        else if (a == "--out") out = next();
        else {
            fprintf(stderr, "unknown argument %s\n", a.c_str());
            return 2;
        }
    }
    lqt_params p;
    if (lqt_params_preset(preset.c_str(), &p)) return die("preset");
    if (stop < 0) stop = numS1 * bcl + bcl;  // OneD.cc:78
    if (out.empty()) {
        // OneD.cc:282-309 file names by flag set
        std::string base = p.iks_scale != 1.0 ? "LQT1" : "CON";
        std::string iso = p.iso ? (base == "LQT1" ? "_wISO" : "_ISO") : "";
        out = "Outputs/VentOneDResults_" + base + iso + ".dat";
    }
    const int n = 100;
    std::vector<uint8_t> ct(n);
    for (int i = 0; i < n; ++i) ct[i] = i < 25 ? LQT_ENDO : (i < 60 ? LQT_MCELL : LQT_EPI);
    lqt_strand_desc d;
    memset(&d, 0, sizeof(d));
    d.ncell = n;
    d.dx = 0.15;
    d.bcl = bcl;
    d.stop_time = stop;
    d.stim_start = 50.0;
    d.stim_duration = 1.0;
    d.stim_strength = -52.0;
    d.nstim = 3;
    d.rec_every = 10;
    d.rec_from = 0.0;
    d.x0 = n * 0.15 + 20.0;
    d.measure_beat = numS1;  // CV in beat floor(t/BCL) == numS1 (OneD.cc:419)
    d.vrec_every = 10;       // OneD.cc:402
    lqt_strands* s = nullptr;
    if (lqt_strands_create(1, &d, ct.data(), nullptr, &p, 1, fp32 ? LQT_FP32 : LQT_FP64, dt, &s))
        return die("lqt_strands_create");
    if (lqt_strands_run(s)) return die("lqt_strands_run");
    int64_t rows = 0;
    double t0 = 0;
    lqt_strands_vtrace(s, 0, nullptr, &t0, &rows);
    std::vector<double> v((size_t)rows * n);
    if (lqt_strands_vtrace(s, 0, v.data(), &t0, &rows)) return die("lqt_strands_vtrace");
    if (!make_dirs(parent_dir(out))) {
        fprintf(stderr, "cannot create the folder of %s\n", out.c_str());
        return 1;
    }
    FILE* f = fopen(out.c_str(), "w+");
    if (!f) {
        fprintf(stderr, "cannot open %s\n", out.c_str());
        return 1;
    }
    // row r was recorded at step 10*r; replay the double time accumulator
    double time = 0.0;
    int64_t step = 0, r = 0;
    for (; r < rows; time += dt, ++step) {
        if (step % 10) continue;
        fprintf(f, "%4.10f\t", time);
        for (int j = 0; j < n; ++j) fprintf(f, "%4.10f\t", v[(size_t)r * n + j]);
        fprintf(f, "\n");
        ++r;
    }
    fclose(f);
    std::vector<double> act(n), apd(n);
    std::vector<int> ead(n);
    lqt_strands_cell_results(s, 0, act.data(), apd.data(), ead.data());
    if (act[10] >= 0 && act[90] >= 0) {
        printf("start = %g\nend = %g\nduration = %g\n", act[10], act[90], act[90] - act[10]);
        printf("CV = %g m/s.\n", 0.15 * 80 / (act[90] - act[10]));
    } else {
        printf("Conduction failure!\n");
    }
    printf("All done.\n");
    lqt_strands_destroy(s);
    return 0;
}
