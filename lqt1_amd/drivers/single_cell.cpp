// lqt_single_cell: the single-cell pacing program of SingleCell.cc:38-287 on
// the GPU through the C ABI.  Same protocol (BCL 1000 ms, 100 beats,
// -53 pA/pF for 1 ms at 50 ms, dt 0.02 ms, MCELL, CON+ISO1) and the same
// output files:
//   Outputs/VentriSingleCellResults_LQT1_ISO_MCELL.dat  "t V ICaL IKs" per step
//   Outputs/VentriSingleCellStats.dat                   last-beat statistics
//
//   lqt_single_cell [--celltype ENDO|MCELL|EPI] [--preset NAME] [--beats N]
//                   [--bcl MS] [--fp32] [--outdir DIR] [--no-trace]
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
    std::string preset = "cpu_default", outdir = "Outputs";
    int celltype = LQT_MCELL, beats = 100;
    double bcl = 1000.0, dt = 0.02;
    bool fp32 = false, trace = true;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&]() -> const char* { return i + 1 < argc ? argv[++i] : ""; };
        if (a == "--celltype") {
            std::string c = next();
            celltype = c == "ENDO" ? LQT_ENDO : c == "EPI" ? LQT_EPI : LQT_MCELL;
        } else if (a == "--preset") preset = next();
        else if (a == "--beats") beats = atoi(next());
        else if (a == "--bcl") bcl = atof(next());
        else if (a == "--fp32") fp32 = true;
        else if (a == "--outdir") outdir = next();
        else if (a == "--no-trace") trace = false;
        else {
            fprintf(stderr, "unknown argument %s\n", a.c_str());
            return 2;
        }
    }
    lqt_params p;
    if (lqt_params_preset(preset.c_str(), &p)) return die("preset");
    const uint8_t ct = (uint8_t)celltype;
    lqt_cells* c = nullptr;
    if (lqt_cells_create(1, &ct, &p, 1, fp32 ? LQT_FP32 : LQT_FP64, dt, &c)) return die("lqt_cells_create");
    lqt_pacing pace = {bcl, 50.0, 1.0, -53.0, beats};
    const int64_t nsteps = (int64_t)(beats * bcl / dt) + 2;
    std::vector<double> tr(trace ? 3 * nsteps : 0);
    int64_t ntr = 0;
    double stats[10];
    if (lqt_cells_pace(c, &pace, stats, 0, 1, trace ? tr.data() : nullptr, trace ? nsteps : 0, &ntr))
        return die("lqt_cells_pace");
    if (!make_dirs(outdir)) {
        fprintf(stderr, "cannot create %s\n", outdir.c_str());
        return 1;
    }
    if (trace) {
        std::string f = outdir + "/VentriSingleCellResults_LQT1_ISO_MCELL.dat";
        FILE* df = fopen(f.c_str(), "w+");
        if (!df) {
            fprintf(stderr, "cannot open %s\n", f.c_str());
            return 1;
        }
        double time = 0.0;  // the protocol's double accumulator
        for (int64_t k = 0; k < ntr; ++k, time += dt)
            fprintf(df, "%4.10f\t%4.10f\t%4.10f\t%4.10f\t\n", time, tr[3 * k], tr[3 * k + 1], tr[3 * k + 2]);
        fclose(df);
    }
    const char* names[10] = {"Resting membrane potential = %.5f mV\n", "Maximum dV/dt (UpstrokeVelocity_max) = %.5f mV/ms\n",
                             "Overshoot = %.5f mV\n", "Amplitude = %.5f mV\n", "APD20 = %.5f ms\n", "APD25 = %.5f ms\n",
                             "APD50 = %.5f ms\n", "APD75 = %.5f ms\n", "APD80 = %.5f ms\n", "APD90 = %.5f ms\n"};
    std::string f = outdir + "/VentriSingleCellStats.dat";
    FILE* af = fopen(f.c_str(), "w+");
    if (!af) {
        fprintf(stderr, "cannot open %s\n", f.c_str());
        return 1;
    }
    for (int j = 0; j < 10; ++j) {
        if (af) fprintf(af, names[j], stats[j]);
        printf(names[j], stats[j]);
    }
    if (af) fclose(af);
    lqt_cells_destroy(c);
    return 0;
}
