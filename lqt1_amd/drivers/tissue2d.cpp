// lqt_tissue2d: the real-slice 2D program of new_11.cu:1262-1860 on the GPU
// through the C ABI.  Reads the slice ("x y fx fy label" per line, labels
// 0-20, new_11.cu:1296-1316), applies the reference's label rules
// (new_11.cu:1460-1489: labels 3-4 -> 2 and 5-6 -> 7; labels 13-16 stay only
// inside the M-island rectangle x in [550, 689), y in [150, 487), otherwise
// 13-14 -> 12 and 15-16 -> 17; type 0 is updated as ENDO and held at
// -100 mV, :1247-1248), marks the S2 box (x in [300, 400), y in [200, 350),
// original label 11-13, :1470), reads the stimulus layer z == 157 of the
// stimulation profile (:1720-1746), paces -52 pA/pF for 3 ms at the start
// of every BCL (:1785-1790; the profile's activation times are not used
// there), and writes Outputs2Dreal/ap_<ms>.vtk every 10 ms (:142-181).
// The S2 itself is commented out in the reference; --s2-time switches it on
// (-104 pA/pF for 3 ms on the S2 box).
//
//   lqt_tissue2d [--geometry FILE] [--stim FILE] [--dims X Y] [--dx MM]
//                [--time MS] [--bcl MS] [--preset NAME] [--fp64]
//                [--dfibre D] [--dcross D] [--outdir DIR] [--vtk-every MS]
//                [--maps FILE]  (activation time and APD90 per cell: "x y act apd90", of the first
//                beat whose upstroke starts in the window; -1 where none)
//                [--maps-window FROM TO]  (ms of the tissue clock, default 0 to one BCL;
//                tracking stops at TO)
//                [--layer Z] [--island X0 X1 Y0 Y1] [--s2-time MS]
//                [--labels-out FILE]
#include <stdio.h>
#include <stdlib.h>

#include <string>
#include <vector>

#include "../../include/lqt.h"
#include "outdir.hpp"

static int die(const char* what) {
    fprintf(stderr, "%s: %s\n", what, lqt_last_error());
    return 1;
}

// labels of GPU3D.cu:865-871 / new_11.cu Itot: EPI {7-10, 17-20},
// MCELL {4-6, 14-16}, everything else ENDO
static uint8_t celltype_of(int l) {
    if ((l >= 7 && l <= 10) || (l >= 17 && l <= 20)) return LQT_EPI;
    if ((l >= 4 && l <= 6) || (l >= 14 && l <= 16)) return LQT_MCELL;
    return LQT_ENDO;
}

int main(int argc, char** argv) {
    std::string geo = "157layerwithfibervector.txt", stim = "stimulationprofile.cnd", preset = "gpu2d",
                outdir = "Outputs2Dreal", labels_out, maps;
    int X = 690, Y = 489, layer = 157;       // new_11.cu:29-30, :1734
    double dx = 0.2, dt = 0.02;              // new_11.cu:62-63
    double tmax = 30000.0, bcl = 1000.0, vtk_every = 10.0;  // :1751-1755
    double maps_from = 0.0, maps_to = -1.0;  // the --maps window; TO < 0: one BCL
    double dfib = 0.154, dcross = 0.0385;    // D1, D2 of new_11.cu:49-51
    int isl[4] = {550, 689, 150, 487};       // new_11.cu:1484
    const int s2box[6] = {300, 400, 200, 350, 11, 13};  // new_11.cu:1470
    double s2_time = -1.0;
    bool fp64 = false;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&]() -> const char* { return i + 1 < argc ? argv[++i] : "0"; };
        if (a == "--geometry") geo = next();
        else if (a == "--stim") stim = next();
        else if (a == "--dims") {
            X = atoi(next());
            Y = atoi(next());
        } else if (a == "--dx") dx = atof(next());
        else if (a == "--time") tmax = atof(next());
        else if (a == "--bcl") bcl = atof(next());
        else if (a == "--preset") preset = next();
        else if (a == "--fp64") fp64 = true;
        else if (a == "--dfibre") dfib = atof(next());
        else if (a == "--dcross") dcross = atof(next());
        else if (a == "--outdir") outdir = next();
        else if (a == "--vtk-every") vtk_every = atof(next());
        else if (a == "--layer") layer = atoi(next());
        else if (a == "--island") {
            for (int k = 0; k < 4; ++k) isl[k] = atoi(next());
        } else if (a == "--s2-time") s2_time = atof(next());
        else if (a == "--labels-out") labels_out = next();
        else if (a == "--maps") maps = next();
        else if (a == "--maps-window") {
            maps_from = atof(next());
            maps_to = atof(next());
        }
        else {
            fprintf(stderr, "unknown argument %s\n", a.c_str());
            return 2;
        }
    }
    lqt_params p;
    if (lqt_params_preset(preset.c_str(), &p)) return die("preset");

    printf("Reading file...\n");
    FILE* f = fopen(geo.c_str(), "r");
    if (!f) {
        fprintf(stderr, "cannot open %s\n", geo.c_str());
        return 1;
    }
    const int64_t ncell = (int64_t)X * Y;
    std::vector<uint8_t> label(ncell, LQT_NONE);
    std::vector<int> rule_label(ncell, -1);
    std::vector<float> fib((size_t)ncell * 3, 0.0f);
    std::vector<int64_t> s2cells;
    int x, y, l;
    float fx, fy;
    int64_t num = 0;
    while (fscanf(f, "%d %d %f %f %d", &x, &y, &fx, &fy, &l) == 5) {
        if (x < 0 || x >= X || y < 0 || y >= Y) {
            fprintf(stderr, "%s: point (%d,%d) outside the %dx%d grid\n", geo.c_str(), x, y, X, Y);
            return 1;
        }
        if ((x == 0 || x == X - 1 || y == 0 || y == Y - 1) && l != 0)
            printf("Tissue exists in geometry boundary!\n");
        const int64_t c = (int64_t)x * Y + y;
        // S2 box on the label as read (new_11.cu:1470)
        if (x >= s2box[0] && x < s2box[1] && y >= s2box[2] && y < s2box[3] && l <= s2box[5] && l >= s2box[4])
            s2cells.push_back(c);
        int t5 = l;
        if (t5 >= 3 && t5 <= 6) t5 = t5 <= 4 ? 2 : 7;
        if (t5 >= 13 && t5 <= 16) {
            const bool inside = x >= isl[0] && x < isl[1] && y >= isl[2] && y < isl[3];
            if (!inside) t5 = t5 <= 14 ? 12 : 17;
        }
        rule_label[c] = t5;
        label[c] = t5 == 0 ? (uint8_t)(LQT_ENDO | LQT_HOLD) : celltype_of(t5);
        fib[3 * c] = fx;
        fib[3 * c + 1] = fy;
        ++num;
    }
    fclose(f);
    printf("There are %lld ventricular points.\n", (long long)num);
    if (!labels_out.empty()) {
        FILE* o = fopen(labels_out.c_str(), "w");
        if (!o) {
            fprintf(stderr, "cannot open %s\n", labels_out.c_str());
            return 1;
        }
        for (int64_t c = 0; c < ncell; ++c)
            if (rule_label[c] >= 0)
                fprintf(o, "%lld %lld %d %d\n", (long long)(c / Y), (long long)(c % Y), rule_label[c], (int)label[c]);
        fclose(o);
    }
    lqt_geometry g;
    g.nx = X;
    g.ny = Y;
    g.nz = 1;
    g.dx = dx;
    g.label = label.data();
    g.fibre = fib.data();
    g.d_fibre = dfib;
    g.d_cross = dcross;
    lqt_tissue* t = nullptr;
    if (lqt_tissue_create(&g, &p, fp64 ? LQT_FP64 : LQT_FP32, dt, &t)) return die("lqt_tissue_create");

    printf("Reading stimulation profile... ");
    std::vector<int64_t> sites;
    FILE* fs = fopen(stim.c_str(), "r");
    if (fs) {
        int t1, t2, t3;
        float t4, t5, t6, t7;
        char rest[256];
        while (fscanf(fs, "%d %d %d %f %f %f %f %255s", &t1, &t2, &t3, &t4, &t5, &t6, &t7, rest) == 8) {
            if (t3 != layer || t1 < 0 || t1 >= X || t2 < 0 || t2 >= Y) continue;
            const int64_t c = (int64_t)t1 * Y + t2;
            if (label[c] != LQT_NONE) sites.push_back(c);
        }
        fclose(fs);
    }
    printf("There are %lld stimulus points.\n\n", (long long)sites.size());
    // set_Istim only while the in-cycle time is below 82 ms (new_11.cu:1785)
    lqt_stim_protocol sp = {bcl, 3.0, -52.0, s2_time, s2_time >= 0 ? 3.0 : 0.0, -104.0, 82.0};
    std::vector<float> off(sites.size(), 0.0f);
    if (lqt_tissue_set_stimulus(t, &sp, (int64_t)sites.size(), sites.data(), off.data(), (int64_t)s2cells.size(),
                                s2cells.data()))
        return die("lqt_tissue_set_stimulus");

    // the trackers hold one beat: track from maps_from to maps_to (one BCL by default)
    if (maps_to < 0.0) maps_to = maps_from + bcl;
    if (!maps.empty() && lqt_tissue_track(t, maps_from, maps_to)) return die("lqt_tissue_track");

    const int64_t per_vtk = (int64_t)(vtk_every / dt);
    const int64_t per_ms = (int64_t)(1 / dt);
    double time = 0.0;
    int64_t count = 0;
    while (time <= tmax) {
        if (per_vtk > 0 && count % per_vtk == 0) {
            printf("Progress = %g%%\n", 100.0 * time / tmax);
            if (!make_dirs(outdir)) {
                fprintf(stderr, "cannot create %s\n", outdir.c_str());
                return 1;
            }
            const std::string fn = outdir + "/ap_" + std::to_string((long long)(count / per_ms)) + ".vtk";
            if (lqt_tissue_write_vtk(t, fn.c_str(), 1)) return die("vtk");
            int64_t gd[3];
            lqt_tissue_v_guard(t, gd);
            if (gd[1]) printf("UP POINT NUMBER = %lld (voxel-steps so far); time = %g\n", (long long)gd[1], time);
        }
        // run to the next output step without passing tmax
        const int64_t chunk = per_vtk > 0 ? per_vtk - count % per_vtk : 1000;
        double tt = time;
        int64_t k = 0;
        while (k < chunk && tt <= tmax) {
            tt += dt;
            ++k;
        }
        if (lqt_tissue_step(t, k)) return die("lqt_tissue_step");
        time = tt;
        count += k;
    }
    if (!maps.empty()) {
        // activation time and APD90 of every tissue cell, in the compacted
        // (x, y) order of the label grid
        int64_t nvox = 0, nghost = 0;
        if (lqt_tissue_info(t, &nvox, &nghost)) return die("lqt_tissue_info");
        std::vector<double> act((size_t)nvox), apd((size_t)nvox);
        if (lqt_tissue_track_results(t, act.data(), apd.data())) return die("lqt_tissue_track_results");
        if (!make_dirs(parent_dir(maps))) {
            fprintf(stderr, "cannot create the folder of %s\n", maps.c_str());
            return 1;
        }
        FILE* o = fopen(maps.c_str(), "w");
        if (!o) {
            fprintf(stderr, "cannot open %s\n", maps.c_str());
            return 1;
        }
        int64_t k = 0;
        for (int xx = 0; xx < X; ++xx)
            for (int yy = 0; yy < Y; ++yy)
                if (label[(int64_t)xx * Y + yy] != LQT_NONE) {
                    fprintf(o, "%d %d %.2f %.2f\n", xx, yy, act[k], apd[k]);
                    ++k;
                }
        fclose(o);
    }
    lqt_tissue_destroy(t);
    printf("All done.\n");
    return 0;
}
