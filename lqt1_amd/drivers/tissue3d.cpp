// lqt_tissue3d: the whole-ventricle program of GPU3D.cu:1250-1860 on the
// GPU through the C ABI.  Reads the reference's geometry file (one voxel per
// line "x y z fx fy fz label", labels 0-20, GPU3D.cu:1276-1297), applies its
// label rules (GPU3D.cu:1447-1462: labels 4-6 become 2 when the fibre's y
// component is at most 4, which every unit fibre is, else 7; labels 14-16
// stay only inside the M-island sphere, distance < 75 voxels from
// (597, 336, 336), otherwise 14 -> 12 and 15-16 -> 17), maps labels to cell
// types (EPI 7-10 and 17-20, MCELL 4-6 and 14-16, else ENDO,
// GPU3D.cu:865-871), marks the S2 box (x in [300, 687), y in [120, 486),
// z in [420, 534), GPU3D.cu:1438), reads the stimulation profile
// ("x y z f f f t_act str", t_act in s, GPU3D.cu:1691-1721), paces at BCL
// 500 ms with -52 pA/pF for 3 ms from each site's activation time while the
// in-cycle time is below 82 ms (GPU3D.cu:1757-1762), writes
// Output3D/ap_<ms>.vtk every 10 ms (every 3rd voxel, -100 outside,
// GPU3D.cu:123-163) and reports the device-side V range guard
// (GPU3D.cu:1796-1829) instead of copying V to the host every step.  The
// S2 itself is commented out in the reference; --s2-time switches it on
// (-104 pA/pF for 3 ms on the S2 box).
//
//   lqt_tissue3d --geometry FILE --dims X Y Z [--stim FILE] [--dx MM]
//                [--time MS] [--bcl MS] [--preset NAME] [--fp64]
//                [--dfibre D] [--dcross D] [--outdir DIR] [--vtk-every MS]
//                [--maps FILE]  (activation time and APD90 per voxel: "x y z act apd90", of the first
//                beat whose upstroke starts in the window; -1 where none)
//                [--maps-window FROM TO]  (ms of the tissue clock, default 0 to one BCL;
//                tracking stops at TO)
//                [--vtk-stride N] [--island X Y Z R] [--s2-box X0 X1 Y0 Y1 Z0 Z1]
//                [--s2-time MS] [--window MS] [--labels-out FILE]
#include <math.h>
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

// labels of GPU3D.cu:865-871
static uint8_t celltype_of(int l) {
    if ((l >= 7 && l <= 10) || (l >= 17 && l <= 20)) return LQT_EPI;
    if ((l >= 4 && l <= 6) || (l >= 14 && l <= 16)) return LQT_MCELL;
    return LQT_ENDO;
}

int main(int argc, char** argv) {
    std::string geo, stim, preset = "gpu3d", outdir = "Output3D", labels_out, maps;
    int X = 690, Y = 489, Z = 537;  // GPU3D.cu:28-30
    double dx = 0.2, dt = 0.02, tmax = 60000.0, bcl = 500.0, vtk_every = 10.0;  // :41-42, :1725-1728
    double dfib = 0.154, dcross = 0.154;  // GPU3D.cu:37-38 (isotropic)
    double island[4] = {597, 336, 336, 75};         // 199*3, 112*3, 112*3, 25*3 (GPU3D.cu:1454-1456)
    int s2box[6] = {300, 687, 120, 486, 420, 534};  // 100*3 .. 229*3 etc. (GPU3D.cu:1438)
    double s2_time = -1.0, window = 82.0, maps_from = 0.0, maps_to = -1.0;
    int vtk_stride = 3;
    bool fp64 = false;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&]() -> const char* { return i + 1 < argc ? argv[++i] : "0"; };
        if (a == "--geometry") geo = next();
        else if (a == "--stim") stim = next();
        else if (a == "--dims") {
            X = atoi(next());
            Y = atoi(next());
            Z = atoi(next());
        } else if (a == "--dx") dx = atof(next());
        else if (a == "--time") tmax = atof(next());
        else if (a == "--bcl") bcl = atof(next());
        else if (a == "--preset") preset = next();
        else if (a == "--fp64") fp64 = true;
        else if (a == "--dfibre") dfib = atof(next());
        else if (a == "--dcross") dcross = atof(next());
        else if (a == "--outdir") outdir = next();
        else if (a == "--vtk-every") vtk_every = atof(next());
        else if (a == "--vtk-stride") vtk_stride = atoi(next());
        else if (a == "--island") {
            for (int k = 0; k < 4; ++k) island[k] = atof(next());
        } else if (a == "--s2-box") {
            for (int k = 0; k < 6; ++k) s2box[k] = atoi(next());
        } else if (a == "--s2-time") s2_time = atof(next());
        else if (a == "--window") window = atof(next());
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
    if (geo.empty() || vtk_stride < 1) {
        fprintf(stderr, "usage: lqt_tissue3d --geometry FILE --dims X Y Z [--stim FILE] ...\n");
        return 2;
    }
    lqt_params p;
    if (lqt_params_preset(preset.c_str(), &p)) return die("preset");

    printf("Reading file...\n");
    FILE* f = fopen(geo.c_str(), "r");
    if (!f) {
        fprintf(stderr, "cannot open %s\n", geo.c_str());
        return 1;
    }
    const int64_t ncell = (int64_t)X * Y * Z;
    std::vector<uint8_t> label(ncell, LQT_NONE);
    std::vector<int> rule_label(ncell, -1);
    std::vector<float> fib((size_t)ncell * 3, 0.0f);
    std::vector<int64_t> s2cells;
    int x, y, z, l;
    double fx, fy, fz;
    int64_t num = 0;
    while (fscanf(f, "%d %d %d %lf %lf %lf %d", &x, &y, &z, &fx, &fy, &fz, &l) == 7) {
        if (x < 0 || x >= X || y < 0 || y >= Y || z < 0 || z >= Z) {
            fprintf(stderr, "%s: voxel (%d,%d,%d) outside the %dx%dx%d grid\n", geo.c_str(), x, y, z, X, Y, Z);
            return 1;
        }
        if ((x == 0 || x == X - 1 || y == 0 || y == Y - 1 || z == 0 || z == Z - 1) && l != 0)
            printf("Tissue exists in geometry boundary!\n");
        const int64_t c = ((int64_t)x * Y + y) * Z + z;
        if (x >= s2box[0] && x < s2box[1] && y >= s2box[2] && y < s2box[3] && z >= s2box[4] && z < s2box[5])
            s2cells.push_back(c);
        int t7 = l;
        // GPU3D.cu:1447-1451 compares the fibre's y component (t5), not the label
        if (t7 >= 4 && t7 <= 6) t7 = fy <= 4 ? 2 : 7;
        const double dist = sqrt(pow(x - island[0], 2.0) + pow(y - island[1], 2.0) + pow(z - island[2], 2.0));
        if (t7 >= 14 && t7 <= 16 && !(dist < island[3])) t7 = t7 <= 14 ? 12 : 17;
        rule_label[c] = t7;
        label[c] = celltype_of(t7);
        // the fibre enters the tensor in single precision (GPU3D.cu reads it as float)
        fib[3 * c] = (float)fx;
        fib[3 * c + 1] = (float)fy;
        fib[3 * c + 2] = (float)fz;
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
                fprintf(o, "%lld %lld %lld %d %d\n", (long long)(c / ((int64_t)Y * Z)), (long long)(c / Z % Y),
                        (long long)(c % Z), rule_label[c], (int)label[c]);
        fclose(o);
    }
    lqt_geometry g;
    g.nx = X;
    g.ny = Y;
    g.nz = Z;
    g.dx = dx;
    g.label = label.data();
    g.fibre = fib.data();
    g.d_fibre = dfib;
    g.d_cross = dcross;
    lqt_tissue* t = nullptr;
    if (lqt_tissue_create(&g, &p, fp64 ? LQT_FP64 : LQT_FP32, dt, &t)) return die("lqt_tissue_create");
    int64_t nvox = 0, nghost = 0;
    lqt_tissue_info(t, &nvox, &nghost);
    printf("%lld tissue voxels, %lld ghost cells.\n", (long long)nvox, (long long)nghost);

    printf("Reading stimulation profile... ");
    std::vector<int64_t> sites;
    std::vector<float> offs;
    if (!stim.empty()) {
        FILE* fs = fopen(stim.c_str(), "r");
        if (!fs) {
            fprintf(stderr, "cannot open %s\n", stim.c_str());
            return 1;
        }
        int t1, t2, t3;
        float t4, t5, t6, t7;
        char rest[256];
        while (fscanf(fs, "%d %d %d %f %f %f %f %255s", &t1, &t2, &t3, &t4, &t5, &t6, &t7, rest) == 8) {
            if (t1 < 0 || t1 >= X || t2 < 0 || t2 >= Y || t3 < 0 || t3 >= Z) continue;
            const int64_t c = ((int64_t)t1 * Y + t2) * Z + t3;
            if (label[c] == LQT_NONE) continue;  // GPU3D.cu:1704
            sites.push_back(c);
            offs.push_back(1000.0f * t7);  // float product, as set_Istim forms it (:403)
        }
        fclose(fs);
    }
    printf("There are %lld stimulus points.\n\n", (long long)sites.size());
    lqt_stim_protocol sp = {bcl, 3.0, -52.0, s2_time, s2_time >= 0 ? 3.0 : 0.0, -104.0, window};
    if (lqt_tissue_set_stimulus(t, &sp, (int64_t)sites.size(), sites.data(), offs.data(), (int64_t)s2cells.size(),
                                s2cells.data()))
        return die("lqt_tissue_set_stimulus");

    // the trackers hold one beat: track from maps_from to maps_to (one BCL by default)
    if (maps_to < 0.0) maps_to = maps_from + bcl;
    if (!maps.empty() && lqt_tissue_track(t, maps_from, maps_to)) return die("lqt_tissue_track");

    const int64_t per_vtk = (int64_t)(vtk_every / dt + 0.5);
    const int64_t per_ms = (int64_t)(1.0 / dt + 0.5);
    double time = 0.0;
    int64_t step = 0;
    while (time <= tmax) {
        if (per_vtk > 0 && step % per_vtk == 0) {
            printf("Progress = %g%%\n", 100.0 * time / tmax);
            if (!make_dirs(outdir)) {
                fprintf(stderr, "cannot create %s\n", outdir.c_str());
                return 1;
            }
            const std::string fn = outdir + "/ap_" + std::to_string((long long)(step / per_ms)) + ".vtk";
            if (lqt_tissue_write_vtk(t, fn.c_str(), vtk_stride)) return die("vtk");
            int64_t gd[3];
            lqt_tissue_v_guard(t, gd);
            if (gd[1]) printf("UP POINT NUMBER = %lld (voxel-steps so far)\n", (long long)gd[1]);
            if (gd[0]) printf("DW POINT NUMBER = %lld (voxel-steps so far)\n", (long long)gd[0]);
        }
        // run to the next output step without passing tmax
        const int64_t chunk = per_vtk > 0 ? per_vtk - step % per_vtk : 1000;
        double tt = time;
        int64_t k = 0;
        while (k < chunk && tt <= tmax) {
            tt += dt;
            ++k;
        }
        if (lqt_tissue_step(t, k)) return die("lqt_tissue_step");
        time = tt;
        step += k;
    }
    if (!maps.empty()) {
        // activation time and APD90 of every tissue voxel, in the compacted
        // (x, y, z) order of the label grid
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
        for (int x = 0; x < X; ++x)
            for (int y = 0; y < Y; ++y)
                for (int z = 0; z < Z; ++z)
                    if (label[((int64_t)x * Y + y) * Z + z] != LQT_NONE) {
                        fprintf(o, "%d %d %d %.2f %.2f\n", x, y, z, act[k], apd[k]);
                        ++k;
                    }
        fclose(o);
    }
    lqt_tissue_destroy(t);
    printf("All done.\n");
    return 0;
}
