// lqt_tissue3d: the whole-ventricle program of GPU3D.cu:1250-1860 on the
// GPU through the C ABI.  Reads the reference's geometry file (one voxel per
// line "x y z fx fy fz label", labels 1-20) and stimulation profile
// ("x y z f f f t_act str", t_act in s), paces at BCL with -52 pA/pF for
// 3 ms from each site's activation time, writes Output3D/ap_<ms>.vtk every
// 10 ms (every 3rd voxel, -100 outside, GPU3D.cu:123-163) and reports the
// device-side V range guard (GPU3D.cu:1796-1829) instead of copying V to the
// host every step.
//
//   lqt_tissue3d --geometry FILE --dims X Y Z [--stim FILE] [--dx MM]
//                [--time MS] [--bcl MS] [--preset NAME] [--fp64]
//                [--dfibre D] [--dcross D] [--outdir DIR] [--vtk-every MS]
#include <stdio.h>
#include <stdlib.h>

#include <string>

#include "../../include/lqt.h"

static int die(const char* what) {
    fprintf(stderr, "%s: %s\n", what, lqt_last_error());
    return 1;
}

int main(int argc, char** argv) {
    std::string geo, stim, preset = "gpu3d", outdir = "Output3D";
    int X = 690, Y = 489, Z = 537;  // GPU3D.cu:28-30
    double dx = 0.2, dt = 0.02, tmax = 60000.0, bcl = 500.0, vtk_every = 10.0;
    double dfib = 0.154, dcross = 0.154;  // GPU3D.cu:37-38 (isotropic)
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
        else {
            fprintf(stderr, "unknown argument %s\n", a.c_str());
            return 2;
        }
    }
    if (geo.empty()) {
        fprintf(stderr, "usage: lqt_tissue3d --geometry FILE --dims X Y Z [--stim FILE] ...\n");
        return 2;
    }
    lqt_params p;
    if (lqt_params_preset(preset.c_str(), &p)) return die("preset");
    lqt_tissue* t = nullptr;
    printf("Reading file...\n");
    if (lqt_tissue_create_from_file(geo.c_str(), X, Y, Z, dx, dfib, dcross, &p, fp64 ? LQT_FP64 : LQT_FP32, dt, &t))
        return die("lqt_tissue_create_from_file");
    int64_t nvox = 0, nghost = 0;
    lqt_tissue_info(t, &nvox, &nghost);
    printf("There are %lld ventricular points (%lld ghost cells).\n", (long long)nvox, (long long)nghost);
    lqt_stim_protocol sp = {bcl, 3.0, -52.0, 0.0, 0.0, -104.0};
    if (!stim.empty()) {
        if (lqt_tissue_set_stimulus_file(t, &sp, stim.c_str())) return die("stimulation profile");
    }
    const int64_t per_vtk = (int64_t)(vtk_every / dt + 0.5);
    double time = 0.0;
    int64_t step = 0;
    while (time <= tmax) {
        if (per_vtk > 0 && step % per_vtk == 0) {
            printf("Progress = %g%%\n", 100.0 * time / tmax);
            std::string f = outdir + "/ap_" + std::to_string((long long)(step / (int64_t)(1.0 / dt + 0.5))) + ".vtk";
            if (lqt_tissue_write_vtk(t, f.c_str(), 3)) return die("vtk");
            int64_t g[3];
            lqt_tissue_v_guard(t, g);
            if (g[1]) printf("UP POINT NUMBER = %lld (voxel-steps so far)\n", (long long)g[1]);
            if (g[0]) printf("DW POINT NUMBER = %lld (voxel-steps so far)\n", (long long)g[0]);
        }
        int64_t chunk = per_vtk > 0 ? per_vtk - step % per_vtk : 1000;
        // do not step past tmax: the loop runs while time <= tmax
        double tt = time;
        int64_t k = 0;
        while (k < chunk && tt <= tmax) {
            tt += dt;
            ++k;
        }
        lqt_status s = lqt_tissue_step(t, k);
        if (s) return die("lqt_tissue_step");
        time = tt;
        step += k;
    }
    lqt_tissue_destroy(t);
    printf("All done.\n");
    return 0;
}
