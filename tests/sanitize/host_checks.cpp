// Host-only entry points of liblqt under AddressSanitizer and
// UndefinedBehaviorSanitizer: the parameter presets and SUPREP fractions,
// the compaction and ghost mirror map (3D shells, a 2D ring, a box corner),
// the QT measurement on synthetic pseudo-ECGs, and the readers of the
// reference's voxel list and stimulation profile (well-formed, trailing
// newline, voxel outside the grid, empty, missing, site outside the grid).
// No GPU is touched.  Prints one line per check; exits non-zero on a failed
// check (the sanitizers abort on their own findings).
//
//     host_checks <scratch dir>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "lqt.h"

static int failures = 0;
#define CHECK(cond, ...)                        \
    do {                                        \
        if (!(cond)) {                          \
            printf("FAIL %s:%d ", __FILE__, __LINE__); \
            printf(__VA_ARGS__);                \
            printf("\n");                       \
            ++failures;                         \
        }                                       \
    } while (0)

static void presets() {
    const char* names[] = {"cpu_default", "gpu3d", "gpu2d", "config1", "config3"};
    for (const char* n : names) {
        lqt_params p;
        CHECK(lqt_params_preset(n, &p) == LQT_OK, "preset %s", n);
    }
    lqt_params p;
    CHECK(lqt_params_preset("none", &p) == LQT_ERR_ARG, "unknown preset accepted");
    CHECK(strlen(lqt_last_error()) > 0, "no error message");
    CHECK(lqt_params_preset(nullptr, &p) != LQT_OK, "null preset name accepted");
    for (double d : {0.0, 10.0, 24.9, 25.0, 100.0, 300.0}) {
        double w = -1, r = -1;
        CHECK(lqt_suprep_fractions(d, &w, &r) == LQT_OK && w >= 0 && r >= 0, "suprep %g", d);
    }
    printf("presets ok\n");
}

// shell of a sphere (radius r_in..r_out in cells) on an n^3 grid
static std::vector<uint8_t> shell(int nx, int ny, int nz, double r_in, double r_out) {
    std::vector<uint8_t> lab((size_t)nx * ny * nz, LQT_NONE);
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y)
            for (int z = 0; z < nz; ++z) {
                const double dx = x - (nx - 1) / 2.0, dy = y - (ny - 1) / 2.0, dz = z - (nz - 1) / 2.0;
                const double r = sqrt(dx * dx + dy * dy + dz * dz);
                if (r >= r_in && r <= r_out) lab[((size_t)x * ny + y) * nz + z] = r < (r_in + r_out) / 2 ? LQT_ENDO : LQT_EPI;
            }
    return lab;
}

static void ghosts() {
    struct Case { int nx, ny, nz; double a, b; } cases[] = {{9, 8, 7, 1.5, 3.6}, {16, 15, 13, 3.0, 7.0},
                                                           {12, 12, 1, 2.0, 5.0}, {3, 3, 3, 0.0, 9.0}};
    for (const Case& c : cases) {
        std::vector<uint8_t> lab = shell(c.nx, c.ny, c.nz, c.a, c.b);
        std::vector<int32_t> g(lab.size());
        int64_t nvox = -1, ngh = -1;
        CHECK(lqt_geometry_ghost_map(c.nx, c.ny, c.nz, lab.data(), g.data(), &nvox, &ngh) == LQT_OK, "ghost map");
        int64_t t = 0, gh = 0;
        for (size_t i = 0; i < lab.size(); ++i) {
            t += lab[i] != LQT_NONE;
            if (g[i] < 0) {
                ++gh;
                CHECK(-g[i] - 1 < nvox, "ghost source out of range");
            }
            if (lab[i] != LQT_NONE) CHECK(g[i] > 0, "tissue cell without index");
        }
        CHECK(t == nvox && gh == ngh, "counts %lld %lld vs %lld %lld", (long long)t, (long long)gh,
              (long long)nvox, (long long)ngh);
    }
    CHECK(lqt_geometry_ghost_map(0, 4, 4, nullptr, nullptr, nullptr, nullptr) != LQT_OK, "empty grid accepted");
    printf("ghost maps ok\n");
}

static void qt() {
    // Gaussian T wave at 300 ms (sigma 20) after a QRS at 60 ms, sampled every 0.2 ms
    std::vector<double> phi(5000);
    for (size_t j = 0; j < phi.size(); ++j) {
        const double t = j * 0.2;
        phi[j] = 0.3 * exp(-0.5 * pow((t - 60.0) / 3.0, 2)) + 0.05 * exp(-0.5 * pow((t - 300.0) / 20.0, 2));
    }
    double q = 0;
    CHECK(lqt_ecg_qt(phi.data(), (int64_t)phi.size(), 0.0, 0.2, 50.0, 80.0, 1000.0, &q) == LQT_OK, "qt");
    const double tend = 300.0 + 20.0 * sqrt(-2.0 * log(0.05));
    CHECK(fabs(q - (tend - 50.0)) <= 0.2 + 1e-9, "qt %g vs %g", q, tend - 50.0);
    CHECK(lqt_ecg_qt(phi.data(), (int64_t)phi.size(), 0.0, 0.2, 50.0, 80.0, 320.0, &q) == LQT_OK && q == -1.0,
          "qt without a T end %g", q);
    CHECK(lqt_ecg_qt(phi.data(), 0, 0.0, 0.2, 50.0, 80.0, 1000.0, &q) == LQT_OK && q == -1.0, "empty trace %g", q);
    CHECK(lqt_ecg_qt(nullptr, 5, 0.0, 0.2, 50.0, 80.0, 1000.0, &q) == LQT_ERR_ARG, "null trace accepted");
    printf("qt ok\n");
}

static std::string write(const std::string& dir, const char* name, const char* text) {
    const std::string p = dir + "/" + name;
    FILE* f = fopen(p.c_str(), "w");
    if (f) {
        fputs(text, f);
        fclose(f);
    }
    return p;
}

static void readers(const std::string& dir) {
    const int nx = 4, ny = 3, nz = 5;
    std::vector<uint8_t> lab(nx * ny * nz);
    std::vector<float> fib(3 * lab.size());
    int64_t nv = -1;
    std::string p = write(dir, "geo_ok.txt", "0 0 0 1 0 0 5\n1 2 4 0 1 0 9\n3 2 4 0 0 1 17");
    CHECK(lqt_geometry_read_file(p.c_str(), nx, ny, nz, lab.data(), fib.data(), &nv) == LQT_OK && nv == 3, "read %lld",
          (long long)nv);
    CHECK(lab[0] != LQT_NONE && lab[(1 * ny + 2) * nz + 4] != LQT_NONE && lab[1] == LQT_NONE, "labels");
    CHECK(fib[3 * ((1 * ny + 2) * nz + 4) + 1] == 1.0f, "fibre");
    p = write(dir, "geo_nl.txt", "0 0 0 1 0 0 5\n1 2 4 0 1 0 9\n");
    CHECK(lqt_geometry_read_file(p.c_str(), nx, ny, nz, lab.data(), nullptr, &nv) == LQT_OK && nv == 2,
          "trailing newline %lld", (long long)nv);
    p = write(dir, "geo_out.txt", "0 0 0 1 0 0 5\n4 0 0 1 0 0 5\n");
    CHECK(lqt_geometry_read_file(p.c_str(), nx, ny, nz, lab.data(), fib.data(), &nv) == LQT_ERR_ARG, "outside grid");
    p = write(dir, "geo_empty.txt", "");
    CHECK(lqt_geometry_read_file(p.c_str(), nx, ny, nz, lab.data(), fib.data(), &nv) == LQT_ERR_IO, "empty");
    p = write(dir, "geo_junk.txt", "0 0 zero 1 0 0 5\n");
    CHECK(lqt_geometry_read_file(p.c_str(), nx, ny, nz, lab.data(), fib.data(), &nv) == LQT_ERR_IO, "junk");
    CHECK(lqt_geometry_read_file((dir + "/missing.txt").c_str(), nx, ny, nz, lab.data(), fib.data(), &nv) == LQT_ERR_IO,
          "missing file");
    // stimulation profile: one site outside the grid is skipped
    p = write(dir, "stim.cnd", "0 0 0 1 1 1 0.001 s\n9 9 9 1 1 1 0.002 s\n3 2 4 1 1 1 0.0025 s\n");
    int64_t n = -1;
    CHECK(lqt_stimulus_read_file(p.c_str(), nx, ny, nz, 0, nullptr, nullptr, &n) == LQT_OK && n == 2, "count %lld",
          (long long)n);
    std::vector<int64_t> cells(2);
    std::vector<float> offs(2);
    CHECK(lqt_stimulus_read_file(p.c_str(), nx, ny, nz, 2, cells.data(), offs.data(), &n) == LQT_OK, "sites");
    CHECK(cells[0] == 0 && cells[1] == (3 * ny + 2) * nz + 4 && offs[0] == 1.0f && offs[1] == 2.5f, "site values");
    CHECK(lqt_stimulus_read_file(p.c_str(), nx, ny, nz, 1, cells.data(), offs.data(), &n) == LQT_ERR_ARG, "cap");
    printf("readers ok\n");
}

int main(int argc, char** argv) {
    const std::string dir = argc > 1 ? argv[1] : ".";
    presets();
    ghosts();
    qt();
    readers(dir);
    printf("%s\n", failures ? "host checks FAILED" : "host checks ok");
    return failures ? 1 : 0;
}
