/* The FP64 oracle (oracle/tpord_oracle.c) under AddressSanitizer and
 * UndefinedBehaviorSanitizer on fixed scenarios: three cells for 3000 steps
 * with a stimulus, the single-cell protocol for 2 beats, a 30-cell strand
 * for one 300 ms beat with its V rows and pseudo-ECG, an S2 candidate
 * paused and resumed, the beat tracker on a synthetic trace, and a small
 * anisotropic shell tissue (ghost map, tensor, 40 steps).  Every result is
 * written as raw doubles to the output file, for the test to compare with
 * the same scenarios through the unsanitised liboracle.so.
 *
 *     oracle_checks <output file>
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static FILE *out;
static void put(const double *x, long n) { fwrite(x, sizeof(double), (size_t)n, out); }
static void put1(double x) { put(&x, 1); }

int main(int argc, char **argv) {
    if (argc < 2) return 2;
    out = fopen(argv[1], "wb");
    if (!out) return 2;
    or_params p;
    or_params_default(&p);

    /* 1. cells */
    const int types[3] = {OR_ENDO, OR_MCELL, OR_EPI};
    for (int c = 0; c < 3; ++c) {
        double s[OR_NSTATE];
        or_cell_init(s, -87.5);
        double t = 0.0;
        for (int k = 0; k < 3000; ++k) {
            const double tc = t - floor(t / 1000.0) * 1000.0;
            or_cell_update(s, types[c], &p, 0.02, (tc >= 50.0 && tc < 51.0) ? -53.0 : 0.0, 0.0, NULL);
            t += 0.02;
        }
        put(s, OR_NSTATE);
    }

    /* 2. single-cell protocol, 2 beats at BCL 500 */
    {
        double stats[10];
        double trace[600] = {0};
        const int64_t n = or_single_cell(OR_MCELL, &p, 500.0, 2, 0.02, -53.0, 1.0, 50.0, stats, trace, 100, 600);
        put1((double)n);
        put(stats, 10);
        put(trace, 600);
    }

    /* 3. strand: 30 cells, one beat of 300 ms */
    enum { NC = 30 };
    uint8_t ct[NC];
    double coeff[NC];
    for (int i = 0; i < NC; ++i) {
        ct[i] = i < 8 ? OR_ENDO : (i < 18 ? OR_MCELL : OR_EPI);
        coeff[i] = 0.154;
    }
    {
        const int64_t cap = (int64_t)(300.0 / 0.02 / 10) + 3;
        double *vrec = (double *)calloc((size_t)cap * NC, sizeof(double));
        double *rt = (double *)calloc((size_t)cap, sizeof(double));
        double act[NC], apd[NC], fin[NC * OR_NSTATE];
        int ead[NC];
        int64_t rows = 0;
        const int64_t steps = or_strand_run2(NC, ct, coeff, &p, 0.15, 0.02, 300.0, 300.0, -52.0, 1.0, 50.0, 3, vrec,
                                             rt, 10, cap, &rows, 0, act, apd, ead, fin);
        put1((double)steps);
        put1((double)rows);
        put(act, NC);
        put(apd, NC);
        for (int i = 0; i < NC; ++i) put1((double)ead[i]);
        put(fin, NC * OR_NSTATE);
        double *phi = (double *)calloc((size_t)rows, sizeof(double));
        or_pseudo_ecg(rows, NC, vrec, 0.15, NC * 0.15 + 20.0, phi);
        put(phi, rows);
        free(phi);
        free(vrec);
        free(rt);
    }

    /* 4. S2 candidate (OneD_VW.cc), paused at 250 ms and resumed */
    {
        or_s2 s2 = {300.0, 20.0, 3.0, -104.0, 10, 3};
        double s[NC * OR_NSTATE];
        for (int i = 0; i < NC; ++i) or_cell_init(s + i * OR_NSTATE, -87.5);
        double time = 0.0, vmax[2] = {-100.0, -100.0};
        int64_t step = 0, rows = 0;
        or_strand_vw(NC, ct, coeff, &p, 0.15, 0.02, 300.0, 600.0, -52.0, 1.0, 50.0, 3, &s2, s, &time, &step, 250.0,
                     vmax, NULL, NULL, 0, 1, 0, &rows);
        put1(time);
        or_strand_vw(NC, ct, coeff, &p, 0.15, 0.02, 300.0, 600.0, -52.0, 1.0, 50.0, 3, &s2, s, &time, &step, 1e300,
                     vmax, NULL, NULL, 0, 1, 0, &rows);
        put1(time);
        put1((double)step);
        put(vmax, 2);
        put(s, NC * OR_NSTATE);
    }

    /* 5. tracker on a synthetic AP */
    {
        or_track tr;
        or_track_init(&tr);
        for (int k = 0; k < 40000; ++k) {
            const double t = k * 0.02;
            double v = -85.0;
            if (t >= 60.0 && t < 61.0) v = -85.0 + 125.0 * (t - 60.0);
            else if (t >= 61.0 && t < 360.0) v = 40.0 - 112.5 * (t - 61.0) / 299.0;
            or_track_step(&tr, t, v);
        }
        double a = 0, d = 0;
        const int e = or_track_result(&tr, &a, &d);
        put1(a);
        put1(d);
        put1((double)e);
    }

    /* 6. tissue: anisotropic shell on 9 x 8 x 7, 40 steps */
    {
        const int nx = 9, ny = 8, nz = 7, ncell = nx * ny * nz;
        uint8_t mask[9 * 8 * 7];
        float fib[3 * 9 * 8 * 7];
        int32_t g[9 * 8 * 7];
        for (int x = 0; x < nx; ++x)
            for (int y = 0; y < ny; ++y)
                for (int z = 0; z < nz; ++z) {
                    const int c = (x * ny + y) * nz + z;
                    const double r = sqrt(pow(x - 4.0, 2) + pow(y - 3.5, 2) + pow(z - 3.0, 2));
                    mask[c] = r >= 1.2 && r <= 3.6;
                    fib[3 * c] = (float)cos(0.3 * z);
                    fib[3 * c + 1] = (float)sin(0.3 * z);
                    fib[3 * c + 2] = 0.0f;
                }
        const int64_t nvox = or_ghost_map(nx, ny, nz, mask, g);
        put1((double)nvox);
        for (int c = 0; c < ncell; ++c) put1((double)g[c]);
        float *vf = (float *)calloc((size_t)nvox * 3, sizeof(float));
        uint8_t *vt = (uint8_t *)calloc((size_t)nvox, 1);
        int64_t k = 0;
        for (int c = 0; c < ncell; ++c)
            if (mask[c]) {
                memcpy(vf + 3 * k, fib + 3 * c, 3 * sizeof(float));
                vt[k] = (c % 3 == 0) ? OR_EPI : OR_ENDO;
                ++k;
            }
        double *dten = (double *)calloc((size_t)nvox * 6, sizeof(double));
        double *drift = (double *)calloc((size_t)nvox * 3, sizeof(double));
        or_tissue_tensor(nx, ny, nz, g, nvox, vf, 0.154, 0.0385, 0.2, dten, drift);
        double *s = (double *)calloc((size_t)nvox * OR_NSTATE, sizeof(double));
        for (int64_t i = 0; i < nvox; ++i) or_cell_init(s + i * OR_NSTATE, -87.5);
        double *ist = (double *)calloc((size_t)nvox, sizeof(double));
        for (int64_t i = 0; i < nvox; i += 5) ist[i] = -80.0;
        double *vb = (double *)calloc((size_t)nvox, sizeof(double));
        double *du = (double *)calloc((size_t)nvox, sizeof(double));
        for (int st = 0; st < 40; ++st) or_tissue_step(nx, ny, nz, g, nvox, dten, drift, 0.2, 0.02, vt, &p, s, ist, vb, du);
        put(dten, nvox * 6);
        put(drift, nvox * 3);
        put(s, nvox * OR_NSTATE);
        free(vf); free(vt); free(dten); free(drift); free(s); free(ist); free(vb); free(du);
    }
    fclose(out);
    printf("oracle checks ok\n");
    return 0;
}
