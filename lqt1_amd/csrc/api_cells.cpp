// C ABI: errors, parameter presets, device queries and the independent-cell
// handle (the Cell/TPORd object interface of SingleCell/Cell.cc:22-42 and
// SingleCell/TPORd.cc:137-376).
#include <math.h>
#include <string.h>

#include <vector>

#include "common.hpp"

namespace lqt {

static thread_local char g_err[1024] = "";

lqt_status set_error(lqt_status code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

Consts<double> derive_consts(const lqt_params& p) {
    Consts<double> c;
    for (int cls = 0; cls < 3; ++cls) {
        const bool epi = cls == CLS_EPI, mid = cls == CLS_M;
        double GKs = 0.0034 * p.sca;
        if (epi) GKs *= 1.4;
        if (p.iks_scale != 1.0) GKs = p.iks_scale * GKs;
        if (p.suprep) {
            double fw, fr;
            lqt_suprep_fractions(p.dose_fmol, &fw, &fr);
            const double GKs_rep = epi ? 1.4 * 0.0034 : 0.0034;
            GKs = fw * GKs + fr * GKs_rep;
        }
        if (p.iso) GKs *= 3.2;
        c.gks[cls] = GKs;
        double PCa = 0.0001;
        if (p.iso) PCa *= p.pca_iso;
        if (epi) PCa *= 1.2;
        if (mid) PCa *= 2.5;
        c.pca[cls] = PCa;
    }
    c.dss_shift = p.iso ? 12.0 : 0.0;
    c.fss_shift = p.iso ? 8.0 : 0.0;
    c.clamp = p.clamp;
    return c;
}

Consts<float> to_float(const Consts<double>& c) {
    Consts<float> f;
    for (int i = 0; i < 3; ++i) {
        f.gks[i] = (float)c.gks[i];
        f.pca[i] = (float)c.pca[i];
    }
    f.dss_shift = (float)c.dss_shift;
    f.fss_shift = (float)c.fss_shift;
    f.clamp = c.clamp;
    return f;
}

}  // namespace lqt

using namespace lqt;

extern "C" {

const char* lqt_last_error(void) { return g_err; }

lqt_status lqt_suprep_fractions(double dose, double* frac_wt, double* frac_rep) {
    LQT_CHECK(frac_wt && frac_rep && dose >= 0, LQT_ERR_ARG, "lqt_suprep_fractions: bad argument");
    // GPU3D.cu:893-900: power law above 25 fmol, geometric below; the
    // replacing fraction is 2.018*tanh(0.0059*dose) written with exp.
    double fw = 1.0;
    if (dose >= 25)
        fw = 3.129 * pow(dose, -0.560);
    else
        fw = pow(0.974, dose);
    *frac_wt = fw;
    *frac_rep = 2.018 * (1.0 - exp(-2.0 * dose * 0.0059)) / (1.0 + exp(-2.0 * dose * 0.0059));
    return LQT_OK;
}

lqt_status lqt_params_preset(const char* name, lqt_params* out) {
    LQT_CHECK(name && out, LQT_ERR_ARG, "lqt_params_preset: null argument");
    lqt_params p;
    p.iks_scale = 1.0;
    p.iso = 1;
    p.pca_iso = 1.6;
    p.suprep = 0;
    p.dose_fmol = 150.0;
    p.sca = 1.0;
    p.clamp = 0;
    p.v0 = -87.5;
    if (!strcmp(name, "cpu_default")) {
        // TPORd.cc:30-36 shipped flags: CON + ISO1
    } else if (!strcmp(name, "gpu3d")) {
        // GPU3D.cu:18-22: LQT1 + ISO1 + SUPREP(100 fmol), PCa x4, V0 -86.2
        p.iks_scale = 0.05;
        p.pca_iso = 4.0;
        p.suprep = 1;
        p.dose_fmol = 100.0;
        p.v0 = -86.2;
    } else if (!strcmp(name, "gpu2d")) {
        // new_11.cu:18-21: LQT1, ISO off, V0 -86.2
        p.iks_scale = 0.05;
        p.iso = 0;
        p.v0 = -86.2;
    } else if (!strcmp(name, "config1")) {
        // control strand, ISO off
        p.iso = 0;
    } else if (!strcmp(name, "config3")) {
        // IKs x0.2 with ISO on, PCa x4 as in GPU3D.cu:767
        p.iks_scale = 0.2;
        p.pca_iso = 4.0;
    } else {
        return set_error(LQT_ERR_ARG, "lqt_params_preset: unknown preset '%s'", name);
    }
    *out = p;
    return LQT_OK;
}

int lqt_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

lqt_status lqt_device_synchronize(void) {
    LQT_HIP(hipDeviceSynchronize());
    return LQT_OK;
}

lqt_status lqt_bench_copy(int64_t bytes, int reps, double* us_per_copy) {
    LQT_CHECK(bytes >= 16 && bytes % 16 == 0 && reps > 0 && us_per_copy, LQT_ERR_ARG,
              "lqt_bench_copy: bytes must be a positive multiple of 16");
    void *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    lqt_status rc = [&]() -> lqt_status {
        LQT_HIP(hipMalloc(&a, bytes));
        LQT_HIP(hipMalloc(&b, bytes));
        LQT_HIP(hipStreamCreate(&st));
        LQT_HIP(hipEventCreate(&e0));
        LQT_HIP(hipEventCreate(&e1));
        launch_fill(true, a, bytes / 4, 1.0, st);
        launch_copy(b, a, bytes, st);  // first touch
        LQT_HIP(hipEventRecord(e0, st));
        for (int i = 0; i < reps; ++i) launch_copy(b, a, bytes, st);
        LQT_HIP(hipEventRecord(e1, st));
        LQT_HIP(hipGetLastError());
        LQT_HIP(hipStreamSynchronize(st));
        float ms = 0;
        LQT_HIP(hipEventElapsedTime(&ms, e0, e1));
        *us_per_copy = 1000.0 * ms / reps;
        return LQT_OK;
    }();
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (st) hipStreamDestroy(st);
    hipFree(a);
    hipFree(b);
    return rc;
}

lqt_status lqt_device_name(char* buf, int len) {
    LQT_CHECK(buf && len > 0, LQT_ERR_ARG, "lqt_device_name: bad buffer");
    int dev = 0;
    LQT_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    LQT_HIP(hipGetDeviceProperties(&prop, dev));
    snprintf(buf, len, "%s", prop.gcnArchName);
    return LQT_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// cells

struct lqt_cells {
    int n = 0;
    int64_t pitch = 0;
    bool fp32 = false;
    double dt = 0.02;
    void* r = nullptr;         // NR * pitch of R
    double* d = nullptr;       // ND * pitch
    void* v = nullptr;         // n of R
    uint8_t* cls = nullptr;
    int* pidx = nullptr;       // nullptr when one parameter set
    void* consts = nullptr;    // nparams Consts<R>
    PaceTrack* track = nullptr;
    double* dbuf = nullptr;    // 2n doubles of staging for istim / dvgap
    double* trace = nullptr;
    int64_t trace_cap = 0;
    hipStream_t stream = nullptr;
    int lanes = 0;             // lqt_cells_set_lanes (0: automatic)
    int ncu = 0;               // compute units of the device
};

static size_t rsize(const lqt_cells* c) { return c->fp32 ? sizeof(float) : sizeof(double); }

// Lanes per cell of the next launch.  More lanes per cell shorten the
// step while the wavefronts do not crowd the SIMDs (measured with the
// single-cell update at 8,192-49,152 cells on 256 compute units): four
// lanes, a 256-lane workgroup of four wavefronts per 64 cells, up to one
// workgroup per compute unit, and in FP32 up to two (two four-lane
// wavefronts on a SIMD still issue fewer instructions than one single-lane
// wavefront: 1.71 against 2.15 us per step); an FP64 four-lane wavefront
// needs more than half of a SIMD's registers, so a second workgroup waits
// for the first (7.88 us) and two lanes (6.2 us against 10.0 on one) are
// used up to two 128-lane workgroups per compute unit.  Beyond that every
// SIMD is busy either way and the single-lane update, which issues the
// fewest instructions in total, is fastest.
static int lanes_per_cell(const lqt_cells* c) {
    if (c->lanes > 0) return c->lanes;
    const int64_t groups = ((int64_t)c->n + 63) / 64;
    if (groups <= (int64_t)c->ncu) return 4;
    if (groups <= 2 * (int64_t)c->ncu) return c->fp32 ? 4 : 2;
    return 1;
}

extern "C" {

lqt_status lqt_cells_create(int n, const uint8_t* celltype, const lqt_params* params, int nparams,
                            lqt_precision prec, double dt, lqt_cells** out) {
    LQT_CHECK(n > 0 && celltype && params && out, LQT_ERR_ARG, "lqt_cells_create: bad argument");
    LQT_CHECK(nparams == 1 || nparams == n, LQT_ERR_ARG, "lqt_cells_create: nparams must be 1 or n");
    // state slots are addressed by 32-bit byte offsets of up to 8 bytes per cell
    LQT_CHECK(n <= (1 << 29) - 64, LQT_ERR_ARG, "lqt_cells_create: at most 2^29 - 64 cells per handle");
    LQT_CHECK(dt > 0, LQT_ERR_ARG, "lqt_cells_create: dt must be > 0");
    LQT_CHECK(prec == LQT_FP64 || dt <= LQT_FP32_MAX_DT, LQT_ERR_ARG, "lqt_cells_create: FP32 needs dt <= %g ms",
              LQT_FP32_MAX_DT);
    LQT_CHECK(prec == LQT_FP32 || prec == LQT_FP64, LQT_ERR_ARG, "lqt_cells_create: bad precision");
    int ndev = 0;
    LQT_CHECK(hipGetDeviceCount(&ndev) == hipSuccess && ndev > 0, LQT_ERR_HIP,
              "lqt_cells_create: no HIP device available");
    int dev = 0, ncu = 0;
    LQT_CHECK(hipGetDevice(&dev) == hipSuccess &&
                  hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && ncu > 0,
              LQT_ERR_HIP, "lqt_cells_create: cannot read the device's compute unit count");
    lqt_cells* c = new lqt_cells();
    c->n = n;
    c->ncu = ncu;
    c->pitch = ((int64_t)n + 63) / 64 * 64;
    c->fp32 = prec == LQT_FP32;
    c->dt = dt;
    lqt_status st;
#define TRY(x)                            \
    if ((st = (x)) != LQT_OK) {           \
        lqt_cells_destroy(c);             \
        return st;                        \
    }
    TRY(([&]() -> lqt_status {
        LQT_HIP(hipStreamCreate(&c->stream));
        LQT_HIP(hipMalloc(&c->r, rsize(c) * NR * c->pitch));
        LQT_HIP(hipMalloc((void**)&c->d, sizeof(double) * ND * c->pitch));
        LQT_HIP(hipMalloc(&c->v, rsize(c) * c->pitch));
        LQT_HIP(hipMalloc((void**)&c->dbuf, sizeof(double) * 2 * c->pitch));
        LQT_HIP(hipMalloc((void**)&c->track, sizeof(PaceTrack) * c->pitch));
        std::vector<uint8_t> cl(n);
        for (int i = 0; i < n; ++i) cl[i] = (uint8_t)cell_class(celltype[i]);
        lqt_status s2 = upload(&c->cls, cl.data(), n);
        if (s2) return s2;
        std::vector<Consts<double>> cd(nparams);
        for (int i = 0; i < nparams; ++i) cd[i] = derive_consts(params[i]);
        if (c->fp32) {
            std::vector<Consts<float>> cf(nparams);
            for (int i = 0; i < nparams; ++i) cf[i] = to_float(cd[i]);
            s2 = upload((Consts<float>**)&c->consts, cf.data(), nparams);
        } else {
            s2 = upload((Consts<double>**)&c->consts, cd.data(), nparams);
        }
        if (s2) return s2;
        if (nparams == n) {
            std::vector<int> pi(n);
            for (int i = 0; i < n; ++i) pi[i] = i;
            s2 = upload(&c->pidx, pi.data(), n);
            if (s2) return s2;
        }
        // initial state of TPORd::init with the parameter set's V0
        std::vector<double> s40((size_t)n * 40);
        for (int i = 0; i < n; ++i) {
            memcpy(&s40[(size_t)i * 40], kInit40, sizeof(kInit40));
            s40[(size_t)i * 40] = params[nparams == n ? i : 0].v0;
        }
        return lqt_cells_set_state(c, s40.data());
    }()));
#undef TRY
    *out = c;
    return LQT_OK;
}

lqt_status lqt_cells_destroy(lqt_cells* c) {
    if (!c) return LQT_OK;
    if (c->stream) hipStreamSynchronize(c->stream);
    hipFree(c->r);
    hipFree(c->d);
    hipFree(c->v);
    hipFree(c->cls);
    hipFree(c->pidx);
    hipFree(c->consts);
    hipFree(c->track);
    hipFree(c->dbuf);
    hipFree(c->trace);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return LQT_OK;
}

lqt_status lqt_cells_step(lqt_cells* c, int nsteps, const double* istim, const double* dvgap) {
    LQT_CHECK(c && nsteps >= 0, LQT_ERR_ARG, "lqt_cells_step: bad argument");
    if (nsteps == 0) return LQT_OK;
    double* di = nullptr;
    double* dg = nullptr;
    if (istim) {
        di = c->dbuf;
        LQT_HIP(hipMemcpyAsync(di, istim, sizeof(double) * c->n, hipMemcpyHostToDevice, c->stream));
    }
    if (dvgap) {
        dg = c->dbuf + c->pitch;
        LQT_HIP(hipMemcpyAsync(dg, dvgap, sizeof(double) * c->n, hipMemcpyHostToDevice, c->stream));
    }
    launch_cells_step(c->fp32, lanes_per_cell(c), c->n, c->r, c->d, c->pitch, c->v, c->cls, c->pidx, c->consts,
                      nsteps, c->dt, di, dg, c->stream);
    LQT_HIP(hipGetLastError());
    LQT_HIP(hipStreamSynchronize(c->stream));
    return LQT_OK;
}

lqt_status lqt_cells_set_lanes(lqt_cells* c, int lanes) {
    LQT_CHECK(c && (lanes == 0 || lanes == 1 || lanes == 2 || lanes == 4), LQT_ERR_ARG,
              "lqt_cells_set_lanes: lanes must be 0 (automatic), 1, 2 or 4");
    c->lanes = lanes;
    return LQT_OK;
}

lqt_status lqt_cells_get_lanes(const lqt_cells* c, int* lanes) {
    LQT_CHECK(c && lanes, LQT_ERR_ARG, "lqt_cells_get_lanes: bad argument");
    *lanes = lanes_per_cell(c);
    return LQT_OK;
}

lqt_status lqt_cells_step_async(lqt_cells* c, int nsteps, const double* istim, const double* dvgap, void* stream) {
    LQT_CHECK(c && nsteps >= 0, LQT_ERR_ARG, "lqt_cells_step_async: bad argument");
    if (nsteps == 0) return LQT_OK;
    launch_cells_step(c->fp32, lanes_per_cell(c), c->n, c->r, c->d, c->pitch, c->v, c->cls, c->pidx, c->consts,
                      nsteps, c->dt, istim, dvgap, stream ? (hipStream_t)stream : c->stream);
    LQT_HIP(hipGetLastError());
    return LQT_OK;
}

lqt_status lqt_cells_get_state(lqt_cells* c, double* out) {
    LQT_CHECK(c && out, LQT_ERR_ARG, "lqt_cells_get_state: bad argument");
    const size_t rs = rsize(c);
    std::vector<unsigned char> r(rs * NR * c->pitch), v(rs * c->pitch);
    std::vector<double> d(ND * c->pitch);
    LQT_HIP(hipStreamSynchronize(c->stream));
    LQT_HIP(hipMemcpy(r.data(), c->r, r.size(), hipMemcpyDeviceToHost));
    LQT_HIP(hipMemcpy(v.data(), c->v, v.size(), hipMemcpyDeviceToHost));
    LQT_HIP(hipMemcpy(d.data(), c->d, sizeof(double) * d.size(), hipMemcpyDeviceToHost));
    auto rget = [&](int64_t idx) -> double {
        return c->fp32 ? (double)((const float*)r.data())[idx] : ((const double*)r.data())[idx];
    };
    for (int i = 0; i < c->n; ++i) {
        for (int j = 0; j < 40; ++j) {
            const Ref40 m = kRef40[j];
            double val = 1.0;
            if (m.kind == 0) val = c->fp32 ? (double)((const float*)v.data())[i] : ((const double*)v.data())[i];
            else if (m.kind == 1) val = rget((int64_t)m.slot * c->pitch + i);
            else if (m.kind == 2) val = d[(size_t)m.slot * c->pitch + i];
            out[(size_t)i * 40 + j] = val;
        }
    }
    return LQT_OK;
}

lqt_status lqt_cells_set_state(lqt_cells* c, const double* in) {
    LQT_CHECK(c && in, LQT_ERR_ARG, "lqt_cells_set_state: bad argument");
    const size_t rs = rsize(c);
    std::vector<unsigned char> r(rs * NR * c->pitch, 0), v(rs * c->pitch, 0);
    std::vector<double> d(ND * c->pitch, 0.0);
    auto rset = [&](int64_t idx, double val) {
        if (c->fp32) ((float*)r.data())[idx] = (float)val;
        else ((double*)r.data())[idx] = val;
    };
    for (int i = 0; i < c->n; ++i) {
        for (int j = 0; j < 40; ++j) {
            const Ref40 m = kRef40[j];
            const double val = in[(size_t)i * 40 + j];
            if (m.kind == 0) {
                if (c->fp32) ((float*)v.data())[i] = (float)val;
                else ((double*)v.data())[i] = val;
            } else if (m.kind == 1) {
                rset((int64_t)m.slot * c->pitch + i, val);
            } else if (m.kind == 2) {
                d[(size_t)m.slot * c->pitch + i] = val;
            }
        }
    }
    LQT_HIP(hipStreamSynchronize(c->stream));
    LQT_HIP(hipMemcpy(c->r, r.data(), r.size(), hipMemcpyHostToDevice));
    LQT_HIP(hipMemcpy(c->v, v.data(), v.size(), hipMemcpyHostToDevice));
    LQT_HIP(hipMemcpy(c->d, d.data(), sizeof(double) * d.size(), hipMemcpyHostToDevice));
    return LQT_OK;
}

lqt_status lqt_cells_get_v(lqt_cells* c, double* out) {
    LQT_CHECK(c && out, LQT_ERR_ARG, "lqt_cells_get_v: bad argument");
    LQT_HIP(hipStreamSynchronize(c->stream));
    if (c->fp32) {
        std::vector<float> v(c->n);
        LQT_HIP(hipMemcpy(v.data(), c->v, sizeof(float) * c->n, hipMemcpyDeviceToHost));
        for (int i = 0; i < c->n; ++i) out[i] = v[i];
    } else {
        LQT_HIP(hipMemcpy(out, c->v, sizeof(double) * c->n, hipMemcpyDeviceToHost));
    }
    return LQT_OK;
}

lqt_status lqt_cells_write_state_file(lqt_cells* c, int i, const char* path) {
    LQT_CHECK(c && path && i >= 0 && i < c->n, LQT_ERR_ARG, "lqt_cells_write_state_file: bad argument");
    std::vector<double> all((size_t)c->n * 40);
    lqt_status st = lqt_cells_get_state(c, all.data());
    if (st) return st;
    FILE* f = fopen(path, "w");
    LQT_CHECK(f, LQT_ERR_IO, "cannot open %s", path);
    for (int j = 0; j < 40; ++j) fprintf(f, "%4.10f\n", all[(size_t)i * 40 + j]);
    fclose(f);
    return LQT_OK;
}

lqt_status lqt_cells_read_state_file(lqt_cells* c, int i, const char* path) {
    LQT_CHECK(c && path && i >= 0 && i < c->n, LQT_ERR_ARG, "lqt_cells_read_state_file: bad argument");
    FILE* f = fopen(path, "r");
    LQT_CHECK(f, LQT_ERR_IO, "cannot open %s", path);
    double vals[40];
    for (int j = 0; j < 40; ++j) {
        if (fscanf(f, "%lf", &vals[j]) != 1) {
            fclose(f);
            return set_error(LQT_ERR_IO, "%s: expected 40 values", path);
        }
    }
    fclose(f);
    std::vector<double> all((size_t)c->n * 40);
    lqt_status st = lqt_cells_get_state(c, all.data());
    if (st) return st;
    memcpy(&all[(size_t)i * 40], vals, sizeof(vals));
    return lqt_cells_set_state(c, all.data());
}

lqt_status lqt_cells_pace(lqt_cells* c, const lqt_pacing* pace, double* stats, int trace_cell, int trace_every,
                          double* vtrace, int64_t trace_cap, int64_t* ntrace) {
    LQT_CHECK(c && pace && pace->bcl > 0 && pace->nbeats > 0, LQT_ERR_ARG, "lqt_cells_pace: bad argument");
    // tracker reset (SingleCell.cc keeps its statistics uninitialised until
    // the last beat starts; zero is used here)
    LQT_HIP(hipMemsetAsync(c->track, 0, sizeof(PaceTrack) * c->n, c->stream));
    if (vtrace) {
        LQT_CHECK(trace_cell >= 0 && trace_cell < c->n && trace_every > 0 && trace_cap > 0, LQT_ERR_ARG,
                  "lqt_cells_pace: bad trace arguments");
        if (c->trace_cap < trace_cap) {
            hipFree(c->trace);
            c->trace = nullptr;
            LQT_HIP(hipMalloc((void**)&c->trace, sizeof(double) * 3 * trace_cap));
            c->trace_cap = trace_cap;
        }
    }
    PaceArgs a;
    a.bcl = pace->bcl;
    a.stim_start = pace->stim_start;
    a.stim_dur = pace->stim_duration;
    a.stim_strength = pace->stim_strength;
    a.dt = c->dt;
    a.trace_cell = trace_cell;
    a.trace_every = vtrace ? trace_every : 0;
    a.trace_cap = trace_cap;
    a.trace = vtrace ? c->trace : nullptr;
    const double numS1 = pace->nbeats;  // SingleCell.cc:45 compares against a double
    const double bcl = a.bcl, s0 = a.stim_start, sd = a.stim_dur, ddt = c->dt;
    const double stopTime = numS1 * bcl;
    // SingleCell.cc:95: for (time = 0; time < stopTime; time += dt), split
    // into launches of at most `chunk` steps.  The host replays the same
    // double accumulator and evaluates the loop's time tests with the
    // reference's expressions (SingleCell.cc:98-106 stimulus window,
    // :144-147 start of the last beat's statistics, :149 the last beat),
    // handing each launch its steps as integers.  A launch also ends before
    // a (kPaceRanges + 1)-th stimulus window would open in it.
    const int chunk = 20000;
    double time = 0.0;
    int64_t step = 0;
    while (time < stopTime) {
        a.time0 = time;
        a.nstim = 0;
        a.rest_step = -1;
        a.last0 = a.last1 = 0;
        a.trace_next = INT32_MAX;
        a.trace_slot = 0;
        double t = time;
        int k = 0;
        bool prev_on = false, in_last = false;
        while (k < chunk && t < stopTime) {
            const double tc = t - floor(t / bcl) * bcl;
            const bool on = tc >= s0 && tc < s0 + sd;
            if (on && !prev_on) {
                if (a.nstim == kPaceRanges) break;
                a.stim_on[a.nstim] = k;
                a.stim_off[a.nstim] = chunk;
                ++a.nstim;
            }
            if (!on && prev_on) a.stim_off[a.nstim - 1] = k;
            prev_on = on;
            if (floor((t - s0) / bcl) == numS1 - 2 && floor((t + ddt - s0) / bcl) == numS1 - 1) a.rest_step = k;
            const bool last = floor((t - s0) / bcl) == numS1 - 1;
            if (last && !in_last) a.last0 = k;
            if (last) a.last1 = k + 1;
            in_last = last;
            if (a.trace_every > 0 && a.trace_next == INT32_MAX && (step + k) % a.trace_every == 0) {
                a.trace_next = k;
                a.trace_slot = (step + k) / a.trace_every;
            }
            t += ddt;
            ++k;
        }
        if (prev_on) a.stim_off[a.nstim - 1] = k;
        a.nsteps = k;
        launch_cells_pace(c->fp32, lanes_per_cell(c), c->n, c->r, c->d, c->pitch, c->v, c->cls, c->pidx, c->consts,
                          c->dt, c->track, a, c->stream);
        LQT_HIP(hipGetLastError());
        time = t;
        step += k;
    }
    LQT_HIP(hipStreamSynchronize(c->stream));
    std::vector<PaceTrack> tr(c->n);
    LQT_HIP(hipMemcpy(tr.data(), c->track, sizeof(PaceTrack) * c->n, hipMemcpyDeviceToHost));
    if (stats) {
        for (int i = 0; i < c->n; ++i) {
            double* s = stats + (size_t)i * 10;
            s[0] = tr[i].rest;
            s[1] = tr[i].dvdt_max;
            s[2] = tr[i].overshoot;
            s[3] = tr[i].amplitude;
            for (int j = 0; j < 6; ++j) s[4 + j] = tr[i].apd[j];
        }
    }
    if (vtrace) {
        const int64_t n = std::min<int64_t>(trace_cap, (step + trace_every - 1) / trace_every);
        LQT_HIP(hipMemcpy(vtrace, c->trace, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
        if (ntrace) *ntrace = n;
    }
    return LQT_OK;
}

}  // extern "C"
