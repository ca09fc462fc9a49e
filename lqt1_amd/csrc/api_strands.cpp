// C ABI: batched 1D strands (OneD.cc:40-448 for many strands at once) in
// one persistent launch captured in a HIP graph, plus the per-strand
// measurements: pseudo-ECG QT, APD90 per cell type, EAD flag and CV.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "common.hpp"
#include "strands.hpp"

namespace lqt {
int strands_occupancy(bool fp32, int block, bool split);
}
using namespace lqt;

constexpr int NMODES = 4;

struct lqt_strands {
    int n = 0;
    int64_t ncells = 0, pitch = 0;
    bool fp32 = true;
    double dt = 0.02;
    int block = 512;  // largest block of the active layout
    // launch layouts (NMODES):
    //   0 = one strand per workgroup, one persistent kernel per block-size
    //       class (strands of 1..8 wavefronts) as parallel graph branches;
    //   1 = one strand per workgroup, one kernel at the largest block;
    //   2 = strands packed on whole wavefronts into 512-lane workgroups
    //       (strands of similar step counts together), one kernel;
    //   3 = as 0, but strands of up to 4 wavefronts run split, each cell on
    //       two lanes of two wavefronts (one per current group), at twice
    //       the block size.
    struct Layout {
        int block, grid, first, count;  // bundles [first, first + count)
        bool split;
    };
    std::vector<Layout> layout[NMODES];
    int mode = 0;
    bool tuned = false;
    double tune_ms[NMODES] = {0, 0, 0, 0};
    hipStream_t side[8] = {};
    hipEvent_t fork = nullptr, join[8] = {};
    hipGraph_t graph_m[NMODES] = {nullptr, nullptr, nullptr, nullptr};
    hipGraphExec_t exec_m[NMODES] = {nullptr, nullptr, nullptr, nullptr};
    std::vector<lqt_strand_desc> desc;
    std::vector<uint8_t> label;  // per cell enum CellType
    std::vector<StrandDev> hdev;
    std::vector<Bundle> hbund;   // the bundles of all modes
    std::vector<int64_t> s1r_h;  // S1 step ranges of all strands
    // device
    StrandDev* ddesc = nullptr;
    int64_t* s1r = nullptr;
    Bundle* bund = nullptr;
    int* queue = nullptr;  // one work counter per layout entry
    uint8_t* cls = nullptr;
    double* coeff = nullptr;
    void* consts = nullptr;
    double* v0 = nullptr;
    float* phi = nullptr;
    int64_t* nphi = nullptr;
    double* phit0 = nullptr;
    double* vrec = nullptr;
    int64_t* nvrec = nullptr;
    double* vrect0 = nullptr;
    int64_t vrec_total = 0;
    double* act = nullptr;
    double* apd = nullptr;
    int* ead = nullptr;
    void* fr = nullptr;
    double* fd = nullptr;
    void* fv = nullptr;
    int64_t* steps = nullptr;
    double* vmax = nullptr;
    int64_t phi_total = 0;
    hipStream_t stream = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double last_ms = 0;
    bool ran = false;
    // a run enqueued by lqt_strands_run_async whose results have not been
    // copied to the host yet, and the stream it went to
    bool pending = false;
    hipStream_t pending_stream = nullptr;
    // host copies of the results of the last run
    std::vector<double> h_act, h_apd, h_phit0, h_vrect0;
    std::vector<int64_t> h_nvrec;
    std::vector<int> h_ead;
    std::vector<int64_t> h_nphi, h_steps;
    std::vector<double> h_vmax;
    std::vector<float> h_phi;
};

static size_t rsz(const lqt_strands* s) { return s->fp32 ? sizeof(float) : sizeof(double); }

template <class R>
static void fill_sargs(lqt_strands* s, StrandsArgs<R>& a) {
    a.nbundles = 0;
    a.bundles = s->bund;
    a.desc = s->ddesc;
    a.s1r = s->s1r;
    a.queue = s->queue;
    a.cls = s->cls;
    a.coeff = s->coeff;
    a.consts = (const Consts<R>*)s->consts;
    a.v0 = s->v0;
    a.dt = (R)s->dt;
    a.dt_d = s->dt;
    a.phi = s->phi;
    a.nphi = s->nphi;
    a.phit0 = s->phit0;
    a.vrec = s->vrec;
    a.nvrec = s->nvrec;
    a.vrect0 = s->vrect0;
    a.act = s->act;
    a.apd = s->apd;
    a.ead = s->ead;
    a.fin = SoA<R>{(R*)s->fr, s->fd, s->pitch};
    a.finv = (R*)s->fv;
    a.steps = s->steps;
    a.vmax = s->vmax;
}

static lqt_status launch_one(lqt_strands* s, int mode, int li, hipStream_t st) {
    const lqt_strands::Layout& L = s->layout[mode][li];
    if (s->fp32) {
        StrandsArgs<float> a;
        fill_sargs(s, a);
        a.nbundles = L.count;
        a.bundles = s->bund + L.first;
        a.queue = s->queue + li;
        launch_strands(true, L.block, L.split, L.grid, &a, st);
    } else {
        StrandsArgs<double> a;
        fill_sargs(s, a);
        a.nbundles = L.count;
        a.bundles = s->bund + L.first;
        a.queue = s->queue + li;
        launch_strands(false, L.block, L.split, L.grid, &a, st);
    }
    LQT_HIP(hipGetLastError());
    return LQT_OK;
}

// The whole run of a mode: zero the work counters (a kernel, no memset
// node), then every layout entry as a parallel branch of the same graph.
static lqt_status launch_run(lqt_strands* s, int mode, hipStream_t st) {
    const int nl = (int)s->layout[mode].size();
    launch_zero_int(s->queue, nl, st);
    LQT_HIP(hipEventRecord(s->fork, st));
    for (int li = 1; li < nl; ++li) {
        LQT_HIP(hipStreamWaitEvent(s->side[li], s->fork, 0));
        lqt_status e = launch_one(s, mode, li, s->side[li]);
        if (e) return e;
        LQT_HIP(hipEventRecord(s->join[li], s->side[li]));
    }
    lqt_status e = launch_one(s, mode, 0, st);
    if (e) return e;
    for (int li = 1; li < nl; ++li) LQT_HIP(hipStreamWaitEvent(st, s->join[li], 0));
    return LQT_OK;
}

static lqt_status graph_for(lqt_strands* s, int mode) {
    if (s->exec_m[mode]) return LQT_OK;
    LQT_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    lqt_status e = launch_run(s, mode, s->stream);
    hipGraph_t g = nullptr;
    hipError_t ce = hipStreamEndCapture(s->stream, &g);
    if (e) return e;
    LQT_HIP(ce);
    s->graph_m[mode] = g;
    LQT_HIP(hipGraphInstantiate(&s->exec_m[mode], g, nullptr, nullptr, 0));
    return LQT_OK;
}

// One run of a layout enqueued on stream st between the timing events.
static lqt_status enqueue_mode(lqt_strands* s, int mode, hipStream_t st) {
    lqt_status e = graph_for(s, mode);
    if (e) return e;
    LQT_HIP(hipEventRecord(s->e0, st));
    LQT_HIP(hipGraphLaunch(s->exec_m[mode], st));
    LQT_HIP(hipEventRecord(s->e1, st));
    return LQT_OK;
}

static lqt_status run_mode(lqt_strands* s, int mode, double* ms_out) {
    lqt_status e = enqueue_mode(s, mode, s->stream);
    if (e) return e;
    LQT_HIP(hipStreamSynchronize(s->stream));
    float ms = 0;
    LQT_HIP(hipEventElapsedTime(&ms, s->e0, s->e1));
    *ms_out = ms;
    return LQT_OK;
}

// QT of one strand from its pseudo-ECG: Q at the stimulus of the measured
// beat; the T wave is the largest |phi| after the QRS (which ends 10 ms
// after the last cell activates); T end is the first sample after the T
// peak where |phi| falls to 5 % of that peak.
template <class T>
static double qt_from_ecg(const T* phi, int64_t n, double t0, double dts, double t_q, double t_qrs_end,
                          double t_beat_end) {
    int64_t ipk = -1;
    double pk = 0;
    for (int64_t j = 0; j < n; ++j) {
        const double t = t0 + j * dts;
        if (t < t_qrs_end || t >= t_beat_end) continue;
        if (fabs(phi[j]) > pk) {
            pk = fabs(phi[j]);
            ipk = j;
        }
    }
    if (ipk < 0 || pk <= 0) return -1.0;
    for (int64_t j = ipk; j < n; ++j) {
        const double t = t0 + j * dts;
        if (t >= t_beat_end) break;
        if (fabs(phi[j]) <= 0.05 * pk) return t - t_q;
    }
    return -1.0;
}

// Step counts of `for (time = 0; time <= stop; time += dt)` (OneD.cc:315)
// for every stop time.  The accumulator runs through the same values for
// every strand, so it is run once up to the largest stop and each count is
// a binary search in it.
static std::vector<int64_t> step_counts(const std::vector<double>& stops, double dt,
                                        std::vector<double>* seq_out = nullptr) {
    double tmax = 0.0;
    for (double x : stops) tmax = std::max(tmax, x);
    std::vector<double> seq;  // time at the start of step k
    double t = 0.0;
    while (t <= tmax) {
        seq.push_back(t);
        t += dt;
    }
    std::vector<int64_t> out(stops.size());
    for (size_t i = 0; i < stops.size(); ++i)
        out[i] = std::upper_bound(seq.begin(), seq.end(), stops[i]) - seq.begin();
    if (seq_out) seq_out->swap(seq);
    return out;
}

// The loop's time tests (OneD.cc:337-347 and OneD_VW.cc:291-301), each in
// its own statements so that no multiply-add is fused: the same roundings
// as the reference's x86 build.
static bool s1_window(double t, const StrandDev& d) {
    const double fl = floor(t / d.bcl);
    const double base = fl * d.bcl;
    const double tc = t - base;
    return tc >= d.stim_start && tc < d.stim_start + d.stim_dur;
}

// First step index in [0, n) whose time satisfies a predicate that is false
// and then true along the accumulator (n if never).
template <class P>
static int64_t first_step(const std::vector<double>& seq, int64_t n, P pred) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (pred(seq[mid])) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

// Step-index form of the strand's time tests over its nsteps steps (seq:
// the accumulator's times, at least nsteps of them).  The S1 test is
// periodic: each beat's window is searched near its nominal start, where
// the in-cycle time rises monotonically, and the scan never goes back.
static void strand_events(StrandDev& h, const std::vector<double>& seq, std::vector<int64_t>& s1r) {
    const int64_t n = h.nsteps;
    const double dt = n > 1 ? seq[1] - seq[0] : 0.02;
    h.s1_first = (int64_t)s1r.size() / 2;
    int64_t cnt = 0, kscan = 0;
    if (n > 0 && h.stim_dur > 0) {
        const double tlast = seq[n - 1];
        const int64_t nb = (int64_t)floor(tlast / h.bcl) + 1;
        for (int64_t b = 0; b <= nb; ++b) {
            // the window opens at in-cycle time max(stim_start, 0): tc >= 0
            const double t_open = b * h.bcl + std::max(0.0, h.stim_start);
            if (t_open - 4 * dt > tlast) break;
            int64_t k = std::max<int64_t>(kscan, first_step(seq, n, [&](double t) { return t >= t_open - 4 * dt; }));
            bool on = false;
            int64_t k_on = 0;
            for (; k < n && seq[k] < t_open + h.stim_dur + 4 * dt; ++k) {
                const bool w = s1_window(seq[k], h);
                if (w && !on) k_on = k;
                if (!w && on) {
                    s1r.push_back(k_on);
                    s1r.push_back(k);
                    ++cnt;
                }
                on = w;
            }
            if (on) {  // the window runs to the scan's end: extend it step by step
                while (k < n && s1_window(seq[k], h)) ++k;
                s1r.push_back(k_on);
                s1r.push_back(k);
                ++cnt;
            }
            kscan = k;
        }
    }
    h.s1_count = cnt;
    const double s2_end = h.s2_start + h.s2_dur;
    h.k_wmax = first_step(seq, n, [&](double t) { return t - h.s2_base >= h.s2_start; });
    h.k_s2a = h.s2_dur > 0 ? h.k_wmax : n;
    h.k_s2b = h.s2_dur > 0 ? first_step(seq, n, [&](double t) { return !(t - h.s2_base < s2_end); }) : n;
    h.k_mb0 = first_step(seq, n, [&](double t) { return floor(t / h.bcl) >= h.measure_beat; });
    h.k_mb1 = first_step(seq, n, [&](double t) { return floor(t / h.bcl) >= h.measure_beat + 1.0; });
    const int64_t k_rec = first_step(seq, n, [&](double t) { return t >= h.rec_from; });
    auto next_multiple = [&](int64_t k, int every) -> int64_t {
        if (every <= 0 || k >= n) return INT64_MAX;
        return (k + every - 1) / every * every;
    };
    h.k_rec0 = next_multiple(k_rec, h.rec_every);
    h.k_vrec0 = next_multiple(k_rec, h.vrec_every);
}

static int strand_waves(const StrandDev& d) { return (d.ncell + 63) / 64; }

static Bundle empty_bundle() {
    Bundle b;
    memset(&b, 0, sizeof(b));
    for (int w = 0; w < MAXS; ++w) b.wslot[w] = -1;
    return b;
}

static void bundle_add(Bundle& b, int sidx, int wave0, int waves) {
    const int slot = b.n++;
    b.sidx[slot] = sidx;
    b.wave0[slot] = wave0;
    for (int w = wave0; w < wave0 + waves; ++w) b.wslot[w] = slot;
}

static int bundle_waves(const Bundle& b, const std::vector<StrandDev>& d) {
    int u = 0;
    for (int i = 0; i < b.n; ++i) u = std::max(u, b.wave0[i] + strand_waves(d[b.sidx[i]]));
    return u;
}

static void set_kmax(std::vector<Bundle>& bs, const std::vector<StrandDev>& d) {
    for (Bundle& b : bs) {
        int64_t m = 0, kb = INT64_MAX;
        for (int i = 0; i < b.n; ++i) {
            const StrandDev& h = d[b.sidx[i]];
            m = std::max(m, h.nsteps);
            kb = std::min({kb, h.k_mb0, h.k_wmax, h.k_rec0, h.k_vrec0});
        }
        b.kmax = m + 2;
        b.kbook = kb;
    }
}

// Layout 2: strands in decreasing step count are placed first-fit into
// 8-wavefront bundles whose longest strand runs at most 1/0.7 times as many
// steps (so that lanes of finished strands do not idle for long); bundles
// are then ordered by their work, largest first.
static std::vector<Bundle> pack_bundles(const std::vector<StrandDev>& d) {
    const int n = (int)d.size();
    std::vector<int> o(n);
    std::iota(o.begin(), o.end(), 0);
    std::stable_sort(o.begin(), o.end(), [&](int a, int b) {
        if (d[a].nsteps != d[b].nsteps) return d[a].nsteps > d[b].nsteps;
        return strand_waves(d[a]) > strand_waves(d[b]);
    });
    std::vector<Bundle> done, open;
    std::vector<int> used;
    std::vector<int64_t> top;
    for (int i : o) {
        const int w = strand_waves(d[i]);
        // bundles whose longest strand is too long for this one (and every
        // later one) are closed
        for (size_t k = 0; k < open.size();) {
            if ((double)d[i].nsteps < 0.7 * (double)top[k]) {
                done.push_back(open[k]);
                open.erase(open.begin() + k);
                used.erase(used.begin() + k);
                top.erase(top.begin() + k);
            } else {
                ++k;
            }
        }
        size_t k = 0;
        while (k < open.size() && !(used[k] + w <= MAXS && open[k].n < MAXS)) ++k;
        if (k == open.size()) {
            open.push_back(empty_bundle());
            used.push_back(0);
            top.push_back(d[i].nsteps);
        }
        bundle_add(open[k], i, used[k], w);
        used[k] += w;
    }
    done.insert(done.end(), open.begin(), open.end());
    set_kmax(done, d);
    std::stable_sort(done.begin(), done.end(), [&](const Bundle& a, const Bundle& b) {
        return (double)a.kmax * bundle_waves(a, d) > (double)b.kmax * bundle_waves(b, d);
    });
    return done;
}

extern "C" {

lqt_status lqt_strands_create(int nstrands, const lqt_strand_desc* desc, const uint8_t* celltype,
                              const double* coeff, const lqt_params* params, int nparams, lqt_precision prec,
                              double dt, lqt_strands** out) {
    LQT_CHECK(nstrands > 0 && desc && celltype && params && nparams > 0 && out && dt > 0, LQT_ERR_ARG,
              "lqt_strands_create: bad argument");
    LQT_CHECK(prec == LQT_FP64 || dt <= LQT_FP32_MAX_DT, LQT_ERR_ARG, "lqt_strands_create: FP32 needs dt <= %g ms",
              LQT_FP32_MAX_DT);
    int ndev = 0;
    LQT_CHECK(hipGetDeviceCount(&ndev) == hipSuccess && ndev > 0, LQT_ERR_HIP,
              "lqt_strands_create: no HIP device available");
    int64_t ncells = 0;
    int maxl = 0;
    for (int i = 0; i < nstrands; ++i) {
        const lqt_strand_desc& d = desc[i];
        LQT_CHECK(d.ncell >= 2 && d.ncell <= 512, LQT_ERR_ARG, "strand %d: ncell %d outside 2..512", i, d.ncell);
        LQT_CHECK(d.cell_offset >= 0 && d.param_index >= 0 && d.param_index < nparams, LQT_ERR_ARG,
                  "strand %d: bad cell_offset or param_index", i);
        LQT_CHECK(d.dx > 0 && d.bcl > 0 && d.stop_time >= 0 && d.rec_every >= 0 && d.nstim >= 0, LQT_ERR_ARG,
                  "strand %d: bad dx, bcl, stop_time, rec_every or nstim", i);
        LQT_CHECK(d.s2_duration >= 0 && d.s2_first >= 0 && d.s2_count >= 0, LQT_ERR_ARG,
                  "strand %d: bad S2 duration or cells", i);
        // the kernel carries step indices and sample intervals in 32 bits
        LQT_CHECK(d.stop_time / dt < (double)(1 << 30) && d.rec_every < (1 << 30) && d.vrec_every < (1 << 30),
                  LQT_ERR_ARG, "strand %d: 2^30 steps or more (stop_time / dt), or a sample interval as long", i);
        // the pseudo-ECG weight 1/(x_j - x0)^2 needs the electrode off every cell
        bool on_cell = false;
        for (int j = 0; j < d.ncell && d.rec_every > 0; ++j) on_cell |= (j * d.dx - d.x0) * (j * d.dx - d.x0) == 0.0;
        LQT_CHECK(!on_cell, LQT_ERR_ARG, "strand %d: pseudo-ECG electrode x0 = %g mm lies on a cell", i, d.x0);
        ncells = std::max<int64_t>(ncells, (int64_t)d.cell_offset + d.ncell);
        maxl = std::max(maxl, d.ncell);
    }
    // the final-state slots are addressed by 32-bit byte offsets
    LQT_CHECK(ncells <= (1 << 29) - 64, LQT_ERR_ARG, "lqt_strands_create: at most 2^29 - 64 cells in all");
    lqt_strands* s = new lqt_strands();
    s->n = nstrands;
    s->ncells = ncells;
    s->pitch = (ncells + 63) / 64 * 64;
    s->fp32 = prec == LQT_FP32;
    s->dt = dt;
    s->desc.assign(desc, desc + nstrands);
    s->label.assign(celltype, celltype + ncells);
    lqt_status st = [&]() -> lqt_status {
        LQT_HIP(hipStreamCreate(&s->stream));
        for (int i = 0; i < 8; ++i) {
            LQT_HIP(hipStreamCreate(&s->side[i]));
            LQT_HIP(hipEventCreateWithFlags(&s->join[i], hipEventDisableTiming));
        }
        LQT_HIP(hipEventCreateWithFlags(&s->fork, hipEventDisableTiming));
        LQT_HIP(hipEventCreate(&s->e0));
        LQT_HIP(hipEventCreate(&s->e1));
        // per strand device descriptors and sample capacities
        s->hdev.resize(nstrands);
        int64_t off = 0, voff = 0;
        for (int i = 0; i < nstrands; ++i) {
            const lqt_strand_desc& d = desc[i];
            StrandDev& h = s->hdev[i];
            h.ncell = d.ncell;
            h.cell_off = d.cell_offset;
            h.pidx = d.param_index;
            h.nstim = d.nstim;
            h.rec_every = d.rec_every;
            h.measure_beat = d.measure_beat;
            h.dx = d.dx;
            h.bcl = d.bcl;
            h.stop_time = d.stop_time;
            h.stim_start = d.stim_start;
            h.stim_dur = d.stim_duration;
            h.stim_strength = d.stim_strength;
            h.rec_from = d.rec_from;
            h.x0 = d.x0;
            h.s2_base = d.s2_base;
            h.s2_start = d.s2_start;
            h.s2_dur = d.s2_duration;
            h.s2_strength = d.s2_strength;
            h.s2_first = d.s2_first;
            h.s2_count = d.s2_count;
            int64_t cap = 0;
            if (d.rec_every > 0 && d.stop_time >= d.rec_from)
                cap = (int64_t)((d.stop_time - std::max(0.0, d.rec_from)) / dt / d.rec_every) + 3;
            h.phi_off = off;
            h.phi_cap = cap;
            off += cap;
            h.vrec_every = d.vrec_every;
            h.pad_ = 0;
            int64_t vcap = 0;
            if (d.vrec_every > 0 && d.stop_time >= d.rec_from)
                vcap = (int64_t)((d.stop_time - std::max(0.0, d.rec_from)) / dt / d.vrec_every) + 3;
            h.vrec_off = voff;
            h.vrec_cap = vcap;
            voff += vcap * d.ncell;
        }
        s->phi_total = off;
        s->vrec_total = voff;
        {
            std::vector<double> stops(nstrands);
            for (int i = 0; i < nstrands; ++i) stops[i] = desc[i].stop_time;
            std::vector<double> seq;
            const std::vector<int64_t> ns = step_counts(stops, dt, &seq);
            for (int i = 0; i < nstrands; ++i) {
                s->hdev[i].nsteps = ns[i];
                strand_events(s->hdev[i], seq, s->s1r_h);
            }
        }
        // Layout 1: every strand at the largest block, most work first (step
        // count x wavefronts).  Layout 0: strands grouped by their own
        // wavefront count (one block-size class each), most work first inside
        // a class.  Layout 2: packed bundles.
        const std::vector<StrandDev>& H = s->hdev;
        auto waves = [&](int i) { return strand_waves(H[i]); };
        auto cost = [&](int i) { return (double)H[i].nsteps * waves(i); };
        std::vector<int> o1(nstrands);
        std::iota(o1.begin(), o1.end(), 0);
        std::stable_sort(o1.begin(), o1.end(), [&](int a, int b) { return cost(a) > cost(b); });
        std::vector<int> o0 = o1;
        std::stable_sort(o0.begin(), o0.end(), [&](int a, int b) { return waves(a) > waves(b); });
        int dev = 0;
        LQT_HIP(hipGetDevice(&dev));
        hipDeviceProp_t prop;
        LQT_HIP(hipGetDeviceProperties(&prop, dev));
        auto grid_of = [&](int block, bool split, int count) {
            const int occ = std::max(1, strands_occupancy(s->fp32, block, split));
            return std::max(1, std::min(count, occ * prop.multiProcessorCount));
        };
        // layouts 0 and 3: one persistent kernel per wavefront-count class
        for (int mode : {0, 3}) {
            for (int i = 0; i < nstrands;) {
                int j = i;
                while (j < nstrands && waves(o0[j]) == waves(o0[i])) ++j;
                // FP64 splits only strands of up to 2 wavefronts: a split
                // workgroup of 6 or 8 wavefronts holds two per SIMD, and the
                // FP64 update does not fit in their 256 registers each
                const bool split = mode == 3 && waves(o0[i]) <= (s->fp32 ? 4 : 2);
                const int b = 64 * waves(o0[i]) * (split ? 2 : 1);
                s->layout[mode].push_back({b, grid_of(b, split, j - i), (int)s->hbund.size(), j - i, split});
                for (int k = i; k < j; ++k) {
                    Bundle bd = empty_bundle();
                    bundle_add(bd, o0[k], 0, waves(o0[k]));
                    s->hbund.push_back(bd);
                }
                i = j;
            }
        }
        const int bmax = std::min(512, (maxl + 63) / 64 * 64);
        s->layout[1].push_back({bmax, grid_of(bmax, false, nstrands), (int)s->hbund.size(), nstrands, false});
        for (int k = 0; k < nstrands; ++k) {
            Bundle bd = empty_bundle();
            bundle_add(bd, o1[k], 0, waves(o1[k]));
            s->hbund.push_back(bd);
        }
        {
            std::vector<Bundle> pk = pack_bundles(H);
            s->layout[2].push_back({512, grid_of(512, false, (int)pk.size()), (int)s->hbund.size(), (int)pk.size(),
                                    false});
            s->hbund.insert(s->hbund.end(), pk.begin(), pk.end());
        }
        set_kmax(s->hbund, H);
        s->block = bmax;
        std::vector<uint8_t> cl(ncells);
        for (int64_t i = 0; i < ncells; ++i) cl[i] = (uint8_t)cell_class(celltype[i]);
        std::vector<double> co(ncells, 0.154);
        if (coeff) co.assign(coeff, coeff + ncells);
        std::vector<double> v0(nparams);
        for (int i = 0; i < nparams; ++i) v0[i] = params[i].v0;
        lqt_status e;
        if ((e = upload(&s->ddesc, s->hdev.data(), nstrands))) return e;
        if (s->s1r_h.empty()) s->s1r_h.assign(2, 0);  // a valid pointer even with no S1
        if ((e = upload(&s->s1r, s->s1r_h.data(), s->s1r_h.size()))) return e;
        if ((e = upload(&s->bund, s->hbund.data(), s->hbund.size()))) return e;  // all layouts' bundles
        if ((e = upload(&s->queue, (const int*)nullptr, 8))) return e;
        if ((e = upload(&s->cls, cl.data(), ncells))) return e;
        if ((e = upload(&s->coeff, co.data(), ncells))) return e;
        if ((e = upload(&s->v0, v0.data(), nparams))) return e;
        std::vector<Consts<double>> cd(nparams);
        for (int i = 0; i < nparams; ++i) cd[i] = derive_consts(params[i]);
        if (s->fp32) {
            std::vector<Consts<float>> cf(nparams);
            for (int i = 0; i < nparams; ++i) cf[i] = to_float(cd[i]);
            if ((e = upload((Consts<float>**)&s->consts, cf.data(), nparams))) return e;
        } else {
            if ((e = upload((Consts<double>**)&s->consts, cd.data(), nparams))) return e;
        }
        LQT_HIP(hipMalloc((void**)&s->phi, sizeof(float) * std::max<int64_t>(1, s->phi_total)));
        LQT_HIP(hipMalloc((void**)&s->nphi, sizeof(int64_t) * nstrands));
        LQT_HIP(hipMalloc((void**)&s->phit0, sizeof(double) * nstrands));
        LQT_HIP(hipMalloc((void**)&s->vrec, sizeof(double) * std::max<int64_t>(1, s->vrec_total)));
        LQT_HIP(hipMalloc((void**)&s->nvrec, sizeof(int64_t) * nstrands));
        LQT_HIP(hipMalloc((void**)&s->vrect0, sizeof(double) * nstrands));
        LQT_HIP(hipMalloc((void**)&s->steps, sizeof(int64_t) * nstrands));
        LQT_HIP(hipMalloc((void**)&s->vmax, sizeof(double) * 2 * nstrands));
        LQT_HIP(hipMalloc((void**)&s->act, sizeof(double) * s->pitch));
        LQT_HIP(hipMalloc((void**)&s->apd, sizeof(double) * s->pitch));
        LQT_HIP(hipMalloc((void**)&s->ead, sizeof(int) * s->pitch));
        LQT_HIP(hipMalloc(&s->fr, rsz(s) * NR * s->pitch));
        LQT_HIP(hipMalloc((void**)&s->fd, sizeof(double) * ND * s->pitch));
        LQT_HIP(hipMalloc(&s->fv, rsz(s) * s->pitch));
        return LQT_OK;
    }();
    if (st) {
        lqt_strands_destroy(s);
        return st;
    }
    *out = s;
    return LQT_OK;
}

lqt_status lqt_strands_destroy(lqt_strands* s) {
    if (!s) return LQT_OK;
    if (s->pending && s->pending_stream) hipStreamSynchronize(s->pending_stream);
    if (s->stream) hipStreamSynchronize(s->stream);
    for (int m = 0; m < NMODES; ++m) {
        if (s->exec_m[m]) hipGraphExecDestroy(s->exec_m[m]);
        if (s->graph_m[m]) hipGraphDestroy(s->graph_m[m]);
    }
    for (int i = 0; i < 8; ++i) {
        if (s->side[i]) hipStreamDestroy(s->side[i]);
        if (s->join[i]) hipEventDestroy(s->join[i]);
    }
    if (s->fork) hipEventDestroy(s->fork);
    hipFree(s->ddesc);
    hipFree(s->s1r);
    hipFree(s->bund);
    hipFree(s->queue);
    hipFree(s->cls);
    hipFree(s->coeff);
    hipFree(s->consts);
    hipFree(s->v0);
    hipFree(s->phi);
    hipFree(s->nphi);
    hipFree(s->phit0);
    hipFree(s->vrec);
    hipFree(s->nvrec);
    hipFree(s->vrect0);
    hipFree(s->act);
    hipFree(s->apd);
    hipFree(s->ead);
    hipFree(s->fr);
    hipFree(s->fd);
    hipFree(s->fv);
    hipFree(s->steps);
    hipFree(s->vmax);
    if (s->e0) hipEventDestroy(s->e0);
    if (s->e1) hipEventDestroy(s->e1);
    if (s->stream) hipStreamDestroy(s->stream);
    delete s;
    return LQT_OK;
}

// Results of the enqueued run to the host (waits for its stream first).
static lqt_status fetch(lqt_strands* s) {
    if (!s->pending) return LQT_OK;
    s->pending = false;
    LQT_HIP(hipStreamSynchronize(s->pending_stream));
    float ms = 0;
    LQT_HIP(hipEventElapsedTime(&ms, s->e0, s->e1));
    s->last_ms = ms;
    s->h_act.resize(s->ncells);
    s->h_apd.resize(s->ncells);
    s->h_ead.resize(s->ncells);
    s->h_nphi.resize(s->n);
    s->h_steps.resize(s->n);
    s->h_phit0.resize(s->n);
    s->h_vrect0.resize(s->n);
    s->h_nvrec.resize(s->n);
    s->h_phi.resize(std::max<int64_t>(1, s->phi_total));
    s->h_vmax.resize(2 * (size_t)s->n);
    LQT_HIP(hipMemcpy(s->h_vmax.data(), s->vmax, sizeof(double) * 2 * s->n, hipMemcpyDeviceToHost));
    LQT_HIP(hipMemcpy(s->h_act.data(), s->act, sizeof(double) * s->ncells, hipMemcpyDeviceToHost));
    LQT_HIP(hipMemcpy(s->h_apd.data(), s->apd, sizeof(double) * s->ncells, hipMemcpyDeviceToHost));
    LQT_HIP(hipMemcpy(s->h_ead.data(), s->ead, sizeof(int) * s->ncells, hipMemcpyDeviceToHost));
    LQT_HIP(hipMemcpy(s->h_nphi.data(), s->nphi, sizeof(int64_t) * s->n, hipMemcpyDeviceToHost));
    LQT_HIP(hipMemcpy(s->h_steps.data(), s->steps, sizeof(int64_t) * s->n, hipMemcpyDeviceToHost));
    LQT_HIP(hipMemcpy(s->h_phit0.data(), s->phit0, sizeof(double) * s->n, hipMemcpyDeviceToHost));
    LQT_HIP(hipMemcpy(s->h_vrect0.data(), s->vrect0, sizeof(double) * s->n, hipMemcpyDeviceToHost));
    LQT_HIP(hipMemcpy(s->h_nvrec.data(), s->nvrec, sizeof(int64_t) * s->n, hipMemcpyDeviceToHost));
    if (s->phi_total > 0)
        LQT_HIP(hipMemcpy(s->h_phi.data(), s->phi, sizeof(float) * s->phi_total, hipMemcpyDeviceToHost));
    s->ran = true;
    return LQT_OK;
}

lqt_status lqt_strands_run_async(lqt_strands* s, void* stream) {
    LQT_CHECK(s, LQT_ERR_ARG, "lqt_strands_run_async: null handle");
    LQT_CHECK(!s->pending, LQT_ERR_STATE, "lqt_strands_run_async: the previous run's results were not read yet");
    hipStream_t st = stream ? (hipStream_t)stream : s->stream;
    lqt_status e = enqueue_mode(s, s->mode, st);
    if (e) return e;
    s->pending = true;
    s->pending_stream = st;
    s->ran = false;
    return LQT_OK;
}

lqt_status lqt_strands_run(lqt_strands* s) {
    LQT_CHECK(s, LQT_ERR_ARG, "lqt_strands_run: null handle");
    if (s->pending) {  // an earlier asynchronous run: finish it first
        lqt_status e = fetch(s);
        if (e) return e;
    }
    lqt_status e = lqt_strands_run_async(s, nullptr);
    if (e) return e;
    return fetch(s);
}

lqt_status lqt_strands_results(lqt_strands* s, lqt_strand_result* out) {
    LQT_CHECK(s && out, LQT_ERR_ARG, "lqt_strands_results: bad argument");
    {
        lqt_status e_ = fetch(s);
        if (e_) return e_;
    }
    LQT_CHECK(s->ran, LQT_ERR_STATE, "lqt_strands_results: call lqt_strands_run first");
    for (int i = 0; i < s->n; ++i) {
        const lqt_strand_desc& d = s->desc[i];
        lqt_strand_result r;
        memset(&r, 0, sizeof(r));
        double sum[3] = {0, 0, 0};
        int cnt[3] = {0, 0, 0};
        double af = 1e300, al = -1e300;
        int ead = 0;
        for (int j = 0; j < d.ncell; ++j) {
            const int64_t c = (int64_t)d.cell_offset + j;
            const int cls = s->label[c] == LQT_MCELL ? 1 : (s->label[c] == LQT_EPI ? 2 : 0);
            if (s->h_apd[c] >= 0) {
                sum[cls] += s->h_apd[c];
                cnt[cls] += 1;
            }
            if (s->h_act[c] >= 0) {
                af = std::min(af, s->h_act[c]);
                al = std::max(al, s->h_act[c]);
            }
            ead |= s->h_ead[c] ? 1 : 0;
        }
        for (int k = 0; k < 3; ++k) r.apd90[k] = cnt[k] ? sum[k] / cnt[k] : -1.0;
        r.ead = ead;
        r.act_first = af < 1e300 ? af : -1.0;
        r.act_last = al > -1e300 ? al : -1.0;
        // CV between cells at 10 % and 90 % of the strand (cells 10 and 90
        // of OneD.cc:422-434 for 100 cells)
        const int i10 = d.ncell / 10, i90 = (9 * d.ncell) / 10;
        const double a10 = s->h_act[d.cell_offset + i10], a90 = s->h_act[d.cell_offset + i90];
        r.cv = (a10 >= 0 && a90 > a10) ? d.dx * (i90 - i10) / (a90 - a10) : -1.0;
        const double t_q = d.measure_beat * d.bcl + d.stim_start;
        const double t_qrs_end = (r.act_last >= 0 ? r.act_last : t_q) + 10.0;
        const double t_end = (d.measure_beat + 1) * d.bcl;
        r.qt = qt_from_ecg(s->h_phi.data() + s->hdev[i].phi_off, s->h_nphi[i], s->h_phit0[i], d.rec_every * s->dt,
                           t_q, t_qrs_end, t_end);
        r.steps = s->h_steps[i];
        r.vmax_first = s->h_vmax[2 * (size_t)i];
        r.vmax_last = s->h_vmax[2 * (size_t)i + 1];
        out[i] = r;
    }
    return LQT_OK;
}

lqt_status lqt_strands_cell_results(lqt_strands* s, int i, double* act, double* apd90, int* ead) {
    LQT_CHECK(s && i >= 0 && i < s->n, LQT_ERR_ARG, "lqt_strands_cell_results: bad argument");
    {
        lqt_status e_ = fetch(s);
        if (e_) return e_;
    }
    LQT_CHECK(s->ran, LQT_ERR_STATE, "lqt_strands_cell_results: call lqt_strands_run first");
    const lqt_strand_desc& d = s->desc[i];
    for (int j = 0; j < d.ncell; ++j) {
        const int64_t c = (int64_t)d.cell_offset + j;
        if (act) act[j] = s->h_act[c];
        if (apd90) apd90[j] = s->h_apd[c];
        if (ead) ead[j] = s->h_ead[c];
    }
    return LQT_OK;
}

lqt_status lqt_strands_ecg(lqt_strands* s, int i, double* phi, double* t0, int64_t* n) {
    LQT_CHECK(s && i >= 0 && i < s->n && n, LQT_ERR_ARG, "lqt_strands_ecg: bad argument");
    {
        lqt_status e_ = fetch(s);
        if (e_) return e_;
    }
    LQT_CHECK(s->ran, LQT_ERR_STATE, "lqt_strands_ecg: call lqt_strands_run first");
    *n = s->h_nphi[i];
    if (t0) *t0 = s->h_phit0[i];
    if (phi)
        for (int64_t j = 0; j < s->h_nphi[i]; ++j) phi[j] = s->h_phi[s->hdev[i].phi_off + j];
    return LQT_OK;
}

lqt_status lqt_strands_vtrace(lqt_strands* s, int i, double* v, double* t0, int64_t* nrows) {
    LQT_CHECK(s && i >= 0 && i < s->n && nrows, LQT_ERR_ARG, "lqt_strands_vtrace: bad argument");
    {
        lqt_status e_ = fetch(s);
        if (e_) return e_;
    }
    LQT_CHECK(s->ran, LQT_ERR_STATE, "lqt_strands_vtrace: call lqt_strands_run first");
    *nrows = s->h_nvrec[i];
    if (t0) *t0 = s->h_vrect0[i];
    if (v && s->h_nvrec[i] > 0) {
        const int64_t cnt = s->h_nvrec[i] * s->desc[i].ncell;
        LQT_HIP(hipMemcpy(v, s->vrec + s->hdev[i].vrec_off, sizeof(double) * cnt, hipMemcpyDeviceToHost));
    }
    return LQT_OK;
}

lqt_status lqt_strands_final_state(lqt_strands* s, int i, double* out) {
    LQT_CHECK(s && out && i >= 0 && i < s->n, LQT_ERR_ARG, "lqt_strands_final_state: bad argument");
    {
        lqt_status e_ = fetch(s);
        if (e_) return e_;
    }
    LQT_CHECK(s->ran, LQT_ERR_STATE, "lqt_strands_final_state: call lqt_strands_run first");
    const lqt_strand_desc& d = s->desc[i];
    const int64_t c0 = d.cell_offset, n = d.ncell;
    const size_t rs = rsz(s);
    // the strand's cells are contiguous in every slot: one copy per slot
    std::vector<unsigned char> buf(8 * (size_t)n);
    for (int m = 0; m < 40; ++m) {
        const Ref40 r = kRef40[m];
        if (r.kind == 0 || r.kind == 1) {
            const unsigned char* src = r.kind == 0 ? (const unsigned char*)s->fv + rs * c0
                                                   : (const unsigned char*)s->fr + rs * ((int64_t)r.slot * s->pitch + c0);
            LQT_HIP(hipMemcpy(buf.data(), src, rs * n, hipMemcpyDeviceToHost));
            for (int64_t j = 0; j < n; ++j)
                out[(size_t)j * 40 + m] = s->fp32 ? (double)((const float*)buf.data())[j] : ((const double*)buf.data())[j];
        } else if (r.kind == 2) {
            LQT_HIP(hipMemcpy(buf.data(), s->fd + (int64_t)r.slot * s->pitch + c0, sizeof(double) * n,
                              hipMemcpyDeviceToHost));
            for (int64_t j = 0; j < n; ++j) out[(size_t)j * 40 + m] = ((const double*)buf.data())[j];
        } else {
            for (int64_t j = 0; j < n; ++j) out[(size_t)j * 40 + m] = 1.0;  // hsp, jp: never updated
        }
    }
    return LQT_OK;
}

lqt_status lqt_strands_last_ms(lqt_strands* s, double* ms) {
    LQT_CHECK(s && ms, LQT_ERR_ARG, "lqt_strands_last_ms: bad argument");
    lqt_status e = fetch(s);
    if (e) return e;
    *ms = s->last_ms;
    return LQT_OK;
}

lqt_status lqt_ecg_qt(const double* phi, int64_t n, double t0, double dts, double t_q, double t_qrs_end,
                      double t_end, double* qt) {
    LQT_CHECK(phi && qt && n >= 0 && dts > 0, LQT_ERR_ARG, "lqt_ecg_qt: bad argument");
    *qt = qt_from_ecg(phi, n, t0, dts, t_q, t_qrs_end, t_end);
    return LQT_OK;
}

lqt_status lqt_strands_block(const lqt_strands* s, int* block) {
    LQT_CHECK(s && block, LQT_ERR_ARG, "lqt_strands_block: bad argument");
    int b = 0;
    for (const auto& L : s->layout[s->mode]) b = std::max(b, L.block);
    *block = b;
    return LQT_OK;
}

lqt_status lqt_strands_autotune(lqt_strands* s, int trial_steps, int* mode, double* ms) {
    LQT_CHECK(s && trial_steps > 0, LQT_ERR_ARG, "lqt_strands_autotune: bad argument");
    {
        lqt_status e_ = fetch(s);  // an asynchronous run still in flight finishes first
        if (e_) return e_;
    }
    // The batch with every stop time scaled by the same factor, so that the
    // longest strand runs trial_steps steps and the strands keep their
    // relative lengths (the order in which they finish decides the layout),
    // timed in every layout.
    double smax = 0.0;
    for (const auto& d : s->hdev) smax = std::max(smax, d.stop_time);
    const double f = std::min(1.0, trial_steps * s->dt / std::max(smax, s->dt));
    std::vector<StrandDev> cut = s->hdev;
    std::vector<double> stops(s->n);
    for (int i = 0; i < s->n; ++i) stops[i] = cut[i].stop_time * f;
    const std::vector<int64_t> ns = step_counts(stops, s->dt);
    for (int i = 0; i < s->n; ++i) {
        cut[i].stop_time = stops[i];
        cut[i].nsteps = ns[i];
    }
    std::vector<Bundle> cb = s->hbund;
    set_kmax(cb, cut);
    LQT_HIP(hipMemcpy(s->ddesc, cut.data(), sizeof(StrandDev) * s->n, hipMemcpyHostToDevice));
    LQT_HIP(hipMemcpy(s->bund, cb.data(), sizeof(Bundle) * cb.size(), hipMemcpyHostToDevice));
    double t[NMODES] = {0, 0, 0, 0};
    lqt_status e = LQT_OK;
    for (int m = 0; m < NMODES && !e; ++m) {
        double w = 0;
        e = run_mode(s, m, &w);  // first replay: warm
        if (!e) e = run_mode(s, m, &t[m]);
    }
    LQT_HIP(hipMemcpy(s->ddesc, s->hdev.data(), sizeof(StrandDev) * s->n, hipMemcpyHostToDevice));
    LQT_HIP(hipMemcpy(s->bund, s->hbund.data(), sizeof(Bundle) * s->hbund.size(), hipMemcpyHostToDevice));
    if (e) return e;
    int best = 0;
    for (int m = 1; m < NMODES; ++m)
        if (t[m] < t[best]) best = m;
    s->mode = best;
    for (int m = 0; m < NMODES; ++m) s->tune_ms[m] = t[m];
    s->tuned = true;
    s->ran = false;
    if (mode) *mode = s->mode;
    if (ms)
        for (int m = 0; m < NMODES; ++m) ms[m] = t[m];
    return LQT_OK;
}

lqt_status lqt_strands_set_mode(lqt_strands* s, int mode) {
    LQT_CHECK(s && mode >= 0 && mode < NMODES, LQT_ERR_ARG, "lqt_strands_set_mode: mode must be 0, 1, 2 or 3");
    s->mode = mode;
    return LQT_OK;
}

}  // extern "C"
