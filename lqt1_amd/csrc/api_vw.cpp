// C ABI: vulnerable-window search of OneD_VW.cc:199-449 with every S2
// timing a strand of one batched launch.
//
// The reference re-runs the whole 5-beat strand for one S2 time after
// another and moves the S2 time by its search rule.  Every candidate run is
// independent of the others, so the candidates a search may visit are run
// together and the search rule is then replayed over their results:
//  * modes 2 and 3 (step right / left by `step`): the next `batch` times of
//    the sequence left, left +/- step, ... (accumulated in double as the
//    reference accumulates them) in one launch; the first vulnerable one
//    ends the search;
//  * mode 1 (bisection): the whole bisection tree below the current
//    interval, `depth` levels at a time (2^depth - 1 strands), walked with
//    the reference's rule afterwards.
// The visited sequence (S2 time, maxEndo, maxEpi) is the one the
// reference's loop would print.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "common.hpp"

namespace {

struct Cand {
    double s2, vmax_first, vmax_last;
};

// One launch: a copy of the base strand per S2 time.
lqt_status run_batch(const lqt_strand_desc* base, const uint8_t* celltype, const double* coeff,
                     const lqt_params* params, lqt_precision prec, double dt, const std::vector<double>& s2,
                     std::vector<Cand>& out, double* ms) {
    const int n = (int)s2.size();
    const int L = base->ncell;
    std::vector<lqt_strand_desc> d((size_t)n, *base);
    std::vector<uint8_t> ct((size_t)n * L);
    std::vector<double> co((size_t)n * L);
    for (int i = 0; i < n; ++i) {
        d[i].cell_offset = i * L;
        d[i].param_index = 0;
        d[i].s2_start = s2[i];
        d[i].rec_every = 0;
        d[i].vrec_every = 0;
        memcpy(&ct[(size_t)i * L], celltype + base->cell_offset, L);
        for (int j = 0; j < L; ++j) co[(size_t)i * L + j] = coeff ? coeff[base->cell_offset + j] : 0.154;
    }
    lqt_strands* h = nullptr;
    lqt_status e = lqt_strands_create(n, d.data(), ct.data(), co.data(), params, 1, prec, dt, &h);
    if (e) return e;
    e = lqt_strands_run(h);
    std::vector<lqt_strand_result> r((size_t)n);
    if (!e) e = lqt_strands_results(h, r.data());
    double t = 0;
    if (!e) e = lqt_strands_last_ms(h, &t);
    lqt_strands_destroy(h);
    if (e) return e;
    *ms += t;
    out.resize(n);
    for (int i = 0; i < n; ++i) out[i] = {s2[i], r[i].vmax_first, r[i].vmax_last};
    return LQT_OK;
}

// OneD_VW.cc:410: unidirectional block, one end above -boundary and the
// other not.
inline bool vulnerable(const Cand& c, double b) { return (c.vmax_first + b) * (c.vmax_last + b) <= 0; }

}  // namespace

extern "C" {

lqt_status lqt_vw_search(const lqt_strand_desc* base, const uint8_t* celltype, const double* coeff,
                         const lqt_params* params, lqt_precision prec, double dt, const lqt_vw_desc* vw,
                         lqt_vw_result* out, double* visited_s2, double* visited_vmax, int64_t cap) {
    LQT_CHECK(base && celltype && params && vw && out && dt > 0, LQT_ERR_ARG, "lqt_vw_search: bad argument");
    LQT_CHECK(vw->mode >= 1 && vw->mode <= 3, LQT_ERR_ARG, "lqt_vw_search: mode must be 1, 2 or 3");
    LQT_CHECK(vw->step > 0 && vw->batch >= 1 && vw->max_candidates >= 1, LQT_ERR_ARG,
              "lqt_vw_search: step, batch and max_candidates must be positive");
    LQT_CHECK(base->s2_duration > 0, LQT_ERR_ARG, "lqt_vw_search: the base strand needs an S2 (s2_duration > 0)");
    LQT_CHECK(vw->mode != 1 || vw->left <= vw->right, LQT_ERR_ARG, "lqt_vw_search: bisection needs left <= right");
    memset(out, 0, sizeof(*out));
    std::vector<Cand> visited;
    const double b = vw->boundary;
    bool found = false;
    lqt_status e = LQT_OK;
    double ms = 0;

    if (vw->mode == 2 || vw->mode == 3) {
        // OneD_VW.cc:405-423 / 427-443: s2 = left, then left +/- step
        double next = vw->left;
        while (!found && (int64_t)visited.size() < vw->max_candidates) {
            const int64_t room = vw->max_candidates - (int64_t)visited.size();
            const int nb = (int)std::min<int64_t>(vw->batch, room);
            std::vector<double> s2((size_t)nb);
            for (int i = 0; i < nb; ++i) {
                s2[i] = next;
                next = vw->mode == 2 ? next + vw->step : next - vw->step;
            }
            std::vector<Cand> res;
            if ((e = run_batch(base, celltype, coeff, params, prec, dt, s2, res, &ms))) return e;
            out->launches += 1;
            out->simulated += nb;
            for (const Cand& c : res) {
                visited.push_back(c);
                if (vulnerable(c, b)) {
                    found = true;
                    break;
                }
            }
        }
    } else {
        // OneD_VW.cc:396-404 and the loop test of :214: bisection while
        // right - left >= step or right == left
        double left = vw->left, right = vw->right;
        const int depth = std::max(1, std::min(12, (int)floor(log2((double)vw->batch + 1.0))));
        bool done = false;
        while (!done && (int64_t)visited.size() < vw->max_candidates) {
            // the tree below (left, right): node k has children 2k+1 (the
            // S2 was too early: left moves up) and 2k+2 (too late: right
            // moves down); nodes whose interval fails the loop test are
            // not run
            const int nn = (1 << depth) - 1;
            std::vector<double> L((size_t)nn), R((size_t)nn);
            std::vector<char> live((size_t)nn, 0);
            std::vector<int> slot((size_t)nn, -1);
            std::vector<double> s2;
            L[0] = left;
            R[0] = right;
            live[0] = 1;
            for (int k = 0; k < nn; ++k) {
                if (!live[k]) continue;
                if (!(R[k] - L[k] >= vw->step || R[k] == L[k])) {
                    live[k] = 0;
                    continue;
                }
                const double m = 0.5 * (L[k] + R[k]);
                slot[k] = (int)s2.size();
                s2.push_back(m);
                if (2 * k + 2 < nn) {
                    L[2 * k + 1] = m; R[2 * k + 1] = R[k]; live[2 * k + 1] = 1;
                    L[2 * k + 2] = L[k]; R[2 * k + 2] = m; live[2 * k + 2] = 1;
                }
            }
            if (s2.empty()) break;  // the interval is below the resolution: no VW
            std::vector<Cand> res;
            if ((e = run_batch(base, celltype, coeff, params, prec, dt, s2, res, &ms))) return e;
            out->launches += 1;
            out->simulated += (int64_t)s2.size();
            int k = 0;
            for (;;) {
                if (k >= nn) {
                    // below the evaluated tree: continue from this interval
                    break;
                }
                if (slot[k] < 0) {  // loop test failed: no VW in range
                    done = true;
                    break;
                }
                const Cand& c = res[slot[k]];
                visited.push_back(c);
                const double m = c.s2;
                if (c.vmax_first + b < 0 && c.vmax_last + b < 0) {
                    left = m;
                    k = 2 * k + 1;
                } else if (c.vmax_first + b > 0 && c.vmax_last + b > 0) {
                    right = m;
                    k = 2 * k + 2;
                } else {
                    found = true;
                    done = true;
                    break;
                }
                if ((int64_t)visited.size() >= vw->max_candidates) {
                    done = true;
                    break;
                }
            }
        }
    }
    out->found = found ? 1 : 0;
    out->visited = (int64_t)visited.size();
    out->device_ms = ms;
    if (!visited.empty()) {
        const Cand& c = visited.back();
        out->s2 = c.s2;
        out->vmax_first = c.vmax_first;
        out->vmax_last = c.vmax_last;
    }
    for (int64_t i = 0; i < (int64_t)visited.size() && i < cap; ++i) {
        if (visited_s2) visited_s2[i] = visited[i].s2;
        if (visited_vmax) {
            visited_vmax[2 * i] = visited[i].vmax_first;
            visited_vmax[2 * i + 1] = visited[i].vmax_last;
        }
    }
    return LQT_OK;
}

}  // extern "C"
