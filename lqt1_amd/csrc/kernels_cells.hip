// Independent-cell kernels: one TPORd cell per lane, state kept in
// registers across the steps of one launch.
#include "internal.hpp"

namespace lqt {

// nsteps updates with fixed Istim / dVgap_dt per cell (the Cell interface
// of SingleCell/Cell.cc:22-42: setIstim, setDVgap_dt, update).
template <class R>
__global__ __launch_bounds__(256) void cells_step_kernel(int n, SoA<R> st, R* __restrict__ v,
                                                         const uint8_t* __restrict__ cls,
                                                         const int* __restrict__ pidx,
                                                         const Consts<R>* __restrict__ consts, int nsteps,
                                                         R dt, const double* __restrict__ ist,
                                                         const double* __restrict__ dvgap) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    CellState<R> s;
    load_state(s, st, k);
    const int c = cls[k];
    const Consts<R> cc = consts[pidx ? pidx[k] : 0];
    const R I = ist ? R(ist[k]) : R(0);
    const R G = dvgap ? R(dvgap[k]) : R(0);
    R vk = v[k];
    for (int i = 0; i < nsteps; ++i) vk = tpord_update<R, false>(s, vk, c, cc, dt, I, G, nullptr);
    v[k] = vk;
    store_state(s, st, k);
}

// The pacing protocol of SingleCell.cc:95-215 for every cell: stimulus in
// the host-evaluated step ranges of the in-cycle time window, then the
// last-beat statistics (rest, dV/dt max, overshoot, APDx).
template <class R>
__global__ __launch_bounds__(256) void cells_pace_kernel(int n, SoA<R> st, R* __restrict__ v,
                                                         const uint8_t* __restrict__ cls,
                                                         const int* __restrict__ pidx,
                                                         const Consts<R>* __restrict__ consts, R dt,
                                                         PaceTrack* __restrict__ track, PaceArgs a) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    CellState<R> s;
    load_state(s, st, k);
    const int c = cls[k];
    const Consts<R> cc = consts[pidx ? pidx[k] : 0];
    PaceTrack tr = track[k];
    R vk = v[k];
    double time = a.time0;
    const double ddt = a.dt;
    static const double frac[6] = {0.20, 0.25, 0.50, 0.75, 0.80, 0.90};
    int r = 0;
    int tnext = a.trace_next;
    int64_t tslot = a.trace_slot;
    for (int i = 0; i < a.nsteps; ++i, time += ddt) {
        while (r < a.nstim && i >= a.stim_off[r]) ++r;
        const R Ist = (r < a.nstim && i >= a.stim_on[r]) ? R(a.stim_strength) : R(0);
        Diag dg;
        vk = tpord_update<R, true>(s, vk, c, cc, dt, Ist, R(0), &dg);
        const double V = (double)vk;
        if (i == tnext) {
            if (a.trace && k == a.trace_cell && tslot < a.trace_cap) {
                a.trace[3 * tslot] = V;
                a.trace[3 * tslot + 1] = dg.ICaL;
                a.trace[3 * tslot + 2] = dg.IKs;
            }
            tnext += a.trace_every;
            ++tslot;
        }
        if (i == a.rest_step) {
            tr.rest = V;
            tr.oldV = V;
            tr.peakfound = 0;
            tr.dvdt_max = -1;
        }
        if (i >= a.last0 && i < a.last1) {
            if (V > tr.oldV) {
                tr.oldV = V;
                const double absdvdt = fabs(dg.dV_dt);
                if (absdvdt > tr.dvdt_max) tr.dvdt_max = absdvdt;
            } else if (!tr.peakfound) {
                tr.peakfound = 1;
                tr.overshoot = tr.oldV;
                tr.amplitude = tr.overshoot - tr.rest;
                for (int j = 0; j < 6; ++j) tr.repo[j] = tr.overshoot - frac[j] * tr.amplitude;
                tr.oldV = V;
            } else {
                for (int j = 0; j < 6; ++j) {
                    if (tr.oldV >= tr.repo[j] && V <= tr.repo[j]) {
                        tr.apd[j] = time - floor(time / a.bcl) * a.bcl - a.stim_start - a.stim_dur;
                        break;
                    }
                }
                tr.oldV = V;
            }
        }
    }
    v[k] = vk;
    store_state(s, st, k);
    track[k] = tr;
}

void launch_cells_step(bool fp32, int n, void* r, double* d, int64_t pitch, void* v, const uint8_t* cls,
                       const int* pidx, const void* consts, int nsteps, double dt, const double* ist,
                       const double* dvgap, hipStream_t st) {
    const int B = 256;
    const int G = (n + B - 1) / B;
    if (fp32) {
        SoA<float> s{(float*)r, d, pitch};
        cells_step_kernel<float><<<G, B, 0, st>>>(n, s, (float*)v, cls, pidx, (const Consts<float>*)consts,
                                                  nsteps, (float)dt, ist, dvgap);
    } else {
        SoA<double> s{(double*)r, d, pitch};
        cells_step_kernel<double><<<G, B, 0, st>>>(n, s, (double*)v, cls, pidx, (const Consts<double>*)consts,
                                                   nsteps, dt, ist, dvgap);
    }
}

void launch_cells_pace(bool fp32, int n, void* r, double* d, int64_t pitch, void* v, const uint8_t* cls,
                       const int* pidx, const void* consts, double dt, PaceTrack* track,
                       const PaceArgs& a, hipStream_t st) {
    const int B = 256;
    const int G = (n + B - 1) / B;
    if (fp32) {
        SoA<float> s{(float*)r, d, pitch};
        cells_pace_kernel<float><<<G, B, 0, st>>>(n, s, (float*)v, cls, pidx, (const Consts<float>*)consts,
                                                  (float)dt, track, a);
    } else {
        SoA<double> s{(double*)r, d, pitch};
        cells_pace_kernel<double><<<G, B, 0, st>>>(n, s, (double*)v, cls, pidx, (const Consts<double>*)consts,
                                                   dt, track, a);
    }
}

}  // namespace lqt
