// Independent-cell kernels: one TPORd cell per lane (or per two or four
// lanes), state kept in registers across the steps of one launch.
//
// Every lane of a wavefront that holds at least one cell computes: lanes
// past the last cell run a copy of one of the wavefront's own cells
// (reading its inputs before the wavefront writes anything, writing
// nothing).  A wavefront whose live lanes all lie in one 16-lane quarter
// steps about 1.3x slower than a full one on gfx950 (measured with the
// pacing protocol: 9.2 against 6.9 us per FP64 step on two lanes per cell,
// 2.55 against 1.95 us FP32, 11.3 against 10.2 us FP64 on one lane, for
// 1-16 cells against 64 or more); the idle lanes cost nothing to fill.
#include "internal.hpp"

namespace lqt {

// The cell that lane k computes (k itself if k < n) for a wavefront that
// holds at least one cell: past the last cell, a cell of the same wavefront.
__device__ __forceinline__ int src_cell(int k, int n) {
    const int w = k & ~63;
    return k < n ? k : w + (k - w) % (n - w);
}

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
    if ((k & ~63) >= n) return;  // a wavefront without cells
    const int q = src_cell(k, n);
    CellState<R> s;
    load_state(s, st, q);
    const int c = cls[q];
    const Consts<R> cc = consts[pidx ? pidx[q] : 0];
    const R I = ist ? R(ist[q]) : R(0);
    const R G = dvgap ? R(dvgap[q]) : R(0);
    R vk = v[q];
    for (int i = 0; i < nsteps; ++i) vk = tpord_update<R, false>(s, vk, c, cc, dt, I, G, nullptr);
    if (k < n) {
        v[k] = vk;
        store_state(s, st, k);
    }
}

// What the pacing protocol records after step i (V = V after it): the
// trace sample of cell k and the last-beat statistics (rest, dV/dt max,
// overshoot, APDx), SingleCell.cc:117-139 and :144-215.
struct PaceState {
    int r;          // stimulus range of the launch that step i is in or before
    int tnext;      // step of the next trace sample
    int64_t tslot;  // its slot
};

template <class R>
__device__ __forceinline__ R pace_istim(PaceState& ps, const PaceArgs& a, int i) {
    while (ps.r < a.nstim && i >= a.stim_off[ps.r]) ++ps.r;
    return (ps.r < a.nstim && i >= a.stim_on[ps.r]) ? R(a.stim_strength) : R(0);
}

__device__ __forceinline__ void pace_record(PaceTrack& tr, PaceState& ps, const PaceArgs& a, int k, int i,
                                            double time, double V, const Diag& dg) {
    if (i == ps.tnext) {
        if (a.trace && k == a.trace_cell && ps.tslot < a.trace_cap) {
            a.trace[3 * ps.tslot] = V;
            a.trace[3 * ps.tslot + 1] = dg.ICaL;
            a.trace[3 * ps.tslot + 2] = dg.IKs;
        }
        ps.tnext += a.trace_every;
        ++ps.tslot;
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
            constexpr double frac[6] = {0.20, 0.25, 0.50, 0.75, 0.80, 0.90};
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
    if ((k & ~63) >= n) return;  // a wavefront without cells
    const int q = src_cell(k, n);
    CellState<R> s;
    load_state(s, st, q);
    const int c = cls[q];
    const Consts<R> cc = consts[pidx ? pidx[q] : 0];
    PaceTrack tr = track[q];
    R vk = v[q];
    double time = a.time0;
    PaceState ps{0, a.trace_next, a.trace_slot};
    for (int i = 0; i < a.nsteps; ++i, time += a.dt) {
        const R Ist = pace_istim<R>(ps, a, i);
        Diag dg;
        vk = tpord_update<R, true>(s, vk, c, cc, dt, Ist, R(0), &dg);
        pace_record(tr, ps, a, k, i, time, (double)vk, dg);  // k >= n is never the trace cell
    }
    if (k < n) {
        v[k] = vk;
        store_state(s, st, k);
        track[k] = tr;
    }
}

// ---------------------------------------------------------------------------
// Split forms: each cell on two lanes of a 128-lane workgroup, lane l of the
// first wavefront advancing the gates of current group A (tp_group_a) of cell
// 64 * blockIdx.x + l, lane l of the second those of group B (tp_group_b).
// The two swap their groups' currents through LDS at a barrier every step
// and then finish the step with the same arithmetic, so both hold the state
// of the single-lane update bit for bit.  The dispatcher puts the two
// wavefronts on different SIMDs, and each issues about 0.6 of the single-lane
// instructions: with few cells, where a step costs one lane's dependent chain
// whatever the number of lanes beside it, the step is that much shorter.
// The first wavefront writes V, the trackers and the trace.

// One step of a split cell; xa / xb are this step's exchange buffers (the
// caller alternates two so that one barrier per step suffices).  Every lane
// of the workgroup calls it.
template <class R, bool DIAG>
__device__ __forceinline__ R split_update(CellState<R>& s, R v, int half, int l, int c, const Consts<R>& cc, R dt,
                                          R Ist, R dVgap, R (*xa)[64], R (*xb)[64], Diag* dg) {
    Pre<R> P;
    GroupA<R> GA;
    GroupB<R> GB;
    if (half == 0) {
        P = tp_pre<R, 1>(s, v, dt);
        GA = tp_group_a(s, v, c, cc, dt, P);
        xa[0][l] = GA.INa; xa[1][l] = GA.INaL; xa[2][l] = GA.Ito; xa[3][l] = GA.IKs; xa[4][l] = GA.INaK;
    } else {
        P = tp_pre<R, 2>(s, v, dt);
        GB = tp_group_b(s, v, c, cc, dt, P);
        xb[0][l] = GB.ICaL; xb[1][l] = GB.ICaNa; xb[2][l] = GB.ICaK; xb[3][l] = GB.IKr;
        xb[4][l] = GB.IK1; xb[5][l] = GB.INaCa_i; xb[6][l] = GB.INaCa_ss; xb[7][l] = GB.IpCa;
        xb[8][l] = GB.ICab; xb[9][l] = GB.Jrel; xb[10][l] = GB.INab; xb[11][l] = GB.IKb;
    }
    __syncthreads();
    if (half == 0) {
        GB.ICaL = xb[0][l]; GB.ICaNa = xb[1][l]; GB.ICaK = xb[2][l]; GB.IKr = xb[3][l];
        GB.IK1 = xb[4][l]; GB.INaCa_i = xb[5][l]; GB.INaCa_ss = xb[6][l]; GB.IpCa = xb[7][l];
        GB.ICab = xb[8][l]; GB.Jrel = xb[9][l]; GB.INab = xb[10][l]; GB.IKb = xb[11][l];
    } else {
        GA.INa = xa[0][l]; GA.INaL = xa[1][l]; GA.Ito = xa[2][l]; GA.IKs = xa[3][l]; GA.INaK = xa[4][l];
    }
    return tp_finish<R, DIAG>(s, v, c, cc, dt, Ist, dVgap, P, GA, GB, dg);
}

template <class R>
__global__ __launch_bounds__(128) void cells_step_split_kernel(int n, SoA<R> st, R* __restrict__ v,
                                                               const uint8_t* __restrict__ cls,
                                                               const int* __restrict__ pidx,
                                                               const Consts<R>* __restrict__ consts, int nsteps,
                                                               R dt, const double* __restrict__ ist,
                                                               const double* __restrict__ dvgap) {
    __shared__ R xa[2][5][64];
    __shared__ R xb[2][12][64];
    const int half = threadIdx.x >= 64u;
    const int l = threadIdx.x & 63;
    const int k = blockIdx.x * 64 + l;  // the grid holds ceil(n / 64) workgroups
    const int q = src_cell(k, n);
    CellState<R> s;
    load_state(s, st, q);
    const int c = cls[q];
    const Consts<R> cc = consts[pidx ? pidx[q] : 0];
    const R I = ist ? R(ist[q]) : R(0);
    const R G = dvgap ? R(dvgap[q]) : R(0);
    R vk = v[q];
    for (int i = 0; i < nsteps; ++i) {
        const int b = i & 1;
        vk = split_update<R, false>(s, vk, half, l, c, cc, dt, I, G, xa[b], xb[b], nullptr);
    }
    if (k < n) {
        if (half == 0) v[k] = vk;
        store_state_half(s, st, k, half);
    }
}

template <class R>
__global__ __launch_bounds__(128) void cells_pace_split_kernel(int n, SoA<R> st, R* __restrict__ v,
                                                               const uint8_t* __restrict__ cls,
                                                               const int* __restrict__ pidx,
                                                               const Consts<R>* __restrict__ consts, R dt,
                                                               PaceTrack* __restrict__ track, PaceArgs a) {
    __shared__ R xa[2][5][64];
    __shared__ R xb[2][12][64];
    const int half = threadIdx.x >= 64u;
    const int l = threadIdx.x & 63;
    const int k = blockIdx.x * 64 + l;  // the grid holds ceil(n / 64) workgroups
    const int q = src_cell(k, n);
    CellState<R> s;
    load_state(s, st, q);
    const int c = cls[q];
    const Consts<R> cc = consts[pidx ? pidx[q] : 0];
    PaceTrack tr = track[q];
    R vk = v[q];
    double time = a.time0;
    PaceState ps{0, a.trace_next, a.trace_slot};
    for (int i = 0; i < a.nsteps; ++i, time += a.dt) {
        const R Ist = pace_istim<R>(ps, a, i);
        Diag dg;
        vk = split_update<R, true>(s, vk, half, l, c, cc, dt, Ist, R(0), xa[i & 1], xb[i & 1], &dg);
        if (half == 0) pace_record(tr, ps, a, k, i, time, (double)vk, dg);  // k >= n is never the trace cell
    }
    if (k < n) {
        if (half == 0) {
            v[k] = vk;
            track[k] = tr;
        }
        store_state_half(s, st, k, half);
    }
}

// ---------------------------------------------------------------------------
// Four-lane forms: each cell on four lanes of a 256-lane workgroup, one per
// wavefront, lane l of wavefront q advancing part (q & 1) + 1 of current
// group A (q < 2) or B (q >= 2) of cell 64 * blockIdx.x + l (tp_group_a /
// tp_group_b with PART 1 or 2); the four swap their parts' currents through
// LDS at one barrier per step and finish the step with the same arithmetic,
// so all four hold the state of the single-lane update bit for bit.  With
// one such workgroup per compute unit each wavefront has a SIMD of its own.
// Wavefront 0 writes V, the trackers and the trace.

template <class R>
struct QuadX {  // one step's currents of the 64 cells of a workgroup
    R a1[3][64];  // INa, INaL, IKs
    R a2[2][64];  // Ito, INaK
    R b1[4][64];  // ICaL, ICaNa, ICaK, Jrel
    R b2[8][64];  // IKr, IK1, INaCa_i, INaCa_ss, IpCa, ICab, INab, IKb
};

// One step of a cell on four lanes; X is this step's exchange buffer (the
// caller alternates two).  Every lane of the workgroup calls it.
template <class R, bool DIAG>
__device__ __forceinline__ R quad_update(CellState<R>& s, R v, int q, int l, int c, const Consts<R>& cc, R dt,
                                         R Ist, R dVgap, QuadX<R>& X, Diag* dg) {
    Pre<R> P;
    if (q == 0) {
        P = tp_pre<R, 1>(s, v, dt);
        const GroupA<R> g = tp_group_a<R, 1>(s, v, c, cc, dt, P);
        X.a1[0][l] = g.INa; X.a1[1][l] = g.INaL; X.a1[2][l] = g.IKs;
    } else if (q == 1) {
        P = tp_pre<R, 1>(s, v, dt);
        const GroupA<R> g = tp_group_a<R, 2>(s, v, c, cc, dt, P);
        X.a2[0][l] = g.Ito; X.a2[1][l] = g.INaK;
    } else if (q == 2) {
        P = tp_pre<R, 2>(s, v, dt);
        const GroupB<R> g = tp_group_b<R, 1>(s, v, c, cc, dt, P);
        X.b1[0][l] = g.ICaL; X.b1[1][l] = g.ICaNa; X.b1[2][l] = g.ICaK; X.b1[3][l] = g.Jrel;
    } else {
        P = tp_pre<R, 2>(s, v, dt);
        const GroupB<R> g = tp_group_b<R, 2>(s, v, c, cc, dt, P);
        X.b2[0][l] = g.IKr; X.b2[1][l] = g.IK1; X.b2[2][l] = g.INaCa_i; X.b2[3][l] = g.INaCa_ss;
        X.b2[4][l] = g.IpCa; X.b2[5][l] = g.ICab; X.b2[6][l] = g.INab; X.b2[7][l] = g.IKb;
    }
    __syncthreads();
    GroupA<R> GA;
    GroupB<R> GB;
    GA.INa = X.a1[0][l]; GA.INaL = X.a1[1][l]; GA.IKs = X.a1[2][l]; GA.Ito = X.a2[0][l]; GA.INaK = X.a2[1][l];
    GB.ICaL = X.b1[0][l]; GB.ICaNa = X.b1[1][l]; GB.ICaK = X.b1[2][l]; GB.Jrel = X.b1[3][l];
    GB.IKr = X.b2[0][l]; GB.IK1 = X.b2[1][l]; GB.INaCa_i = X.b2[2][l]; GB.INaCa_ss = X.b2[3][l];
    GB.IpCa = X.b2[4][l]; GB.ICab = X.b2[5][l]; GB.INab = X.b2[6][l]; GB.IKb = X.b2[7][l];
    return tp_finish<R, DIAG>(s, v, c, cc, dt, Ist, dVgap, P, GA, GB, dg);
}

template <class R>
__global__ __launch_bounds__(256) void cells_step_quad_kernel(int n, SoA<R> st, R* __restrict__ v,
                                                              const uint8_t* __restrict__ cls,
                                                              const int* __restrict__ pidx,
                                                              const Consts<R>* __restrict__ consts, int nsteps,
                                                              R dt, const double* __restrict__ ist,
                                                              const double* __restrict__ dvgap) {
    __shared__ QuadX<R> X[2];
    const int q = threadIdx.x >> 6;
    const int l = threadIdx.x & 63;
    const int k = blockIdx.x * 64 + l;  // the grid holds ceil(n / 64) workgroups
    const int src = src_cell(k, n);
    CellState<R> s;
    load_state(s, st, src);
    const int c = cls[src];
    const Consts<R> cc = consts[pidx ? pidx[src] : 0];
    const R I = ist ? R(ist[src]) : R(0);
    const R G = dvgap ? R(dvgap[src]) : R(0);
    R vk = v[src];
    for (int i = 0; i < nsteps; ++i) vk = quad_update<R, false>(s, vk, q, l, c, cc, dt, I, G, X[i & 1], nullptr);
    if (k < n) {
        if (q == 0) v[k] = vk;
        store_state_quarter(s, st, k, q);
    }
}

template <class R>
__global__ __launch_bounds__(256) void cells_pace_quad_kernel(int n, SoA<R> st, R* __restrict__ v,
                                                              const uint8_t* __restrict__ cls,
                                                              const int* __restrict__ pidx,
                                                              const Consts<R>* __restrict__ consts, R dt,
                                                              PaceTrack* __restrict__ track, PaceArgs a) {
    __shared__ QuadX<R> X[2];
    const int q = threadIdx.x >> 6;
    const int l = threadIdx.x & 63;
    const int k = blockIdx.x * 64 + l;  // the grid holds ceil(n / 64) workgroups
    const int src = src_cell(k, n);
    CellState<R> s;
    load_state(s, st, src);
    const int c = cls[src];
    const Consts<R> cc = consts[pidx ? pidx[src] : 0];
    PaceTrack tr = track[src];
    R vk = v[src];
    double time = a.time0;
    PaceState ps{0, a.trace_next, a.trace_slot};
    for (int i = 0; i < a.nsteps; ++i, time += a.dt) {
        const R Ist = pace_istim<R>(ps, a, i);
        Diag dg;
        vk = quad_update<R, true>(s, vk, q, l, c, cc, dt, Ist, R(0), X[i & 1], &dg);
        if (q == 0) pace_record(tr, ps, a, k, i, time, (double)vk, dg);  // k >= n is never the trace cell
    }
    if (k < n) {
        if (q == 0) {
            v[k] = vk;
            track[k] = tr;
        }
        store_state_quarter(s, st, k, q);
    }
}

// Grid and block of each layout: one lane per cell in 256-lane workgroups;
// two or four lanes per cell in workgroups of 64 cells.
template <class R>
static void launch_step_r(int lanes, int n, SoA<R> s, R* v, const uint8_t* cls, const int* pidx,
                          const Consts<R>* consts, int nsteps, R dt, const double* ist, const double* dvgap,
                          hipStream_t st) {
    const int groups = (n + 63) / 64;
    if (lanes == 4)
        cells_step_quad_kernel<R><<<groups, 256, 0, st>>>(n, s, v, cls, pidx, consts, nsteps, dt, ist, dvgap);
    else if (lanes == 2)
        cells_step_split_kernel<R><<<groups, 128, 0, st>>>(n, s, v, cls, pidx, consts, nsteps, dt, ist, dvgap);
    else
        cells_step_kernel<R><<<(n + 255) / 256, 256, 0, st>>>(n, s, v, cls, pidx, consts, nsteps, dt, ist, dvgap);
}

template <class R>
static void launch_pace_r(int lanes, int n, SoA<R> s, R* v, const uint8_t* cls, const int* pidx,
                          const Consts<R>* consts, R dt, PaceTrack* track, const PaceArgs& a, hipStream_t st) {
    const int groups = (n + 63) / 64;
    if (lanes == 4)
        cells_pace_quad_kernel<R><<<groups, 256, 0, st>>>(n, s, v, cls, pidx, consts, dt, track, a);
    else if (lanes == 2)
        cells_pace_split_kernel<R><<<groups, 128, 0, st>>>(n, s, v, cls, pidx, consts, dt, track, a);
    else
        cells_pace_kernel<R><<<(n + 255) / 256, 256, 0, st>>>(n, s, v, cls, pidx, consts, dt, track, a);
}

void launch_cells_step(bool fp32, int lanes, int n, void* r, double* d, int64_t pitch, void* v,
                       const uint8_t* cls, const int* pidx, const void* consts, int nsteps, double dt,
                       const double* ist, const double* dvgap, hipStream_t st) {
    if (fp32)
        launch_step_r<float>(lanes, n, SoA<float>{(float*)r, d, pitch}, (float*)v, cls, pidx,
                             (const Consts<float>*)consts, nsteps, (float)dt, ist, dvgap, st);
    else
        launch_step_r<double>(lanes, n, SoA<double>{(double*)r, d, pitch}, (double*)v, cls, pidx,
                              (const Consts<double>*)consts, nsteps, dt, ist, dvgap, st);
}

void launch_cells_pace(bool fp32, int lanes, int n, void* r, double* d, int64_t pitch, void* v,
                       const uint8_t* cls, const int* pidx, const void* consts, double dt, PaceTrack* track,
                       const PaceArgs& a, hipStream_t st) {
    if (fp32)
        launch_pace_r<float>(lanes, n, SoA<float>{(float*)r, d, pitch}, (float*)v, cls, pidx,
                             (const Consts<float>*)consts, (float)dt, track, a, st);
    else
        launch_pace_r<double>(lanes, n, SoA<double>{(double*)r, d, pitch}, (double*)v, cls, pidx,
                              (const Consts<double>*)consts, dt, track, a, st);
}

}  // namespace lqt
