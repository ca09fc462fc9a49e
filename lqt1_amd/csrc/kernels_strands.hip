// Persistent batched-strand kernel.  A workgroup takes one bundle of
// strands at a time from a work counter (each strand on whole wavefronts,
// cells on lanes, state in registers for the whole run) and advances every
// strand of the bundle with the loop of OneD.cc:315-436: Jacobi snapshot of
// V in LDS (one barrier per step), no-flux ends by copy, S1 on the first cells while the in-cycle
// time of the double time accumulator is in its window (or, while that
// window is closed, the S2 of OneD_VW.cc:291-301), then the TPORd update.
// The time tests arrive as step indices the host evaluated on the same
// accumulator (StrandDev), so a step compares integers.  The pseudo-ECG of
// PlotScript/plot1DECG_compare.py:89-105 is reduced across each strand every
// rec_every steps (wavefront shuffles, then one LDS slot per wavefront,
// summed by the strand's first lane one step later).
#include "strands.hpp"

namespace lqt {

template <bool B>
struct BoolTag {
    static constexpr bool value = B;
};

template <class R>
__device__ __forceinline__ R wave_sum(R x) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

// FP32: two waves per SIMD (state, trackers and step constants fit in 256
// registers).  FP64 needs about 470 registers: blocks of up to 256 lanes
// run one wave per SIMD and keep them; a strand of 5 to 8 wavefronts (257
// to 512 cells) has two waves on some SIMDs and spills about half of its
// registers to scratch, so its FP64 step is slower (it stays correct).
//
// SPLIT: every cell runs on two lanes of two wavefronts, which the
// dispatcher places on different SIMDs: the first B/2 lanes advance the
// gates of current group A (tp_group_a), the other B/2 the gates of group B
// (tp_group_b), and the two swap their groups' currents through LDS at the
// step's barrier (the one that also publishes V); both then finish the step
// with the same
// arithmetic, so they hold the same V and concentrations bit for bit and
// the result is that of the single-lane update.  A step's dependent chain
// per wavefront is about 0.6 of the single-lane one.  The bundle describes
// the first half; outputs (trackers, pseudo-ECG, V rows) come from it.
template <class R, int B, bool SPLIT>
__global__ __launch_bounds__(B) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? 2 : 1))) void strands_kernel(
    StrandsArgs<R> A) {
    constexpr int NH = SPLIT ? B / 2 : B;  // lanes of one half
    constexpr int NW = NH / 64;            // wavefronts of one half (the bundle's)
    // one V segment per strand: its cells plus a mirrored end on each side
    __shared__ R vbuf[2][NH + 2 * MAXS];
    __shared__ float part[2][NW];
    __shared__ R xa[2][SPLIT ? 5 : 1][SPLIT ? NH : 1];   // group A currents of each cell
    __shared__ R xb[2][SPLIT ? 12 : 1][SPLIT ? NH : 1];  // group B currents of each cell
    __shared__ int s_next;
    const int half = SPLIT ? (int)(threadIdx.x >= (unsigned)NH) : 0;
    const int tid = (int)threadIdx.x - half * NH;  // lane within the half
    const int wid = tid >> 6;
    const bool out_half = half == 0;               // the half that writes outputs

    for (;;) {
        if (threadIdx.x == 0) s_next = atomicAdd(A.queue, 1);
        __syncthreads();
        const int q = s_next;
        __syncthreads();
        if (q >= A.nbundles) break;
        const Bundle& Q = A.bundles[q];
        // Step indices and sample counts are carried in 32 bits: the host
        // refuses strands of 2^30 steps or more (and sample intervals of as
        // many), so no index or increment below reaches INT32_MAX, which
        // stands for "never" (the host's INT64_MAX).
        const int kmax = (int)Q.kmax;
        const int slot = Q.wslot[wid];
        const bool has = slot >= 0;
        const int sidx = has ? Q.sidx[slot] : 0;
        const int w0 = has ? Q.wave0[slot] : 0;
        const StrandDev D = A.desc[sidx];
        const int n = has ? D.ncell : 0;
        const int j = tid - 64 * w0;        // cell of this lane in its strand
        const bool live = has && j < n;
        const bool lead = out_half && has && j == 0;  // writes the strand's own outputs
        const int seg = 64 * w0 + 2 * slot; // LDS segment of the strand
        const int p = seg + 1 + j;          // this cell's LDS slot
        const int nws = (n + 63) >> 6;      // wavefronts of the strand
        const int cell = D.cell_off + (live ? j : 0);

        // diffusion coefficient and its derivative (OneD.cc:108-131)
        const double dxd = D.dx;
        double Di = 0.0, dDi = 0.0;
        if (live) {
            Di = A.coeff[cell];
            if (n == 1) dDi = 0.0;
            else if (j == 0) dDi = (A.coeff[cell + 1] - A.coeff[cell]) / dxd;
            else if (j == n - 1) dDi = (A.coeff[cell] - A.coeff[cell - 1]) / dxd;
            else dDi = (A.coeff[cell + 1] - A.coeff[cell - 1]) / (2.0 * dxd);
        }
        const R RD = (R)Di, RdD = (R)dDi;
        const R inv_dx = (R)(1.0 / dxd), inv_2dx = (R)(1.0 / (2.0 * dxd)), inv_dx2 = (R)(1.0 / (dxd * dxd));
        // pseudo-ECG weight of this cell: one-sided at the ends
        const int c = live ? A.cls[cell] : 0;
        const Consts<R> cc = A.consts[D.pidx];

        // TPORd::init (TPORd.cc:142-198)
        CellState<R> s;
        s.sm = 0; s.sh = 1; s.sj = 1; s.mL = 0; s.hL = 1; s.hLp = 1;
        s.a = 0; s.iF = 1; s.iS = 1; s.ap = 0; s.iFp = 1; s.iSp = 1;
        s.d = 0; s.ff = 1; s.fs = 1; s.fcaf = 1; s.fcas = 1; s.jca = 1; s.nca = 0; s.ffp = 1; s.fcafp = 1;
        s.xrf = 0; s.xrs = 0; s.xs1 = 0; s.xs2 = 0; s.xk1 = 1; s.Jrelnp = 0; s.Jrelp = 0;
        s.nass = 7; s.kss = 145; s.cai = (R)1.0e-4; s.cass = (R)1.0e-4;
        s.nai = 7; s.ki = 145; s.cansr = 1.2; s.cajsr = 1.2; s.CaMKt = 0;
        R v = (R)A.v0[D.pidx];

        const bool s2_cell = j >= D.s2_first && j < D.s2_first + D.s2_count;
        auto k32 = [](int64_t x) { return x > (int64_t)INT32_MAX ? INT32_MAX : (int)x; };
        const int nsteps = k32(D.nsteps), k_s2a = k32(D.k_s2a), k_s2b = k32(D.k_s2b);
        const int s1_first = (int)D.s1_first, s1_count = (int)D.s1_count;
        // the next S1 range (host-evaluated decisions)
        int s1i = 0, s1_on = INT32_MAX, s1_off = INT32_MAX;
        if (has && s1_count > 0) {
            s1_on = k32(A.s1r[2 * (int64_t)s1_first]);
            s1_off = k32(A.s1r[2 * (int64_t)s1_first + 1]);
        }
        // Bookkeeping state: the beat tracker, the peak V after the S2 start
        // (OneD_VW.cc:206-207,390-394; only cells 0 and n-1 report it), the
        // pseudo-ECG weight of this cell (one-sided at the ends) and the
        // sample counters and steps.  Only the steps from Q.kbook on use
        // them; they are set just before those steps (book_init), so that
        // nothing of them is held while the earlier steps run.
        CellTrack tr;
        double wmax;
        R wphi;
        int nrec, nvrow, fin_idx, pend_idx, next_rec, next_vrec;
        int pend, fin_slot;
        auto book_init = [&]() {
            tr.act = -1; tr.rep = -1; tr.vrest = 0; tr.vmax = -1e30; tr.vmin = 1e30; tr.flags = 0;
            wmax = -100.0;
            wphi = R(0);
            if (live) {
                const double r = (j * dxd - D.x0) * (j * dxd - D.x0);
                wphi = (R)((j == 0 || j == n - 1) ? 1.0 / r : 1.0 / (2.0 * r));
            }
            nrec = 0;
            nvrow = 0;
            pend = 0;       // a sample waits for its reduction
            fin_slot = -1;  // slot whose per-wave partials wait for the lead lane
            fin_idx = 0;
            pend_idx = 0;
            next_rec = k32(D.k_rec0);
            next_vrec = k32(D.k_vrec0);
        };
        double time = 0.0;
        int b = 0;
        // V^0 into LDS with mirrored ends
        if (live && out_half) {
            vbuf[0][p] = v;
            if (j == 0) vbuf[0][seg] = v;
            if (j == n - 1) vbuf[0][seg + n + 1] = v;
        }
        // One barrier per step.  The gates and currents of step k need only
        // the cell's own V^k; V^k of the neighbours (written into vbuf[b] at
        // the end of step k-1) and, split, the other half's currents (xa/xb
        // slot b) are read after the barrier, by the diffusion term and the
        // epilogue.  vbuf, xa and xb alternate between two slots, so step
        // k+1 writes the slot nobody reads in step k.
        // The loop body as one step function of two forms: before the
        // bundle's first bookkeeping step (Q.kbook: the earliest measured
        // beat, pseudo-ECG or V-row sample or end-cell peak of its strands)
        // the trackers and samples cannot fire and their code is left out
        // (about 7 % of a step); from there on the full form runs.  Both
        // forms do the same arithmetic on V and the state.
        auto step = [&](int k, auto book) {
            constexpr bool BOOK = decltype(book)::value;
            const bool run = has && k < nsteps;
            if (k >= s1_off) {  // past this S1 range: load the next one
                ++s1i;
                s1_on = s1i < s1_count ? k32(A.s1r[2 * (int64_t)(s1_first + s1i)]) : INT32_MAX;
                s1_off = s1i < s1_count ? k32(A.s1r[2 * (int64_t)(s1_first + s1i) + 1]) : INT32_MAX;
            }
            const bool act = run && live;
            Pre<R> P;
            GroupA<R> GA;
            GroupB<R> GB;
            if (act) {
                if constexpr (!SPLIT) {
                    P = tp_pre(s, v, A.dt);
                    GA = tp_group_a(s, v, c, cc, A.dt, P);
                    GB = tp_group_b(s, v, c, cc, A.dt, P);
                } else if (half == 0) {
                    P = tp_pre<R, 1>(s, v, A.dt);
                    GA = tp_group_a(s, v, c, cc, A.dt, P);
                    xa[b][0][tid] = GA.INa; xa[b][1][tid] = GA.INaL; xa[b][2][tid] = GA.Ito;
                    xa[b][3][tid] = GA.IKs; xa[b][4][tid] = GA.INaK;
                } else {
                    P = tp_pre<R, 2>(s, v, A.dt);
                    GB = tp_group_b(s, v, c, cc, A.dt, P);
                    xb[b][0][tid] = GB.ICaL; xb[b][1][tid] = GB.ICaNa; xb[b][2][tid] = GB.ICaK;
                    xb[b][3][tid] = GB.IKr; xb[b][4][tid] = GB.IK1; xb[b][5][tid] = GB.INaCa_i;
                    xb[b][6][tid] = GB.INaCa_ss; xb[b][7][tid] = GB.IpCa; xb[b][8][tid] = GB.ICab;
                    xb[b][9][tid] = GB.Jrel; xb[b][10][tid] = GB.INab; xb[b][11][tid] = GB.IKb;
                }
            }
            __syncthreads();
            if constexpr (BOOK) {
                // finish the sample whose per-wave partials were written last step
                if (fin_slot >= 0 && lead) {
                    float tot = 0.f;
                    for (int w = 0; w < nws; ++w) tot += part[fin_slot][w0 + w];
                    A.phi[D.phi_off + fin_idx] = tot;
                }
                fin_slot = -1;
                // reduce the sample recorded at the end of the previous step
                if (pend) {
                    const R dv = live ? vbuf[b][p + 1] - vbuf[b][p - 1] : R(0);
                    const R tot = wave_sum<R>(-wphi * dv);
                    const int sl = (int)(pend_idx & 1);
                    if ((tid & 63) == 0) part[sl][wid] = (float)tot;
                    fin_slot = sl;
                    fin_idx = pend_idx;
                    pend = 0;
                }
            }
            if (act) {
                if constexpr (SPLIT) {
                    if (half == 0) {
                        GB.ICaL = xb[b][0][tid]; GB.ICaNa = xb[b][1][tid]; GB.ICaK = xb[b][2][tid];
                        GB.IKr = xb[b][3][tid]; GB.IK1 = xb[b][4][tid]; GB.INaCa_i = xb[b][5][tid];
                        GB.INaCa_ss = xb[b][6][tid]; GB.IpCa = xb[b][7][tid]; GB.ICab = xb[b][8][tid];
                        GB.Jrel = xb[b][9][tid]; GB.INab = xb[b][10][tid]; GB.IKb = xb[b][11][tid];
                    } else {
                        GA.INa = xa[b][0][tid]; GA.INaL = xa[b][1][tid]; GA.Ito = xa[b][2][tid];
                        GA.IKs = xa[b][3][tid]; GA.INaK = xa[b][4][tid];
                    }
                }
                const R vl = vbuf[b][p - 1], vr = vbuf[b][p + 1], vc = vbuf[b][p];
                const R first = (j == 0 || j == n - 1) ? (vr - vl) * inv_dx : (vr - vl) * inv_2dx;
                const R second = (vr + vl - R(2) * vc) * inv_dx2;
                const R dvgap = RdD * first + RD * second;
                R ist = R(0);
                if (k >= s1_on) {  // in-cycle time in the S1 window
                    if (j < D.nstim) ist = (R)D.stim_strength;
                } else if (k >= k_s2a && k < k_s2b && s2_cell) {
                    ist = (R)D.s2_strength;
                }
                v = tp_finish<R, false>(s, v, c, cc, A.dt, ist, dvgap, P, GA, GB, nullptr);
            }
            if (BOOK && run && out_half) {  // trackers, peaks and samples
                if (live) {
                    if (k >= k32(D.k_wmax) && (double)v >= wmax) wmax = (double)v;

                    // beat tracker, in the measured beat only (floor(t/BCL) == k)
                    if (k >= k32(D.k_mb0) && k < k32(D.k_mb1)) {
                        const double vf = (double)v;
                        if (!(tr.flags & 1)) {
                            tr.flags |= 1;
                            tr.vrest = vf;
                            // still depolarised from the previous beat: it
                            // failed to repolarise
                            if (vf > -60.0) tr.flags |= 16;
                        }
                        if (!(tr.flags & 2)) {
                            if (vf >= -30.0) {
                                tr.flags |= 2;
                                tr.act = time;
                                tr.vmax = vf;
                            }
                        } else if (!(tr.flags & 4)) {
                            if (vf > tr.vmax) tr.vmax = vf;
                            const double lvl = tr.vmax - 0.9 * (tr.vmax - tr.vrest);
                            if (vf <= lvl) {
                                tr.flags |= 4;
                                tr.rep = time;
                            } else if (time >= tr.act + 100.0) {
                                // phase 2/3: a rise of more than 2 mV above the
                                // lowest V since act+100 ms is an EAD
                                if (vf < tr.vmin) tr.vmin = vf;
                                if (vf > tr.vmin + 2.0) tr.flags |= 8;
                            }
                        }
                    }
                }
                if (k == next_rec) {  // k % rec_every == 0 and time >= rec_from
                    next_rec += D.rec_every;
                    if (nrec < k32(D.phi_cap)) {
                        if (nrec == 0 && lead) A.phit0[sidx] = time;
                        pend = 1;
                        pend_idx = nrec++;
                    }
                }
                if (k == next_vrec) {
                    next_vrec += D.vrec_every;
                    if (nvrow < k32(D.vrec_cap)) {
                        if (live) A.vrec[D.vrec_off + nvrow * (int64_t)n + j] = (double)v;
                        if (nvrow == 0 && lead) A.vrect0[sidx] = time;
                        ++nvrow;
                    }
                }
            }
            // V^{k+1} into the other slot with mirrored ends
            if (live && out_half) {
                vbuf[b ^ 1][p] = v;
                if (j == 0) vbuf[b ^ 1][seg] = v;
                if (j == n - 1) vbuf[b ^ 1][seg + n + 1] = v;
            }
            b ^= 1;
        };
        const int kb = Q.kbook < (int64_t)kmax ? (int)Q.kbook : kmax;
        int k = 0;
        for (; k < kb; ++k, time += A.dt_d) step(k, BoolTag<false>{});
        book_init();
        for (; k < kmax; ++k, time += A.dt_d) step(k, BoolTag<true>{});

        if (live) {
            if (out_half) {
                A.act[cell] = (tr.flags & 2) ? tr.act : -1.0;
                A.apd[cell] = (tr.flags & 4) ? tr.rep - tr.act : -1.0;
                int e = 0;
                if (tr.flags & 8) e |= 1;
                if (((tr.flags & 2) && !(tr.flags & 4)) || (tr.flags & 16)) e |= 2;
                A.ead[cell] = e;
                A.finv[cell] = v;
                if (j == 0) A.vmax[2 * (int64_t)sidx] = wmax;
                if (j == n - 1) A.vmax[2 * (int64_t)sidx + 1] = wmax;
            }
            if (SPLIT) store_state_half(s, A.fin, cell, half);
            else store_state(s, A.fin, cell);
        }
        if (lead) {
            A.nvrec[sidx] = nvrow;
            A.nphi[sidx] = nrec;
            A.steps[sidx] = D.nsteps;
        }
        __syncthreads();
    }
}

__global__ void zero_int_kernel(int* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

template <class R>
static void launch_b(int block, bool split, int grid, const StrandsArgs<R>& a, hipStream_t st) {
    if (split) {
        // FP64: split blocks of up to 256 lanes only (the host asks for no
        // other); larger ones would hold two waves per SIMD and spill
        if constexpr (sizeof(R) == 8) {
            if (block == 128) strands_kernel<R, 128, true><<<grid, 128, 0, st>>>(a);
            else strands_kernel<R, 256, true><<<grid, 256, 0, st>>>(a);
        } else {
            switch (block) {
                case 128: strands_kernel<R, 128, true><<<grid, 128, 0, st>>>(a); break;
                case 256: strands_kernel<R, 256, true><<<grid, 256, 0, st>>>(a); break;
                case 384: strands_kernel<R, 384, true><<<grid, 384, 0, st>>>(a); break;
                default: strands_kernel<R, 512, true><<<grid, 512, 0, st>>>(a); break;
            }
        }
        return;
    }
    switch (block) {
        case 64: strands_kernel<R, 64, false><<<grid, 64, 0, st>>>(a); break;
        case 128: strands_kernel<R, 128, false><<<grid, 128, 0, st>>>(a); break;
        case 192: strands_kernel<R, 192, false><<<grid, 192, 0, st>>>(a); break;
        case 256: strands_kernel<R, 256, false><<<grid, 256, 0, st>>>(a); break;
        case 320: strands_kernel<R, 320, false><<<grid, 320, 0, st>>>(a); break;
        case 384: strands_kernel<R, 384, false><<<grid, 384, 0, st>>>(a); break;
        case 448: strands_kernel<R, 448, false><<<grid, 448, 0, st>>>(a); break;
        default: strands_kernel<R, 512, false><<<grid, 512, 0, st>>>(a); break;
    }
}

// block: lanes of the workgroup (for a split launch both halves, 2 x 64 x
// the bundle's wavefronts, 128..512).
void launch_strands(bool fp32, int block, bool split, int grid, const void* args, hipStream_t st) {
    if (fp32) launch_b<float>(block, split, grid, *(const StrandsArgs<float>*)args, st);
    else launch_b<double>(block, split, grid, *(const StrandsArgs<double>*)args, st);
}

void launch_zero_int(int* p, int n, hipStream_t st) {
    zero_int_kernel<<<(n + 255) / 256, 256, 0, st>>>(p, n);
}

// Resident workgroups per CU of the strand kernel for a block size.
int strands_occupancy(bool fp32, int block, bool split) {
    int nb = 0;
    hipError_t e = hipSuccess;
#define OCC(R, BB, SP)                                                                                     \
    case BB: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, strands_kernel<R, BB, SP>, BB, 0); break;
    if (split) {
        if (fp32) {
            switch (block) { OCC(float, 128, true) OCC(float, 256, true) OCC(float, 384, true) default: OCC(float, 512, true) }
        } else {
            switch (block) { OCC(double, 128, true) default: OCC(double, 256, true) }
        }
    } else if (fp32) {
        switch (block) { OCC(float, 64, false) OCC(float, 128, false) OCC(float, 192, false) OCC(float, 256, false)
                         OCC(float, 320, false) OCC(float, 384, false) OCC(float, 448, false) default: OCC(float, 512, false) }
    } else {
        switch (block) { OCC(double, 64, false) OCC(double, 128, false) OCC(double, 192, false) OCC(double, 256, false)
                         OCC(double, 320, false) OCC(double, 384, false) OCC(double, 448, false) default: OCC(double, 512, false) }
    }
#undef OCC
    return e == hipSuccess ? nb : 1;
}

}  // namespace lqt
