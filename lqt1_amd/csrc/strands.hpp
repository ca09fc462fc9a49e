// Device-side description of a batch of 1D strands.
#pragma once

#include "internal.hpp"

namespace lqt {

struct StrandDev {
    int ncell, cell_off, pidx, nstim;
    int rec_every, measure_beat, vrec_every, pad_;
    double dx, bcl, stop_time, stim_start, stim_dur, stim_strength, rec_from, x0;
    int64_t phi_off, phi_cap;
    int64_t vrec_off, vrec_cap;  // rows
    // S2 (OneD_VW.cc:291-301): only while the S1 window is closed
    double s2_base, s2_start, s2_dur, s2_strength;
    int s2_first, s2_count;
    // steps of `for (time = 0; time <= stop_time; time += dt)`, counted on
    // the host with the same double accumulator
    int64_t nsteps;
    // The time-dependent decisions of the loop as step indices, evaluated on
    // the host on the same accumulator with the loop's own expressions: S1
    // is on in the ranges s1r[s1_first .. s1_first + s1_count) (pairs
    // [on, off)); S2 window [k_s2a, k_s2b); end-cell peaks from k_wmax;
    // measured beat [k_mb0, k_mb1); first pseudo-ECG and V-row samples
    // k_rec0, k_vrec0 (then every rec_every / vrec_every steps).
    int64_t s1_first, s1_count;
    int64_t k_s2a, k_s2b, k_wmax, k_mb0, k_mb1, k_rec0, k_vrec0;
};

// Strands that share one workgroup: strand s occupies the whole wavefronts
// wave0[s] .. wave0[s] + ceil(ncell/64) - 1, so every wavefront belongs to
// one strand and its shuffles and partial sums stay inside that strand.
// All strands of a bundle advance in lock-step on one time accumulator;
// the loop runs kmax = max(nsteps) + 2 iterations (one to reduce the last
// ECG sample, one to write it out).
constexpr int MAXS = 8;
struct Bundle {
    int n;
    int sidx[MAXS];
    int wave0[MAXS];
    int wslot[MAXS];  // strand slot of wavefront w, -1 if unused
    int64_t kmax;
    // first step at which any of its strands can track, sample or record:
    // min over them of k_mb0, k_wmax, k_rec0, k_vrec0 (the kernel leaves
    // that code out of the steps before it)
    int64_t kbook;
};

// Per-cell beat tracker: -30 mV activation (OneD.cc:422), running peak,
// 90 % repolarisation and the EAD test, all within beat `measure_beat`; a
// cell whose first V of the beat is above -60 mV has not repolarised from
// the previous beat and counts as failing to repolarise.
struct CellTrack {
    double act, rep;
    double vrest, vmax, vmin;
    int flags;  // bit0 entered beat, bit1 activated, bit2 repolarised, bit3 EAD,
                // bit4 entered the beat depolarised (V > -60 mV)
};

template <class R>
struct StrandsArgs {
    int nbundles;
    const Bundle* bundles;     // this launch's bundles, processing order
    int* queue;                // work counter, zeroed before each run
    const StrandDev* desc;
    const int64_t* s1r;        // S1 step ranges of all strands, pairs [on, off)
    const uint8_t* cls;        // per cell class
    const double* coeff;       // per cell D (mm^2/ms)
    const Consts<R>* consts;
    const double* v0;          // per parameter set initial V
    R dt;
    double dt_d;
    float* phi;                // pseudo-ECG samples
    int64_t* nphi;             // per strand sample count
    double* phit0;             // per strand time of the first sample
    double* vrec;              // recorded V rows
    int64_t* nvrec;            // per strand recorded rows
    double* vrect0;            // per strand time of the first row
    double* act;               // per cell
    double* apd;               // per cell APD90
    int* ead;                  // per cell flags: 1 EAD, 2 no repolarisation
    SoA<R> fin;                // final state, pitch = total cells
    R* finv;
    int64_t* steps;            // per strand step count
    double* vmax;              // per strand [2]: peak V of cell 0 / cell n-1 once time - s2_base >= s2_start
};

void launch_strands(bool fp32, int block, bool split, int grid, const void* args, hipStream_t st);
void launch_zero_int(int* p, int n, hipStream_t st);

}  // namespace lqt
