// Host-side declarations shared by the kernel translation units and the
// C ABI implementation.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "state.hpp"

namespace lqt {

// Per-cell statistics tracker of SingleCell.cc:144-215, kept in device
// memory between launches.
struct PaceTrack {
    double oldV, rest, dvdt_max, overshoot, amplitude;
    double repo[6];  // 20, 25, 50, 75, 80, 90 %
    double apd[6];
    int peakfound;
    int pad;
};

// One launch of the pacing protocol.  Its time tests (SingleCell.cc:95-215)
// arrive as step indices of this launch that the host evaluated with the
// protocol's own expressions on the same double time accumulator, so the
// kernel compares integers instead of dividing in double every step.
constexpr int kPaceRanges = 4;
struct PaceArgs {
    double bcl, stim_start, stim_dur, stim_strength;
    double dt;          // step of the double time accumulator
    double time0;       // time at the first step of this launch
    int nsteps;
    int nstim;          // stimulus on in steps [stim_on[r], stim_off[r]), r < nstim
    int stim_on[kPaceRanges], stim_off[kPaceRanges];
    int rest_step;      // the step at which the last beat's statistics start, or -1
    int last0, last1;   // steps [last0, last1) lie in the last beat
    int trace_cell, trace_every;
    int trace_next;     // first step of this launch that records the trace, or INT32_MAX
    int64_t trace_slot; // trace slot of that step
    int64_t trace_cap;
    double* trace;      // device, or nullptr
};

// Per-voxel beat tracker of the tissue (activation and APD90 within a
// window of the tissue clock), updated after each step while enabled.
struct VoxTrack {
    double act, rep;     // step times (ms) of the -30 mV up-crossing and of 90 % repolarisation
    double vrest, vmax;  // V at the window's first step; peak V since activation
    int flags;           // bit0 window entered, bit1 activated, bit2 repolarised
    int pad;
};

template <class R>
struct TissueArgs {
    int64_t nvox;
    const int32_t* vpos;       // padded grid index of voxel k
    const uint8_t* cls;        // cell class of voxel k
    const R* vold;             // dense padded V grid, step n
    R* vnew;                   // dense padded V grid, step n+1
    int32_t sx, sy;            // padded strides of x and y (z stride is 1)
    const R* fd;               // chunk-major, 6 slots: fibre f (3), divergence of D (3)
    R dcross, ddiff;           // D_cross, D_fibre - D_cross
    R inv_h2, inv_4h2, inv_2h; // 1/dx^2, 1/(4dx^2), 1/(2dx)
    R* sr;                     // state, chunk-major: NR slots of R per chunk
    typename SlowStore<R>::type* sd;  // state, chunk-major: ND slow slots per chunk (double, or
                                     // float offsets from sref in FP32)
    SlowRef sref;
    const Consts<R>* consts;
    R dt;
    // stimulus of this step
    int stim_on;
    double tc;                 // in-cycle time
    double stim_dur;
    R stim_strength;
    const float* stim_off;
    int s2_on;
    R s2_strength;
    const uint8_t* s2;
    unsigned long long* guard; // counts: V < -90, V > 200, non-finite
    const int32_t* gout;       // ghost cells (padded index) listed by source voxel
    const int32_t* gwave;      // first gout entry of each wavefront's voxels
};

// kernels_cells.hip
// lanes: 1, 2 or 4 lanes per cell (the same results bit for bit; more lanes
// per cell shorten a step when there are few cells)
void launch_cells_step(bool fp32, int lanes, int n, void* r, double* d, int64_t pitch, void* v,
                       const uint8_t* cls, const int* pidx, const void* consts, int nsteps, double dt,
                       const double* ist, const double* dvgap, hipStream_t st);
void launch_cells_pace(bool fp32, int lanes, int n, void* r, double* d, int64_t pitch, void* v,
                       const uint8_t* cls, const int* pidx, const void* consts, double dt, PaceTrack* track,
                       const PaceArgs& a, hipStream_t st);

// kernels_tissue.hip
// Launch shapes of the fused tissue step (block size, XCD-aware map, waves
// per SIMD the registers are allocated for, stencil staged in LDS); entry 0
// is the default.
struct TissueVariant {
    int block, xcd, waves, lds;
};
constexpr int kNumTissueVariants = 8;
extern const TissueVariant kTissueVariants[kNumTissueVariants];
void launch_tissue_step(bool fp32, bool aniso, bool flat, int variant, const void* args, int64_t nvox,
                        hipStream_t st);
void launch_ghost_fill(bool fp32, void* grid, const int32_t* dst, const int32_t* src, int64_t n,
                       hipStream_t st);
void launch_fill(bool fp32, void* p, int64_t n, double value, hipStream_t st);
void launch_track(bool fp32, int64_t nvox, const int32_t* vpos, const void* grid, double time, VoxTrack* tr,
                  hipStream_t st);
void launch_track_reset(int64_t nvox, VoxTrack* tr, hipStream_t st);
void launch_copy(void* dst, const void* src, int64_t bytes, hipStream_t st);

}  // namespace lqt
