/*
 * lqt.h - C ABI of the MI355X (gfx950) TPORd monodomain solver.
 *
 * Drop-in boundary for the reference's hot path.  Each entry point names
 * the reference interface it replaces (paths relative to the reference
 * repository).  Conventions: opaque handles; every call returns an
 * lqt_status (0 = LQT_OK) and leaves a message for lqt_last_error() on
 * failure; buffers passed in are host memory owned by the caller; device
 * memory is owned by the handle; one handle must not be used from two
 * threads at once, distinct handles are independent; units are the
 * reference's (mV, ms, mm, mm^2/ms, pA/pF, mM, fmol).
 */
#ifndef LQT_H
#define LQT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int lqt_status;
#define LQT_OK 0
#define LQT_ERR_ARG 1      /* invalid argument or size                     */
#define LQT_ERR_HIP 2      /* HIP runtime error (no GPU, launch failure)   */
#define LQT_ERR_IO 3       /* file could not be read or written           */
#define LQT_ERR_RANGE 4    /* non-finite membrane potential detected      */
#define LQT_ERR_STATE 5    /* call out of order                           */

/* LQT_FP32 keeps gates, V and the fast concentrations in float and the slow
 * concentrations in double; its handles accept dt up to LQT_FP32_MAX_DT ms
 * (the slow gates' Rush-Larsen step uses a series valid to dt/tau = 0.125,
 * and every such gate has tau >= 0.8 ms).  LQT_FP64 takes any dt > 0. */
typedef enum { LQT_FP64 = 0, LQT_FP32 = 1 } lqt_precision;
#define LQT_FP32_MAX_DT 0.1

/* enum CellType of SingleCell/Cell.cc:20 (values kept).  In a tissue label
 * grid, LQT_HOLD may be OR'ed onto a type: the voxel is updated as that
 * type and its V is then reset to -100 mV every step (the type-0 voxels of
 * new_11.cu:1247-1248); such voxels are left out of the V guard's counts. */
enum { LQT_LVEPI = 0, LQT_LVENDO = 1, LQT_EPI = 2, LQT_ENDO = 3, LQT_MCELL = 4, LQT_HOLD = 0x80, LQT_NONE = 255 };

/* Runtime form of the reference's compile-time model flags
 * (SingleCell/TPORd.cc:30-36,45; GPU3D.cu:18-22). */
typedef struct {
    double iks_scale;  /* LQT1: GKs x iks_scale (TPORd.cc:705-707 uses 0.05)   */
    int iso;           /* ISO1 beta-adrenergic flag (TPORd.cc:583,589,629,731) */
    double pca_iso;    /* PCa multiplier under ISO1 (1.6 TPORd.cc, 4 GPU3D.cu) */
    int suprep;        /* SUPREP gene therapy (TPORd.cc:709-728)               */
    double dose_fmol;  /* SUPREP dose in fmol                                  */
    double sca;        /* GKs base multiplier `sca` (TPORd.cc:45,696)          */
    int clamp;         /* CLAMP: voltage held (TPORd.cc:929-933)               */
    double v0;         /* initial V: -87.5 (TPORd.cc:144), -86.2 (GPU3D.cu:213) */
} lqt_params;

/* Named parameter sets reproducing each reference build:
 * "cpu_default" (TPORd.cc shipped flags CON+ISO1), "gpu3d" (GPU3D.cu:18-22),
 * "gpu2d" (new_11.cu:18-21), "config1" (CON, ISO off), "config3"
 * (IKs x0.2, ISO on, PCa x4). */
lqt_status lqt_params_preset(const char *name, lqt_params *out);

/* SUPREP dose -> (fraction of suppressed native GKs, fraction of the
 * replacing GKs), GPU3D.cu:885-904 / fit_dose_dependent.py:245-248. */
lqt_status lqt_suprep_fractions(double dose_fmol, double *frac_wt, double *frac_rep);

const char *lqt_last_error(void);
int lqt_device_count(void);
lqt_status lqt_device_synchronize(void);
/* Name of the GPU the library runs on (gcnArchName) into buf. */
lqt_status lqt_device_name(char *buf, int len);
/* No counterpart in the reference: times `reps` launches of a plain
 * streaming copy of `bytes` bytes (device to device) and returns the mean
 * time per copy; the reference point for HBM bandwidth and counters. */
lqt_status lqt_bench_copy(int64_t bytes, int reps, double *us_per_copy);

/* ---------------------------------------------------------------------------
 * Cells: n independent TPORd cells on the GPU, one, two or four lanes per cell.
 * Replaces the Cell/TPORd object interface (SingleCell/Cell.cc:22-42,
 * SingleCell/TPORd.cc:137-376) used by SingleCell.cc and OneD.cc:
 * setIstim / setDVgap_dt / update / getV / outputAllStates / readinAllStates.
 */
typedef struct lqt_cells lqt_cells;

/* celltype[n] in enum CellType; params[nparams] with nparams 1 or n;
 * n <= 2^29 - 64 (and at most 2^29 - 64 cells in all for a strand batch). */
lqt_status lqt_cells_create(int n, const uint8_t *celltype, const lqt_params *params, int nparams,
                            lqt_precision prec, double dt, lqt_cells **out);
lqt_status lqt_cells_destroy(lqt_cells *c);
/* nsteps calls of TPORd::update() per cell with fixed Istim / dVgap_dt
 * (host arrays of n, or NULL for 0).  Replaces setIstim+setDVgap_dt+update. */
lqt_status lqt_cells_step(lqt_cells *c, int nsteps, const double *istim, const double *dvgap);
/* Same, enqueued on the caller's HIP stream (hipStream_t as void*, NULL for
 * the handle's own stream) without waiting: istim and dvgap are DEVICE
 * arrays of n doubles (or NULL), read while the kernel runs.  The handle's
 * other calls use its own stream: the caller synchronises `stream` before
 * reading states back or stepping again from another stream. */
lqt_status lqt_cells_step_async(lqt_cells *c, int nsteps, const double *istim, const double *dvgap, void *stream);
/* Lanes per cell of lqt_cells_step / lqt_cells_step_async / lqt_cells_pace
 * (no counterpart in the reference, whose cells are objects stepped one
 * after another): 2 or 4 lanes of as many wavefronts share a cell's ionic
 * currents (two groups, or each group in two parts) and exchange them every
 * step, with the same results bit for bit and a shorter dependent chain per
 * step; 1 runs one lane per cell; 0 (the default) chooses by size: 4 while
 * ceil(n / 64) <= compute units (n <= 16,384 on MI355X) and in FP32 up to
 * twice that, 2 in FP64 up to twice that, else 1.  lqt_cells_get_lanes
 * reports what the next call uses. */
lqt_status lqt_cells_set_lanes(lqt_cells *c, int lanes);
lqt_status lqt_cells_get_lanes(const lqt_cells *c, int *lanes);
/* All 40 states per cell in the order of TPORd::outputAllStates
 * (TPORd.cc:200-252), out[n*40]. */
lqt_status lqt_cells_get_state(lqt_cells *c, double *out);
lqt_status lqt_cells_set_state(lqt_cells *c, const double *in);
lqt_status lqt_cells_get_v(lqt_cells *c, double *out);
/* Text state file of TPORd::outputAllStates (%4.10f, one value per line)
 * for cell i, and its reader TPORd::readinAllStates (TPORd.cc:254-353). */
lqt_status lqt_cells_write_state_file(lqt_cells *c, int i, const char *path);
lqt_status lqt_cells_read_state_file(lqt_cells *c, int i, const char *path);

/* Single-cell pacing protocol of SingleCell.cc:38-287 for all n cells at once
 * (each with its own parameter set): S1 pacing, then the statistics of the
 * last beat.  stats[n*10] = rest, dV/dt max, overshoot, amplitude, APD20,
 * APD25, APD50, APD75, APD80, APD90 (SingleCell.cc:220-232).  If vtrace is
 * not NULL, (V, ICaL, IKs) of cell `trace_cell` every `trace_every` steps
 * (the columns of SingleCell.cc:117-139) are written to vtrace[3*trace_cap];
 * *ntrace receives the number of samples. */
typedef struct {
    double bcl, stim_start, stim_duration, stim_strength;
    int nbeats;
} lqt_pacing;
lqt_status lqt_cells_pace(lqt_cells *c, const lqt_pacing *pace, double *stats,
                          int trace_cell, int trace_every, double *vtrace, int64_t trace_cap,
                          int64_t *ntrace);

/* ---------------------------------------------------------------------------
 * Tissue: monodomain 1D/2D/3D tissue with the reference's explicit update
 * (GPU3D.cu:1724-1842 main loop, calc_du :370-385 / new_11.cu:352-417,
 * Itot :425-1220, new_u_v :1222-1230).  Grid index (x*ny + y)*nz + z.
 */
typedef struct {
    int nx, ny, nz;
    double dx;              /* mm                                           */
    const uint8_t *label;   /* nx*ny*nz enum CellType, LQT_NONE outside     */
    const float *fibre;     /* nx*ny*nz*3 unit fibre vectors, or NULL       */
    double d_fibre;         /* mm^2/ms along the fibre (D1 in new_11.cu)   */
    double d_cross;         /* mm^2/ms across (D2)                         */
} lqt_geometry;

/* S1 train of GPU3D.cu:1757-1762 / set_Istim :397-409: site k receives
 * `strength` while the in-cycle time t - floor(t/bcl)*bcl lies in
 * [offset_k, offset_k + duration) and, if window > 0, below window
 * (GPU3D.cu:1757 only calls set_Istim while the in-cycle time is below
 * 82 ms); optional S2 of set_S2Istim :411-422: `s2_strength` on the S2
 * sites while t in [s2_time, s2_time + s2_duration). */
typedef struct {
    double bcl, duration, strength;
    double s2_time, s2_duration, s2_strength;  /* s2_duration 0 = no S2 */
    double window;                             /* ms; 0 = no limit */
} lqt_stim_protocol;

typedef struct lqt_tissue lqt_tissue;

/* The padded grid, (nx+2)(ny+2)(nz+2) cells, must stay below 2^30 cells in
 * FP32 and 2^29 in FP64 (the step kernel addresses V by 32-bit byte
 * offsets); the reference's 690 x 489 x 537 grid is 1.8e8. */
lqt_status lqt_tissue_create(const lqt_geometry *geo, const lqt_params *params, lqt_precision prec,
                             double dt, lqt_tissue **out);
lqt_status lqt_tissue_destroy(lqt_tissue *t);
/* Reads the reference's voxel list "x y z fx fy fz label" (GPU3D.cu:1276-1297,
 * labels 1-20 mapped as GPU3D.cu:865-871) into a tissue of the given grid. */
lqt_status lqt_tissue_create_from_file(const char *path, int nx, int ny, int nz, double dx,
                                       double d_fibre, double d_cross, const lqt_params *params,
                                       lqt_precision prec, double dt, lqt_tissue **out);
lqt_status lqt_tissue_info(const lqt_tissue *t, int64_t *nvox, int64_t *nghost);
/* Host only: the voxel list reader of lqt_tissue_create_from_file
 * (GPU3D.cu:1276-1297) into caller grids label[nx*ny*nz] (LQT_NONE where no
 * voxel is listed) and fibre[nx*ny*nz*3] (may be NULL); *nvox (may be NULL)
 * receives the tissue voxel count.  A voxel outside the grid is refused. */
lqt_status lqt_geometry_read_file(const char *path, int nx, int ny, int nz, uint8_t *label, float *fibre,
                                  int64_t *nvox);
/* Host only: the stimulation profile reader of lqt_tissue_set_stimulus_file
 * (GPU3D.cu:1691-1721): the grid indices of the sites inside the grid, in
 * file order, and their offsets 1000 * t_act in ms; *n receives the count
 * (call with cells and offsets_ms NULL to query; cap is their length). */
lqt_status lqt_stimulus_read_file(const char *path, int nx, int ny, int nz, int64_t cap, int64_t *cells,
                                  float *offsets_ms, int64_t *n);
/* Host only (no device needed): compaction and ghost mirror map of a label
 * grid as built by lqt_tissue_create (GPU3D.cu:1301-1333, 1487-1604).
 * g[nx*ny*nz] receives k+1 for tissue voxel k (x,y,z order), -(k+1) for a
 * ghost cell mirroring voxel k, 0 elsewhere. */
lqt_status lqt_geometry_ghost_map(int nx, int ny, int nz, const uint8_t *label, int32_t *g,
                                  int64_t *nvox, int64_t *nghost);
/* Stimulus sites are grid indices; offsets in ms (1000*t_act of the
 * stimulation profile, GPU3D.cu:1691-1721). */
lqt_status lqt_tissue_set_stimulus(lqt_tissue *t, const lqt_stim_protocol *p, int64_t nsites,
                                   const int64_t *site_cells, const float *site_offset_ms,
                                   int64_t ns2, const int64_t *s2_cells);
/* Reads stimulationprofile.cnd ("x y z f f f t_act[s] s", GPU3D.cu:1691-1721). */
lqt_status lqt_tissue_set_stimulus_file(lqt_tissue *t, const lqt_stim_protocol *p, const char *path);
/* Advance nsteps steps of dt and wait for them.  Returns LQT_ERR_RANGE if V
 * became non-finite. */
lqt_status lqt_tissue_step(lqt_tissue *t, int64_t nsteps);
/* Same, enqueued on the caller's HIP stream (hipStream_t passed as void*,
 * NULL for the tissue's own stream); returns without waiting for the device
 * and records no timing events.  The non-finite check of lqt_tissue_step is
 * then the caller's, through lqt_tissue_v_guard. */
lqt_status lqt_tissue_step_async(lqt_tissue *t, int64_t nsteps, void *stream);
lqt_status lqt_tissue_time(const lqt_tissue *t, double *time_ms, int64_t *step);
/* V of every tissue voxel in compacted (x,y,z) order, out[nvox]. */
lqt_status lqt_tissue_read_v(lqt_tissue *t, double *out);
/* All 40 states of tissue voxel k in the 40-line order, out[40].  An FP32
 * tissue holds V, the gates and the fast concentrations in float and the five
 * slow concentrations (nai, ki, cansr, cajsr, CaMKt) as float offsets from
 * reference values (the initial values at creation, or those a checkpoint
 * carries after lqt_tissue_load), so a value written here comes back rounded
 * to that precision. */
lqt_status lqt_tissue_read_state(lqt_tissue *t, int64_t k, double *out);
lqt_status lqt_tissue_write_state(lqt_tissue *t, int64_t k, const double *in);
/* Counts of V < -90 mV, V > 200 mV and non-finite V (GPU3D.cu:1796-1829),
 * reduced on the device. */
lqt_status lqt_tissue_v_guard(lqt_tissue *t, int64_t counts[3]);
/* Activation and APD90 maps.  While enabled, every step that starts at a
 * tissue-clock time t in [t_from, t_to) updates a tracker per voxel with the
 * voxel's new V: activation time = t of the first step whose V >= -30 mV
 * (the criterion of OneD.cc:422), APD90 = time from activation to the first
 * later step whose V <= Vmax - 0.9 (Vmax - Vrest), with Vmax the peak since
 * activation and Vrest the V after the window's first step.  No counterpart
 * in the reference's tissue programs, which write V maps only.  Calling
 * lqt_tissue_track resets the trackers; t_to <= t_from switches them off.
 * The trackers are not part of a checkpoint (lqt_tissue_save/load). */
lqt_status lqt_tissue_track(lqt_tissue *t, double t_from, double t_to);
/* act[nvox], apd90[nvox] in ms (either may be NULL), -1 where the voxel has
 * not activated / not repolarised within the window. */
lqt_status lqt_tissue_track_results(lqt_tissue *t, double *act, double *apd90);
/* Legacy VTK STRUCTURED_POINTS of V on every `stride`-th voxel, -100 outside
 * (GPU3D.cu:123-163 writes stride 3). */
lqt_status lqt_tissue_write_vtk(lqt_tissue *t, const char *path, int stride);
/* Binary checkpoint of time and full state as stored (with the FP32 slow
 * concentrations' reference values); restores into a tissue of the same
 * geometry and precision, also from checkpoints of the earlier format with
 * those concentrations in double. */
lqt_status lqt_tissue_save(lqt_tissue *t, const char *path);
lqt_status lqt_tissue_load(lqt_tissue *t, const char *path);
/* Launch shapes of the fused step kernel (workgroup size; whether
 * workgroups take voxel ranges in the XCD-aware order; waves per SIMD the
 * registers are allocated for; whether a wavefront whose voxels are one run
 * along z stages its stencil columns in LDS).  Shape 0 is the default; a 2D
 * sheet always runs shape 0.  No counterpart in the reference, which
 * launches <<<ceil(n/512), 512>>> (GPU3D.cu:1757-1790). */
lqt_status lqt_tissue_variants(int *count);
lqt_status lqt_tissue_variant_info(int i, int *block, int *xcd_map, int *waves, int *lds_stage);
lqt_status lqt_tissue_set_variant(lqt_tissue *t, int i);
lqt_status lqt_tissue_get_variant(const lqt_tissue *t, int *i);
/* Autotuner: runs every launch shape for trial_steps steps from the current
 * state (state, V, guard and clock are restored after each trial, so the
 * simulation is unchanged; activation trackers are held meanwhile), twice,
 * in opposite orders after a warm-up, keeps the fastest and returns it in
 * *best; us[count] (may be NULL) receives each shape's mean time per step
 * over its two trials. */
lqt_status lqt_tissue_autotune(lqt_tissue *t, int trial_steps, int *best, double *us);
/* Mean device time per step of the last call of lqt_tissue_step, in
 * microseconds: one pair of HIP events on the tissue's stream brackets all
 * of that call's launches (one fused step kernel per step, plus the tracker
 * kernel on tracked steps), divided by the step count.  The step kernel
 * refreshes the ghost cells itself, so *us_ghost_kernel is 0. */
lqt_status lqt_tissue_last_kernel_us(const lqt_tissue *t, double *us_step_kernel, double *us_ghost_kernel);
/* Algorithmic HBM bytes of one fused step kernel launch (per voxel state,
 * V and static data, read and written once; per ghost cell its index read
 * and V written). */
lqt_status lqt_tissue_bytes_per_step(const lqt_tissue *t, double *bytes);

/* ---------------------------------------------------------------------------
 * Strands: many 1D transmural strands (OneD.cc:40-448) in one persistent
 * launch, each with its own length, cell types, parameters and pacing.
 * Per strand the update is OneD.cc:315-398 (Jacobi snapshot of V, no-flux
 * ends by copy, S1 on the first cells); the pseudo-ECG of
 * PlotScript/plot1DECG_compare.py:82-105 is reduced on the device.
 */
typedef struct {
    int ncell;            /* 2..512 cells                                   */
    int cell_offset;      /* first entry in the celltype / coeff arrays     */
    int param_index;      /* entry in the params array                      */
    double dx;            /* mm                                             */
    double bcl, stop_time;          /* ms; loop is time <= stop_time        */
    double stim_start, stim_duration, stim_strength;
    int nstim;            /* cells 0..nstim-1 are stimulated                */
    int rec_every;        /* pseudo-ECG sample every rec_every steps        */
    double rec_from;      /* ms; ECG recorded for time >= rec_from          */
    double x0;            /* electrode position (mm from cell 0), on no cell */
    int measure_beat;     /* beat k = floor(t/bcl) in which act/APD/EAD are taken */
    int vrec_every;       /* V of every cell every vrec_every steps from rec_from
                             (the output file of OneD.cc:402-415); 0 = off     */
    /* S2 of OneD_VW.cc:291-301: looked at only while the S1 window is
     * closed; cells s2_first .. s2_first+s2_count-1 get s2_strength while
     * time - s2_base lies in [s2_start, s2_start + s2_duration).  The
     * reference: base (numS1-1)*BCL, duration 3, strength -104, cells 10..12.
     * s2_duration 0 = no S2. */
    double s2_base, s2_start, s2_duration, s2_strength;
    int s2_first, s2_count;
} lqt_strand_desc;

typedef struct {
    double qt;                 /* ms, from the pseudo-ECG                   */
    double apd90[3];           /* ms, mean over ENDO, MCELL, EPI cells      */
    int ead;                   /* 1 if any cell shows an EAD or fails to repolarise
                                  (within the beat, or from the beat before: V above
                                  -60 mV at the measured beat's first step) */
    double act_first, act_last;/* ms, earliest/latest -30 mV crossing      */
    double cv;                 /* m/s between cells at 10% and 90% of length */
    int64_t steps;
    double vmax_first, vmax_last; /* peak V (mV) of cell 0 / cell n-1 over the
                                     steps with time - s2_base >= s2_start, from
                                     -100 (maxEndo/maxEpi of OneD_VW.cc:206-207,390-394) */
} lqt_strand_result;

typedef struct lqt_strands lqt_strands;

/* celltype/coeff indexed by desc.cell_offset + i; coeff may be NULL (0.154).
 * A strand has 2..512 cells and fewer than 2^30 steps (stop_time / dt), and
 * its sample intervals are below 2^30 steps; LQT_ERR_ARG otherwise. */
lqt_status lqt_strands_create(int nstrands, const lqt_strand_desc *desc, const uint8_t *celltype,
                              const double *coeff, const lqt_params *params, int nparams,
                              lqt_precision prec, double dt, lqt_strands **out);
lqt_status lqt_strands_destroy(lqt_strands *s);
/* Runs every strand to its stop time in one persistent launch (captured in
 * a HIP graph on first use).  Resets the strands to the initial state first. */
lqt_status lqt_strands_run(lqt_strands *s);
/* Same run enqueued on the caller's HIP stream (hipStream_t as void*, NULL
 * for the handle's own stream) without waiting.  The first call of the
 * result functions below (or lqt_strands_last_ms, lqt_strands_autotune,
 * lqt_strands_run) waits for that stream and copies the results; a second
 * lqt_strands_run_async before then is refused (LQT_ERR_STATE). */
lqt_status lqt_strands_run_async(lqt_strands *s, void *stream);
lqt_status lqt_strands_results(lqt_strands *s, lqt_strand_result *out);
/* Per-cell activation time, APD90 and EAD flag of strand i (ncell each). */
lqt_status lqt_strands_cell_results(lqt_strands *s, int i, double *act, double *apd90, int *ead);
/* Pseudo-ECG samples of strand i; *n receives the count (call with
 * phi == NULL to query). */
lqt_status lqt_strands_ecg(lqt_strands *s, int i, double *phi, double *t0, int64_t *n);
/* Recorded V rows of strand i (rows*ncell, row-major), time of row 0;
 * *nrows receives the count (call with v == NULL to query). */
lqt_status lqt_strands_vtrace(lqt_strands *s, int i, double *v, double *t0, int64_t *nrows);
/* Final 40-state vectors of strand i (ncell*40). */
lqt_status lqt_strands_final_state(lqt_strands *s, int i, double *out);
/* Device time of the last lqt_strands_run (ms, HIP events). */
lqt_status lqt_strands_last_ms(lqt_strands *s, double *ms);
/* Largest work-group size of the active launch layout (64..512). */
lqt_status lqt_strands_block(const lqt_strands *s, int *block);
/* Launch layouts: mode 0 runs one strand per workgroup, one persistent
 * kernel per block-size class (strands of 1..8 wavefronts, each class at its
 * own block size, all classes as parallel branches of one graph); mode 1 runs
 * one strand per workgroup at the largest block; mode 2 packs strands of
 * similar step counts on whole wavefronts into 512-lane workgroups; mode 3 is
 * mode 0 with every strand of up to 4 wavefronts (256 cells) split: each cell
 * on two lanes of two wavefronts, one per half of the ionic currents, which
 * exchange their currents every step (same results bit for bit, a shorter
 * step when the strands leave SIMDs idle).  The autotuner times every layout
 * on the batch with all stop times scaled so that the longest strand runs
 * trial_steps steps, and keeps the fastest (ms[4] receives the trial
 * times). */
lqt_status lqt_strands_autotune(lqt_strands *s, int trial_steps, int *mode, double *ms);
lqt_status lqt_strands_set_mode(lqt_strands *s, int mode);

/* ---------------------------------------------------------------------------
 * Vulnerable-window search of OneD_VW.cc:199-449 (its MODE1/2/3), each S2
 * timing a strand of one batched launch.  `base` is the strand of one
 * candidate (cells, coefficients, S1 and the S2 stimulus; its s2_start is
 * replaced per candidate).  A candidate is vulnerable when
 * (vmax_first + boundary) * (vmax_last + boundary) <= 0 (OneD_VW.cc:410).
 *   mode 2: s2 = left, left + step, ... (double accumulation, :405-423);
 *           the first vulnerable one is the result;
 *   mode 3: the same going down (:427-443);
 *   mode 1: bisection of [left, right] (:396-404) while right - left >= step
 *           or right == left, S2 too early (both peaks below -boundary)
 *           moving left up, too late (both above) moving right down.
 * Modes 2/3 run `batch` candidates per launch; mode 1 runs the bisection
 * tree floor(log2(batch+1)) levels deep per launch.  The reference loops
 * forever when no candidate is vulnerable: here the search stops after
 * max_candidates visited candidates with found = 0 ("No VM!").
 */
typedef struct {
    int mode;                  /* 1, 2 or 3 */
    double left, right;        /* ms (OneD_VW.cc:201-202: 260, 260)         */
    double step;               /* ms (0.1)                                  */
    double boundary;           /* mV (10, :204)                             */
    int batch;                 /* strands per launch                        */
    int64_t max_candidates;
} lqt_vw_desc;

typedef struct {
    int found;                 /* "VM found!"                               */
    double s2;                 /* S2 start of the last visited candidate    */
    double vmax_first, vmax_last; /* its maxEndo / maxEpi                   */
    int64_t visited;           /* candidates the reference's loop runs      */
    int64_t simulated;         /* strands run on the device                 */
    int launches;
    double device_ms;
} lqt_vw_result;

/* visited_s2[cap] and visited_vmax[2*cap] (either may be NULL) receive the
 * visited sequence, as the reference prints it (s2, maxEndo, maxEpi). */
lqt_status lqt_vw_search(const lqt_strand_desc *base, const uint8_t *celltype, const double *coeff,
                         const lqt_params *params, lqt_precision prec, double dt, const lqt_vw_desc *vw,
                         lqt_vw_result *out, double *visited_s2, double *visited_vmax, int64_t cap);

/* ---------------------------------------------------------------------------
 * Surrogate: no counterpart in the reference (it has no learned component;
 * its dose sweeps are separate 1D runs post-processed in Python,
 * PlotScript/plot1DECG_multiple.py:55-99).  An MLP n_in -> hidden -> hidden
 * -> n_reg + n_bce with GELU; BF16 operands on the matrix cores with FP32
 * accumulation; squared error on the first n_reg outputs (weighted by a
 * mask), binary cross-entropy with logits on the last n_bce outputs; Adam
 * (beta 0.9/0.999, eps 1e-8) on FP32 master weights.
 */
typedef struct lqt_mlp lqt_mlp;
lqt_status lqt_mlp_create(int n_in, int hidden, int n_reg, int n_bce, uint64_t seed, lqt_mlp **out);
lqt_status lqt_mlp_destroy(lqt_mlp *m);
/* Unpadded FP32 masters: w1[hidden*n_in], b1[hidden], w2[hidden*hidden],
 * b2[hidden], w3[n_out*hidden], b3[n_out] (row = output unit). */
lqt_status lqt_mlp_get_params(lqt_mlp *m, float *w1, float *b1, float *w2, float *b2, float *w3, float *b3);
lqt_status lqt_mlp_set_params(lqt_mlp *m, const float *w1, const float *b1, const float *w2, const float *b2,
                              const float *w3, const float *b3);
/* y[n*n_out] = network(x[n*n_in]) (raw outputs; the BCE outputs are logits). */
lqt_status lqt_mlp_forward(lqt_mlp *m, int64_t n, const float *x, float *y);
/* No counterpart in the reference.  For callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses
 * that forbid hipMalloc): the forward pass of lqt_mlp_forward on device
 * buffers, allocating nothing.  x[n*n_in] and y[n*n_out] are device memory
 * on the handle's device, which must be current; the work is enqueued on
 * `stream` (a hipStream_t, NULL for the null stream) and not waited for.
 * The intermediate buffers live in a fixed 8 MiB device array of the
 * library (one per GPU, held until the process ends), so n is limited to
 * floor(8 MiB / (192 + 4 hidden) / 32) * 32 rows, less alignment: 11,904
 * rows at hidden 128; larger n is refused with LQT_ERR_ARG.  Calls that
 * overlap on one GPU share that array and race.  Same kernels, same bits
 * as lqt_mlp_forward. */
lqt_status lqt_mlp_forward_noalloc(lqt_mlp *m, int64_t n, const float *x, float *y, void *stream);
/* Loss and parameter gradients of one batch of n rows, no update. */
lqt_status lqt_mlp_grads(lqt_mlp *m, int n, const float *x, const float *y, const float *mask, float *loss,
                         float *gw1, float *gb1, float *gw2, float *gb2, float *gw3, float *gb3);
/* epochs passes of shuffled batches with Adam; loss_out[epochs*(n/batch)]. */
lqt_status lqt_mlp_train(lqt_mlp *m, int64_t n, const float *x, const float *y, const float *mask, int epochs,
                         int batch, double lr, uint64_t seed, float *loss_out);

/* Surrogate-steered sweep over an IKs-scale x dose x ISO grid of one strand
 * shape (no counterpart in the reference).  The network (trained on inputs
 * [2 iks - 1, 2 iso - 1, 2 dose - 1, (bcl - 1000)/500, (ncell - 225)/175,
 * (mfrac - 0.35)/0.15] and outputs [(QT - 350)/100, APD90 x3 likewise, EAD
 * logit]) predicts every point; each point whose predicted QT is within
 * band_ms of qt_threshold or whose EAD probability lies in (p_lo, p_hi) is
 * re-run as an OneD.cc strand (S1 -52 pA/pF, 1 ms, cells 0-2, from 50 ms;
 * dx 0.15 mm, D 0.154; ENDO (1 - mfrac)/2, then MCELL mfrac, then EPI;
 * SUPREP at dose * dose_max_fmol; PCa x pca_iso under ISO, 0 meaning the
 * TPORd.cc value 1.6) by the batched solver, whose QT and EAD replace the
 * prediction. */
typedef struct {
    double bcl;            /* ms                                          */
    int ncell;             /* strand length, 2..512                       */
    double mfrac;          /* M-cell fraction                             */
    int nbeats;            /* QT, APD90 and EAD of the last beat          */
    double qt_threshold;   /* ms: long QT when QT > qt_threshold          */
    double band_ms;
    double p_lo, p_hi;
    double dose_max_fmol;  /* dose 1 = this many fmol                     */
    int chunk;             /* strands per solver launch; 0 = 4096         */
    int autotune_steps;    /* launch-layout trial steps per chunk; 0 = 2000 (a chunk
                              of at most one strand of <= 128 cells per CU runs
                              the split layout without trials) */
    double pca_iso;        /* PCa multiplier under ISO (TPORd.cc:629 1.6; GPU3D.cu:767 4) */
} lqt_sweep_desc;

typedef struct {
    int64_t points, rechecked;
    double ead_agree;      /* re-checked points whose predicted EAD call was the solver's */
    double t_surrogate_s;  /* inputs, network and decoding                */
    double t_recheck_s;    /* building, tuning and running the re-check strands */
} lqt_sweep_stats;

/* Grid point (a, b, c) = (iks[a], dose[b], iso[c]) at index (a*n_dose + b)*n_iso + c.
 * Per point: qt (ms), ead and long_qt decisions, rechecked flag, and
 * qt_err = predicted - solver QT on re-checked points (NaN elsewhere).  A
 * re-checked point whose strand has no T end within the beat (QT not
 * measurable, lqt_strand_result.qt = -1) gets qt = NaN; an EAD counts as
 * long QT: long_qt = ead || !(qt <= qt_threshold). */
lqt_status lqt_surrogate_sweep(lqt_mlp *net, int n_iks, const double *iks, int n_dose, const double *dose,
                               int n_iso, const int *iso, const lqt_sweep_desc *desc, lqt_precision prec,
                               double *qt, uint8_t *ead, uint8_t *long_qt, uint8_t *rechecked, double *qt_err,
                               lqt_sweep_stats *stats);

/* Host only: QT interval of a pseudo-ECG trace as lqt_strands_results
 * computes it (the reference never defines QT): samples
 * phi[j] at t0 + j*dts; Q at t_q; T peak = largest |phi| in
 * [t_qrs_end, t_end); T end = first later sample with |phi| <= 5 % of the
 * T peak.  *qt = T end - t_q, or -1 if there is no T end. */
lqt_status lqt_ecg_qt(const double *phi, int64_t n, double t0, double dts, double t_q,
                      double t_qrs_end, double t_end, double *qt);

#ifdef __cplusplus
}
#endif
#endif
