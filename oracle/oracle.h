/*
 * FP64 CPU oracle for the TPORd monodomain hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links or calls
 * this code; it is loaded by tests/, by the smoke check and by the CPU
 * baseline leg of the benchmark, always as the reference that results are
 * compared with.
 *
 * It restates, in plain C, the arithmetic of the reference's CPU path:
 *   SingleCell/TPORd.cc:142-198   (init)      -> or_cell_init
 *   SingleCell/TPORd.cc:379-1021  (update)    -> or_cell_update
 *   SingleCell.cc:95-232          (protocol)  -> or_single_cell
 *   OneD.cc:100-131,315-436       (strand)    -> or_strand_run
 *   OneD_VW.cc:239-398            (S2 candidate) -> or_strand_vw
 *   GPU3D.cu:258-308,370-385,1222-1230, new_11.cu:352-417 (tissue stencil)
 *                                             -> or_tissue_*
 *   GPU3D.cu:1487-1680            (ghost map) -> or_ghost_map
 * Compile-time flags of the reference (CON/LQT1/ISO1/SUPREP) become the
 * runtime or_params below.
 */
#ifndef LQT_ORACLE_H
#define LQT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* State vector in the reference's 40-line order (TPORd.cc:200-252). */
enum {
    OR_V = 0, OR_NAI, OR_NASS, OR_KI, OR_KSS, OR_CAI, OR_CASS, OR_CANSR, OR_CAJSR,
    OR_SM, OR_SH, OR_SJ, OR_HSP, OR_JP, OR_ML, OR_HL, OR_HLP,
    OR_A, OR_IF, OR_IS, OR_AP, OR_IFP, OR_ISP,
    OR_D, OR_FF, OR_FS, OR_FCAF, OR_FCAS, OR_JCA, OR_NCA, OR_FFP, OR_FCAFP,
    OR_XRF, OR_XRS, OR_XS1, OR_XS2, OR_XK1, OR_JRELNP, OR_JRELP, OR_CAMKT,
    OR_NSTATE
};

/* enum CellType of SingleCell/Cell.cc:20 */
enum { OR_LVEPI = 0, OR_LVENDO = 1, OR_EPI = 2, OR_ENDO = 3, OR_MCELL = 4 };

typedef struct {
    double iks_scale;   /* LQT1 factor (TPORd.cc:705-707 uses 0.05); 1.0 = off  */
    int    iso;         /* ISO1 flag (TPORd.cc:583-589,628-630,731-733)        */
    double pca_iso;     /* PCa multiplier under ISO1: 1.6 CPU, 4 in GPU3D.cu   */
    int    suprep;      /* SUPREP gene-therapy block (TPORd.cc:709-728)        */
    double dose_fmol;   /* SUPREP dose in fmol                                 */
    double sca;         /* GKs base multiplier `sca` (TPORd.cc:45,696)         */
    int    clamp;       /* CLAMP flag: voltage held (TPORd.cc:28,929-933)      */
} or_params;

/* Diagnostics of the last update (the getters of TPORd.cc:358-376). */
typedef struct {
    double Iion, dV_dt, ICaL, IKs, IKr, IK1, INa;
} or_diag;

void or_params_default(or_params *p);  /* the shipped CPU build: CON + ISO1 */
void or_cell_init(double *s, double v0);
void or_cell_update(double *s, int celltype, const or_params *p, double dt,
                    double ist, double dvgap, or_diag *diag);
/* n cells with a stride of OR_NSTATE; one update each (no coupling). */
void or_cells_update(int n, double *s, const uint8_t *celltype, const or_params *p,
                     double dt, const double *ist, const double *dvgap);

/* SingleCell.cc:38-287.  stats[10] = rest, dvdt_max, overshoot, amplitude,
 * APD20, APD25, APD50, APD75, APD80, APD90.  If trace != NULL, V of every
 * `trace_every`-th step is written there (at most trace_cap values).
 * Returns the number of steps run. */
int64_t or_single_cell(int celltype, const or_params *p, double bcl, int nbeats,
                       double dt, double stim_strength, double stim_dur,
                       double stim_start, double *stats,
                       double *trace, int trace_every, int64_t trace_cap);

/* OneD.cc:40-448.  One strand of n cells.  The loop is the reference's
 * `for (time = 0; time <= stop; time += dt)` with the same double
 * accumulator.  V of all cells is written every `rec_every` steps into
 * vrec[row*n + i] (rows at most rec_cap); rec_t receives the time.
 * act[i] = first time cell i reaches -30 mV in beat `act_beat`
 * (floor(time/BCL) == act_beat, OneD.cc:419-436), -1 if never.
 * Returns the number of steps run; *nrows the number of recorded rows. */
int64_t or_strand_run(int n, const uint8_t *celltype, const double *coeff,
                      const or_params *p, double dx, double dt, double bcl,
                      double stop_time, double stim_strength, double stim_dur,
                      double stim_start, int nstim_cells,
                      double *vrec, double *rec_t, int rec_every, int64_t rec_cap,
                      int64_t *nrows, int act_beat, double *act,
                      double *final_state);

/* Per-cell beat measurements inside beat k = floor(t/BCL) (the reference
 * defines only the -30 mV activation, OneD.cc:422; APD90 and the EAD test
 * are the library's own definitions, restated here independently of the
 * kernel):  rest = V at the first step of the beat; act = first t with
 * V >= -30 mV; vmax = running peak after act; repolarised at the first t
 * with V <= vmax - 0.9 (vmax - rest), APD90 = t_rep - act; EAD when, from
 * act + 100 ms until repolarisation, V rises more than 2 mV above its
 * running minimum; flag 2 when activated but not repolarised. */
typedef struct {
    double act, rep;
    double vrest, vmax, vmin;
    int flags; /* bit0 entered, bit1 activated, bit2 repolarised, bit3 EAD,
                  bit4 entered depolarised (V > -60 mV at the first step) */
} or_track;
void or_track_init(or_track *t);
void or_track_step(or_track *t, double time, double v);
/* apd = t_rep - act or -1; returns the EAD flags (1 EAD, 2 no repolarisation:
 * within the beat, or V above -60 mV at its first step). */
int or_track_result(const or_track *t, double *act, double *apd);

/* or_strand_run plus the beat measurements of every cell in beat
 * `act_beat` (apd[n], ead[n]; either may be NULL). */
int64_t or_strand_run2(int n, const uint8_t *celltype, const double *coeff,
                       const or_params *p, double dx, double dt, double bcl,
                       double stop_time, double stim_strength, double stim_dur,
                       double stim_start, int nstim,
                       double *vrec, double *rec_t, int rec_every, int64_t rec_cap,
                       int64_t *nrows, int act_beat, double *act, double *apd, int *ead,
                       double *final_state);

/* S2 of OneD_VW.cc:291-301: while the S1 window is closed and
 * time - base lies in [start, start + duration), cells first ..
 * first+count-1 get `strength` (the reference: base (numS1-1)*BCL, start the
 * candidate s2startTime, duration 3 ms, strength -104, cells 10..12). */
typedef struct {
    double base, start, duration, strength;
    int first, count;
} or_s2;

/* OneD_VW.cc:239-398 for one S2 candidate, resumable: advances the caller's
 * state s[n*40] from (*time, *step) while time <= stop_time and
 * time < pause_at, and leaves (*time, *step) at the first step not run.
 * vmax[2] (in/out; the reference starts both at -100) takes the peak V of
 * cell 0 and cell n-1 after each update whose time - base >= start
 * (OneD_VW.cc:390-394).  V rows are recorded as OneD_VW.cc:370-385 writes
 * them: steps with step % rec_every == 0 inside beat rec_beat
 * (floor(time/BCL) == rec_beat; -1 = every beat).  Returns *step. */
int64_t or_strand_vw(int n, const uint8_t *celltype, const double *coeff,
                     const or_params *p, double dx, double dt, double bcl,
                     double stop_time, double stim_strength, double stim_dur,
                     double stim_start, int nstim, const or_s2 *s2,
                     double *s, double *time, int64_t *step, double pause_at,
                     double *vmax, double *vrec, double *rec_t, int rec_every,
                     int rec_beat, int64_t rec_cap, int64_t *nrows);

/* Pseudo-ECG of PlotScript/plot1DECG_compare.py:89-105 over T rows of L cells. */
void or_pseudo_ecg(int64_t T, int L, const double *v, double dx, double x0, double *phi);

/* ---- tissue ------------------------------------------------------------ */

/* Ghost map of GPU3D.cu:1301-1333 and 1487-1604 on an nx*ny*nz grid with
 * x slowest and z fastest (index = (x*ny + y)*nz + z); with nz == 1 the 2D
 * sheet rule of new_11.cu:1535-1604 (8 neighbours).  mask[c] != 0 marks
 * tissue.  On return g[c] = k+1 for the k-th tissue voxel (x,y,z order),
 * -(k+1) for a ghost cell mirroring tissue voxel k, 0 otherwise.
 * Returns the number of tissue voxels. */
int64_t or_ghost_map(int nx, int ny, int nz, const uint8_t *mask, int32_t *g);

/* Per-voxel diffusion tensor D = Dc*I + (Df-Dc) f f^T (new_11.cu:1496-1499
 * generalised to 3D) and its one-sided divergence b_j = sum_i dD_ij/dx_i
 * (new_11.cu:381-407).  dten[6k..] = Dxx,Dyy,Dzz,Dxy,Dxz,Dyz, drift[3k..]. */
void or_tissue_tensor(int nx, int ny, int nz, const int32_t *g, int64_t nvox,
                      const float *fibre, double d_fibre, double d_cross,
                      double dx, double *dten, double *drift);

/* du[k] = div(D grad V) at tissue voxel k with the 19-point stencil; V of a
 * neighbour is read through g (tissue, ghost mirror, or -90 mV outside,
 * GPU3D.cu:258-308).  With nz == 1 the sheet is 2D and the z terms are
 * absent (new_11.cu:352-417, 9-point). */
void or_tissue_du(int nx, int ny, int nz, const int32_t *g, int64_t nvox,
                  const double *dten, const double *drift, double dx,
                  const double *v, double *du);

/* One explicit step of the whole tissue: du from V^n, then the cell update
 * of every voxel (OpenMP over voxels).  s has OR_NSTATE doubles per voxel;
 * ist[k] is the stimulus current of voxel k for this step.  celltype[k]
 * with bit 0x80 set: updated as celltype & 0x7f, then V = -100 mV (the
 * type-0 voxels of new_11.cu:1247-1248). */
void or_tissue_step(int nx, int ny, int nz, const int32_t *g, int64_t nvox,
                    const double *dten, const double *drift, double dx, double dt,
                    const uint8_t *celltype, const or_params *p,
                    double *s, const double *ist, double *scratch_v, double *scratch_du);

int or_num_threads(void);
void or_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
