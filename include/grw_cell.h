/*
 * grw_cell.h -- whole-cell driver over the release-unit library: the global
 * 37-state model (fcn.c) integrated by the Prince-Dormand RK5(4)6M of
 * rk_pd54m.c, with one Monte Carlo sub-interval of every FRU per RK stage.
 *
 * This is the caller side of the boundary in include/grw_fru.h: it performs,
 * through that ABI, the same sequence of calls the reference's rk54pd makes
 * into mpi_master.c.  Protocol switches that the reference fixes at compile
 * time (parameters.h:92-163) are fields of GrwProtocol.
 */
#ifndef GRW_CELL_H
#define GRW_CELL_H

#include <stdint.h>

#include "grw_fru.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  /* current-clamp stimulus, parameters.h:114-117 */
  double pulse_duration, pulse_amplitude, period, shift;
  /* voltage clamp, parameters.h:120-125 (vclamp != 0 enables) */
  int32_t vclamp;
  int32_t chf; /* heart-failure scalings, parameters.h:150-156 */
  double vclamp_duration, vclamp_set, vclamp_shift, vclamp_hold, vclamp_period;
  /* SERCA Vmax scale on vmaxf and vmaxr (fcn.c:218-219); 1 = reference */
  double serca_scale;
  /* integrator controls, parameters.h:95-98 */
  double time_step, step_min, step_max, tolrk;
  /* voltage clamp with a prepulse, parameters.h:128-137 (vppclamp != 0
     enables; fcn.c:485-516, 825-838) */
  int32_t vppclamp;
  /* voltage clamp for I-V studies, parameters.h:141-147 (ivclamp != 0
     enables; fcn.c:521-535); the test potential is iv_set + iv_n * iv_step */
  int32_t ivclamp;
  double vpp_ppset, vpp_ppduration, vpp_shift, vpp_set, vpp_duration, vpp_hold, vpp_period;
  double iv_duration, iv_period, iv_shift, iv_set, iv_hold, iv_step;
  /* the I-V step number (caf.c:31,140,234: iv_n, counted up at every
     iv_period restart) */
  int32_t iv_n;
  int32_t reserved;
} GrwProtocol;

/* Per-run counters of the integrator. */
typedef struct {
  int64_t attempts, rejects, forced;
} GrwRkStats;

void grw_protocol_default(GrwProtocol *p);

/* fcn.c:29-1072 on the host: derivatives F[37] of the global state for the
 * release-unit fluxes J = {Jxfer, Jtr, ICa, Ito2}; cur[24] (may be NULL) is
 * filled when keepc is nonzero.  Mutates state[1] below -90 mV and state[0]
 * under voltage clamp, as the reference does. */
int grw_fcn(const GrwProtocol *p, double t, double state[37], double F[37], double cur[24],
            int keepc, const double J[4]);

/* rk_pd54m.c:29-345 with the release units in ctx.  Integrates state[37] from
 * t to tstep starting with step oldh; returns the next step size in *newh and
 * the currents of the first accepted stage in cur[24] (fcn.c:1046-1071).
 * errw[37] are the error weights (initialize_state.c:274-310).  If log is not
 * NULL, each accepted step appends (t, V, Cai, ICa) until log_cap rows. */
int grw_rk54pd(grw_ctx *ctx, const GrwProtocol *p, const double errw[37], double t, double tstep,
               double oldh, double state[37], double cur[24], double *newh, double *log,
               int log_cap, int *log_n, GrwRkStats *st);

/* caf.c:150-251: from t0, nsteps calls of rk54pd of length time_step each; after
 * every call row k of states[nsteps][37], currents[nsteps][24] (currents of the
 * call's first stage, i.e. at its start) and other[nsteps][12]
 * (send_calc_fru_avg) is written.  Any output pointer may be NULL.  *h carries
 * the step size across calls (start with step_max). */
int grw_run_cell(grw_ctx *ctx, const GrwProtocol *p, const double errw[37], double t0, int nsteps,
                 double state[37], double *h, double *states, double *currents, double *other,
                 double *log, int log_cap, int *log_n, GrwRkStats *st);

/* grw_run_cell with the integrator on the device: the same results and
 * arguments, but rk54pd and fcn run in the tails of the Monte Carlo launches
 * and the host only enqueues batches of RK attempts (no host round trip per
 * sub-interval).  Uses the four-lanes-per-FRU kernel. */
int grw_run_cell_device(grw_ctx *ctx, const GrwProtocol *p, const double errw[37], double t0,
                        int nsteps, double state[37], double *h, double *states,
                        double *currents, double *other, double *log, int log_cap, int *log_n,
                        GrwRkStats *st);

/* The Monte Carlo kernel of the device integrator (grw_run_cell_device,
 * grw_batch_run):
 *   GRW_LAYOUT_QUAD   four lanes per FRU (one per cleft), 64 FRUs per block,
 *                     registers as the compiler allocates them (the default);
 *   GRW_LAYOUT_WAVES  single-cell contexts only: 16 FRUs per block whose four
 *                     wavefronts take the LCC, the two halves of the RyRs and
 *                     the Heun step of every micro-step, exchanging through LDS;
 *   GRW_LAYOUT_QUAD3, GRW_LAYOUT_QUAD4  the four-lane kernel with its registers
 *                     capped for three or four wavefronts per SIMD (latency
 *                     hiding for batches, against spills);
 *   GRW_LAYOUT_QUEUE  the four-lane kernel's work as a queue of 16-FRU items
 *                     that each wavefront takes one after another, then one
 *                     block per cell for the cell's fcn and integrator stage;
 *                     switched off: grw_set_layout refuses it (a device run of
 *                     it did not finish and the cause is not known);
 *   GRW_LAYOUT_WAVE1  the four-lane kernel's work with one wavefront (16 FRUs)
 *                     per block, so that a wavefront's registers are free for
 *                     the next block as soon as its own FRUs are done, then one
 *                     block per cell for the cell's fcn and integrator stage;
 *   GRW_LAYOUT_WAVE1_3, GRW_LAYOUT_WAVE1_4  the same with its registers
 *                     capped for three or four wavefronts per SIMD;
 *   GRW_LAYOUT_QUADP  single-cell contexts only: the four-lane kernel whose
 *                     Monte Carlo blocks leave their partial sums without the
 *                     last-block election, and the cell's fcn block collects
 *                     them as they arrive and runs the integrator stage;
 *   GRW_LAYOUT_REFILL one wavefront per 64 FRUs whose 16 quads each take the
 *                     block's next FRU as soon as their own is done (so that
 *                     FRU lanes do not wait for the slowest FRU of a fixed
 *                     group), then one block per cell for the cell's fcn and
 *                     integrator stage;
 *   GRW_LAYOUT_REFILL_3  the same with its registers capped for three
 *                     wavefronts per SIMD.
 * Every layout gives the same results bit for bit; the speed differs with the
 * workload (grw_autotune_cell / grw_autotune_batch time them).  No counterpart
 * in the reference. */
#define GRW_LAYOUT_QUAD 1
#define GRW_LAYOUT_WAVES 2
#define GRW_LAYOUT_QUAD3 3
#define GRW_LAYOUT_QUAD4 4
#define GRW_LAYOUT_QUEUE 5
#define GRW_LAYOUT_WAVE1 6
#define GRW_LAYOUT_WAVE1_3 7
#define GRW_LAYOUT_QUADP 8
#define GRW_LAYOUT_WAVE1_4 9
#define GRW_LAYOUT_REFILL 10
#define GRW_LAYOUT_REFILL_3 11
int grw_set_layout(grw_ctx *ctx, int layout);

/* Result of an autotuning pass: the choice (now set on the context) and every
 * candidate timed, with the wall time of its trial run. */
#define GRW_TUNE_MAX 16
typedef struct {
  int32_t layout, graphs; /* the fastest candidate (grw_set_layout, grw_set_graphs) */
  int32_t ncand, reserved;
  int32_t cand_layout[GRW_TUNE_MAX], cand_graphs[GRW_TUNE_MAX];
  double cand_ms[GRW_TUNE_MAX];
} GrwTune;

/* Autotuner of the device integrator for a single cell: times trial_steps
 * output steps of grw_run_cell_device from (t0, state, h) under each of
 * GRW_LAYOUT_QUADP, GRW_LAYOUT_QUAD, GRW_LAYOUT_WAVES, GRW_LAYOUT_QUAD3 and
 * GRW_LAYOUT_WAVE1, with and without graph replays, every trial from a
 * snapshot of the context's device state (a warm-up step before each);
 * restores the snapshot, so the context's FRU state, beats and counters are as
 * they were, and sets the fastest candidate (a later candidate must beat the
 * one kept by more than 1%: ties keep the first, GRW_LAYOUT_QUADP without
 * graphs).
 * The layouts give identical results, so the choice changes only the speed.
 * No counterpart in the reference. */
int grw_autotune_cell(grw_ctx *ctx, const GrwProtocol *p, const double errw[37], double t0,
                      const double state[37], double h, int trial_steps, GrwTune *out);

/* Per-beat statistics of the device-driven runs of ctx since its state was
 * last loaded or seeded (the reference writes only 1 ms samples,
 * data_io.c:22-140; these come from every accepted RK step).  Beat b starts at
 * the stimulus t_s = shift + b * period and holds the accepted steps ending in
 * [t_s, t_s + period).  Each record is GRW_BEAT_STATS doubles:
 *   0 t_s (ms)             1 V at the t_s output sample (rest, mV)
 *   2 peak V               3 time of peak V - t_s
 *   4 APD90: time - t_s at which V, after its peak, falls through
 *     V_peak - 0.9 (V_peak - V_rest), linear between accepted steps (NaN if
 *     not within the beat, or if the peak is under 20 mV above rest)
 *   5 peak Cai (mM)        6 time of peak Cai - t_s
 *   7 Cai at t_s           8 peak (most negative) ICa of the FRU flux (uA/uF)
 *   9 CaNSR at t_s (SR load, mM)
 *  10 mean CaJSR at t_s (JSR load, send_calc_fru_avg output 1, mM)
 *  11 accepted steps in the beat
 *  12 Ca entered through the LCCs over the beat, -ICa Acap / (2 Vmyo F) (mM of
 *     myoplasm; fcn.c:1058), 13 Ca released by the RyRs over the beat, the
 *     whole-cell JRyR of send_calc_fru_avg (mpi_master.c:178; mM of myoplasm),
 *     both trapezoids over the accepted steps (the first from the step before
 *     t_s); 13 / 12 is the EC-coupling gain
 *  14 peak JRyR (mM/ms)
 * Under a voltage-clamp protocol the beats are its periods: t_s = vclamp_shift
 * + b * vclamp_period (vpp_shift / vpp_period with the prepulse, iv_shift /
 * iv_period for the I-V steps, which take precedence in that order).  Writes up to cap finished beats (and, if
 * include_current, the beat in progress) to out and their count to *n. */
#define GRW_BEAT_STATS 15
int grw_get_beats(grw_ctx *ctx, double *out, int cap, int *n, int include_current);

#ifdef __cplusplus
}
#endif
#endif
