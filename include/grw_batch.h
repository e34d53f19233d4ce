/*
 * grw_batch.h -- many independent cells on one GPU.
 *
 * A batch context holds any number of cells, each with its own FRU count,
 * master seed, parameter scales and protocol (pacing, SERCA scale, step
 * control), all started from one initial condition.  Every RK stage of every
 * cell runs in the same launches: the Monte Carlo sub-interval of all FRUs of
 * all cells, each cell's flux reduction, and each cell's global ODE step on
 * the device (the integrator of grw_run_cell_device, one control block per
 * cell).  Cells advance independently: each takes its own RK steps and a
 * launch does nothing for a cell that has finished.  The reference runs one
 * cell per MPI job (caf.c); this has no counterpart there.
 *
 * The per-FRU calls of grw_fru.h act on single-cell contexts only; a batch
 * context accepts grw_destroy, grw_get_stats, grw_reset_stats, grw_set_timing
 * and the calls below.
 */
#ifndef GRW_BATCH_H
#define GRW_BATCH_H

#include <stdint.h>

#include "grw_cell.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int32_t ncells;
  int32_t precision;  /* GRW_FP64 or GRW_MIXED */
  int32_t device;     /* HIP device ordinal */
  int32_t beat_cap;   /* beat records kept per cell (grw_batch_get_beats) */
} grw_batch_config;

typedef struct {
  int32_t nfru;        /* FRUs simulated for this cell (parameters.h:43-46) */
  uint32_t seed;       /* master seed a0 of initialize_ran.c:23-36 */
  double nfru_scale;   /* whole-cell scale (caf.c:88); 0 selects 12500 / nfru */
  double k23_scale;    /* RyR opening-rate scale on k23cf (fru_rates_local.c:103); 1 = reference */
  double gamma_scale;  /* LCC mode-switch scale on gammacf (fru_rates_local.c:70); 1 = reference */
  GrwProtocol prot;    /* pacing, clamp, SERCA scale, integrator controls */
  double t0;           /* the cell's start time (ms): grw_batch_load_ic sets its clock there */
} grw_cell_spec;

int grw_batch_create(const grw_batch_config *cfg, const grw_cell_spec *cells, grw_ctx **out);

/* Initial condition of every cell: global states[ncells][37] (initialize_state.c:
 * 25-316; the caller computes each cell's total-Ca entry for its own scale),
 * error weights errw[37] (initialize_state.c:274-310) and nic FRU records
 * (fru [nic][5], ltype [nic][4][2], ryr [nic][4][5], ito2 [nic][4], the layouts
 * of grw_load_state) that FRU j of every cell takes as record j mod nic
 * (initialize_state.c:38-41,62-65).  Seeds every cell's streams from its master
 * seed and restarts its clock at its spec's t0; under the voltage clamp or the
 * I-V clamp V starts at the holding potential (initialize_state.c:222-228). */
int grw_batch_load_ic(grw_ctx *ctx, const double *states, const double errw[37], int nic,
                      const double *fru, const int32_t *ltype, const int32_t *ryr,
                      const int32_t *ito2);

/* Advance cell k by nsteps[k] output steps of its time_step (caf.c:150-251;
 * 0 leaves it idle), continuing from where the previous run stopped.  On
 * return states[ncells][37] (may be NULL) holds the global states and st[k]
 * (may be NULL) accumulates cell k's RK attempts / rejections / forced accepts. */
int grw_batch_run(grw_ctx *ctx, const int32_t *nsteps, double *states, GrwRkStats *st);

/* Autotuner for a batch: as grw_autotune_cell, timing grw_batch_run of
 * trial_steps output steps of every cell under GRW_LAYOUT_QUAD,
 * GRW_LAYOUT_QUAD3, GRW_LAYOUT_QUAD4, GRW_LAYOUT_WAVE1, GRW_LAYOUT_WAVE1_3,
 * GRW_LAYOUT_WAVE1_4, GRW_LAYOUT_REFILL and GRW_LAYOUT_REFILL_3, with and
 * without graph replays, from a snapshot of the batch (which needs device
 * memory for a second copy of its FRU state). */
int grw_autotune_batch(grw_ctx *ctx, int trial_steps, GrwTune *out);

/* Beat records of cell k, as grw_get_beats. */
int grw_batch_get_beats(grw_ctx *ctx, int cell, double *out, int cap, int *n, int include_current);

#ifdef __cplusplus
}
#endif
#endif
