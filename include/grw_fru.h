/*
 * grw_fru.h -- C ABI of the MI355X (gfx950) release-unit Monte Carlo.
 *
 * One context owns the device-resident state of one cell: NFRU functional
 * release units (a JSR and four dyadic clefts each, every cleft with one LCC,
 * five RyRs and one Ito2 channel) and their per-FRU MT19937 streams.  The calls
 * replace the master-side interface of the reference's MPI layer
 * (mpi_master.c) and the per-FRU routines it distributes (simfru_local.c,
 * calc_fru_flux_local.c, calc_fru_avg_local.c); a host global integrator calls
 * them in the same order the reference's rk_pd54m.c does.
 *
 * Conventions: every function returns 0 on success and a negative code on
 * failure (grw_last_error() describes the last failure of the calling thread);
 * nothing aborts.  Array arguments are plain host pointers in the reference's
 * C layouts (row-major, one row per FRU).  Calls on one context are serialised
 * by the caller; separate contexts are independent.
 */
#ifndef GRW_FRU_H
#define GRW_FRU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GRW_OK 0
#define GRW_EINVAL (-1)   /* bad argument or invalid channel state */
#define GRW_EDEVICE (-2)  /* HIP runtime error */
#define GRW_ESTATE (-3)   /* kernel met an unknown channel state */

#define GRW_FP64 0   /* plain FP64 drop-in */
#define GRW_MIXED 1  /* FP32 rates and local-Ca right-hand side, FP64 state/time/thresholds */

typedef struct grw_ctx grw_ctx;

typedef struct {
  int32_t nfru;          /* FRUs simulated: NFRU_sim (parameters.h:43-46, caf.c:86) */
  int32_t precision;     /* GRW_FP64 or GRW_MIXED */
  double nfru_scale;     /* whole-cell scale factor NFRU_scale (caf.c:88), 12500/nfru */
  int32_t block;         /* FRUs per workgroup; 0 selects the tuned default */
  int32_t sequential_sums; /* 1: reduce in the reference's sequential FRU order
                              (calc_fru_flux_local.c:45-70) for bit-parity runs */
  double k23_scale;      /* RyR opening-rate scale on k23cf (fru_rates_local.c:103); 1 = reference */
  double gamma_scale;    /* LCC mode-switch scale on gammacf (fru_rates_local.c:70); 1 = reference */
  int32_t device;        /* HIP device ordinal */
  int32_t lanes_per_fru; /* 4: one lane per cleft (default); 1: one lane per FRU */
} grw_config;

typedef struct {
  uint64_t fru_steps;    /* micro-steps summed over FRUs (simfru_local.c:127 iterations) */
  uint64_t simfru_calls; /* FRU sub-interval calls (FRUs x distrib calls) */
  uint64_t launches;     /* Monte Carlo kernel launches */
  double kernel_ms;      /* summed device time of the timed Monte Carlo launches */
  uint64_t critical_steps; /* sum over launches of the largest FRU-step count in the launch */
  double tail_ms;        /* summed device time of the serial launch tails (integrator) */
  uint64_t timed_launches; /* launches that kernel_ms covers: every launch of the host-driven
                              path; the device integrator times the five launches of one RK
                              attempt in eight, a systematic sample */
  uint64_t slot_steps;   /* sum over wavefronts of their FRU slots times their slowest FRU's
                            micro-steps: fru_steps / slot_steps is the fraction of FRU lanes
                            doing micro-steps while their wavefronts run; not counted (0) by
                            the single-cell polling layout, where it would lengthen every
                            launch's critical path */
} grw_stats;

const char *grw_last_error(void);
void grw_config_default(grw_config *cfg);

/* Allocate device state for cfg->nfru FRUs.  Replaces initialize_mpi
 * (mpi_master.c:625-635): no processes, one device. */
int grw_create(const grw_config *cfg, grw_ctx **out);
/* Replaces end_mpi (mpi_master.c:643-651). */
int grw_destroy(grw_ctx *ctx);

/* Upload FRU state.  Replaces initialize_mpi_state (mpi_master.c:199-361).
 * fru [n][5] (CaJSR, CaSS1..4), ltype [n][4][2] (LCC state 1..12, gate 1|2),
 * ryr [n][4][5] (1..8), ito2 [n][4] (1|2), mti [n] (0..624), mt [n][624]
 * (32-bit words; the reference stores them in unsigned long). */
int grw_load_state(grw_ctx *ctx, const double *fru, const int32_t *ltype, const int32_t *ryr,
                   const int32_t *ito2, const int32_t *mti, const uint32_t *mt);
/* Download FRU state in the same layouts.  Replaces parallel_get_FRUs
 * (mpi_master.c:509-617).  Any pointer may be NULL to skip that array. */
int grw_get_state(grw_ctx *ctx, double *fru, int32_t *ltype, int32_t *ryr, int32_t *ito2,
                  int32_t *mti, uint32_t *mt);
/* Seed every FRU's MT19937 on the device from master seed a0, as
 * initialize_ran.c:23-36 does with a0 = 12350232 (the gcc -O3 reading of its
 * seed formula: seed_j = a0 + 23 j(j+1)/2 + 4093 j^2 + 1 mod 2^32). */
int grw_seed(grw_ctx *ctx, uint32_t a0);

/* Advance every FRU over [t0, t1] with (V, Cai, CaNSR) interpolated linearly
 * from dep0 to depf, then reduce and return {Jxfer, Jtr, ICa, Ito2} scaled to
 * the whole cell, evaluated with depf.  Replaces distrib_simFRU
 * (mpi_master.c:369-469); per FRU it is simfru_local (simfru_local.c:32-278). */
int grw_distrib_simfru(grw_ctx *ctx, double t0, double t1, const double dep0[3],
                       const double depf[3], double out4[4]);
/* Fluxes of the current FRU state at dep = (V, Cai, CaNSR).  Replaces
 * send_calc_fru_flux (mpi_master.c:41-113). */
int grw_calc_flux(grw_ctx *ctx, const double dep[3], double out4[4]);
/* The 12 averaged FRU outputs for global state[37].  Replaces
 * send_calc_fru_avg (mpi_master.c:122-191). */
int grw_calc_avg(grw_ctx *ctx, const double state[37], double other[12]);
/* Checkpoint / roll back FRU state incl. RNG.  Replace send_save_state and
 * send_resume_state (mpi_master.c:477-501, mpi_slave.c:150-226). */
int grw_save_state(grw_ctx *ctx);
int grw_resume_state(grw_ctx *ctx);

/* Per-FRU surface without the whole-cell conversion.
 * grw_simfru_local: simfru_local (simfru_local.c:32-278) on every FRU.
 * grw_calc_fru_flux_local: the 5 sums of calc_fru_flux_local.c:28-77.
 * grw_calc_fru_avg_local: the 11 sums of calc_fru_avg_local.c:27-105. */
int grw_simfru_local(grw_ctx *ctx, double t0, double t1, const double dep0[3],
                     const double depf[3]);
int grw_calc_fru_flux_local(grw_ctx *ctx, double num5[5]);
int grw_calc_fru_avg_local(grw_ctx *ctx, double num11[11]);

/* The five sums of calc_fru_flux_local.c:28-77 without any allocation: an
 * entry point with no counterpart in the reference, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops).
 * Inputs are nfru FRUs in the reference's layouts in DEVICE memory (fru
 * [n][5], ltype [n][4][2], ito2 [n][4]); the sums go to num5 (DEVICE, 5
 * doubles); the kernel is enqueued on stream (a hipStream_t; NULL: the null
 * stream) and the call returns without waiting.  Its reduction keeps its block
 * partials and ticket in a fixed device array of the library (1 MB per GPU,
 * held until the process ends), so nfru is at most 1,048,576 (GRW_EINVAL
 * beyond).  The sums equal grw_calc_fru_flux_local's on the same FRUs bit for
 * bit.  Calls that overlap on one GPU share that array and race: serialise
 * them (one stream, or events). */
int grw_calc_fru_flux_local_noalloc(int nfru, const int32_t *ltype, const int32_t *ito2,
                                    const double *fru, double *num5, void *stream);

int grw_get_stats(grw_ctx *ctx, grw_stats *st);
int grw_reset_stats(grw_ctx *ctx);
/* Enable device timing with HIP events: every launch of the host-driven path
 * (adds a sync per launch), one RK attempt in eight on the device integrator. */
int grw_set_timing(grw_ctx *ctx, int on);
/* Device integrator (grw_run_cell_device, grw_batch_run): launch seven of
 * every eight RK attempts as one replay of a HIP graph captured on first use
 * (the eighth, the one timed, is launched directly).  Results are identical
 * either way.  No counterpart in the reference, whose attempts are host calls
 * (rk_pd54m.c:93-345). */
int grw_set_graphs(grw_ctx *ctx, int on);

#ifdef __cplusplus
}
#endif
#endif
