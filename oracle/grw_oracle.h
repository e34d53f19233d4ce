/*
 * grw_oracle.h -- CPU restatement of the Greenstein-Winslow 2002 release-unit
 * Monte Carlo, used ONLY as test infrastructure.
 *
 * This library is the checker for the HIP implementation in greenstein2002_amd/.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it; nothing in the product path calls it.  Every function names the reference
 * file:line it restates.  It is compiled with gcc -O3 -ffp-contract=off so that
 * its double arithmetic follows the reference's own build (gcc -O3 on x86-64,
 * no FMA contraction) operation for operation.
 *
 * Parity pins (tests/test_oracle_*.py):
 *   mt19937 / seeding  -> std::mt19937 known answer, reference objects, fixtures
 *   rates / fcn_fru    -> reference objects (bit-exact), invariants
 *   simfru step        -> reference objects (bit-exact traces), fixtures
 *   flux / avg sums    -> reference objects, brute force
 *   global fcn         -> reference fcn.o (bit-exact)
 *   flux / avg outputs -> the reference's send_calc_fru_flux, distrib_simFRU and
 *                         send_calc_fru_avg (mpi_master.c, one MPI rank)
 *   rk54pd driver      -> the reference's rk54pd (rk_pd54m.c) call by call: the
 *                         same sub-intervals, saves, rollbacks and states
 */
#ifndef GRW_ORACLE_H
#define GRW_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NCLEFT 4
#define ORC_NRYR 5
#define ORC_MTN 624
#define ORC_NGLOBAL 37
#define ORC_NCUR 24
#define ORC_NOTHER 12

/* One MT19937 stream (mt19937_local.c): 624 32-bit words plus the read index. */
typedef struct {
  uint32_t w[ORC_MTN];
  int32_t mti;
} orc_mt;

/* State of one functional release unit (mpi_slave.c:31-45, one row). */
typedef struct {
  double conc[5];               /* [0] CaJSR, [1..4] CaSS of clefts 0..3 (mM) */
  int32_t lcc[ORC_NCLEFT];      /* LCC gating state 1..12 */
  int32_t yinact[ORC_NCLEFT];   /* voltage gate: 1 inactivated, 2 permissive */
  int32_t ryr[ORC_NCLEFT][ORC_NRYR]; /* RyR state 1..6, virtual 7 (open) / 8 */
  int32_t ito2[ORC_NCLEFT];     /* 1 closed, 2 open */
  orc_mt rng;
} orc_fru;

/* Output of the rate evaluation (fru_rates_local.c:30-44). */
typedef struct {
  double lcc_rate[ORC_NCLEFT][4];
  int32_t lcc_dest[ORC_NCLEFT][4];
  int32_t lcc_len[ORC_NCLEFT];
  double vinact_rate[ORC_NCLEFT];
  double ryr_rate[ORC_NCLEFT][ORC_NRYR][4];
  int32_t ryr_dest[ORC_NCLEFT][ORC_NRYR][4];
  int32_t ryr_len[ORC_NCLEFT][ORC_NRYR];
  double ito2_rate[ORC_NCLEFT];
  double max_rate;
} orc_rates;

/* Per-micro-step record written by orc_simfru when a trace buffer is given. */
typedef struct {
  double t;          /* time at the end of the step */
  double dt;         /* step length */
  double conc[5];    /* concentrations after the Heun update */
  int32_t lcc[ORC_NCLEFT];
  int32_t yinact[ORC_NCLEFT];
  int32_t ryr[ORC_NCLEFT][ORC_NRYR];
  int32_t ito2[ORC_NCLEFT];
  int32_t mti;       /* reference-format read index after the step */
  int32_t pad;
} orc_step;

/* Runtime form of the compile-time protocol switches (parameters.h:107-163). */
typedef struct {
  double pulse_duration, pulse_amplitude, period, shift;       /* :114-117 */
  int32_t vclamp;                                               /* :120 */
  int32_t chf;                                                  /* :150 */
  double vclamp_duration, vclamp_set, vclamp_shift, vclamp_hold, vclamp_period; /* :121-125 */
  double serca_scale;     /* multiplies vmaxf and vmaxr (fcn.c:218-219); 1 = reference */
  double k23_scale;       /* multiplies k23cf (fru_rates_local.c:103); 1 = reference */
  double gamma_scale;     /* multiplies gammacf (fru_rates_local.c:70); 1 = reference */
  int32_t vppclamp;       /* :128-137, clamp with a prepulse */
  int32_t ivclamp;        /* :141-147, I-V steps */
  double vpp_ppset, vpp_ppduration, vpp_shift, vpp_set, vpp_duration, vpp_hold, vpp_period;
  double iv_duration, iv_period, iv_shift, iv_set, iv_hold, iv_step;
  int32_t iv_n;           /* caf.c:31: the I-V step number */
  int32_t pad;
} orc_protocol;

/* ---- MT19937 (mt19937_local.c) ---- */
void orc_mt_seed(orc_mt *m, uint32_t seed);                 /* sgrnd_local :75-92 */
uint32_t orc_mt_next_u32(orc_mt *m);                        /* genrand word */
double orc_mt_uniform(orc_mt *m);                           /* MersenneTwisterOne_local :202-240 */
void orc_mt_block(orc_mt *m, int n, double *out);           /* MersenneTwister_fast :97-154 */
uint32_t orc_seed_for_fru(uint32_t a0, int64_t j);          /* initialize_ran.c:30-34 (gcc -O3 form) */

/* ---- per-FRU kernel pieces ---- */
double orc_fru_rates(orc_fru *f, const double dep[3], orc_rates *r,
                     const orc_protocol *p);                /* fru_rates_local.c:30-798 */
void orc_fcn_fru(const double conc[5], const double dep[3], const double lcc_open[4],
                 const int32_t nryr_open[4], double dconc[5]); /* fcn_fru.c:23-137 */
int orc_simfru(double t0, double t1, const double dep0[3], const double depf[3],
               orc_fru *f, const orc_protocol *p, orc_step *trace, int trace_cap);
                                                            /* simfru_local.c:32-278 */

/* ---- reductions over a block of FRUs (sequential iFRU order) ---- */
void orc_flux_sums(int n, const orc_fru *f, double num[5]); /* calc_fru_flux_local.c:28-77 */
void orc_avg_sums(int n, const orc_fru *f, double num[11]); /* calc_fru_avg_local.c:27-105 */
void orc_flux_currents(const double num[5], const double dep[3], int nfru, double scale,
                       double out4[4]);                     /* mpi_master.c:55-112 */
void orc_avg_outputs(const double num[11], const double state[ORC_NGLOBAL], int nfru,
                     double scale, double other[ORC_NOTHER]); /* mpi_master.c:175-190 */

/* ---- global model ---- */
void orc_protocol_default(orc_protocol *p);                 /* parameters.h defaults */
void orc_fcn(double t, double state[ORC_NGLOBAL], double F[ORC_NGLOBAL],
             double cur[ORC_NCUR], int keepc, double Jxfer, double Jtr, double ICa,
             double Ito2, const orc_protocol *p);           /* fcn.c:29-1072 */
void orc_errweight(double w[ORC_NGLOBAL]);                  /* initialize_state.c:274-310 */
double orc_total_ca(const orc_fru *f, int n, const double state[ORC_NGLOBAL], double scale);
                                                            /* initialize_state.c:233-246 */

/* ---- whole cell (caf.c + rk_pd54m.c + mpi_master.c with one process) ---- */

/* One call orc_rk54pd makes into the FRU layer, in call order: the same record
 * the reference harness writes for rk_pd54m.c's calls (oracle/ref_harness.c). */
enum { ORC_EV_FLUX = 0, ORC_EV_SAVE = 1, ORC_EV_DISTRIB = 2, ORC_EV_RESUME = 3 };
typedef struct {
  int32_t kind, pad;
  double t0, t1;     /* sub-interval (ORC_EV_DISTRIB) */
  double dep[3];     /* dep (ORC_EV_FLUX) or depf (ORC_EV_DISTRIB) */
  double out[4];     /* Jxfer, Jtr, ICa, Ito2 */
} orc_event;

typedef struct {
  int nfru;
  double scale;
  orc_fru *fru;       /* live state */
  orc_fru *hold;      /* rollback copy (mpi_slave.c:40-45) */
  orc_protocol prot;
  double errw[ORC_NGLOBAL];
  /* counters */
  int64_t n_attempts, n_rejects, n_forced, n_fru_steps;
  /* optional call trace of orc_rk54pd (NULL: off); ev_n counts past ev_cap */
  orc_event *ev;
  int32_t ev_cap, ev_n;
} orc_cell;

int orc_cell_init(orc_cell *c, int nfru, double scale, const orc_protocol *p);
void orc_cell_free(orc_cell *c);
void orc_distrib_simfru(orc_cell *c, double t0, double t1, const double dep0[3],
                        const double depf[3], double out4[4]); /* mpi_master.c:369-469 */
void orc_send_calc_fru_flux(orc_cell *c, const double dep[3], double out4[4]); /* :41-113 */
void orc_send_calc_fru_avg(orc_cell *c, const double state[ORC_NGLOBAL],
                           double other[ORC_NOTHER]);      /* :122-191 */
void orc_save_state(orc_cell *c);                           /* mpi_slave.c:192-226 */
void orc_resume_state(orc_cell *c);                         /* mpi_slave.c:150-184 */
/* rk_pd54m.c:29-345.  Optional per-accepted-step log of (t, V, Cai, ICa). */
double orc_rk54pd(orc_cell *c, double t, double tstep, double oldh,
                  double state[ORC_NGLOBAL], double cur[ORC_NCUR],
                  double *log4, int log_cap, int *log_n);

#ifdef __cplusplus
}
#endif
#endif
