/*
 * ref_harness.c -- flat C entry points around the reference's own objects, so
 * that tests can call them through ctypes.
 *
 * Built by oracle/Makefile together with the UNMODIFIED reference sources
 * (simfru_local.c, fru_rates_local.c, fcn_fru.c, mt19937_local.c,
 * initialize_ran.c, calc_fru_flux_local.c, calc_fru_avg_local.c,
 * parameters_fcn_fru.c, fcn.c, initialize_state.c, data_io.c,
 * write_fru_props.c, dynamicFRU.c and the MPI layer rk_pd54m.c, mpi_master.c,
 * mpi_slave.c) compiled in place from the reference tree with the reference's
 * flags (gcc -O3) against the MPICH headers and library of this image.  The
 * result, oracle/_ref/libgrw_ref.so, is test infrastructure only: it pins the
 * CPU restatement in grw_oracle.c and the package's file writers, and is never
 * part of the product.
 *
 * The MPI layer runs as a single rank (an MPICH singleton: MPI_Init without
 * mpiexec, procs = 1), so mpi_master.c gives every FRU to the master
 * (mpi_master.c:289-317) and its reductions are over one process.
 *
 * This file defines the globals the translation units expect from caf.c
 * (caf.c:27-38).  It calls the reference's routines only; it restates none.
 * The link wraps (ld --wrap) distrib_simFRU, send_calc_fru_flux,
 * send_save_state and send_resume_state so that every call rk54pd makes
 * (rk_pd54m.c:97,106,125-210,328) is recorded, in order, into a caller-given
 * event buffer before it returns.
 */
#include <math.h>
#include <mpi.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "parameters.h"

/* caf.c:27-38 */
double time_of_last_change, glob_time_to_stim;
int current_phase;
double errweight[N];
int iv_n;
double NFRU_scale;
double NFRU_sim_factor;
int NFRU_sim;
int my_rank;
int procs;

/* ---- MPI (single rank) ---- */
int ref_mpi_init(void) {
  int flag = 0;
  MPI_Initialized(&flag);
  if (!flag) {
    int argc = 0;
    char **argv = NULL;
    if (MPI_Init(&argc, &argv) != MPI_SUCCESS) return -1;
  }
  MPI_Comm_rank(MPI_COMM_WORLD, &my_rank);
  MPI_Comm_size(MPI_COMM_WORLD, &procs);
  return (my_rank == 0 && procs == 1) ? 0 : -2;
}

/* ---- event log of the calls rk54pd makes ---- */
enum { EV_FLUX = 0, EV_SAVE = 1, EV_DISTRIB = 2, EV_RESUME = 3 };
typedef struct {
  int32_t kind, pad;
  double t0, t1;     /* sub-interval (EV_DISTRIB) */
  double dep[3];     /* FRUdep_states (EV_FLUX) or FRUdep_statesf (EV_DISTRIB) */
  double out[4];     /* Jxfer, Jtr, ICa, Ito2 */
} ref_event;

static ref_event *ev_buf;
static int ev_cap, ev_n;

void ref_trace(ref_event *buf, int cap) {
  ev_buf = buf;
  ev_cap = cap;
  ev_n = 0;
}
int ref_trace_n(void) { return ev_n; }

static ref_event *ev_next(int kind) {
  if (!ev_buf) return NULL;
  if (ev_n >= ev_cap) {
    ++ev_n; /* counted, not stored: the caller sees n > cap */
    return NULL;
  }
  ref_event *e = &ev_buf[ev_n++];
  memset(e, 0, sizeof *e);
  e->kind = kind;
  return e;
}

void __real_distrib_simFRU(double, double, double[Nstates_FRUdep], double[Nstates_FRUdep],
                           double *, double *, double *, double *);
void __real_send_calc_fru_flux(double[Nstates_FRUdep], double *, double *, double *, double *);
void __real_send_save_state(void);
void __real_send_resume_state(void);

void __wrap_distrib_simFRU(double t0, double t1, double dep0[Nstates_FRUdep],
                           double depf[Nstates_FRUdep], double *Jxfer, double *Jtr, double *ICa,
                           double *Ito2) {
  __real_distrib_simFRU(t0, t1, dep0, depf, Jxfer, Jtr, ICa, Ito2);
  ref_event *e = ev_next(EV_DISTRIB);
  if (e) {
    e->t0 = t0;
    e->t1 = t1;
    memcpy(e->dep, depf, sizeof e->dep);
    e->out[0] = *Jxfer; e->out[1] = *Jtr; e->out[2] = *ICa; e->out[3] = *Ito2;
  }
}

void __wrap_send_calc_fru_flux(double dep[Nstates_FRUdep], double *Jxfer, double *Jtr,
                               double *ICa, double *Ito2) {
  __real_send_calc_fru_flux(dep, Jxfer, Jtr, ICa, Ito2);
  ref_event *e = ev_next(EV_FLUX);
  if (e) {
    memcpy(e->dep, dep, sizeof e->dep);
    e->out[0] = *Jxfer; e->out[1] = *Jtr; e->out[2] = *ICa; e->out[3] = *Ito2;
  }
}

void __wrap_send_save_state(void) {
  __real_send_save_state();
  ev_next(EV_SAVE);
}

void __wrap_send_resume_state(void) {
  __real_send_resume_state();
  ev_next(EV_RESUME);
}

/* ---- the master-side surface of mpi_master.c over one rank ---- */

/* initialize_mpi_state (mpi_master.c:199-361) for n <= MAX_LOCAL_NFRU FRUs in
 * the reference's array layout; sets NFRU_sim / NFRU_scale as caf.c:86-88. */
int ref_load_block(int n, double scale, double *fru, int *ltype, int *ryr, int *ito2, int *mti,
                   unsigned long *mt) {
  if (n < 0 || n > MAX_LOCAL_NFRU) return -1;
  if (ref_mpi_init()) return -2;
  static double g_fru[NFRU_sim_max][Nstates_FRU];
  static int g_ltype[NFRU_sim_max][Nclefts_FRU][Nindepstates_LType];
  static int g_ryr[NFRU_sim_max][Nclefts_FRU][NRyRs_per_cleft];
  static int g_ito2[NFRU_sim_max][Nclefts_FRU];
  static int g_mti[NFRU_sim_max];
  static unsigned long g_mt[NFRU_sim_max][mtN + 1];
  memcpy(g_fru, fru, sizeof(double) * n * Nstates_FRU);
  memcpy(g_ltype, ltype, sizeof(int) * n * Nclefts_FRU * Nindepstates_LType);
  memcpy(g_ryr, ryr, sizeof(int) * n * Nclefts_FRU * NRyRs_per_cleft);
  memcpy(g_ito2, ito2, sizeof(int) * n * Nclefts_FRU);
  memcpy(g_mti, mti, sizeof(int) * n);
  memcpy(g_mt, mt, sizeof(unsigned long) * n * (mtN + 1));
  NFRU_sim = n;
  NFRU_scale = scale;
  initialize_mpi_state(g_fru, g_ltype, g_ryr, g_ito2, g_mti, g_mt);
  return 0;
}

/* parallel_get_FRUs (mpi_master.c:509-617) into caller arrays of n FRUs. */
int ref_get_block(int n, double *fru, int *ltype, int *ryr, int *ito2, int *mti,
                  unsigned long *mt) {
  if (n != NFRU_sim || n > MAX_LOCAL_NFRU) return -1;
  static double g_fru[NFRU_sim_max][Nstates_FRU];
  static int g_ltype[NFRU_sim_max][Nclefts_FRU][Nindepstates_LType];
  static int g_ryr[NFRU_sim_max][Nclefts_FRU][NRyRs_per_cleft];
  static int g_ito2[NFRU_sim_max][Nclefts_FRU];
  static int g_mti[NFRU_sim_max];
  static unsigned long g_mt[NFRU_sim_max][mtN + 1];
  parallel_get_FRUs(g_fru, g_ltype, g_ryr, g_ito2, g_mt, g_mti);
  memcpy(fru, g_fru, sizeof(double) * n * Nstates_FRU);
  memcpy(ltype, g_ltype, sizeof(int) * n * Nclefts_FRU * Nindepstates_LType);
  memcpy(ryr, g_ryr, sizeof(int) * n * Nclefts_FRU * NRyRs_per_cleft);
  memcpy(ito2, g_ito2, sizeof(int) * n * Nclefts_FRU);
  memcpy(mti, g_mti, sizeof(int) * n);
  memcpy(mt, g_mt, sizeof(unsigned long) * n * (mtN + 1));
  return 0;
}

void ref_send_calc_fru_flux(double *dep, double *out4) {
  __real_send_calc_fru_flux(dep, &out4[0], &out4[1], &out4[2], &out4[3]);
}

void ref_send_calc_fru_avg(double *state, double *other) { send_calc_fru_avg(state, other); }

void ref_distrib_simfru(double t0, double t1, double *dep0, double *depf, double *out4) {
  __real_distrib_simFRU(t0, t1, dep0, depf, &out4[0], &out4[1], &out4[2], &out4[3]);
}

void ref_save_state(void) { __real_send_save_state(); }
void ref_resume_state(void) { __real_send_resume_state(); }

/* NFRU_sim / NFRU_scale as caf.c:86-88 sets them, for the file writers that
 * loop over NFRU_sim (data_io.c, write_fru_props.c). */
void ref_set_nfru(int n, double scale) {
  NFRU_sim = n;
  NFRU_scale = scale;
}

void ref_set_errweight(const double *w) { memcpy(errweight, w, sizeof(double) * N); }

/* rk54pd (rk_pd54m.c:29-345) on the loaded block. */
double ref_rk54pd(double t, double tstep, double oldh, double *state, double *cur) {
  return rk54pd(t, tstep, oldh, state, cur);
}

/* dynamicFRU's phase decision (dynamicFRU.c:26-150) with NFRU_sim_factor = 1,
 * the reference configuration's factor (parameters.h:43-46, caf.c:89): no FRU
 * is gathered or redistributed.  In: phase, time of last change; out: both,
 * and glob_time_to_stim. */
void ref_dynamic_fru(double t, double V, double pryr_open, double plcc_open, int *phase,
                     double *t_last, double *time_to_stim) {
  static double g_fru[NFRU_sim_max][Nstates_FRU];
  static int g_ltype[NFRU_sim_max][Nclefts_FRU][Nindepstates_LType];
  static int g_ryr[NFRU_sim_max][Nclefts_FRU][NRyRs_per_cleft];
  static int g_ito2[NFRU_sim_max][Nclefts_FRU];
  static int g_mti[NFRU_sim_max];
  static unsigned long g_mt[NFRU_sim_max][mtN + 1];
  double state[N];
  memset(state, 0, sizeof state);
  state[index_V] = V;
  current_phase = *phase;
  time_of_last_change = *t_last;
  NFRU_sim_factor = 1.0;
  glob_time_to_stim = -1.0;
  dynamicFRU(t, state, g_fru, g_ltype, g_ryr, g_ito2, g_mti, g_mt, pryr_open, plcc_open);
  *phase = current_phase;
  *t_last = time_of_last_change;
  *time_to_stim = glob_time_to_stim;
}

/* ---- MT19937 ---- */
void ref_sgrnd(uint32_t seed, int *mti, unsigned long *mt) { sgrnd_local(seed, mti, mt); }
void ref_mt_fast(int *mti, unsigned long *mt, int n, double *out) {
  MersenneTwister_fast(mti, mt, n, out);
}
double ref_mt_one(int *mti, unsigned long *mt) { return MersenneTwisterOne_local(mti, mt); }

/* initialize_ran with the reference's hard-coded a = 12350232 for n FRUs;
 * mt must hold n rows of mtN+1 words. */
void ref_initialize_ran(int n, unsigned long *mt, int *mti) {
  NFRU_sim = n;
  initialize_ran((unsigned long (*)[mtN + 1])mt, mti);
}

/* ---- per-FRU pieces ---- */
double ref_fru_rates(int *ltype, int *ryr, int *ito2, const double *dep, const double *fru,
                     double *lrate, int *lidx, int *llen, double *vrate, double *rrate,
                     int *ridx, int *rlen, double *irate, int *mti, unsigned long *mt) {
  return fru_rates_local((int (*)[Nindepstates_LType])ltype, (int (*)[NRyRs_per_cleft])ryr, ito2,
                         dep, fru, (double (*)[4])lrate, (int (*)[4])lidx, llen, vrate,
                         (double (*)[NRyRs_per_cleft][4])rrate, (int (*)[NRyRs_per_cleft][4])ridx,
                         (int (*)[NRyRs_per_cleft])rlen, irate, mti, mt);
}

void ref_fcn_fru(const double *fru, const double *dep, const double *lopen, const int *nopen,
                 double *d) {
  fcn_fru(0.0, fru, dep, lopen, nopen, d);
}

void ref_simfru(double t0, double t1, double *dep0, double *depf, int *ltype, int *ryr,
                int *ito2, double *fru, int *mti, unsigned long *mt) {
  simfru_local(t0, t1, dep0, depf, (int (*)[Nindepstates_LType])ltype,
               (int (*)[NRyRs_per_cleft])ryr, ito2, fru, mti, mt);
}

/* ---- reductions ---- */
void ref_flux_sums(int n, int *ltype, int *ito2, double *fru, double *num5) {
  calc_fru_flux_local(n, (int (*)[Nclefts_FRU][Nindepstates_LType])ltype,
                      (int (*)[Nclefts_FRU])ito2, (double (*)[Nstates_FRU])fru, num5);
}

void ref_avg_sums(int n, double *fru, int *ltype, int *ryr, int *ito2, double *num11) {
  calc_fru_avg_local(num11, n, (double (*)[Nstates_FRU])fru,
                     (int (*)[Nclefts_FRU][Nindepstates_LType])ltype,
                     (int (*)[Nclefts_FRU][NRyRs_per_cleft])ryr, (int (*)[Nclefts_FRU])ito2);
}

/* ---- global model (reference compile-time protocol) ---- */
void ref_fcn(double t, double *state, double *F, double *cur, int keepc, double Jxfer,
             double Jtr, double ICa, double Ito2) {
  fcn(t, state, F, cur, keepc, Jxfer, Jtr, ICa, Ito2);
}

/* initialize_state reading ic/stable relative to ref_root (the reference tree).
 * Arrays hold n FRUs; mt rows hold mtN+1 words. */
int ref_initialize_state(const char *ref_root, int n, double scale, double *state, double *fru,
                         int *ltype, int *ryr, int *ito2, unsigned long *mt, int *mti,
                         double *errw) {
  char cwd[4096];
  if (!getcwd(cwd, sizeof cwd)) return -1;
  if (chdir(ref_root) != 0) return -2;
  NFRU_sim = n;
  NFRU_scale = scale;
  initialize_state(state, (double (*)[Nstates_FRU])fru,
                   (int (*)[Nclefts_FRU][Nindepstates_LType])ltype,
                   (int (*)[Nclefts_FRU][NRyRs_per_cleft])ryr, (int (*)[Nclefts_FRU])ito2,
                   (unsigned long (*)[mtN + 1])mt, mti);
  memcpy(errw, errweight, sizeof(double) * N);
  return chdir(cwd) == 0 ? 0 : -3;
}
