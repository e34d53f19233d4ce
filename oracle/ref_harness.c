/*
 * ref_harness.c -- flat C entry points around the reference's own hot-path
 * objects, so that tests can call them through ctypes.
 *
 * Built by oracle/Makefile together with the UNMODIFIED reference sources
 * (simfru_local.c, fru_rates_local.c, fcn_fru.c, mt19937_local.c,
 * initialize_ran.c, calc_fru_flux_local.c, calc_fru_avg_local.c,
 * parameters_fcn_fru.c, fcn.c, initialize_state.c, data_io.c, write_fru_props.c)
 * compiled in place from the reference tree with the reference's flags
 * (gcc -O3).  The result, oracle/_ref/libgrw_ref.so, is test infrastructure
 * only: it pins the CPU restatement in grw_oracle.c and the package's file
 * writers, and is never part of the product.
 *
 * This file defines the globals those translation units expect from caf.c
 * (caf.c:27-38) and the master-side flux routine that fcn.c's lastcall needs
 * (restating the single-process branch of mpi_master.c:41-113 over the
 * reference's calc_fru_flux_local).
 */
#include <math.h>
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

/* FRU block the harness's flux routine reduces over (set by ref_set_block). */
static int blk_n;
static int (*blk_ltype)[Nclefts_FRU][Nindepstates_LType];
static int (*blk_ito2)[Nclefts_FRU];
static double (*blk_fru)[Nstates_FRU];

void ref_set_block(int n, int *ltype, int *ito2, double *fru, double scale) {
  blk_n = n;
  blk_ltype = (int (*)[Nclefts_FRU][Nindepstates_LType])ltype;
  blk_ito2 = (int (*)[Nclefts_FRU])ito2;
  blk_fru = (double (*)[Nstates_FRU])fru;
  NFRU_sim = n;
  NFRU_scale = scale;
}

/* Single-process branch of mpi_master.c:41-113 over the block above. */
void send_calc_fru_flux(double dep[Nstates_FRUdep], double *Jxfer, double *Jtr, double *ICa,
                        double *Ito2) {
  double num[6] = {0, 0, 0, 0, 0, 0};
  const double V = dep[index_frudep_V], Cai = dep[index_frudep_Cai];
  const double CaNSR = dep[index_frudep_CaNSR];
  const double VF_over_RT = V / RT_over_F;
  const double VFsq_over_RT = (1000.0 * Faraday) * VF_over_RT;
  const double e1 = exp(VF_over_RT), e2 = e1 * e1;
  double ica, ito2, icanum;
  if (blk_n > 0) calc_fru_flux_local(blk_n, blk_ltype, blk_ito2, blk_fru, num);
  if (fabs(V) < 1.e-6) {
    icanum = num[2] - num[3] * Cao * 0.341;
    ica = PCa * 2.0 * 1000.0 * Faraday * icanum;
    ica = ica / Acap;
    ito2 = ((double)num[4]) * PCl * 1000.0 * Faraday * (Clo - Cli);
    ito2 = ito2 / Acap;
  } else {
    icanum = num[2] * e2 - num[3] * Cao * 0.341;
    ica = PCa * 4.0 * VFsq_over_RT * icanum / (e2 - 1.0);
    ica = ica / Acap;
    ito2 = ((double)num[4]) * PCl * VFsq_over_RT * (Cli - Clo * e1) / (1.0 - e1);
    ito2 = ito2 / Acap;
  }
  *Jxfer = NFRU_scale * ((num[0] - ((double)NFRU_sim * (double)Nclefts_FRU) * Cai) / tauxfer);
  *Jtr = NFRU_scale * ((((double)NFRU_sim) * CaNSR - num[1]) / tautr);
  *ICa = NFRU_scale * ica;
  *Ito2 = NFRU_scale * ito2;
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
