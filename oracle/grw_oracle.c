/*
 * grw_oracle.c -- CPU restatement of the reference hot path (test infrastructure).
 *
 * This is the checker for the HIP path, not part of the product: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  Each function
 * follows the reference file:line named in its comment.  Where the reference's
 * floating-point expression order matters for bit-exactness, the same order is
 * kept (it is written out with explicit parentheses); everything else (data
 * layout, table-driven transitions, control flow) is this file's own.
 *
 * Build: gcc -O3 -ffp-contract=off -fPIC -shared (oracle/Makefile).
 */
#include "grw_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Physical constants: parameters_fcn_fru.c:21-90.  Each initialiser keeps the
 * reference's left-to-right constant expression so the rounded values agree. */
static const double kKo = 4.0, kNao = 138.0, kCao = 2.0, kClo = 150.0, kCli = 20.0;
static const double kFaraday = 96.5;
static const double kRToverF = 8.314 * 310.0 / 96.5;
static const double kAcap = 1.534e-4;
static const double kVNSR = (0.53 * 2.10e-6);
static const double kVmyo = 25.84e-6;
static const double kVJSR = (((double)4) * 0.53 * 0.5 * 0.2 * 1.05e-10);
static const double kVSS = (0.5 * 0.2 * 2.03e-12);
static const double kPCa = (1.5 / 2.8 * 0.2 * 0.9 * 0.9468e-11);
static const double kPCl = 2.65e-15;
static const double kBSLtot = 1.124, kCSQNtot = 13.5, kBSRtot = 0.047;
static const double kKBSL = 0.0087, kKmCSQN = 0.63, kKBSR = 0.00087;
static const double kLTRPNtot = 70.0e-3, kHTRPNtot = 140.0e-3;
static const double kkhtrpn_plus = 20.0, kkhtrpn_minus = 0.066e-3;
static const double kkltrpn_plus = 40.0, kkltrpn_minus = 40.0e-3;
static const double kCMDNtot = 50.0e-3, kEGTAtot = 0.0, kKmCMDN = 2.38e-3, kKmEGTA = 1.5e-4;
static const double kJRyRmax = (0.6 * 1.5 * 1.1 * 3.96);
static const double ktautr = 3.0, ktauxfer = 0.005, ktauss2ss = (10.0 * 0.005);

/* Open-state codes: parameters.h:326-334 */
enum { LCC_O1 = 6, LCC_O2 = 12, Y_OPEN = 2, RYR_O1 = 3, RYR_O2 = 4, RYR_O3 = 7, ITO2_OPEN = 2 };

/* ------------------------------------------------------------------------ */
/* MT19937 (Matsumoto & Nishimura 1998), restating mt19937_local.c.          */

/* sgrnd_local, mt19937_local.c:75-92: init_genrand with multiplier 1812433253. */
void orc_mt_seed(orc_mt *m, uint32_t seed) {
  m->w[0] = seed;
  for (uint32_t i = 1; i < ORC_MTN; ++i)
    m->w[i] = 1812433253u * (m->w[i - 1] ^ (m->w[i - 1] >> 30)) + i;
  m->mti = ORC_MTN;
}

/* The 624-word regeneration of mt19937_local.c:129-140 (N=624, M=397). */
static void mt_regenerate(orc_mt *m) {
  uint32_t *w = m->w;
  for (int k = 0; k < ORC_MTN; ++k) {
    uint32_t y = (w[k] & 0x80000000u) | (w[(k + 1) % ORC_MTN] & 0x7fffffffu);
    w[k] = w[(k + 397) % ORC_MTN] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
  }
  m->mti = 0;
}

/* One tempered word (mt19937_local.c:143-147). */
uint32_t orc_mt_next_u32(orc_mt *m) {
  if (m->mti >= ORC_MTN) mt_regenerate(m);
  uint32_t y = m->w[m->mti++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}

/* MersenneTwisterOne_local, mt19937_local.c:202-240: y / 2^32 in [0,1). */
double orc_mt_uniform(orc_mt *m) { return ((double)orc_mt_next_u32(m)) / 4294967296.0; }

/* MersenneTwister_fast, mt19937_local.c:97-154: n consecutive uniforms.  Its
 * fast branch and its per-draw regeneration branch give the same sequence. */
void orc_mt_block(orc_mt *m, int n, double *out) {
  for (int k = 0; k < n; ++k) out[k] = orc_mt_uniform(m);
}

/* initialize_ran.c:30-34.  The reference computes a_j = a_{j-1} + 23 j and
 * seed_j = abs(a_j + 4093 j^2) + 1 in int.  Built with gcc -O3 (Makefile:20) the
 * signed overflow for j >= 722 is taken as impossible, the arithmetic is carried
 * out wide and abs() is dropped, which leaves
 *   seed_j = (a0 + 23 j (j+1)/2 + 4093 j^2 + 1) mod 2^32
 * (sgrnd_local keeps the low 32 bits, mt19937_local.c:80). */
uint32_t orc_seed_for_fru(uint32_t a0, int64_t j) {
  int64_t v = (int64_t)(int32_t)a0 + 23 * (j * (j + 1) / 2) + 4093 * j * j + 1;
  return (uint32_t)(uint64_t)v;
}

/* ------------------------------------------------------------------------ */
/* Transition rates: fru_rates_local.c:30-798.                               */

/* LCC constants fru_rates_local.c:50-96 */
static const double fL = 0.85, gL = 2.0, fLprime = 0.005, gLprime = 7.0;
static const double bL = 1.9356;
static const double aL = 2.0;
static const double omega = 0.83 * 2.0 * 1.3 * 0.01;
static const double alphacf = 4.0 * 1.2 * 0.416;
static const double betacf = 4.0 * 0.45 * 0.049;
static const double gammacf = 0.83 * 1.9 * 1.3 * 0.31 * 7.5 * 0.09233;
static const double yCa_frac = 0.4;
static const double KdIto2 = 0.1502, kbIto2 = 2.0;
/* RyR constants fru_rates_local.c:100-121 */
static const double k21 = 250.0, k32 = 9.6, k43 = 0.07 / 0.06667 * 13.0, k45 = 0.07;
static const double k52 = 0.001235, k65 = 30.0;
static const double k12cf = 3000.0, k23cf = 10.0 * 30000.0, k34cf = 0.6 * 3000.0;
static const double k54cf = 0.6 * 0.198, k25cf = 10.0 * 300.0, k56cf = 2.0 * 4.0 * 3000.0;
static const double k_rate = 0.00127215, k_rate2 = 3.4188;
static const double threshCa34to7 = 0.0368369379834969;
static const double threshCa56to8 = 0.00011447933531005;
static const double threshMAX = 2.0;
static const double threshMAXCa = 0.0504410547074504;

static inline double ref_max(double a, double b) { return (a < b) ? b : a; } /* parameters.h:320-322 */
static inline double ref_min(double a, double b) { return (a < b) ? a : b; } /* parameters.h:316-318 */

/* Fill one channel's exit list. */
static inline void put2(double *r, int32_t *d, int32_t *len, double tot, double r1, int d1) {
  r[0] = tot; r[1] = r1; d[1] = d1; *len = 2;
}
static inline void put3(double *r, int32_t *d, int32_t *len, double tot, double r1, int d1,
                        double r2, int d2) {
  r[0] = tot; r[1] = r1; d[1] = d1; r[2] = r2; d[2] = d2; *len = 3;
}
static inline void put4(double *r, int32_t *d, int32_t *len, double tot, double r1, int d1,
                        double r2, int d2, double r3, int d3) {
  r[0] = tot; r[1] = r1; d[1] = d1; r[2] = r2; d[2] = d2; r[3] = r3; d[3] = d3; *len = 4;
}

/* LCC mode-switching chain, fru_rates_local.c:164-341.  States 1-5 are C0..C4,
 * 6 is O, 7-11 are CCa0..CCa4, 12 is OCa.  Each case lists the exits in the
 * reference's order and forms the total with the reference's summation order. */
static int lcc_exits(int s, double alpha, double beta, double gam, double *r, int32_t *d,
                     int32_t *len) {
  const double alpha_p = aL * alpha;   /* :133 */
  const double beta_p = beta / (bL * 1.0); /* :134, kappa = 1 */
  const double bL2 = bL * bL, bL3 = bL * bL * bL, bL4 = bL * bL * bL * bL;
  switch (s) {
    case 1: put3(r, d, len, (4.0 * alpha + gam), 4.0 * alpha, 2, gam, 7); return 0;
    case 2: put4(r, d, len, (beta + 3.0 * alpha + aL * gam), beta, 1, 3.0 * alpha, 3,
                 aL * gam, 8); return 0;
    case 3: put4(r, d, len, (2.0 * beta + 2.0 * alpha + (aL * aL) * gam), 2.0 * beta, 2,
                 2.0 * alpha, 4, (aL * aL) * gam, 9); return 0;
    case 4: put4(r, d, len, (3.0 * beta + alpha + (aL * aL * aL) * gam), 3.0 * beta, 3, alpha,
                 5, (aL * aL * aL) * gam, 10); return 0;
    case 5: put4(r, d, len, (4.0 * beta + fL + (aL * aL * aL * aL) * gam), 4.0 * beta, 4, fL,
                 6, (aL * aL * aL * aL) * gam, 11); return 0;
    case 6: put2(r, d, len, gL, gL, 5); return 0;
    case 7: put3(r, d, len, (4.0 * alpha_p + omega), omega, 1, 4.0 * alpha_p, 8); return 0;
    case 8: put4(r, d, len, (beta_p + 3.0 * alpha_p + omega / bL), beta_p, 7, omega / bL, 2,
                 3.0 * alpha_p, 9); return 0;
    case 9: put4(r, d, len, (2.0 * beta_p + 2.0 * alpha_p + omega / bL2), 2.0 * beta_p, 8,
                 omega / bL2, 3, 2.0 * alpha_p, 10); return 0;
    case 10: put4(r, d, len, (3.0 * beta_p + alpha_p + omega / bL3), 3.0 * beta_p, 9,
                  omega / bL3, 4, alpha_p, 11); return 0;
    case 11: put4(r, d, len, (4.0 * beta_p + omega / bL4 + fLprime), 4.0 * beta_p, 10,
                  omega / bL4, 5, fLprime, 12); return 0;
    case 12: put2(r, d, len, gLprime, gLprime, 11); return 0;
    default: return -1; /* :336-339 prints and continues; here it is an error */
  }
}

/* RyR rate constants of one cleft, fru_rates_local.c:375-382. */
typedef struct { double k12, k23, k34, k54, k25, k56; } ryr_k;

/* RyR exits in one of three CaSS regimes (fru_rates_local.c:384-735).  States
 * 7 and 8 are the merged rapid-equilibrium states; depending on the regime a
 * real state is collapsed into them, or a merged state is split back with one
 * uniform draw (:452,476,581).  Returns the (possibly updated) state, or -1. */
static int ryr_exits(int s, int regime, const ryr_k *k, orc_mt *rng, double *r, int32_t *d,
                     int32_t *len) {
  /* regime-dependent destinations of the exits that lead into merged states */
  const int d2_23 = (regime == 3) ? 7 : 3;        /* :406 / :527 / :643 */
  const int d2_25 = (regime == 1) ? 5 : 8;        /* :407 / :528 / :644 */
  const int d4_45 = (regime == 1) ? 5 : 8;        /* :427 / :548 */
  for (;;) {
    switch (s) {
      case 1: put2(r, d, len, k->k12, k->k12, 2); return s;
      case 2: put4(r, d, len, (k21 + k->k23 + k->k25), k21, 1, k->k23, d2_23, k->k25, d2_25);
        return s;
      case 3:
        if (regime == 3) { s = 7; continue; }                       /* :647-659 */
        put3(r, d, len, (k32 + k->k34), k32, 2, k->k34, 4); return s;
      case 4:
        if (regime == 3) { s = 7; continue; }                       /* :661-673 */
        put3(r, d, len, (k43 + k45), k43, 3, k45, d4_45); return s;
      case 5:
        if (regime != 1) { s = 8; continue; }                       /* :551-563, :675-687 */
        put4(r, d, len, (k52 + k->k54 + k->k56), k52, 2, k->k54, 4, k->k56, 6); return s;
      case 6:
        if (regime != 1) { s = 8; continue; }                       /* :565-577, :689-701 */
        put2(r, d, len, k65, k65, 5); return s;
      case 7:
        if (regime == 3) {                                          /* :703-714 */
          double a1 = k->k34 / (k43 + k->k34);
          double a2 = (1.0 - a1) * k32;
          double a3 = a1 * k45;
          put3(r, d, len, (a2 + a3), a2, 2, a3, 8);
          return s;
        }
        /* split 7 into 4 or 3: :450-472 (regime 1), :579-601 (regime 2) */
        s = (orc_mt_uniform(rng) < (k->k34 / (k->k34 + k43))) ? 4 : 3;
        continue;
      case 8:
        if (regime != 1) {                                          /* :603-614, :716-727 */
          double a1 = k65 / (k->k56 + k65);
          double a2 = a1 * k52;
          double a3 = a1 * k->k54;
          put3(r, d, len, (a2 + a3), a2, 2, a3, (regime == 3) ? 7 : 4);
          return s;
        }
        s = (orc_mt_uniform(rng) < (k->k56 / (k->k56 + k65))) ? 6 : 5; /* :474-496 */
        continue;
      default: return -1;
    }
  }
}

double orc_fru_rates(orc_fru *f, const double dep[3], orc_rates *r, const orc_protocol *p) {
  const double V = dep[0];
  const double gcf = (p && p->gamma_scale != 1.0) ? gammacf * p->gamma_scale : gammacf;
  const double k23c = (p && p->k23_scale != 1.0) ? k23cf * p->k23_scale : k23cf;
  int bad = 0;

  const double alpha = alphacf * exp(0.012 * (V - 35.0)); /* :131 */
  const double beta = betacf * exp(-0.05 * (V - 35.0));   /* :132 */
  for (int c = 0; c < ORC_NCLEFT; ++c) {
    double gam = gcf * f->conc[c + 1];                     /* :165 */
    if (lcc_exits(f->lcc[c], alpha, beta, gam, r->lcc_rate[c], r->lcc_dest[c], &r->lcc_len[c]))
      bad = 1;
  }

  /* Voltage-dependent inactivation, :345-369 (nu = 1) */
  const double yinf = 1.0 * (yCa_frac / (1.0 + exp((V + 12.5) / 5.0)) + (1.0 - yCa_frac));
  const double tau = 60.0 + 340.0 / (1.0 + exp((V + 30.0) / 12.0));
  for (int c = 0; c < ORC_NCLEFT; ++c) {
    if (f->yinact[c] == 2) r->vinact_rate[c] = (1.0 - yinf) / tau;
    else if (f->yinact[c] == 1) r->vinact_rate[c] = yinf / tau;
    else bad = 1;
  }

  /* RyR, :374-736 */
  for (int c = 0; c < ORC_NCLEFT; ++c) {
    const double ca = f->conc[c + 1];
    const double sat = ref_min(threshMAX, (ca * ca) / k_rate);
    const double sat2 = ref_min(threshMAX, (ca * ca) / k_rate2);
    ryr_k k = {k12cf * sat2, k23c * sat, k34cf * sat, k54cf * sat, k25cf * sat, k56cf * sat};
    const double cm = ref_min(ca, threshMAXCa);
    const int regime = (cm < threshCa56to8) ? 1 : (cm < threshCa34to7) ? 2 : 3;
    for (int i = 0; i < ORC_NRYR; ++i) {
      int s = ryr_exits(f->ryr[c][i], regime, &k, &f->rng, r->ryr_rate[c][i], r->ryr_dest[c][i],
                        &r->ryr_len[c][i]);
      if (s < 0) bad = 1; else f->ryr[c][i] = s;
    }
  }

  /* Ito2, :740-759 */
  for (int c = 0; c < ORC_NCLEFT; ++c) {
    if (f->ito2[c] == 1) r->ito2_rate[c] = (kbIto2 / KdIto2) * f->conc[c + 1];
    else if (f->ito2[c] == 2) r->ito2_rate[c] = kbIto2;
    else bad = 1;
  }

  /* Largest total exit rate, :761-777 */
  double m = 0.0;
  for (int c = 0; c < ORC_NCLEFT; ++c) {
    m = ref_max(m, r->lcc_rate[c][0]);
    m = ref_max(m, r->vinact_rate[c]);
    m = ref_max(m, r->ito2_rate[c]);
    for (int i = 0; i < ORC_NRYR; ++i) m = ref_max(m, r->ryr_rate[c][i][0]);
  }
  r->max_rate = bad ? -1.0 : m;
  return r->max_rate;
}

/* ------------------------------------------------------------------------ */
/* Local Ca ODE: fcn_fru.c:23-137.                                            */
void orc_fcn_fru(const double conc[5], const double dep[3], const double lcc_open[4],
                 const int32_t nryr_open[4], double dconc[5]) {
  const double V = dep[0], Cai = dep[1], CaNSR = dep[2], CaJSR = conc[0];
  const double JDHconstant = 1.0 / (2.0 * kVSS * kFaraday * 1000.0);
  const double Jtr = (CaNSR - CaJSR) / ktautr;
  double JRyR[4], Jxfer[4], Jss2ss[4], JDHPR[4], bSS[4];
  for (int c = 0; c < 4; ++c) JRyR[c] = kJRyRmax * ((double)nryr_open[c]) * (CaJSR - conc[c + 1]);
  const double JRyRtot = JRyR[0] + JRyR[1] + JRyR[2] + JRyR[3];
  for (int c = 0; c < 4; ++c) Jxfer[c] = (conc[c + 1] - Cai) / ktauxfer;
  for (int c = 0; c < 4; ++c) Jss2ss[c] = (conc[c + 1] - conc[(c + 1) % 4 + 1]) / ktauss2ss;

  const double VF_over_RT = V / kRToverF;
  const double VFsq_over_RT = (1000.0 * kFaraday) * VF_over_RT;
  const double exp_VFRT = exp(2.0 * VF_over_RT);
  if (fabs(V) < 1.e-6) { /* :82-91 first-order expansion */
    for (int c = 0; c < 4; ++c)
      JDHPR[c] = -kPCa * 2.0 * 1000.0 * kFaraday * (conc[c + 1] - kCao * 0.341) * lcc_open[c] *
                 JDHconstant;
  } else { /* :101-105 */
    const double a2 = kPCa * 4.0 * VFsq_over_RT / (exp_VFRT - 1.0) * JDHconstant;
    for (int c = 0; c < 4; ++c)
      JDHPR[c] = -(conc[c + 1] * exp_VFRT - kCao * 0.341) * a2 * lcc_open[c];
  }
  for (int c = 0; c < 4; ++c) { /* :115-126 rapid buffering */
    const double s = conc[c + 1];
    const double a1 = kBSRtot * kKBSR / ((s + kKBSR) * (s + kKBSR));
    const double a2 = kBSLtot * kKBSL / ((s + kKBSL) * (s + kKBSL));
    bSS[c] = 1.0 / (1.0 + a1 + a2);
  }
  const double a1 = kCSQNtot * kKmCSQN / ((CaJSR + kKmCSQN) * (CaJSR + kKmCSQN));
  const double bJSR = 1.0 / (1.0 + a1);
  dconc[0] = bJSR * (Jtr - kVSS / kVJSR * JRyRtot);
  for (int c = 0; c < 4; ++c) /* ring: cleft c receives from cleft c-1 (:133-136) */
    dconc[c + 1] = bSS[c] * (JDHPR[c] + JRyR[c] - Jxfer[c] - Jss2ss[c] + Jss2ss[(c + 3) % 4]);
}

/* ------------------------------------------------------------------------ */
/* simfru_local.c:32-278: advance one FRU from t0 to t1.                     */
static inline double lcc_open_ind(const orc_fru *f, int c) {
  return ((f->lcc[c] == LCC_O1 || f->lcc[c] == LCC_O2) && f->yinact[c] == Y_OPEN) ? 1.0 : 0.0;
}
static inline int ryr_open_count(const orc_fru *f, int c) {
  int n = 0;
  for (int i = 0; i < ORC_NRYR; ++i) {
    int s = f->ryr[c][i];
    n += (s == RYR_O1 || s == RYR_O2 || s == RYR_O3);
  }
  return n;
}

/* Categorical draw (simfru_local.c:217-227, :234-245): the stay probability is
 * 1 - dt*k_tot, the j-th exit has dt*k_j; cumulative compare with '>'. */
static inline int pick_exit(double u, double dt, const double *rate, int len) {
  double acc = 1.0 - dt * rate[0];
  if (!(u > acc)) return 0;
  int j = 1;
  acc += dt * rate[1];
  while (u > acc && j + 1 < len) {
    ++j;
    acc += dt * rate[j];
  }
  return j;
}

int orc_simfru(double t0, double t1, const double dep0[3], const double depf[3], orc_fru *f,
               const orc_protocol *p, orc_step *trace, int trace_cap) {
  double span[3], dep[3];
  for (int i = 0; i < 3; ++i) span[i] = depf[i] - dep0[i];
  const double tint = t1 - t0;
  for (int i = 0; i < 3; ++i) dep[i] = dep0[i];

  double lopen[4];
  int32_t nopen[4];
  for (int c = 0; c < 4; ++c) { lopen[c] = lcc_open_ind(f, c); nopen[c] = ryr_open_count(f, c); }

  orc_rates r;
  double t = t0;
  int nsteps = 0;
  while (t < t1) {
    const double maxr = orc_fru_rates(f, dep, &r, p);
    if (maxr < 0.0) return -1;
    /* step length :135-146 */
    double dt = 0.1 / maxr;
    dt = ref_min(dt, 10.001e-3);
    double cmax = 0.e-3;
    for (int c = 0; c < 4; ++c) cmax = ref_max(cmax, f->conc[c + 1]);
    if (cmax > 1.5e-3) dt = ref_min(dt, 5.001e-3);
    dt = ref_min(dt, t1 - t);

    /* Heun step with frozen channel states :149-173 */
    double d[5], k1[5], y1[5];
    orc_fcn_fru(f->conc, dep, lopen, nopen, d);
    for (int i = 0; i < 5; ++i) { k1[i] = dt * d[i]; y1[i] = f->conc[i] + k1[i]; }
    t = t + dt;
    for (int i = 0; i < 3; ++i) dep[i] = dep0[i] + (t - t0) / tint * span[i];
    orc_fcn_fru(y1, dep, lopen, nopen, d);
    for (int i = 0; i < 5; ++i) f->conc[i] = f->conc[i] + (k1[i] + dt * d[i]) / 2.0;

    /* 32 uniforms, 8 per cleft: LCC, Vinact, 5 RyR, Ito2 (:211, :215-253) */
    double u[32];
    orc_mt_block(&f->rng, 32, u);
    for (int c = 0; c < 4; ++c) {
      const double *uc = u + 8 * c;
      int j = pick_exit(uc[0], dt, r.lcc_rate[c], r.lcc_len[c]);
      if (j) f->lcc[c] = r.lcc_dest[c][j];
      if (uc[1] >= 1.0 - dt * r.vinact_rate[c]) f->yinact[c] = 3 - f->yinact[c];
      int moved = 0;
      for (int i = 0; i < ORC_NRYR; ++i) {
        int jr = pick_exit(uc[2 + i], dt, r.ryr_rate[c][i], r.ryr_len[c][i]);
        if (jr) { f->ryr[c][i] = r.ryr_dest[c][i][jr]; moved = 1; }
      }
      if (uc[7] >= 1.0 - dt * r.ito2_rate[c]) f->ito2[c] = 3 - f->ito2[c];
      /* open indicators :256-276 (RyR recount only when a RyR moved) */
      lopen[c] = lcc_open_ind(f, c);
      if (moved) nopen[c] = ryr_open_count(f, c);
    }

    if (trace && nsteps < trace_cap) {
      orc_step *s = &trace[nsteps];
      s->t = t; s->dt = dt;
      memcpy(s->conc, f->conc, sizeof s->conc);
      memcpy(s->lcc, f->lcc, sizeof s->lcc);
      memcpy(s->yinact, f->yinact, sizeof s->yinact);
      memcpy(s->ryr, f->ryr, sizeof s->ryr);
      memcpy(s->ito2, f->ito2, sizeof s->ito2);
      s->mti = f->rng.mti; s->pad = 0;
    }
    ++nsteps;
  }
  return nsteps;
}

/* ------------------------------------------------------------------------ */
/* Reductions in the reference's sequential order.                          */

/* calc_fru_flux_local.c:28-77 */
void orc_flux_sums(int n, const orc_fru *f, double num[5]) {
  double ica_num = 0.0, oca_num = 0.0;
  int nito2 = 0;
  for (int k = 0; k < n; ++k)
    for (int c = 0; c < 4; ++c) {
      if (lcc_open_ind(&f[k], c) != 0.0) { ica_num = ica_num + f[k].conc[c + 1]; oca_num = oca_num + 1; }
      if (f[k].ito2[c] == ITO2_OPEN) nito2 = nito2 + 1;
    }
  double s_ss = 0.0, s_jsr = 0.0;
  for (int k = 0; k < n; ++k) {
    s_jsr += f[k].conc[0];
    for (int c = 0; c < 4; ++c) s_ss += f[k].conc[c + 1];
  }
  num[0] = s_ss; num[1] = s_jsr; num[2] = ica_num; num[3] = oca_num; num[4] = nito2;
}

/* calc_fru_avg_local.c:27-105 */
void orc_avg_sums(int n, const orc_fru *f, double num[11]) {
  double CaSS = 0.0, CaJSR = 0.0, totSS = 0.0, totJSR = 0.0, JRyR = 0.0, PRyR = 0.0;
  double PNorm = 0.0, PnotV = 0.0, PL = 0.0, PIto2 = 0.0;
  int nready = 0;
  for (int k = 0; k < n; ++k) {
    const double cj = f[k].conc[0];
    CaJSR = CaJSR + cj;
    totJSR = totJSR + cj * (1.0 + kCSQNtot / (kKmCSQN + cj));
    for (int c = 0; c < 4; ++c) {
      const double s = f[k].conc[c + 1];
      CaSS = CaSS + s;
      totSS = totSS + s * (1.0 + kBSRtot / (kKBSR + s) + kBSLtot / (kKBSL + s));
      int nopen = 0;
      for (int i = 0; i < ORC_NRYR; ++i) {
        int st = f[k].ryr[c][i];
        if (st == RYR_O1 || st == RYR_O2 || st == RYR_O3) nopen = nopen + 1;
        if (st == 1 || st == 2) nready = nready + 1;
      }
      JRyR = JRyR + kJRyRmax * ((double)nopen) * (cj - s);
      PRyR = PRyR + ((double)nopen);
      if (f[k].lcc[c] <= 6) PNorm = PNorm + 1.0;
      if (f[k].yinact[c] == Y_OPEN) PnotV = PnotV + 1.0;
      if (lcc_open_ind(&f[k], c) != 0.0) PL = PL + 1.0;
      if (f[k].ito2[c] == ITO2_OPEN) PIto2 = PIto2 + 1.0;
    }
  }
  num[0] = CaSS; num[1] = CaJSR; num[2] = JRyR; num[3] = PRyR; num[4] = (double)nready;
  num[5] = totSS; num[6] = totJSR; num[7] = PNorm; num[8] = PnotV; num[9] = PL; num[10] = PIto2;
}

/* Sums -> whole-cell fluxes and currents, mpi_master.c:55-112 (and :401-468 with
 * the sub-interval's end estimate as dep). */
void orc_flux_currents(const double num[5], const double dep[3], int nfru, double scale,
                       double out4[4]) {
  const double V = dep[0], Cai = dep[1], CaNSR = dep[2];
  const double VF_over_RT = V / kRToverF;
  const double VFsq_over_RT = (1000.0 * kFaraday) * VF_over_RT;
  const double e1 = exp(VF_over_RT);
  const double e2 = e1 * e1;
  double ICa, Ito2;
  if (fabs(V) < 1.e-6) {
    const double icanum = num[2] - num[3] * kCao * 0.341;
    ICa = kPCa * 2.0 * 1000.0 * kFaraday * icanum;
    ICa = ICa / kAcap;
    Ito2 = ((double)num[4]) * kPCl * 1000.0 * kFaraday * (kClo - kCli);
    Ito2 = Ito2 / kAcap;
  } else {
    const double icanum = num[2] * e2 - num[3] * kCao * 0.341;
    ICa = kPCa * 4.0 * VFsq_over_RT * icanum / (e2 - 1.0);
    ICa = ICa / kAcap;
    Ito2 = ((double)num[4]) * kPCl * VFsq_over_RT * (kCli - kClo * e1) / (1.0 - e1);
    Ito2 = Ito2 / kAcap;
  }
  const double Jtr = (((double)nfru) * CaNSR - num[1]) / ktautr;
  const double Jxfer = (num[0] - ((double)nfru * (double)4) * Cai) / ktauxfer;
  out4[0] = scale * Jxfer;
  out4[1] = scale * Jtr;
  out4[2] = scale * ICa;
  out4[3] = scale * Ito2;
}

/* mpi_master.c:175-190 (index map indices.h:43-54; the labels of the last two
 * outputs are swapped in the reference and kept swapped here). */
void orc_avg_outputs(const double num[11], const double state[ORC_NGLOBAL], int nfru,
                     double scale, double other[ORC_NOTHER]) {
  const double Cai = state[6], CaNSR = state[7], LTRPNCa = state[9], HTRPNCa = state[10];
  other[0] = num[0] / (double)(nfru * 4);
  other[1] = num[1] / (double)(nfru);
  other[2] = num[2] * scale * kVSS / kVmyo;
  other[3] = num[3] / (double)(5 * 4 * nfru);
  other[4] = (double)(num[4]) / (double)(5 * 4 * nfru);
  other[8] = 1.e6 * ((num[5] * kVSS + num[6] * kVJSR) * scale + CaNSR * kVNSR +
                     Cai * kVmyo * (1.0 + kCMDNtot / (kKmCMDN + Cai) + kEGTAtot / (kKmEGTA + Cai)) +
                     (LTRPNCa * kLTRPNtot + HTRPNCa * kHTRPNtot) * kVmyo);
  other[5] = num[7] / (double)(nfru * 4);
  other[6] = num[8] / (double)(nfru * 4);
  other[7] = num[9] / (double)(nfru * 4);
  other[9] = num[10] / (double)(nfru * 4);
  other[10] = 1.e6 * num[5] * kVSS * scale;
  other[11] = 1.e6 * num[6] * kVJSR * scale;
}

/* ------------------------------------------------------------------------ */
/* Global model.                                                              */

void orc_protocol_default(orc_protocol *p) {
  memset(p, 0, sizeof *p);
  p->pulse_duration = 1; p->pulse_amplitude = -100; p->period = 1000; p->shift = 10;
  p->vclamp = 0; p->chf = 0;
  p->vclamp_duration = 200.0; p->vclamp_set = 0.0; p->vclamp_shift = 300.0;
  p->vclamp_hold = -80.0; p->vclamp_period = 500.0;
  p->serca_scale = 1.0; p->k23_scale = 1.0; p->gamma_scale = 1.0;
  p->vppclamp = 0; p->vpp_ppset = -40.0; p->vpp_ppduration = 100.0; p->vpp_shift = 10000.0;
  p->vpp_set = 0.0; p->vpp_duration = 300.0; p->vpp_hold = -80.0; p->vpp_period = 500.0;
  p->ivclamp = 0; p->iv_duration = 50.0; p->iv_period = 200.0; p->iv_shift = 10.0;
  p->iv_set = -40.0; p->iv_hold = -100.0; p->iv_step = 10.0; p->iv_n = 0;
}

/* initialize_state.c:274-310 */
void orc_errweight(double w[ORC_NGLOBAL]) {
  for (int i = 0; i < ORC_NGLOBAL; ++i) w[i] = 1.0;
  w[0] = 0; w[4] = 0.1; w[5] = 1 / 140.0; w[6] = 1.e3; w[7] = 2.5;
  w[20] = 0.0; w[30] = 0.0; w[31] = 0.1; w[36] = 0.0;
}

/* initialize_state.c:233-246 */
double orc_total_ca(const orc_fru *f, int n, const double s[ORC_NGLOBAL], double scale) {
  double a1 = 0.0, a2 = 0.0;
  for (int k = 0; k < n; ++k) {
    a1 = a1 + f[k].conc[0] * (1.0 + kCSQNtot / (kKmCSQN + f[k].conc[0]));
    for (int c = 0; c < 4; ++c) {
      const double x = f[k].conc[c + 1];
      a2 = a2 + x * (1.0 + kBSRtot / (kKBSR + x) + kBSLtot / (kKBSL + x));
    }
  }
  return 1.e6 * ((a1 * kVJSR + a2 * kVSS) * scale + s[7] * kVNSR +
                 s[6] * kVmyo * (1.0 + kCMDNtot / (kKmCMDN + s[6]) + kEGTAtot / (kKmEGTA + s[6])) +
                 (s[9] * kLTRPNtot + s[10] * kHTRPNtot) * kVmyo);
}

/* Derivatives of a 10-state Ito1 channel laid out as two rows of five
 * (activation C0..C3,O and inactivated CI0..CI3,OI), fcn.c:846-1003.
 * fw[r][k]: k -> k+1 in row r; bw[r][k]: k -> k-1; vx[r][k]: row r -> other row.
 * Summation order follows the reference's per-state expressions.  The last
 * state (OI) is given by conservation and its derivative by the negative sum. */
static void ito_chain(const double x[2][5], const double fw[2][5], const double bw[2][5],
                      const double vx[2][5], double dx[2][5]) {
  for (int r = 0; r < 2; ++r) {
    const int o = 1 - r;
    for (int k = 0; k < 5; ++k) {
      double out, in;
      if (k == 0) {
        out = (fw[r][0] + vx[r][0]) * x[r][0];
        in = bw[r][1] * x[r][1] + vx[o][0] * x[o][0];
      } else if (k == 4) {
        out = (bw[r][4] + vx[r][4]) * x[r][4];
        in = fw[r][3] * x[r][3] + vx[o][4] * x[o][4];
      } else if (r == 0) { /* C1..C3: (forward + backward + to-inactivated) */
        out = (fw[r][k] + bw[r][k] + vx[r][k]) * x[r][k];
        in = bw[r][k + 1] * x[r][k + 1] + vx[o][k] * x[o][k] + fw[r][k - 1] * x[r][k - 1];
      } else {             /* CI1..CI3: (forward + to-activated + backward) */
        out = (fw[r][k] + vx[r][k] + bw[r][k]) * x[r][k];
        in = bw[r][k + 1] * x[r][k + 1] + vx[o][k] * x[o][k] + fw[r][k - 1] * x[r][k - 1];
      }
      dx[r][k] = in - out;
    }
  }
}

/* Ito1 rate lattice for given activation / inactivation rates (fcn.c:851-881,
 * :932-962).  f and b are the four rate-scaling factors of each channel. */
static void ito_rates(double aa, double ba, double ai, double bi, const double f[4],
                      const double b[4], double fw[2][5], double bw[2][5], double vx[2][5]) {
  fw[0][0] = 4.0 * aa; fw[0][1] = 3.0 * aa; fw[0][2] = 2.0 * aa; fw[0][3] = aa; fw[0][4] = 0;
  fw[1][0] = 4.0 * b[0] * aa; fw[1][1] = 3.0 * b[1] * aa / b[0];
  fw[1][2] = 2.0 * b[2] * aa / b[1]; fw[1][3] = b[3] * aa / b[2]; fw[1][4] = 0;
  bw[0][0] = 0; bw[0][1] = ba; bw[0][2] = 2.0 * ba; bw[0][3] = 3.0 * ba; bw[0][4] = 4.0 * ba;
  bw[1][0] = 0; bw[1][1] = ba / f[0]; bw[1][2] = 2.0 * f[0] * ba / f[1];
  bw[1][3] = 3.0 * f[1] * ba / f[2]; bw[1][4] = 4.0 * f[2] * ba / f[3];
  vx[0][0] = bi; vx[0][1] = f[0] * bi; vx[0][2] = f[1] * bi; vx[0][3] = f[2] * bi; vx[0][4] = f[3] * bi;
  vx[1][0] = ai; vx[1][1] = ai / b[0]; vx[1][2] = ai / b[1]; vx[1][3] = ai / b[2]; vx[1][4] = ai / b[3];
}

/* Unpack 10 Ito1 states starting at state index i0 (indices.h:16-35); the last
 * one is 1 - sum of the others in the reference's order (fcn.c:361,373). */
static void ito_unpack(const double *s, int i0, double x[2][5]) {
  x[0][0] = s[i0]; x[0][1] = s[i0 + 1]; x[0][2] = s[i0 + 2]; x[0][3] = s[i0 + 3]; x[0][4] = s[i0 + 4];
  x[1][0] = s[i0 + 5]; x[1][1] = s[i0 + 6]; x[1][2] = s[i0 + 7]; x[1][3] = s[i0 + 8];
  x[1][4] = 1.0 - x[1][0] - x[1][1] - x[1][2] - x[1][3] - x[0][4] - x[0][0] - x[0][1] - x[0][2] - x[0][3];
}
static void ito_pack(const double dx[2][5], int i0, double *F) {
  F[i0] = dx[0][0]; F[i0 + 1] = dx[0][1]; F[i0 + 2] = dx[0][2]; F[i0 + 3] = dx[0][3];
  F[i0 + 4] = dx[0][4];
  F[i0 + 5] = dx[1][0]; F[i0 + 6] = dx[1][1]; F[i0 + 7] = dx[1][2]; F[i0 + 8] = dx[1][3];
  F[i0 + 9] = -dx[1][0] - dx[1][1] - dx[1][2] - dx[1][3] - dx[0][4] - dx[0][0] - dx[0][1] - dx[0][2] - dx[0][3];
}

/* fcn.c:29-1072: 37-state global RHS with whole-cell Jxfer, Jtr, ICa, Ito2 from
 * the release units.  Mutates state[1] (mNa steady state below -90 mV, :663-669)
 * and state[0] under voltage clamp (:811-814), as the reference does. */
void orc_fcn(double t, double state[ORC_NGLOBAL], double F[ORC_NGLOBAL], double cur[ORC_NCUR],
             int keepc, double Jxfer, double Jtr, double ICa, double Ito2, const orc_protocol *p) {
  /* membrane parameters fcn.c:171-219 */
  const double GKr = 0.029, GKs = 0.035, GK1 = 3.0, GKp = 1.2 * 0.002216, GNa = 12.8;
  const double kNaCa = 0.9 * 0.30, KmNa = 87.5, KmCa = 1.38, KmK1 = 13.0, ksat = 0.2, eta = 0.35;
  const double INaKmax = 1.3 * 0.693, KmNai = 10.0, KmKo = 1.5, IpCamax = 0.6 * 0.05, KmpCa = 0.0005;
  const double GCab = 3.3 * 7.684e-5, GNab = 0.001;
  const double Kfb = 0.26e-3, Krb = 1.8, KSR = 1.0, Nfb = 0.75, Nrb = 0.75;
  const double vmax0 = 1.53 * 137.0e-6;
  const double vmaxf = (p->serca_scale != 1.0) ? vmax0 * p->serca_scale : vmax0;
  const double vmaxr = vmaxf;
  /* Ito1 parameters fcn.c:223-264 */
  const double KvScale = 1.13 * 1.03 * 1.55, Kv43Frac = 0.77;
  const double GKv43 = Kv43Frac * KvScale * 0.1;
  const double PKv14 = (1.0 - Kv43Frac) * KvScale * 4.792933e-7;
  const double f43[4] = {1.8936, 14.224647456, 158.574378389, 142.936645351};
  const double b43[4] = {6.77348, 15.6212705152, 28.7532603313, 524.576206679};
  const double f14[4] = {0.20005, 0.320280, 13.50909223, 1151.7651385};
  const double b14[4] = {2.2300, 12.000299, 5.3701338025, 5.2396395511};
  /* HERG fcn.c:303-320 */
  const double TH = 5.320000001;
  const double A0 = 0.017147641733086, B0 = 0.03304608038835, A1 = 0.03969328381141,
               B1 = -0.04306054163980, A2 = 0.02057448605977, B2 = 0.02617412715118,
               A3 = 0.00134366604423, B3 = -0.02691385498399, A4 = 0.10666316491288,
               B4 = 0.00568908859717, A5 = 0.00646393910049, B5 = -0.04536642959543,
               A6 = 0.00008039374403, B6 = 0.00000069808924;
  const double C2H_to_C3H = TH * 0.02608362043337, C3H_to_C2H = TH * 0.14832978132145;

  double mNa = state[1];
  const double hNa = state[2], jNa = state[3], Nai = state[4], Ki = state[5];
  double Cai = state[6];
  const double CaNSR = state[7], xKs = state[8], LTRPNCa = state[9], HTRPNCa = state[10];
  double k43x[2][5], k14x[2][5];
  ito_unpack(state, 11, k43x);
  ito_unpack(state, 21, k14x);
  const double C1H = state[32], C2H = state[33], C3H = state[34], OH = state[35];
  const double IH = 1.0 - OH - C1H - C2H - C3H;
  double V = state[0];

  /* stimulus :434-443 */
  const double ton = floor((t - p->shift) / p->period) * p->period;
  const double toff = ton + p->pulse_duration;
  double Istim = 0.0;
  if ((t - p->shift) >= ton && (t - p->shift) <= toff) Istim = Istim + p->pulse_amplitude;

  /* voltage clamp :448-483 */
  if (p->vclamp) {
    const double von = floor((t - p->vclamp_shift) / p->vclamp_period) * p->vclamp_period;
    const double voff = von + p->vclamp_duration;
    if (((t - p->vclamp_shift) >= von) && (von >= 0.0) && ((t - p->vclamp_shift) < voff))
      V = p->vclamp_set;
    else
      V = p->vclamp_hold;
  }
  /* voltage clamp with a prepulse :485-516 */
  if (p->vppclamp) {
    const double on = floor((t - p->vpp_shift) / p->vpp_period) * p->vpp_period;
    const double ppoff = on + p->vpp_ppduration;
    const double off = on + p->vpp_duration + p->vpp_ppduration;
    if (((t - p->vpp_shift) >= on) && (on >= 0.0) && ((t - p->vpp_shift) < ppoff)) {
      V = p->vpp_ppset;
    } else if (((t - p->vpp_shift) >= ppoff) && (on >= 0.0) && ((t - p->vpp_shift) < off)) {
      V = p->vpp_set;
    } else if (((t - p->vpp_shift) < on) || (on < 0.0) || ((t - p->vpp_shift) >= off)) {
      V = p->vpp_hold;
    }
  }
  /* voltage clamp for I-V relationship studies :521-535 */
  if (p->ivclamp) {
    const double pstart = (floor((t + 1.e-20) / p->iv_period)) * p->iv_period;
    const double pend = (1.0 + floor((t + 1.e-20) / p->iv_period)) * p->iv_period;
    const double on = p->iv_shift + pstart;
    const double off = ref_min(pend, on + p->iv_duration);
    if ((t >= on) && (t < off)) V = p->iv_set + p->iv_n * p->iv_step;
    else V = p->iv_hold;
  }

  /* reversal potentials :549-559 */
  const double ENa = kRToverF * log(kNao / Nai);
  const double EK = kRToverF * log(kKo / Ki);
  const double EKs = kRToverF * log((kKo + 0.01833 * kNao) / (Ki + 0.01833 * Nai));
  if (Cai <= 0.0) Cai = 1.e-15;
  const double ECa = 0.5 * kRToverF * log(kCao / Cai);

  /* INa, IKr, IKs, Ito1, IK1, INab, IKp :565-619 */
  if (mNa <= 0.0) mNa = 1.e-15;
  const double INa = GNa * (mNa * mNa * mNa) * hNa * jNa * (V - ENa);
  const double fKo = pow(kKo / 4.0, 0.5);
  const double IKr = GKr * fKo * OH * (V - EK);
  const double IKs = GKs * (xKs * xKs) * (V - EKs);
  double IKv43 = GKv43 * k43x[0][4] * (V - EK);
  double VF_over_RT = V / kRToverF;
  const double VFsq_over_RT = (1000.0 * kFaraday) * VF_over_RT;
  const double exp_VFRT = exp(VF_over_RT);
  const double OK14 = k14x[0][4];
  double a1, a2, a3, a4, a5;
  a1 = Ki * exp_VFRT - kKo;
  a2 = exp_VFRT - 1.0;
  const double IKv14_K = (fabs(V) < 1.e-6) ? PKv14 * OK14 * 1000.0 * kFaraday * (Ki - kKo)
                                            : PKv14 * OK14 * VFsq_over_RT * (a1 / a2);
  a1 = Nai * exp_VFRT - kNao;
  const double IKv14_Na = (fabs(V) < 1.e-6) ? 0.02 * PKv14 * OK14 * 1000.0 * kFaraday * (Nai - kNao)
                                             : 0.02 * PKv14 * OK14 * VFsq_over_RT * (a1 / a2);
  const double IKv14 = IKv14_K + IKv14_Na;
  double Ito1 = IKv43 + IKv14;
  const double K1_inf = 1.0 / (0.94 + exp(1.76 / kRToverF * (V - EK)));
  double IK1 = GK1 * kKo / (kKo + KmK1) * K1_inf * (V - EK);
  const double INab = GNab * (V - ENa);
  const double KpV = 1.0 / (1.0 + exp((7.488 - V) / 5.98));
  const double IKp = GKp * KpV * (V - EK);

  /* INaK, INaCa, ICab, IpCa :623-649 */
  VF_over_RT = V / kRToverF;
  const double sigma = (exp(kNao / 67.3) - 1.0) / 7.0;
  a1 = 1.0 + 0.1245 * exp(-0.1 * VF_over_RT);
  a2 = 0.0365 * sigma / exp_VFRT;
  const double fNaK = 1.0 / (a1 + a2);
  a1 = kKo / (kKo + KmKo);
  a2 = 1.0 + pow(KmNai / Nai, 1.5);
  const double INaK = INaKmax * fNaK * (a1 / a2);
  const double exp_eta = exp(eta * VF_over_RT);
  a1 = exp_eta * (Nai * Nai * Nai) * kCao;
  a2 = exp_eta / exp_VFRT * (kNao * kNao * kNao) * Cai;
  a3 = 1.0 + ksat * exp_eta / exp_VFRT;
  a4 = KmCa + kCao;
  a5 = 5000.0 / (KmNa * KmNa * KmNa + kNao * kNao * kNao);
  double INaCa = kNaCa * a5 * (a1 - a2) / (a4 * a3);
  const double ICab = GCab * (V - ECa);
  const double IpCa = IpCamax * Cai / (KmpCa + Cai);

  /* INa gates :654-719 */
  double alpha_m;
  if (fabs(V + 47.13) <= 1.e-4) alpha_m = 0.32 / (0.1 - 0.005 * (V + 47.13));
  else alpha_m = (0.32 * (V + 47.13)) / (1.0 - exp(-0.1 * (V + 47.13)));
  const double beta_m = 0.08 * exp(-V / 11.0);
  double dmNa;
  if (V >= -90.0) {
    dmNa = alpha_m * (1.0 - mNa) - beta_m * mNa;
  } else {
    dmNa = 0.0;
    mNa = alpha_m / (alpha_m + beta_m);
    state[1] = mNa;
  }
  const double alpha_h = 0.135 * exp((80 + V) / -6.8);
  const double beta_h = (V < -38.73809636838782) ? 3.56 * exp(0.079 * V) + 310000 * exp(0.35 * V)
                                                 : 1 / (0.13 * (1 + exp((V + 10.66) / -11.1)));
  double alpha_j = 0;
  if (V < -37.78) {
    a1 = -127140 * exp(0.2444 * V);
    a2 = 3.474E-5 * exp(-0.04391 * V);
    a3 = 1.0 + exp(0.311 * (V + 79.23));
    alpha_j = (a1 - a2) * (V + 37.78) / a3;
  }
  const double beta_j = (V < -39.82600037702883)
                            ? 0.1212 * exp(-0.01052 * V) / (1.0 + exp(-0.1378 * (V + 40.14)))
                            : 0.3 * exp(-2.535E-7 * V) / (1 + exp(-0.1 * (V + 32)));
  const double dhNa = alpha_h * (1.0 - hNa) - beta_h * hNa;
  const double djNa = alpha_j * (1.0 - jNa) - beta_j * jNa;

  /* HERG :721-743 */
  const double C1H_to_C2H = TH * A0 * exp(B0 * V);
  const double C2H_to_C1H = TH * A1 * exp(B1 * V);
  const double C3H_to_OH = TH * A2 * exp(B2 * V);
  const double OH_to_C3H = TH * A3 * exp(B3 * V);
  const double OH_to_IH = TH * A4 * exp(B4 * V);
  const double IH_to_OH = TH * A5 * exp(B5 * V);
  const double C3H_to_IH = TH * A6 * exp(B6 * V);
  const double IH_to_C3H = (OH_to_C3H * IH_to_OH * C3H_to_IH) / (C3H_to_OH * OH_to_IH);
  const double dC1H = C2H_to_C1H * C2H - C1H_to_C2H * C1H;
  const double dC2H = (C1H_to_C2H * C1H + C3H_to_C2H * C3H) - ((C2H_to_C1H + C2H_to_C3H) * C2H);
  const double dC3H = (C2H_to_C3H * C2H + OH_to_C3H * OH + IH_to_C3H * IH) -
                      ((C3H_to_IH + C3H_to_OH + C3H_to_C2H) * C3H);
  const double dOH = (C3H_to_OH * C3H + IH_to_OH * IH) - ((OH_to_C3H + OH_to_IH) * OH);

  /* IKs gate :745-755 */
  const double xKs_inf = 1.0 / (1.0 + exp(-(V - 24.7) / 13.6));
  if (fabs(V - 10) < 1.e-6) {
    a1 = (7.19e-5) / 0.148;
    a2 = (1.31e-4) / 0.0687;
  } else {
    a1 = 7.19e-5 * (V - 10.0) / (1.0 - exp(-0.148 * (V - 10.0)));
    a2 = 1.31e-4 * (V - 10.0) / (exp(0.0687 * (V - 10.0)) - 1.0);
  }
  const double tau_xKs = 1.0 / (a1 + a2);
  const double dxKs = (xKs_inf - xKs) / tau_xKs;

  /* SERCA, troponin, calmodulin :759-776 */
  const double fb = pow(Cai / Kfb, Nfb);
  const double rb = pow(CaNSR / Krb, Nrb);
  double Jup = KSR * (vmaxf * fb - vmaxr * rb) / (1.0 + fb + rb);
  const double dLTRPNCa = kkltrpn_plus * Cai * (1.0 - LTRPNCa) - kkltrpn_minus * LTRPNCa;
  const double dHTRPNCa = kkhtrpn_plus * Cai * (1.0 - HTRPNCa) - kkhtrpn_minus * HTRPNCa;
  const double Jtrpn = kLTRPNtot * dLTRPNCa + kHTRPNtot * dHTRPNCa;
  a1 = kCMDNtot * kKmCMDN / (pow(Cai + kKmCMDN, 2.0));
  a2 = kEGTAtot * kKmEGTA / (pow(Cai + kKmEGTA, 2.0));
  const double beta_i = 1.0 / (1.0 + a1 + a2);

  if (p->chf) { /* :780-786 */
    IK1 = 0.68 * IK1;
    Jup = 0.38 * Jup;
    INaCa = 1.75 * INaCa;
    IKv43 = 0.143 * IKv43;
    Ito1 = IKv43 + IKv14;
  }

  /* concentration and voltage derivatives :791-823 */
  a1 = kAcap / (kVmyo * kFaraday * 1000.0);
  const double dNai = -(INa + INab + 3.0 * (INaCa + INaK) + IKv14_Na) * a1;
  a3 = IKr + IKs + IK1 + IKp;
  const double dKi = -(a3 - 2.0 * INaK + (Ito1 - IKv14_Na) + Istim) * a1;
  a3 = ICab - 2.0 * INaCa + IpCa;
  const double dCai = beta_i * (Jxfer * kVSS / kVmyo - Jup - Jtrpn - a3 * 0.5 * a1);
  const double dCaNSR = Jup * kVmyo / kVNSR - Jtr * kVJSR / kVNSR;
  const double dCaTOT = -(a3 + ICa) * 0.5 * a1 * kVmyo * 1.0e6;
  a1 = INa + ICa + IKr + IKs;
  a2 = IK1 + IKp + INaCa + INaK + Ito1 + Ito2;
  a3 = IpCa + ICab + INab;
  double dV, Itot;
  if (p->vclamp || p->ivclamp || p->vppclamp) { /* :811-838 */
    dV = 0.0;
    state[0] = V;
    Itot = a1 + a2 + a3;
  } else {
    Itot = a1 + a2 + a3 + Istim;
    dV = -Itot;
  }

  /* Ito1 chains :846-1003 */
  double fw[2][5], bw[2][5], vx[2][5], d43[2][5], d14[2][5];
  ito_rates(0.543708 * exp(0.028983 * V), 0.080185 * exp(-0.0468437 * V),
            0.0498424 * exp(-0.000373016 * V), 0.000819482 * exp(0.00000005374 * V), f43, b43,
            fw, bw, vx);
  ito_chain(k43x, fw, bw, vx, d43);
  ito_rates(6.0 * 0.31551 * exp(0.00695 * V), 6.0 * 0.001966 * exp(-0.08527 * V),
            6.0 * 0.0004938, 6.0 * 0.0000176, f14, b14, fw, bw, vx);
  ito_chain(k14x, fw, bw, vx, d14);

  F[0] = dV; F[1] = dmNa; F[2] = dhNa; F[3] = djNa; F[4] = dNai; F[5] = dKi; F[6] = dCai;
  F[7] = dCaNSR; F[8] = dxKs; F[9] = dLTRPNCa; F[10] = dHTRPNCa;
  ito_pack(d43, 11, F);
  ito_pack(d14, 21, F);
  F[31] = dCaTOT; F[32] = dC1H; F[33] = dC2H; F[34] = dC3H; F[35] = dOH;
  F[36] = -dC1H - dC2H - dC3H - dOH;

  if (keepc) { /* :1046-1071, index map indices.h:56-79 */
    cur[0] = INa; cur[1] = IKr; cur[2] = IKs; cur[3] = Ito1; cur[4] = IK1; cur[5] = IKp;
    cur[6] = INaCa; cur[7] = INaK; cur[8] = IpCa; cur[9] = ICab; cur[10] = INab; cur[11] = ICa;
    cur[12] = ICa * kAcap / (-2.0 * kVmyo * kFaraday * 1000.0);
    cur[13] = Jup; cur[14] = Jtrpn; cur[15] = Jtr * kVJSR / kVmyo; cur[16] = Jxfer * kVSS / kVmyo;
    cur[17] = IKv43; cur[18] = IKv14; cur[19] = IKv14_K; cur[20] = IKv14_Na; cur[21] = Ito2;
    cur[22] = Istim; cur[23] = Itot;
  }
}

/* ------------------------------------------------------------------------ */
/* Whole cell with one process: mpi_master.c / mpi_slave.c with procs = 1.   */

int orc_cell_init(orc_cell *c, int nfru, double scale, const orc_protocol *p) {
  memset(c, 0, sizeof *c);
  c->nfru = nfru;
  c->scale = scale;
  c->fru = (orc_fru *)calloc((size_t)nfru, sizeof(orc_fru));
  c->hold = (orc_fru *)calloc((size_t)nfru, sizeof(orc_fru));
  if (!c->fru || !c->hold) return -1;
  if (p) c->prot = *p; else orc_protocol_default(&c->prot);
  orc_errweight(c->errw);
  return 0;
}

void orc_cell_free(orc_cell *c) {
  free(c->fru);
  free(c->hold);
  c->fru = c->hold = 0;
}

void orc_distrib_simfru(orc_cell *c, double t0, double t1, const double dep0[3],
                        const double depf[3], double out4[4]) {
  /* FRUs are independent within a sub-interval (mpi_master.c:409-421 and the
   * slaves each run their own share): with OpenMP they run on the host's
   * threads, and the sums below stay in sequential FRU order, so the result
   * is the same for any thread count. */
  int64_t steps = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : steps)
  for (int k = 0; k < c->nfru; ++k) {
    int n = orc_simfru(t0, t1, dep0, depf, &c->fru[k], &c->prot, 0, 0);
    if (n > 0) steps += n;
  }
  c->n_fru_steps += steps;
  double num[5];
  orc_flux_sums(c->nfru, c->fru, num);
  orc_flux_currents(num, depf, c->nfru, c->scale, out4); /* V from depf: :401-407 */
}

void orc_send_calc_fru_flux(orc_cell *c, const double dep[3], double out4[4]) {
  double num[5];
  orc_flux_sums(c->nfru, c->fru, num);
  orc_flux_currents(num, dep, c->nfru, c->scale, out4);
}

void orc_send_calc_fru_avg(orc_cell *c, const double state[ORC_NGLOBAL], double other[ORC_NOTHER]) {
  double num[11];
  orc_avg_sums(c->nfru, c->fru, num);
  orc_avg_outputs(num, state, c->nfru, c->scale, other);
}

/* mpi_slave.c:150-226: every process copies its own FRUs; here the copies
 * are split over the OpenMP threads the same way. */
void orc_save_state(orc_cell *c) {
#pragma omp parallel for schedule(static)
  for (int k = 0; k < c->nfru; ++k) c->hold[k] = c->fru[k];
}
void orc_resume_state(orc_cell *c) {
#pragma omp parallel for schedule(static)
  for (int k = 0; k < c->nfru; ++k) c->fru[k] = c->hold[k];
}

static void orc_ev(orc_cell *c, int kind, double t0, double t1, const double dep[3],
                   const double out[4]) {
  if (!c->ev) return;
  if (c->ev_n >= c->ev_cap) { ++c->ev_n; return; }
  orc_event *e = &c->ev[c->ev_n++];
  memset(e, 0, sizeof *e);
  e->kind = kind;
  e->t0 = t0;
  e->t1 = t1;
  if (dep) memcpy(e->dep, dep, sizeof e->dep);
  if (out) memcpy(e->out, out, sizeof e->out);
}

/* orc_distrib_simfru plus its trace record */
static void orc_stage(orc_cell *c, double t0, double t1, const double dep0[3],
                      const double depf[3], double out4[4]) {
  orc_distrib_simfru(c, t0, t1, dep0, depf, out4);
  orc_ev(c, ORC_EV_DISTRIB, t0, t1, depf, out4);
}

/* rk_pd54m.c:29-345: Prince-Dormand RK5(4)6M over [t, tstep] with one Monte Carlo
 * sub-interval per stage, FRU rollback on rejection and stimulus alignment. */
double orc_rk54pd(orc_cell *c, double t, double tstep, double oldh, double state[ORC_NGLOBAL],
                  double cur[ORC_NCUR], double *log4, int log_cap, int *log_n) {
  enum { N = ORC_NGLOBAL };
  const double step_min = 1.e-6, step_max = 0.1, tol = 1.e-6, min_tolvalue = 1.e-10;
  const double huge_number = 1e99;
  double F[N], F1[N], y[N], ym[N], k1[N], k2[N], k3[N], k4[N], k5[N], k6[N];
  double dummy[ORC_NCUR];
  double dep0[3], depf[3], J[4] = {0, 0, 0, 0};
  const orc_protocol *p = &c->prot;

  double h = ref_min(oldh, tstep - t);
  h = ref_min(step_max, h);
  double oh = h, start = t;
  int notdone = 1, success = 1, keepc = 1;
  while (notdone) {
    c->n_attempts++;
    dep0[0] = state[0]; dep0[1] = state[6]; dep0[2] = state[7];
    if (success) {
      orc_send_calc_fru_flux(c, dep0, J);
      orc_ev(c, ORC_EV_FLUX, 0, 0, dep0, J);
      orc_fcn(start, state, F1, cur, keepc, J[0], J[1], J[2], J[3], p);
      keepc = 0;
      orc_save_state(c);
      orc_ev(c, ORC_EV_SAVE, 0, 0, 0, 0);
    }
    /* stage 2 */
    for (int i = 0; i < N; ++i) { k1[i] = h * F1[i]; y[i] = state[i] + k1[i] / 5.0; }
    depf[0] = y[0]; depf[1] = y[6]; depf[2] = y[7];
    orc_stage(c, start, start + h / 5.0, dep0, depf, J);
    dep0[0] = state[0]; dep0[1] = state[6]; dep0[2] = state[7];
    orc_fcn(start + h / 5.0, y, F, dummy, 0, J[0], J[1], J[2], J[3], p);
    /* stage 3 */
    for (int i = 0; i < N; ++i) { k2[i] = h * F[i]; y[i] = state[i] + (3.0 / 40.0 * k1[i] + 9.0 / 40.0 * k2[i]); }
    depf[0] = y[0]; depf[1] = y[6]; depf[2] = y[7];
    orc_stage(c, start + h / 5.0, start + h * 3.0 / 10.0, dep0, depf, J);
    dep0[0] = state[0]; dep0[1] = state[6]; dep0[2] = state[7];
    orc_fcn(start + h * 3.0 / 10.0, y, F, dummy, 0, J[0], J[1], J[2], J[3], p);
    /* stage 4 */
    for (int i = 0; i < N; ++i) {
      k3[i] = h * F[i];
      y[i] = state[i] + (3.0 / 10.0 * k1[i] - 9.0 / 10.0 * k2[i] + 6.0 / 5.0 * k3[i]);
    }
    depf[0] = y[0]; depf[1] = y[6]; depf[2] = y[7];
    orc_stage(c, start + h * 3.0 / 10.0, start + h * 3.0 / 5.0, dep0, depf, J);
    dep0[0] = state[0]; dep0[1] = state[6]; dep0[2] = state[7];
    orc_fcn(start + h * 3.0 / 5.0, y, F, dummy, 0, J[0], J[1], J[2], J[3], p);
    /* stage 5 */
    for (int i = 0; i < N; ++i) {
      k4[i] = F[i] * h;
      y[i] = state[i] + 226.0 / 729.0 * k1[i] - 25.0 / 27.0 * k2[i] + 880.0 / 729.0 * k3[i] +
             55.0 / 729.0 * k4[i];
    }
    depf[0] = y[0]; depf[1] = y[6]; depf[2] = y[7];
    orc_stage(c, start + h * 3.0 / 5.0, start + h * 2.0 / 3.0, dep0, depf, J);
    dep0[0] = state[0]; dep0[1] = state[6]; dep0[2] = state[7];
    orc_fcn(start + h * 2.0 / 3.0, y, F, dummy, 0, J[0], J[1], J[2], J[3], p);
    /* stage 6 */
    for (int i = 0; i < N; ++i) {
      k5[i] = h * F[i];
      y[i] = state[i] - 181.0 / 270.0 * k1[i] + 5.0 / 2.0 * k2[i] - 266.0 / 297.0 * k3[i] -
             91.0 / 27.0 * k4[i] + 189.0 / 55.0 * k5[i];
    }
    depf[0] = y[0]; depf[1] = y[6]; depf[2] = y[7];
    orc_stage(c, start + h * 2.0 / 3.0, start + h, dep0, depf, J);
    orc_fcn(start + h, y, F, dummy, 0, J[0], J[1], J[2], J[3], p);
    for (int i = 0; i < N; ++i) k6[i] = h * F[i];

    /* error estimate :225-238 */
    double err = 0.0;
    for (int i = 0; i < N; ++i) {
      y[i] = state[i] + 31.0 / 540.0 * k1[i] + 190.0 / 297.0 * k3[i] - 145.0 / 108.0 * k4[i] +
             351.0 / 220.0 * k5[i] + 1.0 / 20.0 * k6[i];
      ym[i] = state[i] + 19.0 / 216.0 * k1[i] + 1000.0 / 2079.0 * k3[i] - 125.0 / 216.0 * k4[i] +
              81.0 / 88.0 * k5[i] + 5.0 / 56.0 * k6[i];
      double e = fabs((ym[i] - y[i]) * 0.2 * c->errw[i]);
      if (e < huge_number) err = ref_max(e, err);
      else err = huge_number;
    }

    if (err < tol) { /* accept :241-286 */
      for (int i = 0; i < N; ++i) {
        if (fabs(y[i]) < min_tolvalue) y[i] = 0.0;
        state[i] = y[i];
      }
      start = start + h;
      success = 1;
      if (log4 && log_n && *log_n < log_cap) {
        double *L = log4 + 4 * (*log_n);
        L[0] = start; L[1] = state[0]; L[2] = state[6]; L[3] = J[2];
        ++*log_n;
      }
      if (start >= tstep) {
        notdone = 0;
      } else {
        h = ref_min(0.85 * h * pow(tol / err, 0.2), 4.0 * h);
        notdone = 1;
        oh = h;
        double tts = p->shift - fmod(start, p->period);
        if (tts < 0.0) tts = tts + p->period;
        if (tts < h) h = ref_max(step_min, tts - 0.5 * step_min);
      }
    } else {
      if (h <= step_min) { /* forced accept :291-309 */
        for (int i = 0; i < N; ++i) {
          if (fabs(y[i]) < min_tolvalue) y[i] = 0.0;
          state[i] = y[i];
        }
        c->n_forced++;
        start = start + h;
        success = 1;
        if (log4 && log_n && *log_n < log_cap) {
          double *L = log4 + 4 * (*log_n);
          L[0] = start; L[1] = state[0]; L[2] = state[6]; L[3] = J[2];
          ++*log_n;
        }
        if (start >= tstep) notdone = 0;
        else { h = step_min; notdone = 1; }
        oh = h;
      } else { /* reject :310-329 */
        c->n_rejects++;
        h = 0.85 * h * pow(tol / err, 0.2);
        double tts = p->shift - fmod(start, p->period);
        if (tts < 0.0) tts = tts + p->period;
        if (tts < h) h = ref_max(step_min, tts - 0.5 * step_min);
        success = 0;
        notdone = 1;
        orc_resume_state(c);
        orc_ev(c, ORC_EV_RESUME, 0, 0, 0, 0);
      }
      h = ref_min(step_max, ref_max(step_min, h));
      h = ref_min(h, tstep - start);
    }
    oldh = oh;
  }
  return oldh;
}
