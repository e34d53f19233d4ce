// grw_global.h -- the 37-state global model (membrane currents, SR and
// cytosolic Ca, Ito1/HERG Markov chains) that consumes the release units'
// fluxes, usable on the host and on the device.
//
// fcn.c:29-1072 restated with the reference's expression order; the Ito1
// chains are evaluated by one lattice routine instead of 20 hand-written
// derivatives.  Protocol switches that the reference fixes at compile time
// (parameters.h:107-163) are fields of GrwProtocol.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/grw_cell.h"
#include "grw_model.h"

namespace grw {

constexpr int kNglobal = 37;
constexpr int kNcur = 24;

// Ito1 lattice: two rows (activation C0..C3,O / inactivated CI0..CI3,OI).
struct ItoRates {
  double fw[2][5], bw[2][5], vx[2][5];
};

// The lattice constants f, b come with their reciprocals (DivK): every
// division here is by one of them.
GRW_HD void ito_rates(double aa, double ba, double ai, double bi, const DivK f[4], const DivK b[4],
                      ItoRates& q) {
  q.fw[0][0] = 4.0 * aa; q.fw[0][1] = 3.0 * aa; q.fw[0][2] = 2.0 * aa; q.fw[0][3] = aa; q.fw[0][4] = 0;
  q.fw[1][0] = 4.0 * b[0].b * aa; q.fw[1][1] = rdiv(3.0 * b[1].b * aa, b[0]);
  q.fw[1][2] = rdiv(2.0 * b[2].b * aa, b[1]); q.fw[1][3] = rdiv(b[3].b * aa, b[2]); q.fw[1][4] = 0;
  q.bw[0][0] = 0; q.bw[0][1] = ba; q.bw[0][2] = 2.0 * ba; q.bw[0][3] = 3.0 * ba; q.bw[0][4] = 4.0 * ba;
  q.bw[1][0] = 0; q.bw[1][1] = rdiv(ba, f[0]); q.bw[1][2] = rdiv(2.0 * f[0].b * ba, f[1]);
  q.bw[1][3] = rdiv(3.0 * f[1].b * ba, f[2]); q.bw[1][4] = rdiv(4.0 * f[2].b * ba, f[3]);
  q.vx[0][0] = bi; q.vx[0][1] = f[0].b * bi; q.vx[0][2] = f[1].b * bi; q.vx[0][3] = f[2].b * bi;
  q.vx[0][4] = f[3].b * bi;
  q.vx[1][0] = ai; q.vx[1][1] = rdiv(ai, b[0]); q.vx[1][2] = rdiv(ai, b[1]); q.vx[1][3] = rdiv(ai, b[2]);
  q.vx[1][4] = rdiv(ai, b[3]);
}

// Derivatives of the lattice; summation order per state as in fcn.c:883-998.
GRW_HD void ito_chain(const double x[2][5], const ItoRates& q, double dx[2][5]) {
  for (int r = 0; r < 2; ++r) {
    const int o = 1 - r;
    for (int k = 0; k < 5; ++k) {
      double out, in;
      if (k == 0) {
        out = (q.fw[r][0] + q.vx[r][0]) * x[r][0];
        in = q.bw[r][1] * x[r][1] + q.vx[o][0] * x[o][0];
      } else if (k == 4) {
        out = (q.bw[r][4] + q.vx[r][4]) * x[r][4];
        in = q.fw[r][3] * x[r][3] + q.vx[o][4] * x[o][4];
      } else {
        out = (r == 0) ? (q.fw[r][k] + q.bw[r][k] + q.vx[r][k]) * x[r][k]
                       : (q.fw[r][k] + q.vx[r][k] + q.bw[r][k]) * x[r][k];
        in = q.bw[r][k + 1] * x[r][k + 1] + q.vx[o][k] * x[o][k] + q.fw[r][k - 1] * x[r][k - 1];
      }
      dx[r][k] = in - out;
    }
  }
}

GRW_HD void ito_unpack(const double* s, int i0, double x[2][5]) {
  for (int k = 0; k < 5; ++k) x[0][k] = s[i0 + k];
  for (int k = 0; k < 4; ++k) x[1][k] = s[i0 + 5 + k];
  x[1][4] = 1.0 - x[1][0] - x[1][1] - x[1][2] - x[1][3] - x[0][4] - x[0][0] - x[0][1] - x[0][2] - x[0][3];
}

GRW_HD void ito_pack(const double dx[2][5], int i0, double* F) {
  for (int k = 0; k < 5; ++k) F[i0 + k] = dx[0][k];
  for (int k = 0; k < 4; ++k) F[i0 + 5 + k] = dx[1][k];
  F[i0 + 9] = -dx[1][0] - dx[1][1] - dx[1][2] - dx[1][3] - dx[0][4] - dx[0][0] - dx[0][1] - dx[0][2] - dx[0][3];
}

// Transcendental calls of fcn.c:29-1072 whose arguments depend on the state
// or time: their arguments (fcn_prep) and values (fcn_trans).  The host
// evaluates them one by one; the device integrator evaluates each kind as one
// call across the lanes of a wavefront (same function, same argument, same
// value).  Calls with constant arguments stay in fcn_finish; the exp of the
// IK1 factor, whose argument needs EK, follows the logs (fcn_k1_arg).
constexpr int kFcnLog = 4, kFcnPow = 3, kFcnExp = 33;
struct FcnTrans {
  double lg_arg[kFcnLog], pw_base[kFcnPow], pw_exp[kFcnPow], ex_arg[kFcnExp];
  double lg[kFcnLog], pw[kFcnPow], ex[kFcnExp];
  double k1x;         // exp(fcn_k1_arg(V, lg[1])) (the device's; the host computes its own)
  double V, Istim;    // fcn_voltage at the call's time (the device's)
};

// pow(x, 2.0) of fcn.c:774-775.  Compilers fold it into the product on the
// host (exact: the product is the correctly rounded square); the device's math
// library does not fold it, and its pow takes ~250 instructions.
GRW_HD double pow2(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return x * x;
#else
  return pow(x, 2.0);
#endif
}

// fcn.c:448-535: the membrane potential under the voltage-clamp protocols --
// the plain clamp (:448-483), the clamp with a prepulse (:485-516) and the I-V
// steps (:521-535) -- in the reference's order, each overriding V when on.
// div is the division the caller uses (IEEE either way).  Returns whether a
// clamp is on, in which case fcn holds V: dV = 0, state[V] = V and the total
// current leaves out the stimulus (fcn.c:811-838).
template <class Div>
GRW_HD bool clamp_V(double t, const GrwProtocol& p, double& V, Div div) {
  if (p.vclamp) {
    const double von = floor(div(t - p.vclamp_shift, p.vclamp_period)) * p.vclamp_period;
    const double voff = von + p.vclamp_duration;
    if (((t - p.vclamp_shift) >= von) && (von >= 0.0) && ((t - p.vclamp_shift) < voff))
      V = p.vclamp_set;
    else
      V = p.vclamp_hold;
  }
  if (p.vppclamp) {
    const double on = floor(div(t - p.vpp_shift, p.vpp_period)) * p.vpp_period;
    const double ppoff = on + p.vpp_ppduration;
    const double off = on + p.vpp_duration + p.vpp_ppduration;
    const double ts = t - p.vpp_shift;
    if ((ts >= on) && (on >= 0.0) && (ts < ppoff))
      V = p.vpp_ppset;
    else if ((ts >= ppoff) && (on >= 0.0) && (ts < off))
      V = p.vpp_set;
    else if ((ts < on) || (on < 0.0) || (ts >= off))
      V = p.vpp_hold;
  }
  if (p.ivclamp) {
    const double start = (floor(div(t + 1.e-20, p.iv_period))) * p.iv_period;
    const double end = (1.0 + floor(div(t + 1.e-20, p.iv_period))) * p.iv_period;
    const double on = p.iv_shift + start;
    double off = on + p.iv_duration;
    off = (end < off) ? end : off;  // min(period_end, time_iv_clamp_off)
    if ((t >= on) && (t < off))
      V = p.iv_set + p.iv_n * p.iv_step;
    else
      V = p.iv_hold;
  }
  return p.vclamp || p.vppclamp || p.ivclamp;
}

GRW_HD bool clamped(const GrwProtocol& p) { return p.vclamp || p.vppclamp || p.ivclamp; }

struct DivRdiv {
  GRW_HD double operator()(double a, double b) const { return rdiv(a, b); }
};
struct DivIeee {
  GRW_HD double operator()(double a, double b) const { return a / b; }
};

// The membrane potential the transcendental arguments use: the state's, or
// the clamp's under a voltage-clamp protocol.
GRW_HD double fcn_prep_V(double t, const double* state, const GrwProtocol& p) {
  double V = state[0];
  (void)clamp_V(t, p, V, DivRdiv());
  return V;
}

// The arguments of fcn.c's state-dependent exp calls (fcn.c:448-1044), each of the form
// a (V + b), a (V + b) / c with a = +-1, or a VF/RT, with the reference's
// operations: row j holds the form and constants of argument j.  (V + 0 is V
// for the exp, x / c is IEEE division by the constant, and 7.488 - V equals
// -(V - 7.488) exactly.)
struct ExArg {
  int form;  // 0: a (V + b); 1: a (V + b) / c; 2: a VF/RT
  double a, b;
  DivK c;
};
constexpr DivK kNoDiv{1.0, 1.0};
constexpr ExArg kExArg[kFcnExp] = {
    {2, 1.0, 0.0, kNoDiv},                      // 0  VF/RT
    {1, -1.0, -7.488, divk(5.98)},              // 1  (7.488 - V) / 5.98
    {2, -0.1, 0.0, kNoDiv},                     // 2  -0.1 VF/RT
    {2, 0.35, 0.0, kNoDiv},                     // 3  eta VF/RT
    {0, -0.1, 47.13, kNoDiv},                   // 4
    {1, -1.0, 0.0, divk(11.0)},                 // 5  -V / 11
    {1, 1.0, 80.0, divk(-6.8)},                 // 6  (80 + V) / -6.8
    {0, 0.079, 0.0, kNoDiv},                    // 7
    {0, 0.35, 0.0, kNoDiv},                     // 8
    {1, 1.0, 10.66, divk(-11.1)},               // 9
    {0, 0.2444, 0.0, kNoDiv},                   // 10
    {0, -0.04391, 0.0, kNoDiv},                 // 11
    {0, 0.311, 79.23, kNoDiv},                  // 12
    {0, -0.01052, 0.0, kNoDiv},                 // 13
    {0, -0.1378, 40.14, kNoDiv},                // 14
    {0, -2.535E-7, 0.0, kNoDiv},                // 15
    {0, -0.1, 32.0, kNoDiv},                    // 16
    {0, 0.03304608038835, 0.0, kNoDiv},         // 17
    {0, -0.04306054163980, 0.0, kNoDiv},        // 18
    {0, 0.02617412715118, 0.0, kNoDiv},         // 19
    {0, -0.02691385498399, 0.0, kNoDiv},        // 20
    {0, 0.00568908859717, 0.0, kNoDiv},         // 21
    {0, -0.04536642959543, 0.0, kNoDiv},        // 22
    {0, 0.00000069808924, 0.0, kNoDiv},         // 23
    {1, -1.0, -24.7, divk(13.6)},               // 24 -(V - 24.7) / 13.6
    {0, -0.148, -10.0, kNoDiv},                 // 25
    {0, 0.0687, -10.0, kNoDiv},                 // 26
    {0, 0.028983, 0.0, kNoDiv},                 // 27
    {0, -0.0468437, 0.0, kNoDiv},               // 28
    {0, -0.000373016, 0.0, kNoDiv},             // 29
    {0, 0.00000005374, 0.0, kNoDiv},            // 30
    {0, 0.00695, 0.0, kNoDiv},                  // 31
    {0, -0.08527, 0.0, kNoDiv},                 // 32
};

GRW_HD double fcn_ex_arg(int j, double V, double VF_over_RT) {
  const ExArg& r = kExArg[j];
  const double x = r.a * (V + r.b);
  if (r.form == 1) return rdiv(x, r.c);
  return (r.form == 2) ? r.a * VF_over_RT : x;
}

// The log arguments (fcn.c:549-559; Cai floored at 1e-15, :556-557) and the
// pow bases and exponents (fcn.c:631, :759-760).  The divisions are IEEE
// (rdiv, grw_model.h).
GRW_HD double fcn_lg_arg(int k, const double* state) {
  const double Nai = state[4], Ki = state[5];
  const double Cai = (state[6] <= 0.0) ? 1.e-15 : state[6];
  switch (k) {
    case 0: return rdiv(Nao, Nai);
    case 1: return rdiv(Ko, Ki);
    case 2: return rdiv(Ko + 0.01833 * Nao, Ki + 0.01833 * Nai);
    default: return rdiv(Cao, Cai);
  }
}
GRW_HD void fcn_pw_arg(int k, const double* state, double& base, double& expo) {
  const double Nai = state[4];
  const double Cai = (state[6] <= 0.0) ? 1.e-15 : state[6];
  const double CaNSR = state[7];
  const double KmNai = 10.0, Kfb = 0.26e-3, Krb = 1.8, Nfb = 0.75, Nrb = 0.75;
  switch (k) {
    case 0: base = rdiv(KmNai, Nai); expo = 1.5; return;
    case 1: base = rdiv(Cai, divk(Kfb)); expo = Nfb; return;
    default: base = rdiv(CaNSR, divk(Krb)); expo = Nrb; return;
  }
}

// All arguments on one thread (the host fcn; the device integrator computes
// them one per lane beside the calls, fcn_block in grw_fru.hip).
GRW_HD void fcn_prep(double t, const double* state, const GrwProtocol& p, FcnTrans& q) {
  const double V = fcn_prep_V(t, state, p);
  for (int k = 0; k < kFcnLog; ++k) q.lg_arg[k] = fcn_lg_arg(k, state);
  for (int k = 0; k < kFcnPow; ++k) fcn_pw_arg(k, state, q.pw_base[k], q.pw_exp[k]);
  const double VF_over_RT = rdiv(V, divk(RT_over_F));
  for (int j = 0; j < kFcnExp; ++j) q.ex_arg[j] = fcn_ex_arg(j, V, VF_over_RT);
}

GRW_HD void fcn_trans(FcnTrans& q) {
  for (int k = 0; k < kFcnLog; ++k) q.lg[k] = log(q.lg_arg[k]);
  for (int k = 0; k < kFcnPow; ++k) q.pw[k] = pow(q.pw_base[k], q.pw_exp[k]);
  for (int k = 0; k < kFcnExp; ++k) q.ex[k] = exp(q.ex_arg[k]);
}

// fcn.c:29-1072 after its transcendental calls: every exp / log / pow whose
// argument depends on the state is taken from q (fcn_prep + fcn_trans), every
// other operation is done in the reference's order.  The work falls into four
// groups that share no intermediate values -- the membrane currents, the two
// Ito1 lattices, the Na / HERG / IKs gates and the myoplasmic Ca handling --
// and an assembly of the concentration and voltage derivatives from them;
// fcn_finish runs them in turn, the device integrator runs the four groups on
// the four wavefronts of a block (fcn_block in grw_fru.hip), with the same
// values.

// Membrane potential (voltage clamp applied, fcn.c:448-483) and stimulus
// current (fcn.c:434-443) at time t.
GRW_HD void fcn_voltage(double t, const double* state, const GrwProtocol& p, double& V, double& Istim) {
  const double ton = floor((t - p.shift) / p.period) * p.period;  // :434-443
  const double toff = ton + p.pulse_duration;
  Istim = 0.0;
  if ((t - p.shift) >= ton && (t - p.shift) <= toff) Istim = Istim + p.pulse_amplitude;
  V = state[0];
  (void)clamp_V(t, p, V, DivIeee());  // :448-535
}

struct FcnCurrents {
  double INa, IKr, IKs, IKv43, IKv14, IKv14_K, IKv14_Na, IK1, INab, IKp, INaK, INaCa, ICab, IpCa;
};

// The argument of the IK1 factor's exp (fcn.c:612): 1.76 (V - EK) F/RT
// with EK = RT/F log(Ko / Ki).
GRW_HD double fcn_k1_arg(double V, double lg1) {
  const double EK = RT_over_F * lg1;
  return 1.76 / RT_over_F * (V - EK);
}

// The sarcolemmal channel currents before the CHF scaling (fcn.c:485-623);
// k1x is exp(fcn_k1_arg(V, q.lg[1])).  The pumps and exchanger follow in
// fcn_pumps.
GRW_HD void fcn_currents(const double* state, double V, const FcnTrans& q, double k1x, FcnCurrents& c) {
  const double GKr = 0.029, GKs = 0.035, GK1 = 3.0, GKp = 1.2 * 0.002216, GNa = 12.8;
  const double KmK1 = 13.0;
  const double GNab = 0.001;
  const double KvScale = 1.13 * 1.03 * 1.55, Kv43Frac = 0.77;
  const double GKv43 = Kv43Frac * KvScale * 0.1;
  const double PKv14 = (1.0 - Kv43Frac) * KvScale * 4.792933e-7;
  double mNa = state[1];
  const double hNa = state[2], jNa = state[3], Nai = state[4], Ki = state[5];
  const double xKs = state[8], OH = state[35];

  const double ENa = RT_over_F * q.lg[0];
  const double EK = RT_over_F * q.lg[1];
  const double EKs = RT_over_F * q.lg[2];

  if (mNa <= 0.0) mNa = 1.e-15;
  c.INa = GNa * (mNa * mNa * mNa) * hNa * jNa * (V - ENa);
  const double fKo = pow(Ko / 4.0, 0.5);
  c.IKr = GKr * fKo * OH * (V - EK);
  c.IKs = GKs * (xKs * xKs) * (V - EKs);
  c.IKv43 = GKv43 * state[11 + 4] * (V - EK);
  const double VF_over_RT = V / RT_over_F;
  const double VFsq_over_RT = (1000.0 * Faraday) * VF_over_RT;
  const double exp_VFRT = q.ex[0];
  const double OK14 = state[21 + 4];
  double a1, a2;
  a1 = Ki * exp_VFRT - Ko;
  a2 = exp_VFRT - 1.0;
  c.IKv14_K = (fabs(V) < 1.e-6) ? PKv14 * OK14 * 1000.0 * Faraday * (Ki - Ko)
                                : PKv14 * OK14 * VFsq_over_RT * (a1 / a2);
  a1 = Nai * exp_VFRT - Nao;
  c.IKv14_Na = (fabs(V) < 1.e-6) ? 0.02 * PKv14 * OK14 * 1000.0 * Faraday * (Nai - Nao)
                                 : 0.02 * PKv14 * OK14 * VFsq_over_RT * (a1 / a2);
  c.IKv14 = c.IKv14_K + c.IKv14_Na;
  const double K1_inf = 1.0 / (0.94 + k1x);
  c.IK1 = GK1 * Ko / (Ko + KmK1) * K1_inf * (V - EK);
  c.INab = GNab * (V - ENa);
  const double KpV = 1.0 / (1.0 + q.ex[1]);
  c.IKp = GKp * KpV * (V - EK);
}

// INaK, INaCa, ICab and IpCa before the CHF scaling (fcn.c:625-649).
GRW_HD void fcn_pumps(const double* state, double V, const FcnTrans& q, FcnCurrents& c) {
  const double kNaCa = 0.9 * 0.30, KmNa = 87.5, KmCa = 1.38, ksat = 0.2;
  const double INaKmax = 1.3 * 0.693, KmKo = 1.5, IpCamax = 0.6 * 0.05, KmpCa = 0.0005;
  const double GCab = 3.3 * 7.684e-5;
  const double Nai = state[4];
  double Cai = state[6];
  if (Cai <= 0.0) Cai = 1.e-15;
  const double ECa = 0.5 * RT_over_F * q.lg[3];
  const double exp_VFRT = q.ex[0];
  double a1, a2, a3, a4, a5;
  const double sigma = (exp(Nao / 67.3) - 1.0) / 7.0;
  a1 = 1.0 + 0.1245 * q.ex[2];
  a2 = 0.0365 * sigma / exp_VFRT;
  const double fNaK = 1.0 / (a1 + a2);
  a1 = Ko / (Ko + KmKo);
  a2 = 1.0 + q.pw[0];
  c.INaK = INaKmax * fNaK * (a1 / a2);
  const double exp_eta = q.ex[3];
  a1 = exp_eta * (Nai * Nai * Nai) * Cao;
  a2 = exp_eta / exp_VFRT * (Nao * Nao * Nao) * Cai;
  a3 = 1.0 + ksat * exp_eta / exp_VFRT;
  a4 = KmCa + Cao;
  a5 = 5000.0 / (KmNa * KmNa * KmNa + Nao * Nao * Nao);
  c.INaCa = kNaCa * a5 * (a1 - a2) / (a4 * a3);
  c.ICab = GCab * (V - ECa);
  c.IpCa = IpCamax * Cai / (KmpCa + Cai);
}

// One of the two Ito1 lattices (fcn.c:860-998): Kv4.3 (kv14 = 0, F[11..20])
// or Kv1.4 (kv14 = 1, F[21..30]).  Straight-line code with the lattice's
// constants selected, so that two lanes evaluate the two lattices side by
// side; the Kv1.4 inactivation rates have no exp, and a rate times 1.0 is
// the rate.
GRW_HD void fcn_ito_lattice(const double* state, const FcnTrans& q, int kv14, double* F) {
  constexpr DivK f43[4] = {divk(1.8936), divk(14.224647456), divk(158.574378389), divk(142.936645351)};
  constexpr DivK b43[4] = {divk(6.77348), divk(15.6212705152), divk(28.7532603313), divk(524.576206679)};
  constexpr DivK f14[4] = {divk(0.20005), divk(0.320280), divk(13.50909223), divk(1151.7651385)};
  constexpr DivK b14[4] = {divk(2.2300), divk(12.000299), divk(5.3701338025), divk(5.2396395511)};
  const bool k = kv14 != 0;
  DivK f[4], b[4];
  for (int i = 0; i < 4; ++i) {
    f[i] = k ? f14[i] : f43[i];
    b[i] = k ? b14[i] : b43[i];
  }
  const int i0 = k ? 21 : 11;
  const double aa = (k ? 6.0 * 0.31551 : 0.543708) * q.ex[k ? 31 : 27];
  const double ba = (k ? 6.0 * 0.001966 : 0.080185) * q.ex[k ? 32 : 28];
  const double ai = (k ? 6.0 * 0.0004938 : 0.0498424) * (k ? 1.0 : q.ex[29]);
  const double bi = (k ? 6.0 * 0.0000176 : 0.000819482) * (k ? 1.0 : q.ex[30]);
  double x[2][5], dx[2][5];
  ito_unpack(state, i0, x);
  ItoRates iq;
  ito_rates(aa, ba, ai, bi, f, b, iq);
  ito_chain(x, iq, dx);
  ito_pack(dx, i0, F);
}

// The two Ito1 lattices: F[11..30].
GRW_HD void fcn_ito(const double* state, const FcnTrans& q, double* F) {
  fcn_ito_lattice(state, q, 0, F);
  fcn_ito_lattice(state, q, 1, F);
}

// The INa gates and the HERG chain (fcn.c:651-743): F[1..3], F[32..36].
// Below -90 mV the m gate is set to its steady state (fcn.c:663-669): *m_set
// receives that value, else -1.  xKs follows in fcn_xks.
GRW_HD void fcn_gates(const double* state, double V, const FcnTrans& q, double* F, double* m_set) {
  const double TH = 5.320000001;
  const double C2H_to_C3H = TH * 0.02608362043337, C3H_to_C2H = TH * 0.14832978132145;
  double mNa = state[1];
  const double hNa = state[2], jNa = state[3];
  const double C1H = state[32], C2H = state[33], C3H = state[34], OH = state[35];
  const double IH = 1.0 - OH - C1H - C2H - C3H;
  if (mNa <= 0.0) mNa = 1.e-15;
  double a1, a2, a3;
  *m_set = -1.0;

  double alpha_m;
  if (fabs(V + 47.13) <= 1.e-4) alpha_m = 0.32 / (0.1 - 0.005 * (V + 47.13));
  else alpha_m = (0.32 * (V + 47.13)) / (1.0 - q.ex[4]);
  const double beta_m = 0.08 * q.ex[5];
  double dmNa;
  if (V >= -90.0) {
    dmNa = alpha_m * (1.0 - mNa) - beta_m * mNa;
  } else {
    dmNa = 0.0;
    *m_set = alpha_m / (alpha_m + beta_m);
  }
  const double alpha_h = 0.135 * q.ex[6];
  const double beta_h = (V < -38.73809636838782) ? 3.56 * q.ex[7] + 310000 * q.ex[8]
                                                 : 1 / (0.13 * (1 + q.ex[9]));
  double alpha_j = 0;
  if (V < -37.78) {
    a1 = -127140 * q.ex[10];
    a2 = 3.474E-5 * q.ex[11];
    a3 = 1.0 + q.ex[12];
    alpha_j = (a1 - a2) * (V + 37.78) / a3;
  }
  const double beta_j = (V < -39.82600037702883)
                            ? 0.1212 * q.ex[13] / (1.0 + q.ex[14])
                            : 0.3 * q.ex[15] / (1 + q.ex[16]);
  const double dhNa = alpha_h * (1.0 - hNa) - beta_h * hNa;
  const double djNa = alpha_j * (1.0 - jNa) - beta_j * jNa;

  const double C1H_to_C2H = TH * 0.017147641733086 * q.ex[17];
  const double C2H_to_C1H = TH * 0.03969328381141 * q.ex[18];
  const double C3H_to_OH = TH * 0.02057448605977 * q.ex[19];
  const double OH_to_C3H = TH * 0.00134366604423 * q.ex[20];
  const double OH_to_IH = TH * 0.10666316491288 * q.ex[21];
  const double IH_to_OH = TH * 0.00646393910049 * q.ex[22];
  const double C3H_to_IH = TH * 0.00008039374403 * q.ex[23];
  const double IH_to_C3H = (OH_to_C3H * IH_to_OH * C3H_to_IH) / (C3H_to_OH * OH_to_IH);
  const double dC1H = C2H_to_C1H * C2H - C1H_to_C2H * C1H;
  const double dC2H = (C1H_to_C2H * C1H + C3H_to_C2H * C3H) - ((C2H_to_C1H + C2H_to_C3H) * C2H);
  const double dC3H = (C2H_to_C3H * C2H + OH_to_C3H * OH + IH_to_C3H * IH) -
                      ((C3H_to_IH + C3H_to_OH + C3H_to_C2H) * C3H);
  const double dOH = (C3H_to_OH * C3H + IH_to_OH * IH) - ((OH_to_C3H + OH_to_IH) * OH);

  F[1] = dmNa; F[2] = dhNa; F[3] = djNa;
  F[32] = dC1H; F[33] = dC2H; F[34] = dC3H; F[35] = dOH;
  F[36] = -dC1H - dC2H - dC3H - dOH;
}

// The IKs gate xKs (fcn.c:745-755): F[8].
GRW_HD void fcn_xks(const double* state, double V, const FcnTrans& q, double* F) {
  const double xKs = state[8];
  double a1, a2;
  const double xKs_inf = 1.0 / (1.0 + q.ex[24]);
  if (fabs(V - 10) < 1.e-6) {
    a1 = (7.19e-5) / 0.148;
    a2 = (1.31e-4) / 0.0687;
  } else {
    a1 = 7.19e-5 * (V - 10.0) / (1.0 - q.ex[25]);
    a2 = 1.31e-4 * (V - 10.0) / (q.ex[26] - 1.0);
  }
  const double tau_xKs = 1.0 / (a1 + a2);
  F[8] = (xKs_inf - xKs) / tau_xKs;
}

struct FcnCa {
  double Jup, Jtrpn, beta_i;
};

// SERCA uptake before the CHF scaling, troponin and the myoplasmic buffer
// factor (fcn.c:748-778): F[9], F[10].
GRW_HD void fcn_ca(const double* state, const GrwProtocol& p, const FcnTrans& q, double* F, FcnCa& ca) {
  const double KSR = 1.0;
  const double vmax0 = 1.53 * 137.0e-6;
  const double vmaxf = (p.serca_scale != 1.0) ? vmax0 * p.serca_scale : vmax0;
  const double vmaxr = vmaxf;
  double Cai = state[6];
  const double LTRPNCa = state[9], HTRPNCa = state[10];
  if (Cai <= 0.0) Cai = 1.e-15;
  const double fb = q.pw[1];
  const double rb = q.pw[2];
  ca.Jup = KSR * (vmaxf * fb - vmaxr * rb) / (1.0 + fb + rb);
  const double dLTRPNCa = kltrpn_plus * Cai * (1.0 - LTRPNCa) - kltrpn_minus * LTRPNCa;
  const double dHTRPNCa = khtrpn_plus * Cai * (1.0 - HTRPNCa) - khtrpn_minus * HTRPNCa;
  ca.Jtrpn = LTRPNtot * dLTRPNCa + HTRPNtot * dHTRPNCa;
  const double a1 = CMDNtot * KmCMDN / (pow2(Cai + KmCMDN));
  const double a2 = EGTAtot * KmEGTA / (pow2(Cai + KmEGTA));
  ca.beta_i = 1.0 / (1.0 + a1 + a2);
  F[9] = dLTRPNCa;
  F[10] = dHTRPNCa;
}

// The CHF scaling (fcn.c:780-786) of the currents and the uptake, and Ito1.
GRW_HD void fcn_chf(FcnCurrents& c, FcnCa& ca, double& Ito1, bool chf) {
  Ito1 = c.IKv43 + c.IKv14;
  if (chf) {
    c.IK1 = 0.68 * c.IK1;
    ca.Jup = 0.38 * ca.Jup;
    c.INaCa = 1.75 * c.INaCa;
    c.IKv43 = 0.143 * c.IKv43;
    Ito1 = c.IKv43 + c.IKv14;
  }
}

// The terms of the concentration and voltage derivatives (fcn.c:791-822)
// that do not depend on the release units' fluxes, each as far as the
// reference's left-to-right evaluation gets before the first flux enters.
struct FcnAsm {
  double dNai, dKi;         // complete
  double a3ca, mca;         // ICab - 2 INaCa + IpCa, and that times 0.5 Acap/(Vmyo F 1000)
  double nsr;               // Jup Vmyo / VNSR
  double beta_i, Jup, Jtrpn;
  double INa, IKr, IKs;     // the operands of the first voltage sum around ICa
  double p2;                // IK1 + IKp + INaCa + INaK + Ito1 (Ito2 follows)
  double a3v;               // IpCa + ICab + INab
};
constexpr int kFcnAsm = 13;

GRW_HD FcnAsm fcn_asm_pre(FcnCurrents c, FcnCa ca, double Istim, bool chf) {
  double Ito1;
  fcn_chf(c, ca, Ito1, chf);
  const double a1 = Acap / (Vmyo * Faraday * 1000.0);
  FcnAsm s;
  s.dNai = -(c.INa + c.INab + 3.0 * (c.INaCa + c.INaK) + c.IKv14_Na) * a1;
  const double a3 = c.IKr + c.IKs + c.IK1 + c.IKp;
  s.dKi = -(a3 - 2.0 * c.INaK + (Ito1 - c.IKv14_Na) + Istim) * a1;
  s.a3ca = c.ICab - 2.0 * c.INaCa + c.IpCa;
  s.mca = s.a3ca * 0.5 * a1;
  s.nsr = rdiv(ca.Jup * Vmyo, divk(VNSR));
  s.beta_i = ca.beta_i;
  s.Jup = ca.Jup;
  s.Jtrpn = ca.Jtrpn;
  s.INa = c.INa;
  s.IKr = c.IKr;
  s.IKs = c.IKs;
  s.p2 = c.IK1 + c.IKp + c.INaCa + c.INaK + Ito1;
  s.a3v = c.IpCa + c.ICab + c.INab;
  return s;
}

// The rest with the fluxes J = {Jxfer, Jtr, ICa, Ito2}: F[0], F[4..7], F[31];
// returns Itot.  Under a voltage-clamp protocol state[0] takes the clamp
// potential (fcn.c:811-814, 825-838).
GRW_HD double fcn_asm_fin(const FcnAsm& s, double* state, double V, double Istim, const double J[4], bool vclamp,
                          double* F) {
  const double Jxfer = J[0], Jtr = J[1], ICa = J[2], Ito2 = J[3];
  const double a1c = Acap / (Vmyo * Faraday * 1000.0);
  const double dCai = s.beta_i * (rdiv(Jxfer * VSS, divk(Vmyo)) - s.Jup - s.Jtrpn - s.mca);
  const double dCaNSR = s.nsr - rdiv(Jtr * VJSR, divk(VNSR));
  const double dCaTOT = -(s.a3ca + ICa) * 0.5 * a1c * Vmyo * 1.0e6;
  const double a1 = s.INa + ICa + s.IKr + s.IKs;
  const double a2 = s.p2 + Ito2;
  double dV, Itot;
  if (vclamp) {
    dV = 0.0;
    state[0] = V;
    Itot = a1 + a2 + s.a3v;
  } else {
    Itot = a1 + a2 + s.a3v + Istim;
    dV = -Itot;
  }
  F[0] = dV; F[4] = s.dNai; F[5] = s.dKi; F[6] = dCai; F[7] = dCaNSR; F[31] = dCaTOT;
  return Itot;
}

// The CHF scaling, the concentration and voltage derivatives F[0], F[4..7],
// F[31] and the currents record (fcn.c:780-1071).
GRW_HD void fcn_assemble(double* state, double V, double Istim, FcnCurrents c, FcnCa ca, double* F,
                         double* cur, bool keepc, const double J[4], const GrwProtocol& p) {
  const double Itot = fcn_asm_fin(fcn_asm_pre(c, ca, Istim, p.chf != 0), state, V, Istim, J, clamped(p), F);
  if (keepc && cur) {  // :1046-1071
    const double Jxfer = J[0], Jtr = J[1], ICa = J[2], Ito2 = J[3];
    double Ito1;
    fcn_chf(c, ca, Ito1, p.chf != 0);
    cur[0] = c.INa; cur[1] = c.IKr; cur[2] = c.IKs; cur[3] = Ito1; cur[4] = c.IK1; cur[5] = c.IKp;
    cur[6] = c.INaCa; cur[7] = c.INaK; cur[8] = c.IpCa; cur[9] = c.ICab; cur[10] = c.INab; cur[11] = ICa;
    cur[12] = rdiv(ICa * Acap, divk(-2.0 * Vmyo * Faraday * 1000.0));
    cur[13] = ca.Jup; cur[14] = ca.Jtrpn; cur[15] = rdiv(Jtr * VJSR, divk(Vmyo));
    cur[16] = rdiv(Jxfer * VSS, divk(Vmyo));
    cur[17] = c.IKv43; cur[18] = c.IKv14; cur[19] = c.IKv14_K; cur[20] = c.IKv14_Na; cur[21] = Ito2;
    cur[22] = Istim; cur[23] = Itot;
  }
}

GRW_HD void fcn_finish(double t, double* state, double* F, double* cur, bool keepc, const double J[4],
                       const GrwProtocol& p, const FcnTrans& q) {
  double V, Istim, m_set;
  fcn_voltage(t, state, p, V, Istim);
  FcnCurrents c;
  fcn_currents(state, V, q, exp(fcn_k1_arg(V, q.lg[1])), c);
  fcn_pumps(state, V, q, c);
  fcn_ito(state, q, F);
  fcn_gates(state, V, q, F, &m_set);
  fcn_xks(state, V, q, F);
  FcnCa ca;
  fcn_ca(state, p, q, F, ca);
  if (m_set >= 0.0) state[1] = m_set;
  fcn_assemble(state, V, Istim, c, ca, F, cur, keepc, J, p);
}

// fcn.c:29-1072.  J = {Jxfer, Jtr, ICa, Ito2} from the release units.  Mutates
// state[1] below -90 mV (:663-669) and state[0] under voltage clamp (:811-814).
GRW_HD void fcn(double t, double* state, double* F, double* cur, bool keepc, const double J[4],
                const GrwProtocol& p) {
  FcnTrans q;
  fcn_prep(t, state, p, q);
  fcn_trans(q);
  fcn_finish(t, state, F, cur, keepc, J, p, q);
}

}  // namespace grw
