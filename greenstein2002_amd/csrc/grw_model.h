// grw_model.h -- release-unit model: constants, channel-state packing and the
// per-FRU rate / local-Ca arithmetic shared by the gfx950 kernels and the host.
//
// Arithmetic follows the reference expression by expression (same operation
// order, no contraction: every TU including this header is built with
// -ffp-contract=off) so that, in the FP64 form, results equal the reference's
// up to the last-ulp differences of the device exp().  The template parameter R
// is the precision of rate and right-hand-side evaluation: double for the FP64
// drop-in, float for the mixed form.  State, time, step length and transition
// thresholds are always double.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define GRW_HD __host__ __device__ __forceinline__

namespace grw {

constexpr int kNcleft = 4;
constexpr int kNryr = 5;
constexpr int kMtN = 624;
constexpr int kMtStride = 640;  // words per MT row in HBM (2560 B, 128 B aligned)
// Row positions 624..631 (in the row's padding) always hold a second copy of
// words 0..7: every writer of a row stores both copies, so that a window of
// words running past the end of the generation reads on without wrapping.
constexpr int kMtMirror = 8;
static_assert(kMtN + kMtMirror <= kMtStride, "the copy lives in the row padding");

// ---- physical constants: parameters_fcn_fru.c:21-90 (same constant expressions)
constexpr double Ko = 4.0, Nao = 138.0, Cao = 2.0, Clo = 150.0, Cli = 20.0;
constexpr double Faraday = 96.5;
constexpr double RT_over_F = 8.314 * 310.0 / 96.5;
constexpr double Acap = 1.534e-4;
constexpr double VNSR = (0.53 * 2.10e-6);
constexpr double Vmyo = 25.84e-6;
constexpr double VJSR = (((double)4) * 0.53 * 0.5 * 0.2 * 1.05e-10);
constexpr double VSS = (0.5 * 0.2 * 2.03e-12);
constexpr double PCa = (1.5 / 2.8 * 0.2 * 0.9 * 0.9468e-11);
constexpr double PCl = 2.65e-15;
constexpr double BSLtot = 1.124, CSQNtot = 13.5, BSRtot = 0.047;
constexpr double KBSL = 0.0087, KmCSQN = 0.63, KBSR = 0.00087;
constexpr double LTRPNtot = 70.0e-3, HTRPNtot = 140.0e-3;
constexpr double khtrpn_plus = 20.0, khtrpn_minus = 0.066e-3;
constexpr double kltrpn_plus = 40.0, kltrpn_minus = 40.0e-3;
constexpr double CMDNtot = 50.0e-3, EGTAtot = 0.0, KmCMDN = 2.38e-3, KmEGTA = 1.5e-4;
constexpr double JRyRmax = (0.6 * 1.5 * 1.1 * 3.96);
constexpr double tautr = 3.0, tauxfer = 0.005, tauss2ss = (10.0 * 0.005);

// ---- LCC / Vinact / Ito2 constants: fru_rates_local.c:50-96
constexpr double fL = 0.85, gL = 2.0, fLprime = 0.005, gLprime = 7.0;
constexpr double bL = 1.9356, aL = 2.0;
constexpr double bL2 = bL * bL, bL3 = bL * bL * bL, bL4 = bL * bL * bL * bL;
constexpr double omega = 0.83 * 2.0 * 1.3 * 0.01;
constexpr double alphacf = 4.0 * 1.2 * 0.416;
constexpr double betacf = 4.0 * 0.45 * 0.049;
constexpr double gammacf = 0.83 * 1.9 * 1.3 * 0.31 * 7.5 * 0.09233;
constexpr double yCa_frac = 0.4;
constexpr double KdIto2 = 0.1502, kbIto2 = 2.0, kfIto2 = kbIto2 / KdIto2;
// ---- RyR constants: fru_rates_local.c:100-121
constexpr double k21 = 250.0, k32 = 9.6, k43 = 0.07 / 0.06667 * 13.0, k45 = 0.07;
constexpr double k52 = 0.001235, k65 = 30.0;
constexpr double k12cf = 3000.0, k23cf = 10.0 * 30000.0, k34cf = 0.6 * 3000.0;
constexpr double k54cf = 0.6 * 0.198, k25cf = 10.0 * 300.0, k56cf = 2.0 * 4.0 * 3000.0;
constexpr double k_rate = 0.00127215, k_rate2 = 3.4188;
constexpr double threshCa34to7 = 0.0368369379834969;
constexpr double threshCa56to8 = 0.00011447933531005;
constexpr double threshMAX = 2.0;
constexpr double threshMAXCa = 0.0504410547074504;

// min/max macros of parameters.h:316-322 (NaN behaviour included)
template <class T> GRW_HD T rmax(T a, T b) { return (a < b) ? b : a; }
template <class T> GRW_HD T rmin(T a, T b) { return (a < b) ? a : b; }

#if defined(__HIPCC__)
// IEEE maxNum / minNum, one v_max / v_min instruction instead of a compare and
// two selects.  They equal rmax / rmin whenever the first operand is not NaN
// (a NaN second operand is dropped by both) and the operands are not zeros of
// opposite sign.  The Monte Carlo step uses them where that holds: the largest
// exit rate is an accumulator that starts at +0 and takes only larger rates,
// and step lengths and subspace Ca are positive.
__device__ __forceinline__ double vmax(double a, double b) { return __builtin_fmax(a, b); }
__device__ __forceinline__ double vmin(double a, double b) { return __builtin_fmin(a, b); }
__device__ __forceinline__ float vmax(float a, float b) { return __builtin_fmaxf(a, b); }
__device__ __forceinline__ float vmin(float a, float b) { return __builtin_fminf(a, b); }
// The same instruction for operands the compiler cannot prove canonical (LDS
// loads, lane permutes), before which it would insert a canonicalising
// v_max x, x, x each: the two differ only for signalling NaNs, which no
// arithmetic produces.
__device__ __forceinline__ double vmax_nc(double a, double b) {
  double r;
  asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}
__device__ __forceinline__ float vmax_nc(float a, float b) {
  float r;
  asm("v_max_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}
#endif

// Division and exp in the rate and local-Ca arithmetic: IEEE in FP64 (the
// drop-in keeps the reference's operations bit for bit); in the FP32 mixed form
// (device only) the hardware reciprocal and exp2 (v_rcp_f32, v_exp_f32, about
// 1 ulp), whose error is far below the FP32 rounding the mixed form accepts.
//
// On the device the FP64 division is the hardware's IEEE sequence (v_rcp_f64,
// two Newton steps, the quotient and one fma correction) without its scaling
// and fix-up instructions: v_div_scale leaves both operands unchanged and
// v_div_fmas is a plain fma unless an operand is zero, infinite or denormal,
// their exponents lie 768 or more apart, or the numerator is within 2^53 of
// the smallest normal, and v_div_fixup then returns the quotient with the sign
// it already has.  The rate and buffer terms divided here are finite, nonzero
// denominators of moderate size (1 + exp(.), tau >= 60, (c + K)^2, rate sums
// with a positive constant) and numerators that are moderate or +0, so the
// eight operations give IEEE a / b bit for bit.  (The one quotient that can
// see a zero denominator, the GHK factor at V = 0, is replaced by its limit
// form there and never used.)
GRW_HD double rdiv(double a, double b) {
#if defined(__HIP_DEVICE_COMPILE__)
  const double r0 = __builtin_amdgcn_rcp(b);
  const double r1 = fma(r0, fma(-b, r0, 1.0), r0);
  const double r2 = fma(r1, fma(-b, r1, 1.0), r1);
  const double q0 = a * r2;
  return fma(fma(-b, q0, a), r2, q0);
#else
  return a / b;
#endif
}
__device__ __forceinline__ float rdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// Division by a model constant b, with rb = RN(1/b) fixed at compile time.  On
// the device in FP64: q0 = RN(a rb) is within an ulp of a/b, the remainder
// a - b q0 is exact in one fma, and RN(q0 + r rb) is the correctly rounded
// a/b (Markstein's theorem: rb within 2^-53 relative of 1/b, q0 faithful, no
// overflow or underflow, which the model's variables never approach), so the
// result equals IEEE division bit for bit in three operations instead of the
// eleven of the general sequence.  The host divides; FP32 keeps rdiv(float).
struct DivK {
  double b, rb;
};
constexpr DivK divk(double b) { return DivK{b, 1.0 / b}; }

GRW_HD double rdiv(double a, DivK k) {
#if defined(__HIP_DEVICE_COMPILE__)
  const double q0 = a * k.rb;
  return fma(fma(-k.b, q0, a), k.rb, q0);
#else
  return a / k.b;
#endif
}
__device__ __forceinline__ float rdiv(float a, DivK k) { return rdiv(a, (float)k.b); }
GRW_HD double xexp(double x) { return exp(x); }
__device__ __forceinline__ float xexp(float x) { return __expf(x); }

// ---- packed channel word, one per cleft
//   bits 0-3  LCC state 1..12      bit 4  voltage gate (0: y=1 inactivated, 1: y=2)
//   bits 5+3i RyR i state-1 (0..7) bit 20 Ito2 (0: closed, 1: open)
GRW_HD int ch_lcc(uint32_t w) { return (int)(w & 15u); }
GRW_HD int ch_y(uint32_t w) { return (int)((w >> 4) & 1u) + 1; }
GRW_HD int ch_ryr(uint32_t w, int i) { return (int)((w >> (5 + 3 * i)) & 7u) + 1; }
GRW_HD int ch_ito2(uint32_t w) { return (int)((w >> 20) & 1u) + 1; }
GRW_HD uint32_t ch_set_lcc(uint32_t w, int s) { return (w & ~15u) | (uint32_t)s; }
GRW_HD uint32_t ch_flip_y(uint32_t w) { return w ^ (1u << 4); }
GRW_HD uint32_t ch_set_ryr(uint32_t w, int i, int s) {
  const int sh = 5 + 3 * i;
  return (w & ~(7u << sh)) | ((uint32_t)(s - 1) << sh);
}
GRW_HD uint32_t ch_flip_ito2(uint32_t w) { return w ^ (1u << 20); }
GRW_HD uint32_t ch_pack(int lcc, int y, const int* ryr, int ito2) {
  uint32_t w = (uint32_t)lcc | ((uint32_t)(y - 1) << 4) | ((uint32_t)(ito2 - 1) << 20);
  for (int i = 0; i < kNryr; ++i) w |= (uint32_t)(ryr[i] - 1) << (5 + 3 * i);
  return w;
}
// open indicators: simfru_local.c:101-117
// by bit tables: bit s of 0x1040 marks the open LCC states 6 and 12, bit
// s - 1 of 0x4C the open RyR states 3, 4 and 7
GRW_HD bool ch_lcc_open(uint32_t w) {
  return ((0x1040u >> (w & 15u)) & (w >> 4) & 1u) != 0u;
}
GRW_HD int ch_ryr_open(uint32_t w) {
  uint32_t m = 0;
  for (int i = 0; i < kNryr; ++i) m |= ((0x4Cu >> ((w >> (5 + 3 * i)) & 7u)) & 1u) << i;
  return __builtin_popcount(m);
}

// ---- exit list of one channel (at most 3 exits)
template <class R>
struct Exits {  // fully defined whatever the number of exits
  R tot = R(0);
  R r[3] = {R(0), R(0), R(0)};
  int8_t d[3] = {0, 0, 0};
  int8_t n = 0;
};

// LCC mode-switching chain, fru_rates_local.c:164-341.  States 1-5 C0..C4,
// 6 O, 7-11 CCa0..CCa4, 12 OCa.  Totals keep the reference's summation order.
template <class R>
GRW_HD bool lcc_exits(int s, R alpha, R beta, R gam, Exits<R>& e) {
  const R ap = R(aL) * alpha;
  const R bp = beta / R(bL * 1.0);
  switch (s) {
    case 1: e.tot = (R(4.0) * alpha + gam); e.r[0] = R(4.0) * alpha; e.d[0] = 2; e.r[1] = gam; e.d[1] = 7; e.n = 2; return true;
    case 2: e.tot = (beta + R(3.0) * alpha + R(aL) * gam); e.r[0] = beta; e.d[0] = 1; e.r[1] = R(3.0) * alpha; e.d[1] = 3; e.r[2] = R(aL) * gam; e.d[2] = 8; e.n = 3; return true;
    case 3: e.tot = (R(2.0) * beta + R(2.0) * alpha + R(aL * aL) * gam); e.r[0] = R(2.0) * beta; e.d[0] = 2; e.r[1] = R(2.0) * alpha; e.d[1] = 4; e.r[2] = R(aL * aL) * gam; e.d[2] = 9; e.n = 3; return true;
    case 4: e.tot = (R(3.0) * beta + alpha + R(aL * aL * aL) * gam); e.r[0] = R(3.0) * beta; e.d[0] = 3; e.r[1] = alpha; e.d[1] = 5; e.r[2] = R(aL * aL * aL) * gam; e.d[2] = 10; e.n = 3; return true;
    case 5: e.tot = (R(4.0) * beta + R(fL) + R(aL * aL * aL * aL) * gam); e.r[0] = R(4.0) * beta; e.d[0] = 4; e.r[1] = R(fL); e.d[1] = 6; e.r[2] = R(aL * aL * aL * aL) * gam; e.d[2] = 11; e.n = 3; return true;
    case 6: e.tot = R(gL); e.r[0] = R(gL); e.d[0] = 5; e.n = 1; return true;
    case 7: e.tot = (R(4.0) * ap + R(omega)); e.r[0] = R(omega); e.d[0] = 1; e.r[1] = R(4.0) * ap; e.d[1] = 8; e.n = 2; return true;
    case 8: e.tot = (bp + R(3.0) * ap + R(omega / bL)); e.r[0] = bp; e.d[0] = 7; e.r[1] = R(omega / bL); e.d[1] = 2; e.r[2] = R(3.0) * ap; e.d[2] = 9; e.n = 3; return true;
    case 9: e.tot = (R(2.0) * bp + R(2.0) * ap + R(omega / bL2)); e.r[0] = R(2.0) * bp; e.d[0] = 8; e.r[1] = R(omega / bL2); e.d[1] = 3; e.r[2] = R(2.0) * ap; e.d[2] = 10; e.n = 3; return true;
    case 10: e.tot = (R(3.0) * bp + ap + R(omega / bL3)); e.r[0] = R(3.0) * bp; e.d[0] = 9; e.r[1] = R(omega / bL3); e.d[1] = 4; e.r[2] = ap; e.d[2] = 11; e.n = 3; return true;
    case 11: e.tot = (R(4.0) * bp + R(omega / bL4) + R(fLprime)); e.r[0] = R(4.0) * bp; e.d[0] = 10; e.r[1] = R(omega / bL4); e.d[1] = 5; e.r[2] = R(fLprime); e.d[2] = 12; e.n = 3; return true;
    case 12: e.tot = R(gLprime); e.r[0] = R(gLprime); e.d[0] = 11; e.n = 1; return true;
    default: e.tot = R(0); e.n = 0; return false;
  }
}

// RyR rate constants of one cleft (fru_rates_local.c:375-382) and its regime
// (fru_rates_local.c:384, :505): 1 = low CaSS, 2 = middle, 3 = high.
template <class R>
struct RyrK {
  R k12, k23, k34, k54, k25, k56;
  int regime;
};
template <class R>
GRW_HD RyrK<R> ryr_k(R ca, R k23c) {
  RyrK<R> k;
  const R sat = rmin(R(threshMAX), rdiv(ca * ca, divk(k_rate)));
  const R sat2 = rmin(R(threshMAX), rdiv(ca * ca, divk(k_rate2)));
  k.k12 = R(k12cf) * sat2;
  k.k23 = k23c * sat;
  k.k34 = R(k34cf) * sat;
  k.k54 = R(k54cf) * sat;
  k.k25 = R(k25cf) * sat;
  k.k56 = R(k56cf) * sat;
  const R cm = rmin(ca, R(threshMAXCa));
  k.regime = (cm < R(threshCa56to8)) ? 1 : (cm < R(threshCa34to7)) ? 2 : 3;
  return k;
}

// Exits of a RyR whose state is already resolved for the regime (states 3/4
// never occur in regime 3, 5/6 never in regimes 2/3, 7 and 8 never in regime 1).
// fru_rates_local.c:390-448, :511-549, :603-614, :627-645, :703-727.
template <class R>
GRW_HD void ryr_exits(int s, const RyrK<R>& k, Exits<R>& e) {
  const int g = k.regime;
  switch (s) {
    case 1: e.tot = k.k12; e.r[0] = k.k12; e.d[0] = 2; e.n = 1; return;
    case 2: e.tot = (R(k21) + k.k23 + k.k25); e.r[0] = R(k21); e.d[0] = 1; e.r[1] = k.k23; e.d[1] = (g == 3) ? 7 : 3;
      e.r[2] = k.k25; e.d[2] = (g == 1) ? 5 : 8; e.n = 3; return;
    case 3: e.tot = (R(k32) + k.k34); e.r[0] = R(k32); e.d[0] = 2; e.r[1] = k.k34; e.d[1] = 4; e.n = 2; return;
    case 4: e.tot = (R(k43) + R(k45)); e.r[0] = R(k43); e.d[0] = 3; e.r[1] = R(k45); e.d[1] = (g == 1) ? 5 : 8; e.n = 2; return;
    case 5: e.tot = (R(k52) + k.k54 + k.k56); e.r[0] = R(k52); e.d[0] = 2; e.r[1] = k.k54; e.d[1] = 4; e.r[2] = k.k56; e.d[2] = 6; e.n = 3; return;
    case 6: e.tot = R(k65); e.r[0] = R(k65); e.d[0] = 5; e.n = 1; return;
    case 7: {
      const R a1 = rdiv(k.k34, R(k43) + k.k34);
      const R a2 = (R(1.0) - a1) * R(k32);
      const R a3 = a1 * R(k45);
      e.tot = (a2 + a3); e.r[0] = a2; e.d[0] = 2; e.r[1] = a3; e.d[1] = 8; e.n = 2; return;
    }
    default: {  // 8
      const R a1 = rdiv(R(k65), k.k56 + R(k65));
      const R a2 = a1 * R(k52);
      const R a3 = a1 * k.k54;
      e.tot = (a2 + a3); e.r[0] = a2; e.d[0] = 2; e.r[1] = a3; e.d[1] = (g == 3) ? 7 : 4; e.n = 2; return;
    }
  }
}

// Does state s need one uniform to be split before its rates exist?
// fru_rates_local.c:450-496 (regime 1, states 7 and 8) and :579-601 (regime 2, state 7).
GRW_HD constexpr bool ryr_needs_draw(int s, int regime) {
  return (regime == 1 && (s == 7 || s == 8)) || (regime == 2 && s == 7);
}
// Deterministic collapse into the merged states, :551-577 and :647-701.
GRW_HD constexpr int ryr_collapse(int s, int regime) {
  if (regime == 3 && (s == 3 || s == 4)) return 7;
  if (regime != 1 && (s == 5 || s == 6)) return 8;
  return s;
}
// Split of a merged state with uniform u: 7 -> 4|3, 8 -> 6|5.
template <class R>
GRW_HD int ryr_split(int s, double u, const RyrK<R>& k) {
  if (s == 7) return (u < (double)rdiv(k.k34, k.k34 + R(k43))) ? 4 : 3;
  return (u < (double)rdiv(k.k56, k.k56 + R(k65))) ? 6 : 5;
}

// ---- local Ca right-hand side: fcn_fru.c:23-137
template <class R>
GRW_HD void fcn_fru(const R conc[5], R V, R Cai, R CaNSR, const bool lopen[4], const int nopen[4],
                    R d[5]) {
  const R CaJSR = conc[0];
  const R JDHconstant = R(1.0) / (R(2.0) * R(VSS) * R(Faraday) * R(1000.0));
  const R Jtr = (CaNSR - CaJSR) / R(tautr);
  R JRyR[4], Jxfer[4], Jss2ss[4], JDHPR[4];
  for (int c = 0; c < 4; ++c) JRyR[c] = R(JRyRmax) * ((R)nopen[c]) * (CaJSR - conc[c + 1]);
  const R JRyRtot = JRyR[0] + JRyR[1] + JRyR[2] + JRyR[3];
  for (int c = 0; c < 4; ++c) Jxfer[c] = (conc[c + 1] - Cai) / R(tauxfer);
  for (int c = 0; c < 4; ++c) Jss2ss[c] = (conc[c + 1] - conc[(c + 1) % 4 + 1]) / R(tauss2ss);
  const R VF_over_RT = V / R(RT_over_F);
  const R VFsq_over_RT = (R(1000.0) * R(Faraday)) * VF_over_RT;
  const R e = xexp(R(2.0) * VF_over_RT);
  if (fabs(V) < R(1.e-6)) {
    for (int c = 0; c < 4; ++c)
      JDHPR[c] = -R(PCa) * R(2.0) * R(1000.0) * R(Faraday) * (conc[c + 1] - R(Cao) * R(0.341)) *
                 (lopen[c] ? R(1.0) : R(0.0)) * JDHconstant;
  } else {
    const R a2 = R(PCa) * R(4.0) * VFsq_over_RT / (e - R(1.0)) * JDHconstant;
    for (int c = 0; c < 4; ++c)
      JDHPR[c] = -(conc[c + 1] * e - R(Cao) * R(0.341)) * a2 * (lopen[c] ? R(1.0) : R(0.0));
  }
  R bSS[4];
  for (int c = 0; c < 4; ++c) {
    const R s = conc[c + 1];
    const R a1 = R(BSRtot) * R(KBSR) / ((s + R(KBSR)) * (s + R(KBSR)));
    const R a2 = R(BSLtot) * R(KBSL) / ((s + R(KBSL)) * (s + R(KBSL)));
    bSS[c] = R(1.0) / (R(1.0) + a1 + a2);
  }
  const R a1 = R(CSQNtot) * R(KmCSQN) / ((CaJSR + R(KmCSQN)) * (CaJSR + R(KmCSQN)));
  const R bJSR = R(1.0) / (R(1.0) + a1);
  d[0] = bJSR * (Jtr - R(VSS) / R(VJSR) * JRyRtot);
  for (int c = 0; c < 4; ++c)
    d[c + 1] = bSS[c] * (JDHPR[c] + JRyR[c] - Jxfer[c] - Jss2ss[c] + Jss2ss[(c + 3) % 4]);
}

// ---- reduced sums -> whole-cell fluxes (mpi_master.c:55-112); num = {sum CaSS,
// sum CaJSR, sum CaSS over open LCCs, open LCC count, open Ito2 count}
// e1 = exp(V F/RT) (flux_exp), which a caller can evaluate before the sums exist.
GRW_HD double flux_exp(double V) { return exp(rdiv(V, divk(RT_over_F))); }

GRW_HD void flux_currents_e(const double num[5], double V, double e1, double Cai, double CaNSR, int nfru,
                            double scale, double out4[4]) {
  const double VF_over_RT = rdiv(V, divk(RT_over_F));
  const double VFsq_over_RT = (1000.0 * Faraday) * VF_over_RT;
  const double e2 = e1 * e1;
  double ICa, Ito2;
  if (fabs(V) < 1.e-6) {
    const double icanum = num[2] - num[3] * Cao * 0.341;
    ICa = PCa * 2.0 * 1000.0 * Faraday * icanum;
    ICa = rdiv(ICa, divk(Acap));
    Ito2 = ((double)num[4]) * PCl * 1000.0 * Faraday * (Clo - Cli);
    Ito2 = rdiv(Ito2, divk(Acap));
  } else {
    const double icanum = num[2] * e2 - num[3] * Cao * 0.341;
    ICa = PCa * 4.0 * VFsq_over_RT * icanum / (e2 - 1.0);
    ICa = rdiv(ICa, divk(Acap));
    Ito2 = ((double)num[4]) * PCl * VFsq_over_RT * (Cli - Clo * e1) / (1.0 - e1);
    Ito2 = rdiv(Ito2, divk(Acap));
  }
  const double Jtr = rdiv(((double)nfru) * CaNSR - num[1], divk(tautr));
  const double Jxfer = rdiv(num[0] - ((double)nfru * (double)4) * Cai, divk(tauxfer));
  out4[0] = scale * Jxfer;
  out4[1] = scale * Jtr;
  out4[2] = scale * ICa;
  out4[3] = scale * Ito2;
}

GRW_HD void flux_currents(const double num[5], double V, double Cai, double CaNSR, int nfru,
                          double scale, double out4[4]) {
  flux_currents_e(num, V, flux_exp(V), Cai, CaNSR, nfru, scale, out4);
}

// ---- average outputs (mpi_master.c:175-190); labels of the last two outputs
// are swapped in the reference and kept so.
GRW_HD void avg_outputs(const double num[11], const double* state, int nfru, double scale,
                        double other[12]) {
  const double Cai = state[6], CaNSR = state[7], LTRPNCa = state[9], HTRPNCa = state[10];
  other[0] = num[0] / (double)(nfru * 4);
  other[1] = num[1] / (double)(nfru);
  other[2] = num[2] * scale * VSS / Vmyo;
  other[3] = num[3] / (double)(5 * 4 * nfru);
  other[4] = (double)(num[4]) / (double)(5 * 4 * nfru);
  other[8] = 1.e6 * ((num[5] * VSS + num[6] * VJSR) * scale + CaNSR * VNSR +
                     Cai * Vmyo * (1.0 + CMDNtot / (KmCMDN + Cai) + EGTAtot / (KmEGTA + Cai)) +
                     (LTRPNCa * LTRPNtot + HTRPNCa * HTRPNtot) * Vmyo);
  other[5] = num[7] / (double)(nfru * 4);
  other[6] = num[8] / (double)(nfru * 4);
  other[7] = num[9] / (double)(nfru * 4);
  other[9] = num[10] / (double)(nfru * 4);
  other[10] = 1.e6 * num[5] * VSS * scale;
  other[11] = 1.e6 * num[6] * VJSR * scale;
}

}  // namespace grw
