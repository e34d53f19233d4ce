// grw_cell.cpp -- host-side global integrator over the release-unit ABI.
//
// rk54pd follows rk_pd54m.c:29-345 step for step: each RK5(4)6M stage runs one
// Monte Carlo sub-interval of all FRUs through grw_distrib_simfru, accepted steps
// checkpoint the FRUs (grw_save_state), rejected ones roll them back
// (grw_resume_state), and the step size is aligned to the stimulus onsets.
// grw_run_cell is the 1 ms output loop of caf.c:150-251.
#include <cmath>
#include <cstring>

#include "../../include/grw_cell.h"
#include "grw_global.h"

using namespace grw;

namespace {
inline double dmin(double a, double b) { return (a < b) ? a : b; }
inline double dmax(double a, double b) { return (a < b) ? b : a; }
}  // namespace

extern "C" {

void grw_protocol_default(GrwProtocol* p) {
  std::memset(p, 0, sizeof *p);
  p->pulse_duration = 1;
  p->pulse_amplitude = -100;
  p->period = 1000;
  p->shift = 10;
  p->vclamp = 0;
  p->chf = 0;
  p->vclamp_duration = 200.0;
  p->vclamp_set = 0.0;
  p->vclamp_shift = 300.0;
  p->vclamp_hold = -80.0;
  p->vclamp_period = 500.0;
  p->serca_scale = 1.0;
  p->time_step = 1.0;
  p->step_min = 1.e-6;
  p->step_max = 0.1;
  p->tolrk = 1.e-6;
  p->vppclamp = 0;  // parameters.h:128-137
  p->vpp_ppset = -40.0;
  p->vpp_ppduration = 100.0;
  p->vpp_shift = 10000.0;
  p->vpp_set = 0.0;
  p->vpp_duration = 300.0;
  p->vpp_hold = -80.0;
  p->vpp_period = 500.0;
  p->ivclamp = 0;  // parameters.h:141-147
  p->iv_duration = 50.0;
  p->iv_period = 200.0;
  p->iv_shift = 10.0;
  p->iv_set = -40.0;
  p->iv_hold = -100.0;
  p->iv_step = 10.0;
  p->iv_n = 0;
}

int grw_fcn(const GrwProtocol* p, double t, double state[37], double F[37], double cur[24],
            int keepc, const double J[4]) {
  if (!p || !state || !F || !J) return GRW_EINVAL;
  fcn(t, state, F, cur, keepc != 0, J, *p);
  return 0;
}

int grw_rk54pd(grw_ctx* ctx, const GrwProtocol* p, const double errw[37], double t, double tstep,
               double oldh, double state[37], double cur[24], double* newh, double* log,
               int log_cap, int* log_n, GrwRkStats* st) {
  constexpr int N = kNglobal;
  if (!ctx || !p || !errw || !state || !newh) return GRW_EINVAL;
  const double min_tolvalue = 1.e-10, huge_number = 1e99;
  const double step_min = p->step_min, step_max = p->step_max, tol = p->tolrk;
  double F[N], F1[N], y[N], ym[N], k[6][N];
  double dep0[3], depf[3], J[4] = {0, 0, 0, 0};
  double dummy[kNcur];
  int rc;

  double h = dmin(oldh, tstep - t);
  h = dmin(step_max, h);
  double oh = h, start = t;
  bool notdone = true, success = true, keepc = true;
  while (notdone) {
    if (st) st->attempts++;
    dep0[0] = state[0]; dep0[1] = state[6]; dep0[2] = state[7];
    if (success) {
      if ((rc = grw_calc_flux(ctx, dep0, J))) return rc;
      fcn(start, state, F1, cur, keepc, J, *p);
      keepc = false;
      if ((rc = grw_save_state(ctx))) return rc;
    }
    // stages 2..6, each preceded by one Monte Carlo sub-interval (:110-218)
    for (int i = 0; i < N; ++i) { k[0][i] = h * F1[i]; y[i] = state[i] + k[0][i] / 5.0; }
    for (int s = 1; s <= 5; ++s) {
      // sub-interval of stage s+1 with the reference's time expressions
      double ta, tb;
      switch (s) {
        case 1: ta = start; tb = start + h / 5.0; break;
        case 2: ta = start + h / 5.0; tb = start + h * 3.0 / 10.0; break;
        case 3: ta = start + h * 3.0 / 10.0; tb = start + h * 3.0 / 5.0; break;
        case 4: ta = start + h * 3.0 / 5.0; tb = start + h * 2.0 / 3.0; break;
        default: ta = start + h * 2.0 / 3.0; tb = start + h; break;
      }
      depf[0] = y[0]; depf[1] = y[6]; depf[2] = y[7];
      if ((rc = grw_distrib_simfru(ctx, ta, tb, dep0, depf, J))) return rc;
      dep0[0] = state[0]; dep0[1] = state[6]; dep0[2] = state[7];
      fcn(tb, y, F, dummy, false, J, *p);
      for (int i = 0; i < N; ++i) k[s][i] = h * F[i];
      switch (s) {
        case 1:
          for (int i = 0; i < N; ++i) y[i] = state[i] + (3.0 / 40.0 * k[0][i] + 9.0 / 40.0 * k[1][i]);
          break;
        case 2:
          for (int i = 0; i < N; ++i)
            y[i] = state[i] + (3.0 / 10.0 * k[0][i] - 9.0 / 10.0 * k[1][i] + 6.0 / 5.0 * k[2][i]);
          break;
        case 3:
          for (int i = 0; i < N; ++i)
            y[i] = state[i] + 226.0 / 729.0 * k[0][i] - 25.0 / 27.0 * k[1][i] + 880.0 / 729.0 * k[2][i] +
                   55.0 / 729.0 * k[3][i];
          break;
        case 4:
          for (int i = 0; i < N; ++i)
            y[i] = state[i] - 181.0 / 270.0 * k[0][i] + 5.0 / 2.0 * k[1][i] - 266.0 / 297.0 * k[2][i] -
                   91.0 / 27.0 * k[3][i] + 189.0 / 55.0 * k[4][i];
          break;
        default:
          break;
      }
    }
    // truncation error :225-238
    double err = 0.0;
    for (int i = 0; i < N; ++i) {
      y[i] = state[i] + 31.0 / 540.0 * k[0][i] + 190.0 / 297.0 * k[2][i] - 145.0 / 108.0 * k[3][i] +
             351.0 / 220.0 * k[4][i] + 1.0 / 20.0 * k[5][i];
      ym[i] = state[i] + 19.0 / 216.0 * k[0][i] + 1000.0 / 2079.0 * k[2][i] - 125.0 / 216.0 * k[3][i] +
              81.0 / 88.0 * k[4][i] + 5.0 / 56.0 * k[5][i];
      const double e = std::fabs((ym[i] - y[i]) * 0.2 * errw[i]);
      if (e < huge_number) err = dmax(e, err);
      else err = huge_number;
    }
    const bool accept = err < tol;
    if (accept || h <= step_min) {
      for (int i = 0; i < N; ++i) {
        if (std::fabs(y[i]) < min_tolvalue) y[i] = 0.0;
        state[i] = y[i];
      }
      start = start + h;
      success = true;
      if (!accept && st) st->forced++;
      if (log && log_n && *log_n < log_cap) {
        double* L = log + 4 * (*log_n);
        L[0] = start; L[1] = state[0]; L[2] = state[6]; L[3] = J[2];
        ++*log_n;
      }
      if (start >= tstep) {
        notdone = false;
      } else if (accept) {  // :258-285
        h = dmin(0.85 * h * std::pow(tol / err, 0.2), 4.0 * h);
        oh = h;
        double tts = p->shift - std::fmod(start, p->period);
        if (tts < 0.0) tts = tts + p->period;
        if (tts < h) h = dmax(step_min, tts - 0.5 * step_min);
      } else {  // forced accept :291-309
        h = step_min;
      }
      if (!accept) {
        oh = h;
        h = dmin(step_max, dmax(step_min, h));
        h = dmin(h, tstep - start);
      }
    } else {  // reject :310-333
      if (st) st->rejects++;
      h = 0.85 * h * std::pow(tol / err, 0.2);
      double tts = p->shift - std::fmod(start, p->period);
      if (tts < 0.0) tts = tts + p->period;
      if (tts < h) h = dmax(step_min, tts - 0.5 * step_min);
      success = false;
      if ((rc = grw_resume_state(ctx))) return rc;
      h = dmin(step_max, dmax(step_min, h));
      h = dmin(h, tstep - start);
    }
    oldh = oh;
  }
  *newh = oldh;
  return 0;
}

int grw_run_cell(grw_ctx* ctx, const GrwProtocol* p, const double errw[37], double t0, int nsteps,
                 double state[37], double* h, double* states, double* currents, double* other,
                 double* log, int log_cap, int* log_n, GrwRkStats* st) {
  if (!ctx || !p || !errw || !state || !h || nsteps < 0) return GRW_EINVAL;
  double cur[kNcur] = {0};
  double t = t0;
  for (int n = 0; n < nsteps; ++n) {
    const double tstep = t + p->time_step;
    int rc = grw_rk54pd(ctx, p, errw, t, tstep, *h, state, cur, h, log, log_cap, log_n, st);
    if (rc) return rc;
    if (states) std::memcpy(states + (size_t)n * kNglobal, state, sizeof(double) * kNglobal);
    if (currents) std::memcpy(currents + (size_t)n * kNcur, cur, sizeof(double) * kNcur);
    if (other && (rc = grw_calc_avg(ctx, state, other + (size_t)n * 12))) return rc;
    t = t + p->time_step;
  }
  return 0;
}

}  // extern "C"
