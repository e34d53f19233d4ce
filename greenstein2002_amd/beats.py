"""Per-beat statistics of a paced cell, from its accepted RK steps.

The integrator on the device keeps these online (``grw_get_beats``,
include/grw_cell.h); ``beat_stats`` computes the same records with the same
arithmetic from a run's accepted-step log ``(t, V, Cai, ICa)`` and its 1 ms
output samples, so that device records can be checked against host traces and
traces of any other driver (the host integrator, the CPU oracle) can be reduced
the same way.  The reference writes only 1 ms samples (data_io.c:22-140), too
coarse for APD90 or the ICaL peak, and has no such reduction.

Record fields (``FIELDS``): beat b starts at the stimulus t_s = shift + b period
(parameters.h:114-117) and holds the accepted steps ending in [t_s, t_s + period);
under a voltage-clamp protocol its periods are used (``beat_window``).
Ca_in (mM of myoplasm, -ICa Acap / (2 Vmyo F), fcn.c:1058) and Ca_rel (the
whole-cell JRyR of send_calc_fru_avg integrated) are trapezoids over the
accepted steps; Ca_rel / Ca_in is the EC-coupling gain.
"""
from __future__ import annotations

import math

import numpy as np

from . import constants as K

FIELDS = ("t_stim", "V_rest", "V_peak", "t_V_peak", "APD90", "Cai_peak", "t_Cai_peak",
          "Cai_rest", "ICa_peak", "CaNSR_rest", "CaJSR_rest", "steps", "Ca_in", "Ca_rel", "JRyR_peak")
NSTAT = len(FIELDS)
NAN = float("nan")


def beat_window(prot) -> tuple[float, float]:
    """(shift, period) of the beats of a run under ``prot``: the pacing period,
    or the period of the voltage-clamp protocol in force (the I-V steps, the
    clamp with a prepulse, the plain clamp, in that precedence, as the device
    integrator's records use)."""
    if prot.ivclamp:
        return prot.iv_shift, prot.iv_period
    if prot.vppclamp:
        return prot.vpp_shift, prot.vpp_period
    if prot.vclamp:
        return prot.vclamp_shift, prot.vclamp_period
    return prot.shift, prot.period


def ca_influx(ica):
    """fcn.c:1058: ICa (uA/uF) as Ca flux per litre of myoplasm (mM/ms)."""
    return ica * K.Acap / (-2.0 * K.Vmyo * K.Faraday * 1000.0)


def beat_stats(log, row_t, states, other, shift: float, period: float, include_current=True,
               jrel=None):
    """Beat records [n, NSTAT] from accepted steps ``log`` [m, 4] (t_end, V, Cai,
    ICa, in order) and output samples at times ``row_t`` with global states
    [.., 37] and averaged outputs [.., 12].  Rest values come from the sample at
    t_s; APD90 is the time after t_s at which V, past its peak, first falls
    through V_peak - 0.9 (V_peak - V_rest), linear between accepted steps (NaN
    if it does not within the beat or the peak is under 20 mV above rest).
    ``jrel`` (one release flux per log row) fills Ca_rel and JRyR_peak; without
    it they are NaN."""
    rows = {float(t): r for r, t in enumerate(np.asarray(row_t, np.float64))}
    out = []
    cur = -1
    bt = None
    prev_t = prev_V = prev_jin = prev_jrel = NAN
    jr = np.full(len(log), NAN) if jrel is None else np.asarray(jrel, np.float64)
    for (t, V, Cai, ICa), jrl in zip(np.asarray(log, np.float64), jr):
        t, V, Cai, ICa, jrl = float(t), float(V), float(Cai), float(ICa), float(jrl)
        jin = ca_influx(ICa)
        b = int(math.floor((t - shift) / period)) if t >= shift else -1
        seg = prev_t == prev_t
        din = 0.5 * (prev_jin + jin) * (t - prev_t) if seg else 0.0
        drel = 0.5 * (prev_jrel + jrl) * (t - prev_t) if seg else 0.0
        if b != cur:
            if cur >= 0:
                out.append(bt)
            cur = b
            if b >= 0:
                bt = [NAN] * NSTAT
                bt[0] = shift + b * period
                bt[2] = V
                bt[3] = t - bt[0]
                bt[5] = Cai
                bt[6] = t - bt[0]
                bt[8] = ICa
                bt[11] = 1.0
                bt[12] = din
                bt[13] = drel
                bt[14] = jrl
                r = rows.get(bt[0])
                if r is not None:
                    bt[1] = float(states[r][0])
                    bt[7] = float(states[r][6])
                    bt[9] = float(states[r][7])
                    bt[10] = float(other[r][1])
        elif b >= 0:
            if V > bt[2]:
                bt[2] = V
                bt[3] = t - bt[0]
            if Cai > bt[5]:
                bt[5] = Cai
                bt[6] = t - bt[0]
            if ICa < bt[8]:
                bt[8] = ICa
            bt[11] += 1.0
            bt[12] += din
            bt[13] += drel
            if jrl > bt[14]:
                bt[14] = jrl
            if bt[4] != bt[4] and bt[1] == bt[1] and bt[2] - bt[1] > 20.0 and t - bt[0] > bt[3]:
                thr = bt[2] - 0.9 * (bt[2] - bt[1])
                if prev_V >= thr and V < thr:
                    bt[4] = prev_t + (thr - prev_V) * (t - prev_t) / (V - prev_V) - bt[0]
        prev_t, prev_V, prev_jin, prev_jrel = t, V, jin, jrl
    if include_current and cur >= 0:
        out.append(bt)
    return np.array(out, np.float64).reshape(-1, NSTAT)


def as_dicts(records) -> list[dict]:
    return [dict(zip(FIELDS, map(float, r))) for r in np.asarray(records).reshape(-1, NSTAT)]
