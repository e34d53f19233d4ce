"""Dynamic FRU number: a cell simulated with fewer release units while nothing
happens in them (dynamicFRU.c:26-152, called by caf.c:246-250 after every
output step when NFRU_sim_high != NFRU_sim_low, parameters.h:43-46).

The phase decision is the reference's (``phase_decision``, pinned to its
compiled dynamicFRU.c in tests/test_dynamic.py).  The redistribution of FRU
state at a switch is defined here, because the reference's does not define
one: switching up it replicates the low FRUs into all but the last block of the
arrays and then hands the MPI layer only the low count while setting NFRU_sim
to the high one (dynamicFRU.c:110-131), and switching down it averages the
master's copy of the FRU arrays, not gathered from the slaves, over
NFRU_sim_factor - 1 blocks and divides by NFRU_sim_factor (dynamicFRU.c:
136-144).  Here, with factor = nfru_high / nfru_low:

* high -> low: FRU i < nfru_low keeps its channel states and MT19937 stream,
  and its concentrations become the mean over its replicas i + j nfru_low,
  j = 0 .. factor-1, so that the whole-cell Ca of the FRUs (scaled by
  NFRU_scale) is unchanged; FRUs from nfru_low on stay as they were;
* low -> high: FRUs i < nfru_low come back with everything they carry, and
  every replica i >= nfru_low takes the concentrations and channel states of
  FRU i mod nfru_low and keeps its own stream (seeded for its index at the
  start, advanced only while it is simulated).

A ``DynamicCell`` holds one device context per FRU count and moves the FRU
state between them through host memory at a switch (a few per beat); the
global state, clock and step size carry over.  It has the interface of
``Cell`` that ``output.run_caf`` uses, which calls ``dynamic_step`` after each
output step as caf.c does.
"""
from __future__ import annotations

import math

import numpy as np

from .cell import Cell, Protocol

LOW, HIGH = 0, 1


def phase_decision(t: float, V: float, pryr_open: float, plcc_open: float, phase: int, t_last: float,
                   prot: Protocol | None = None):
    """dynamicFRU.c:49-105: returns (phase, time of the last change, whether
    it changed, time to the next stimulus or None when not evaluated).  t is
    the end of the output step, V its membrane potential, pryr_open and
    plcc_open the open fractions of send_calc_fru_avg (outputs 3 and 7)."""
    p = prot or Protocol()
    change = False
    tts = None
    if t - t_last > 1.0:
        tts = p.shift - math.fmod(t, p.period)
        if tts < 0.0:
            tts = tts + p.period
        tfs = math.fmod(t, p.period) - p.shift
        if tfs < 0.0:
            tfs = tfs + p.period
        if phase == LOW:
            if 0.0 <= tts < 100.0:  # a stimulus within 100 ms
                change = True
            if plcc_open > 2e-6 or pryr_open > 6.e-3:
                change = True
        else:
            if V < -80 and plcc_open < 1.e-6 and pryr_open < 2.e-3:
                change = True
            if t - t_last < 1.0:
                change = False
            if 0.0 <= tts < 100.0:
                change = False
            if 0.0 <= tfs < 50.0:  # within 50 ms after a stimulus
                change = False
    if change:
        return (HIGH if phase == LOW else LOW), t, True, tts
    return phase, t_last, False, tts


def average_down(conc_high: np.ndarray, nfru_low: int) -> np.ndarray:
    """Concentrations of the low FRUs: FRU i takes the mean over its replicas
    i + j nfru_low (j = 0 .. factor-1)."""
    f = len(conc_high) // nfru_low
    return conc_high.reshape(f, nfru_low, -1).mean(axis=0)


def replicate_up(low: dict, high: dict, nfru_low: int) -> dict:
    """FRU state of the high phase: FRUs below nfru_low from the low phase in
    full, the replicas their concentrations and channel states with their own
    streams (from ``high``, the high context as it was left)."""
    out = {k: np.array(v, copy=True) for k, v in high.items()}
    n_high = len(out["fru"])
    src = np.arange(n_high) % nfru_low
    for key in ("fru", "ltype", "ryr", "ito2"):
        out[key] = np.asarray(low[key])[src].copy()
    out["mti"][:nfru_low] = low["mti"]
    out["mt"][:nfru_low] = low["mt"]
    return out


class DynamicCell:
    """A cell switching between nfru_low and nfru_high FRUs (whole-cell scales
    12500 / nfru each, parameters.h:43-46), starting in the high phase as
    caf.c:86-88 does."""

    def __init__(self, nfru_low: int, nfru_high: int, seed: int = 1234, precision: str = "fp64",
                 protocol: Protocol | None = None):
        if nfru_low <= 0 or nfru_high % nfru_low != 0:
            raise ValueError("nfru_high must be a multiple of nfru_low")
        self.nfru_low, self.nfru_high = nfru_low, nfru_high
        self.protocol = protocol or Protocol()
        self.cells = {LOW: Cell(nfru_low, seed=seed, precision=precision, protocol=self.protocol),
                      HIGH: Cell(nfru_high, seed=seed, precision=precision, protocol=self.protocol)}
        self.phase = HIGH
        self.t_last = 0.0  # time_of_last_change (caf.c:27), zero at the start
        self.switches: list[tuple[float, int]] = []

    # the interface of Cell that run_caf uses
    @property
    def cell(self) -> Cell:
        return self.cells[self.phase]

    @property
    def fru(self):
        return self.cell.fru

    @property
    def state(self):
        return self.cell.state

    @property
    def t(self):
        return self.cell.t

    def run(self, nsteps: int, **kw):
        return self.cell.run(nsteps, **kw)

    def restart(self, t=None):
        # an I-V restart reloads the initial condition in the current phase's context
        self.cell.restart(t)

    def close(self):
        for c in self.cells.values():
            c.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def dynamic_step(self, tstep: float, state, other) -> bool:
        """caf.c:246-250 after the output step ending at tstep; returns whether
        the phase switched."""
        ph, self.t_last, changed, _ = phase_decision(tstep, float(state[0]), float(other[3]),
                                                     float(other[7]), self.phase, self.t_last, self.protocol)
        if changed:
            self._switch(ph)
            self.switches.append((tstep, ph))
        return changed

    def _switch(self, to: int) -> None:
        src, dst = self.cells[self.phase], self.cells[to]
        g = src.fru.parallel_get_FRUs()
        if to == LOW:
            new = {k: np.array(g[k][: self.nfru_low], copy=True) for k in g}
            new["fru"] = average_down(np.asarray(g["fru"]), self.nfru_low)
        else:
            new = replicate_up(g, dst.fru.parallel_get_FRUs(), self.nfru_low)
        dst.fru.initialize_mpi_state(new["fru"], new["ltype"], new["ryr"], new["ito2"], new["mti"], new["mt"])
        dst.state = np.array(src.state, np.float64)
        dst.t = src.t
        dst.h = src.h
        self.phase = to
