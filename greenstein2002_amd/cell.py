"""Whole-cell runs: the global model integrated around the device-resident FRUs.

``Protocol`` is the runtime form of the reference's compile-time protocol
switches (parameters.h:92-163).  ``Cell`` couples a ``FruCell`` with the
37-state global state and runs the caf.c output loop (caf.c:150-251) through
``grw_run_cell`` / ``grw_rk54pd`` (include/grw_cell.h), which follow
rk_pd54m.c:29-345 on the host and call the FRU library for every RK stage.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import check
from .fru import FruCell
from .ic import errweight, load_stable

NGLOBAL = 37
NCUR = 24
NOTHER = 12
LAYOUT_QUAD, LAYOUT_WAVES = 1, 2  # grw_set_layout


class GrwProtocol(C.Structure):
    _fields_ = [
        ("pulse_duration", C.c_double),
        ("pulse_amplitude", C.c_double),
        ("period", C.c_double),
        ("shift", C.c_double),
        ("vclamp", C.c_int32),
        ("chf", C.c_int32),
        ("vclamp_duration", C.c_double),
        ("vclamp_set", C.c_double),
        ("vclamp_shift", C.c_double),
        ("vclamp_hold", C.c_double),
        ("vclamp_period", C.c_double),
        ("serca_scale", C.c_double),
        ("time_step", C.c_double),
        ("step_min", C.c_double),
        ("step_max", C.c_double),
        ("tolrk", C.c_double),
        ("vppclamp", C.c_int32),
        ("ivclamp", C.c_int32),
        ("vpp_ppset", C.c_double),
        ("vpp_ppduration", C.c_double),
        ("vpp_shift", C.c_double),
        ("vpp_set", C.c_double),
        ("vpp_duration", C.c_double),
        ("vpp_hold", C.c_double),
        ("vpp_period", C.c_double),
        ("iv_duration", C.c_double),
        ("iv_period", C.c_double),
        ("iv_shift", C.c_double),
        ("iv_set", C.c_double),
        ("iv_hold", C.c_double),
        ("iv_step", C.c_double),
        ("iv_n", C.c_int32),
        ("reserved", C.c_int32),
    ]

_INT_FIELDS = ("vclamp", "chf", "vppclamp", "ivclamp", "iv_n", "reserved")


class GrwRkStats(C.Structure):
    _fields_ = [("attempts", C.c_int64), ("rejects", C.c_int64), ("forced", C.c_int64)]


TUNE_MAX = 16  # GRW_TUNE_MAX
LAYOUT_NAMES = {1: "quad", 2: "waves", 3: "quad3", 4: "quad4", 5: "queue", 6: "wave1", 7: "wave1_3", 8: "quadp", 9: "wave1_4",
                10: "refill", 11: "refill_3"}


class GrwTune(C.Structure):
    _fields_ = [("layout", C.c_int32), ("graphs", C.c_int32), ("ncand", C.c_int32), ("reserved", C.c_int32),
                ("cand_layout", C.c_int32 * TUNE_MAX), ("cand_graphs", C.c_int32 * TUNE_MAX),
                ("cand_ms", C.c_double * TUNE_MAX)]

    def as_dict(self) -> dict:
        return {"layout": LAYOUT_NAMES[self.layout], "graphs": bool(self.graphs),
                "candidates": [{"layout": LAYOUT_NAMES[self.cand_layout[k]], "graphs": bool(self.cand_graphs[k]),
                                "ms": round(self.cand_ms[k], 3)} for k in range(self.ncand)]}


def _declare(L):
    if getattr(L, "_grw_cell_declared", False):
        return
    vp = C.c_void_p
    L.grw_protocol_default.argtypes = [C.POINTER(GrwProtocol)]
    L.grw_protocol_default.restype = None
    L.grw_fcn.argtypes = [C.POINTER(GrwProtocol), C.c_double, vp, vp, vp, C.c_int, vp]
    L.grw_fcn.restype = C.c_int
    L.grw_rk54pd.argtypes = [vp, C.POINTER(GrwProtocol), vp, C.c_double, C.c_double, C.c_double, vp,
                             vp, C.POINTER(C.c_double), vp, C.c_int, C.POINTER(C.c_int),
                             C.POINTER(GrwRkStats)]
    L.grw_rk54pd.restype = C.c_int
    L.grw_run_cell.argtypes = [vp, C.POINTER(GrwProtocol), vp, C.c_double, C.c_int, vp,
                               C.POINTER(C.c_double), vp, vp, vp, vp, C.c_int, C.POINTER(C.c_int),
                               C.POINTER(GrwRkStats)]
    L.grw_run_cell.restype = C.c_int
    L.grw_run_cell_device.argtypes = L.grw_run_cell.argtypes
    L.grw_run_cell_device.restype = C.c_int
    L.grw_get_beats.argtypes = [vp, vp, C.c_int, C.POINTER(C.c_int), C.c_int]
    L.grw_get_beats.restype = C.c_int
    L.grw_set_layout.argtypes = [vp, C.c_int]
    L.grw_set_layout.restype = C.c_int
    L.grw_autotune_cell.argtypes = [vp, C.POINTER(GrwProtocol), vp, C.c_double, vp, C.c_double, C.c_int,
                                    C.POINTER(GrwTune)]
    L.grw_autotune_cell.restype = C.c_int
    L._grw_cell_declared = True


def _vp(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


@dataclass
class Protocol:
    """parameters.h:92-163 as runtime values (defaults = the reference's)."""

    pulse_duration: float = 1.0
    pulse_amplitude: float = -100.0
    period: float = 1000.0
    shift: float = 10.0
    vclamp: bool = False
    chf: bool = False
    vclamp_duration: float = 200.0
    vclamp_set: float = 0.0
    vclamp_shift: float = 300.0
    vclamp_hold: float = -80.0
    vclamp_period: float = 500.0
    serca_scale: float = 1.0
    time_step: float = 1.0
    step_min: float = 1.e-6
    step_max: float = 0.1
    tolrk: float = 1.e-6
    # voltage clamp with a prepulse, parameters.h:128-137 (fcn.c:485-516)
    vppclamp: bool = False
    vpp_ppset: float = -40.0
    vpp_ppduration: float = 100.0
    vpp_shift: float = 10000.0
    vpp_set: float = 0.0
    vpp_duration: float = 300.0
    vpp_hold: float = -80.0
    vpp_period: float = 500.0
    # voltage clamp for I-V studies, parameters.h:141-147 (fcn.c:521-535); the
    # test potential is iv_set + iv_n * iv_step, iv_n the step number (caf.c:234)
    ivclamp: bool = False
    iv_duration: float = 50.0
    iv_period: float = 200.0
    iv_shift: float = 10.0
    iv_set: float = -40.0
    iv_hold: float = -100.0
    iv_step: float = 10.0
    iv_n: int = 0
    reserved: int = 0

    @property
    def clamped(self) -> bool:
        return bool(self.vclamp or self.vppclamp or self.ivclamp)

    def initial_V(self, V: float) -> float:
        """initialize_state.c:222-228: V starts at the holding potential under
        the voltage clamp and under the I-V clamp."""
        if self.vclamp:
            V = self.vclamp_hold
        if self.ivclamp:
            V = self.iv_hold
        return V

    def to_c(self) -> GrwProtocol:
        p = GrwProtocol()
        for name, _ in GrwProtocol._fields_:
            v = getattr(self, name)
            setattr(p, name, int(v) if name in _INT_FIELDS else float(v))
        return p


def fcn(t: float, state, J, prot: Protocol | None = None, keepc: bool = True):
    """fcn.c:29-1072 (host).  Returns (F, currents, state after the call)."""
    L = _lib.load()
    _declare(L)
    p = (prot or Protocol()).to_c()
    st = np.array(state, np.float64)
    F = np.zeros(NGLOBAL)
    cur = np.zeros(NCUR)
    J = np.ascontiguousarray(J, np.float64)
    check(L.grw_fcn(C.byref(p), float(t), _vp(st), _vp(F), _vp(cur), int(keepc), _vp(J)))
    return F, cur, st


@dataclass
class RunResult:
    t: np.ndarray                 # output times (ms), one per time_step
    states: np.ndarray            # [n, 37]
    currents: np.ndarray          # [n, 24]
    other: np.ndarray             # [n, 12]
    log: np.ndarray               # [m, 4] accepted steps: t, V, Cai, ICa
    stats: dict = field(default_factory=dict)


class Cell:
    """One myocyte: FruCell + global state + protocol."""

    def __init__(self, nfru: int, seed: int = 1234, precision: str = "fp64",
                 protocol: Protocol | None = None, ic: dict | None = None, block: int = 0,
                 sequential_sums: bool = False, k23_scale: float = 1.0, gamma_scale: float = 1.0,
                 nfru_scale: float | None = None, lanes_per_fru: int = 4):
        self.nfru = nfru
        self.protocol = protocol or Protocol()
        scale = 12500.0 / nfru if nfru_scale is None else nfru_scale
        self.fru = FruCell(nfru, nfru_scale=scale, precision=precision, block=block,
                           sequential_sums=sequential_sums, k23_scale=k23_scale,
                           gamma_scale=gamma_scale, lanes_per_fru=lanes_per_fru)
        self.ic = ic or load_stable(nfru, scale)
        self.seed = seed
        self.errw = errweight()
        self._L = _lib.load()
        _declare(self._L)
        self.layout = "quad"
        self.restart(0.0)

    def restart(self, t: float | None = None) -> None:
        """Back to the initial condition with the same MT19937 seeds, keeping
        the clock (or setting it to t): initialize_ran, initialize_state and
        initialize_mpi_state, as caf.c:232-244 does at every I-V step, with the
        next step size step_max (caf.c:236)."""
        ic = self.ic
        self.state = np.array(ic["state"], np.float64)
        self.state[0] = self.protocol.initial_V(self.state[0])  # initialize_state.c:222-228
        self.fru.seed(self.seed)
        st = self.fru.parallel_get_FRUs()
        self.fru.initialize_mpi_state(ic["fru"], ic["ltype"], ic["ryr"], ic["ito2"], st["mti"], st["mt"])
        if t is not None:
            self.t = float(t)
        self.h = self.protocol.step_max

    def close(self):
        self.fru.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_layout(self, layout: str) -> None:
        """The Monte Carlo kernel of the device integrator (grw_set_layout):
        "quad", four lanes per FRU (the default); "waves", the
        wave-specialised kernel (four wavefronts per 16 FRUs); "quad3" /
        "quad4", the four-lane kernel with registers capped for three / four
        wavefronts per SIMD; "queue", the four-lane kernel's work as a queue
        of 16-FRU items per wavefront, then a per-cell stage kernel (switched
        off); "wave1", the four-lane kernel's work with one wavefront per
        block, then a per-cell stage kernel; "wave1_3" / "wave1_4", the same
        with registers capped for three / four wavefronts per SIMD; "quadp", the
        four-lane kernel whose fcn block collects the partials by polling
        (single cells); "refill", one wavefront per 64 FRUs whose quads take
        the next FRU as soon as theirs is done, then a per-cell stage kernel;
        "refill_3", the same with registers capped for three wavefronts per
        SIMD.  All give the same results bit for bit."""
        codes = {v: k for k, v in LAYOUT_NAMES.items()}
        if layout not in codes:
            raise ValueError("layout must be one of %s" % sorted(codes))
        check(self._L.grw_set_layout(self.fru._h, codes[layout]))
        self.layout = layout

    def autotune(self, trial_steps: int = 10) -> dict:
        """Time trial_steps output steps of the device integrator from the
        current state under every layout, with and without graph replays, and
        keep the fastest (grw_autotune_cell).  The cell's state is as before."""
        p = self.protocol.to_c()
        r = GrwTune()
        check(self._L.grw_autotune_cell(self.fru._h, C.byref(p), _vp(self.errw), float(self.t), _vp(self.state),
                                        float(self.h), int(trial_steps), C.byref(r)))
        self.layout = LAYOUT_NAMES[r.layout]
        return r.as_dict()

    def beats(self, include_current: bool = True) -> np.ndarray:
        """Per-beat statistics of the device-driven runs so far (grw_get_beats;
        fields in ``greenstein2002_amd.beats.FIELDS``)."""
        from .beats import NSTAT

        cap = 4097
        out = np.zeros((cap, NSTAT))
        n = C.c_int(0)
        check(self._L.grw_get_beats(self.fru._h, _vp(out), cap, C.byref(n), int(include_current)))
        return out[: n.value].copy()

    def run(self, nsteps: int, log_cap: int = 200000, outputs: bool = True,
            driver: str = "device") -> RunResult:
        """nsteps calls of rk54pd of length time_step (caf.c:150-251).

        driver="device": the integrator runs on the GPU (grw_run_cell_device);
        driver="host": the host integrator calls the FRU ABI per sub-interval
        (grw_run_cell), exactly the reference's call sequence."""
        if driver not in ("device", "host"):
            raise ValueError("driver must be 'device' or 'host'")
        fn = self._L.grw_run_cell_device if driver == "device" else self._L.grw_run_cell
        p = self.protocol.to_c()
        states = np.zeros((nsteps, NGLOBAL)) if outputs else None
        cur = np.zeros((nsteps, NCUR)) if outputs else None
        other = np.zeros((nsteps, NOTHER)) if outputs else None
        log = np.zeros((max(log_cap, 1), 4))
        nlog = C.c_int(0)
        h = C.c_double(self.h)
        st = GrwRkStats()
        check(fn(self.fru._h, C.byref(p), _vp(self.errw), float(self.t), int(nsteps),
                                   _vp(self.state), C.byref(h), _vp(states), _vp(cur), _vp(other),
                                   _vp(log) if log_cap else None, int(log_cap), C.byref(nlog),
                                   C.byref(st)))
        # the clock advances by adding time_step once per output step, as
        # caf.c:150 and the integrators do, so a run split into pieces restarts
        # each piece at the same time as an unbroken run reaches it
        t_out = np.zeros(nsteps)
        t = float(self.t)
        for k in range(nsteps):
            t += self.protocol.time_step
            t_out[k] = t
        self.t = t
        self.h = h.value
        return RunResult(t=t_out, states=states, currents=cur, other=other, log=log[: nlog.value],
                         stats=dict(attempts=st.attempts, rejects=st.rejects, forced=st.forced))
