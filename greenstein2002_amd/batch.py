"""Many independent cells on one GPU (include/grw_batch.h).

``Batch`` runs cells of different FRU counts, seeds, parameter scales and
protocols together: every RK stage of every cell is one set of launches, and
each cell's global ODE is integrated on the device with its own step sizes.
All cells start from ic/stable (FRU j takes record j mod 1250, the reference's
rewind, initialize_state.c:38-41) with each cell's total-Ca entry computed for
its own scale (initialize_state.c:233-246).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import GRW_FP64, GRW_MIXED, GrwStats, check
from .beats import NSTAT
from .cell import LAYOUT_NAMES, NGLOBAL, GrwProtocol, GrwRkStats, GrwTune, Protocol, _declare as _declare_cell
from .ic import REWIND, errweight, load_stable


class GrwBatchConfig(C.Structure):
    _fields_ = [("ncells", C.c_int32), ("precision", C.c_int32), ("device", C.c_int32),
                ("beat_cap", C.c_int32)]


class GrwCellSpec(C.Structure):
    _fields_ = [("nfru", C.c_int32), ("seed", C.c_uint32), ("nfru_scale", C.c_double),
                ("k23_scale", C.c_double), ("gamma_scale", C.c_double), ("prot", GrwProtocol),
                ("t0", C.c_double)]


@dataclass
class CellSpec:
    nfru: int
    seed: int = 1234
    nfru_scale: float | None = None   # None: 12500 / nfru
    k23_scale: float = 1.0            # RyR opening rate (fru_rates_local.c:103)
    gamma_scale: float = 1.0          # LCC mode switch (fru_rates_local.c:70)
    protocol: Protocol = field(default_factory=Protocol)
    t0: float = 0.0                   # the cell's start time (ms)

    @property
    def scale(self) -> float:
        return 12500.0 / self.nfru if self.nfru_scale is None else float(self.nfru_scale)


def _declare(L):
    if getattr(L, "_grw_batch_declared", False):
        return
    _declare_cell(L)
    vp = C.c_void_p
    L.grw_batch_create.argtypes = [C.POINTER(GrwBatchConfig), vp, C.POINTER(vp)]
    L.grw_batch_load_ic.argtypes = [vp, vp, vp, C.c_int, vp, vp, vp, vp]
    L.grw_batch_run.argtypes = [vp, vp, vp, vp]
    L.grw_batch_get_beats.argtypes = [vp, C.c_int, vp, C.c_int, C.POINTER(C.c_int), C.c_int]
    for n in ("grw_batch_create", "grw_batch_load_ic", "grw_batch_run", "grw_batch_get_beats"):
        getattr(L, n).restype = C.c_int
    L.grw_autotune_batch.argtypes = [vp, C.c_int, C.POINTER(GrwTune)]
    L.grw_autotune_batch.restype = C.c_int
    L._grw_batch_declared = True


def _vp(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Batch:
    """Cells ``specs`` on one GPU, each started from ic/stable at its t0."""

    def __init__(self, specs, precision: str = "fp64", beat_cap: int = 64, device: int = 0):
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        self.specs = list(specs)
        L = _lib.load()
        _declare(L)
        self._L = L
        n = len(self.specs)
        cfg = GrwBatchConfig(ncells=n, precision=GRW_FP64 if precision == "fp64" else GRW_MIXED,
                             device=device, beat_cap=beat_cap)
        arr = (GrwCellSpec * n)()
        for k, sp in enumerate(self.specs):
            arr[k].nfru = int(sp.nfru)
            arr[k].seed = int(sp.seed) & 0xFFFFFFFF
            arr[k].nfru_scale = sp.scale
            arr[k].k23_scale = float(sp.k23_scale)
            arr[k].gamma_scale = float(sp.gamma_scale)
            arr[k].prot = sp.protocol.to_c()
            arr[k].t0 = float(sp.t0)
        h = C.c_void_p()
        check(L.grw_batch_create(C.byref(cfg), C.cast(arr, C.c_void_p), C.byref(h)))
        self._h = h
        self.ncells = n
        self.beat_cap = beat_cap
        self.t = np.array([float(sp.t0) for sp in self.specs])
        self.attempts = np.zeros(n, np.int64)
        self.rejects = np.zeros(n, np.int64)
        base = load_stable(REWIND)
        by_shape = {}  # the global state depends on the FRU count and scale only
        for sp in self.specs:
            if (sp.nfru, sp.scale) not in by_shape:
                by_shape[(sp.nfru, sp.scale)] = load_stable(sp.nfru, sp.scale)["state"]
        states = np.array([by_shape[(sp.nfru, sp.scale)] for sp in self.specs], np.float64)
        self.errw = errweight()
        fru = np.ascontiguousarray(base["fru"], np.float64)
        lt = np.ascontiguousarray(base["ltype"], np.int32)
        ryr = np.ascontiguousarray(base["ryr"], np.int32)
        ito2 = np.ascontiguousarray(base["ito2"], np.int32)
        check(L.grw_batch_load_ic(h, _vp(states), _vp(self.errw), REWIND, _vp(fru), _vp(lt), _vp(ryr),
                                  _vp(ito2)))
        self.state = states

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._L.grw_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):  # pragma: no cover - best effort
        try:
            self.close()
        except Exception:
            pass

    def run(self, nsteps) -> np.ndarray:
        """Advance cell k by nsteps[k] output steps (an int applies to all);
        returns the global states [ncells, 37]."""
        ns = np.broadcast_to(np.asarray(nsteps, np.int32), (self.ncells,)).copy()
        st = (GrwRkStats * self.ncells)()
        out = np.zeros((self.ncells, NGLOBAL))
        check(self._L.grw_batch_run(self._h, _vp(ns), _vp(out), C.cast(st, C.c_void_p)))
        self.attempts += np.array([s.attempts for s in st], np.int64)
        self.rejects += np.array([s.rejects for s in st], np.int64)
        for k, sp in enumerate(self.specs):  # one addition of time_step per output step (caf.c:150)
            for _ in range(int(ns[k])):
                self.t[k] += sp.protocol.time_step
        self.state = out
        return out

    def beats(self, cell: int, include_current: bool = True) -> np.ndarray:
        cap = self.beat_cap + 1
        out = np.zeros((cap, NSTAT))
        n = C.c_int(0)
        check(self._L.grw_batch_get_beats(self._h, int(cell), _vp(out), cap, C.byref(n), int(include_current)))
        return out[: n.value].copy()

    def stats(self) -> dict:
        st = GrwStats()
        check(self._L.grw_get_stats(self._h, C.byref(st)))
        return dict(fru_steps=st.fru_steps, simfru_calls=st.simfru_calls, launches=st.launches,
                    kernel_ms=st.kernel_ms, critical_steps=st.critical_steps, tail_ms=st.tail_ms,
                    timed_launches=st.timed_launches, slot_steps=st.slot_steps)

    def reset_stats(self) -> None:
        check(self._L.grw_reset_stats(self._h))

    def set_timing(self, on: bool) -> None:
        check(self._L.grw_set_timing(self._h, int(bool(on))))

    def set_layout(self, layout: str) -> None:
        """Monte Carlo kernel of the batch ("quad", "quad3", "quad4",
        "queue", "wave1", "wave1_3", "wave1_4", "refill", "refill_3"; see
        Cell.set_layout)."""
        codes = {v: k for k, v in LAYOUT_NAMES.items()}
        if layout not in codes:
            raise ValueError("layout must be one of %s" % sorted(codes))
        check(self._L.grw_set_layout(self._h, codes[layout]))

    def autotune(self, trial_steps: int = 2) -> dict:
        """Time trial_steps output steps of every cell under each layout, with
        and without graph replays, from a snapshot of the batch, and keep the
        fastest (grw_autotune_batch).  The batch's state is as before."""
        r = GrwTune()
        check(self._L.grw_autotune_batch(self._h, int(trial_steps), C.byref(r)))
        return r.as_dict()

    def set_graphs(self, on: bool) -> None:
        """Replay seven of every eight RK attempts of the device integrator
        from one captured HIP graph (identical results)."""
        check(self._L.grw_set_graphs(self._h, int(bool(on))))
