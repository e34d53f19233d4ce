"""Host-side mirror of the reference's master-side FRU interface.

``FruCell`` owns the device-resident release units of one cell and exposes
the calls of ``mpi_master.c`` (``initialize_mpi_state``, ``distrib_simFRU``,
``send_calc_fru_flux``, ``send_calc_fru_avg``, ``send_save_state``,
``send_resume_state``, ``parallel_get_FRUs``) and the per-FRU routines they
distribute (``simfru_local``, ``calc_fru_flux_local``, ``calc_fru_avg_local``)
with the same argument meaning.  Each call goes straight to the HIP library
through its C ABI (``include/grw_fru.h``); nothing here computes on the CPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import GRW_FP64, GRW_MIXED, GrwConfig, GrwStats, check

MTN = 624


def _vp(a: np.ndarray) -> C.c_void_p:
    return a.ctypes.data_as(C.c_void_p)


def _dep(x) -> np.ndarray:
    a = np.ascontiguousarray(x, dtype=np.float64)
    if a.shape != (3,):
        raise ValueError("FRU-dependent global state must be (V, Cai, CaNSR)")
    return a


class FruCell:
    """NFRU release units of one cell on one GPU.

    precision: "fp64" (drop-in) or "mixed" (FP32 rates / local-Ca RHS).
    sequential_sums: reduce in the reference's sequential FRU order (parity runs).
    """

    def __init__(self, nfru: int, nfru_scale: float | None = None, precision: str = "fp64",
                 block: int = 0, sequential_sums: bool = False, k23_scale: float = 1.0,
                 gamma_scale: float = 1.0, device: int = 0, lanes_per_fru: int = 4):
        L = _lib.load()
        cfg = GrwConfig()
        L.grw_config_default(C.byref(cfg))
        cfg.nfru = int(nfru)
        cfg.nfru_scale = float(12500.0 / nfru if nfru_scale is None else nfru_scale)
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        cfg.precision = GRW_FP64 if precision == "fp64" else GRW_MIXED
        cfg.block = int(block)
        cfg.sequential_sums = int(bool(sequential_sums))
        cfg.k23_scale = float(k23_scale)
        cfg.gamma_scale = float(gamma_scale)
        cfg.device = int(device)
        cfg.lanes_per_fru = int(lanes_per_fru)
        h = C.c_void_p()
        check(L.grw_create(C.byref(cfg), C.byref(h)))
        self._L = L
        self._h = h
        self.nfru = int(nfru)
        self.nfru_scale = cfg.nfru_scale
        self.precision = precision

    # ---- lifetime (initialize_mpi / end_mpi)
    def close(self) -> None:
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

    # ---- state transfer (mpi_master.c:199-361, :509-617)
    def initialize_mpi_state(self, fru, ltype, ryr, ito2, mti, mt) -> None:
        n = self.nfru
        fru = np.ascontiguousarray(fru, np.float64).reshape(n, 5)
        ltype = np.ascontiguousarray(ltype, np.int32).reshape(n, 4, 2)
        ryr = np.ascontiguousarray(ryr, np.int32).reshape(n, 4, 5)
        ito2 = np.ascontiguousarray(ito2, np.int32).reshape(n, 4)
        mti = np.ascontiguousarray(mti, np.int32).reshape(n)
        mt = np.ascontiguousarray(np.asarray(mt)[:, :MTN], np.uint32).reshape(n, MTN)
        check(self._L.grw_load_state(self._h, _vp(fru), _vp(ltype), _vp(ryr), _vp(ito2), _vp(mti),
                                     _vp(mt)))

    load_state = initialize_mpi_state

    def parallel_get_FRUs(self) -> dict:
        n = self.nfru
        out = dict(
            fru=np.zeros((n, 5)),
            ltype=np.zeros((n, 4, 2), np.int32),
            ryr=np.zeros((n, 4, 5), np.int32),
            ito2=np.zeros((n, 4), np.int32),
            mti=np.zeros(n, np.int32),
            mt=np.zeros((n, MTN), np.uint32),
        )
        check(self._L.grw_get_state(self._h, _vp(out["fru"]), _vp(out["ltype"]), _vp(out["ryr"]),
                                    _vp(out["ito2"]), _vp(out["mti"]), _vp(out["mt"])))
        return out

    get_state = parallel_get_FRUs

    def seed(self, a0: int) -> None:
        """initialize_ran (initialize_ran.c:23-36) with master seed a0."""
        check(self._L.grw_seed(self._h, int(a0) & 0xFFFFFFFF))

    # ---- the Monte Carlo (mpi_master.c:369-469)
    def distrib_simFRU(self, t0: float, t1: float, dep0, depf) -> np.ndarray:
        out = np.zeros(4)
        check(self._L.grw_distrib_simfru(self._h, float(t0), float(t1), _vp(_dep(dep0)),
                                         _vp(_dep(depf)), _vp(out)))
        return out

    def simfru_local(self, t0: float, t1: float, dep0, depf) -> None:
        check(self._L.grw_simfru_local(self._h, float(t0), float(t1), _vp(_dep(dep0)),
                                       _vp(_dep(depf))))

    # ---- reductions (mpi_master.c:41-191)
    def send_calc_fru_flux(self, dep) -> np.ndarray:
        out = np.zeros(4)
        check(self._L.grw_calc_flux(self._h, _vp(_dep(dep)), _vp(out)))
        return out

    def send_calc_fru_avg(self, state) -> np.ndarray:
        st = np.ascontiguousarray(state, np.float64)
        if st.shape != (37,):
            raise ValueError("global state must have 37 entries")
        out = np.zeros(12)
        check(self._L.grw_calc_avg(self._h, _vp(st), _vp(out)))
        return out

    def calc_fru_flux_local(self) -> np.ndarray:
        out = np.zeros(5)
        check(self._L.grw_calc_fru_flux_local(self._h, _vp(out)))
        return out

    def calc_fru_avg_local(self) -> np.ndarray:
        out = np.zeros(11)
        check(self._L.grw_calc_fru_avg_local(self._h, _vp(out)))
        return out

    # ---- rollback (mpi_master.c:477-501)
    def send_save_state(self) -> None:
        check(self._L.grw_save_state(self._h))

    def send_resume_state(self) -> None:
        check(self._L.grw_resume_state(self._h))

    # ---- counters
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

    def set_graphs(self, on: bool) -> None:
        """Replay seven of every eight RK attempts of the device integrator
        from one captured HIP graph (identical results)."""
        check(self._L.grw_set_graphs(self._h, int(bool(on))))
