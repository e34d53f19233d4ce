"""ctypes bindings for the CPU checker libraries (test infrastructure only).

``libgrw_oracle.so`` is the clean-room restatement in ``oracle/grw_oracle.c``;
``_ref/libgrw_ref.so`` is the reference's own hot-path sources compiled in place
(``oracle/Makefile``).  Only tests, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this module, always as the checker.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "libgrw_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libgrw_ref.so")

MTN = 624
NGLOBAL = 37
NCUR = 24
NOTHER = 12

# struct orc_fru (grw_oracle.h)
FRU_DTYPE = np.dtype(
    [
        ("conc", "<f8", (5,)),
        ("lcc", "<i4", (4,)),
        ("yinact", "<i4", (4,)),
        ("ryr", "<i4", (4, 5)),
        ("ito2", "<i4", (4,)),
        ("mt", "<u4", (MTN,)),
        ("mti", "<i4"),
    ],
    align=True,
)

# struct orc_rates
RATES_DTYPE = np.dtype(
    [
        ("lcc_rate", "<f8", (4, 4)),
        ("lcc_dest", "<i4", (4, 4)),
        ("lcc_len", "<i4", (4,)),
        ("vinact_rate", "<f8", (4,)),
        ("ryr_rate", "<f8", (4, 5, 4)),
        ("ryr_dest", "<i4", (4, 5, 4)),
        ("ryr_len", "<i4", (4, 5)),
        ("ito2_rate", "<f8", (4,)),
        ("max_rate", "<f8"),
    ],
    align=True,
)

# struct orc_step
STEP_DTYPE = np.dtype(
    [
        ("t", "<f8"),
        ("dt", "<f8"),
        ("conc", "<f8", (5,)),
        ("lcc", "<i4", (4,)),
        ("yinact", "<i4", (4,)),
        ("ryr", "<i4", (4, 5)),
        ("ito2", "<i4", (4,)),
        ("mti", "<i4"),
        ("pad", "<i4"),
    ],
    align=True,
)

# struct orc_protocol
PROT_DTYPE = np.dtype(
    [
        ("pulse_duration", "<f8"),
        ("pulse_amplitude", "<f8"),
        ("period", "<f8"),
        ("shift", "<f8"),
        ("vclamp", "<i4"),
        ("chf", "<i4"),
        ("vclamp_duration", "<f8"),
        ("vclamp_set", "<f8"),
        ("vclamp_shift", "<f8"),
        ("vclamp_hold", "<f8"),
        ("vclamp_period", "<f8"),
        ("serca_scale", "<f8"),
        ("k23_scale", "<f8"),
        ("gamma_scale", "<f8"),
        ("vppclamp", "<i4"),
        ("ivclamp", "<i4"),
        ("vpp_ppset", "<f8"),
        ("vpp_ppduration", "<f8"),
        ("vpp_shift", "<f8"),
        ("vpp_set", "<f8"),
        ("vpp_duration", "<f8"),
        ("vpp_hold", "<f8"),
        ("vpp_period", "<f8"),
        ("iv_duration", "<f8"),
        ("iv_period", "<f8"),
        ("iv_shift", "<f8"),
        ("iv_set", "<f8"),
        ("iv_hold", "<f8"),
        ("iv_step", "<f8"),
        ("iv_n", "<i4"),
        ("pad", "<i4"),
    ],
    align=True,
)


# struct orc_event / ref_event: one call the RK driver makes into the FRU layer
EV_FLUX, EV_SAVE, EV_DISTRIB, EV_RESUME = 0, 1, 2, 3
EVENT_DTYPE = np.dtype(
    [
        ("kind", "<i4"),
        ("pad", "<i4"),
        ("t0", "<f8"),
        ("t1", "<f8"),
        ("dep", "<f8", (3,)),
        ("out", "<f8", (4,)),
    ],
    align=True,
)


class OrcCell(C.Structure):
    _fields_ = [
        ("nfru", C.c_int),
        ("scale", C.c_double),
        ("fru", C.c_void_p),
        ("hold", C.c_void_p),
        ("prot", C.c_byte * PROT_DTYPE.itemsize),
        ("errw", C.c_double * NGLOBAL),
        ("n_attempts", C.c_int64),
        ("n_rejects", C.c_int64),
        ("n_forced", C.c_int64),
        ("n_fru_steps", C.c_int64),
        ("ev", C.c_void_p),
        ("ev_cap", C.c_int32),
        ("ev_n", C.c_int32),
    ]


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


_DP = C.c_void_p


class Oracle:
    """The restatement (grw_oracle.c)."""

    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        L = C.CDLL(path)
        self.L = L
        L.orc_mt_seed.argtypes = [_DP, C.c_uint32]
        L.orc_mt_next_u32.argtypes = [_DP]
        L.orc_mt_next_u32.restype = C.c_uint32
        L.orc_mt_uniform.argtypes = [_DP]
        L.orc_mt_uniform.restype = C.c_double
        L.orc_mt_block.argtypes = [_DP, C.c_int, _DP]
        L.orc_seed_for_fru.argtypes = [C.c_uint32, C.c_int64]
        L.orc_seed_for_fru.restype = C.c_uint32
        L.orc_fru_rates.argtypes = [_DP, _DP, _DP, _DP]
        L.orc_fru_rates.restype = C.c_double
        L.orc_fcn_fru.argtypes = [_DP, _DP, _DP, _DP, _DP]
        L.orc_simfru.argtypes = [C.c_double, C.c_double, _DP, _DP, _DP, _DP, _DP, C.c_int]
        L.orc_simfru.restype = C.c_int
        L.orc_flux_sums.argtypes = [C.c_int, _DP, _DP]
        L.orc_avg_sums.argtypes = [C.c_int, _DP, _DP]
        L.orc_flux_currents.argtypes = [_DP, _DP, C.c_int, C.c_double, _DP]
        L.orc_avg_outputs.argtypes = [_DP, _DP, C.c_int, C.c_double, _DP]
        L.orc_protocol_default.argtypes = [_DP]
        L.orc_fcn.argtypes = [C.c_double, _DP, _DP, _DP, C.c_int] + [C.c_double] * 4 + [_DP]
        L.orc_errweight.argtypes = [_DP]
        L.orc_total_ca.argtypes = [_DP, C.c_int, _DP, C.c_double]
        L.orc_total_ca.restype = C.c_double
        L.orc_cell_init.argtypes = [C.POINTER(OrcCell), C.c_int, C.c_double, _DP]
        L.orc_cell_init.restype = C.c_int
        L.orc_cell_free.argtypes = [C.POINTER(OrcCell)]
        L.orc_distrib_simfru.argtypes = [C.POINTER(OrcCell), C.c_double, C.c_double, _DP, _DP, _DP]
        L.orc_send_calc_fru_flux.argtypes = [C.POINTER(OrcCell), _DP, _DP]
        L.orc_send_calc_fru_avg.argtypes = [C.POINTER(OrcCell), _DP, _DP]
        L.orc_save_state.argtypes = [C.POINTER(OrcCell)]
        L.orc_resume_state.argtypes = [C.POINTER(OrcCell)]
        L.orc_rk54pd.argtypes = [C.POINTER(OrcCell), C.c_double, C.c_double, C.c_double, _DP, _DP,
                                 _DP, C.c_int, C.POINTER(C.c_int)]
        L.orc_rk54pd.restype = C.c_double

    # -- protocol
    def protocol(self, **kw):
        p = np.zeros((), PROT_DTYPE)
        self.L.orc_protocol_default(_p(p))
        for k, v in kw.items():
            p[k] = v
        return p

    # -- rng
    def seed_for_fru(self, a0: int, j: int) -> int:
        return int(self.L.orc_seed_for_fru(a0 & 0xFFFFFFFF, j))

    def mt_seed(self, seed: int):
        f = np.zeros((), FRU_DTYPE)
        self.L.orc_mt_seed(C.c_void_p(f.ctypes.data + FRU_DTYPE.fields["mt"][1]), seed & 0xFFFFFFFF)
        return f

    def mt_ptr(self, frus: np.ndarray, k: int = 0):
        return C.c_void_p(frus.ctypes.data + k * FRU_DTYPE.itemsize + FRU_DTYPE.fields["mt"][1])

    def mt_block(self, frus: np.ndarray, n: int, k: int = 0):
        out = np.zeros(n, np.float64)
        self.L.orc_mt_block(self.mt_ptr(frus, k), n, _p(out))
        return out

    def mt_u32(self, frus: np.ndarray, n: int, k: int = 0):
        ptr = self.mt_ptr(frus, k)
        return np.array([self.L.orc_mt_next_u32(ptr) for _ in range(n)], np.uint32)

    # -- kernel pieces
    @staticmethod
    def row(frus: np.ndarray, k: int):
        return C.c_void_p(frus.ctypes.data + k * FRU_DTYPE.itemsize)

    def rates(self, frus: np.ndarray, k: int, dep, prot=None):
        """orc_fru_rates on FRU k of frus (mutates RyR states / RNG like the reference)."""
        r = np.zeros((), RATES_DTYPE)
        dep = np.ascontiguousarray(dep, np.float64)
        m = self.L.orc_fru_rates(self.row(frus, k), _p(dep), _p(r),
                                 _p(prot) if prot is not None else None)
        return m, r

    def fcn_fru(self, conc, dep, lopen, nopen):
        conc = np.ascontiguousarray(conc, np.float64)
        dep = np.ascontiguousarray(dep, np.float64)
        lopen = np.ascontiguousarray(lopen, np.float64)
        nopen = np.ascontiguousarray(nopen, np.int32)
        d = np.zeros(5)
        self.L.orc_fcn_fru(_p(conc), _p(dep), _p(lopen), _p(nopen), _p(d))
        return d

    def simfru(self, t0, t1, dep0, depf, frus: np.ndarray, k: int = 0, prot=None, trace_cap=0):
        """orc_simfru on FRU k of frus; returns (steps, trace records)."""
        dep0 = np.ascontiguousarray(dep0, np.float64)
        depf = np.ascontiguousarray(depf, np.float64)
        tr = np.zeros(max(trace_cap, 1), STEP_DTYPE)
        n = self.L.orc_simfru(t0, t1, _p(dep0), _p(depf), self.row(frus, k),
                              _p(prot) if prot is not None else None,
                              _p(tr) if trace_cap else None, trace_cap)
        return n, tr[: min(max(n, 0), trace_cap)]

    def flux_sums(self, frus):
        out = np.zeros(5)
        self.L.orc_flux_sums(len(frus), _p(frus), _p(out))
        return out

    def avg_sums(self, frus):
        out = np.zeros(11)
        self.L.orc_avg_sums(len(frus), _p(frus), _p(out))
        return out

    def flux_currents(self, num5, dep, nfru, scale):
        num5 = np.ascontiguousarray(num5, np.float64)
        dep = np.ascontiguousarray(dep, np.float64)
        out = np.zeros(4)
        self.L.orc_flux_currents(_p(num5), _p(dep), nfru, scale, _p(out))
        return out

    def avg_outputs(self, num11, state, nfru, scale):
        num11 = np.ascontiguousarray(num11, np.float64)
        state = np.ascontiguousarray(state, np.float64)
        out = np.zeros(NOTHER)
        self.L.orc_avg_outputs(_p(num11), _p(state), nfru, scale, _p(out))
        return out

    def fcn(self, t, state, J4, keepc=1, prot=None):
        if prot is None:
            prot = self.protocol()
        st = np.array(state, np.float64)
        F = np.zeros(NGLOBAL)
        cur = np.zeros(NCUR)
        self.L.orc_fcn(t, _p(st), _p(F), _p(cur), keepc, *[float(x) for x in J4], _p(prot))
        return F, cur, st

    def errweight(self):
        w = np.zeros(NGLOBAL)
        self.L.orc_errweight(_p(w))
        return w

    def total_ca(self, frus, state, scale):
        state = np.ascontiguousarray(state, np.float64)
        return self.L.orc_total_ca(_p(frus), len(frus), _p(state), scale)


class OracleCell:
    """Whole cell on the CPU restatement (caf.c loop over rk_pd54m.c, one process)."""

    def __init__(self, orc: Oracle, frus: np.ndarray, scale: float, prot=None):
        self.o = orc
        self.c = OrcCell()
        if prot is None:
            prot = orc.protocol()
        self.prot = prot
        rc = orc.L.orc_cell_init(C.byref(self.c), len(frus), scale, _p(prot))
        if rc:
            raise MemoryError("orc_cell_init failed")
        self.nfru = len(frus)
        self.fru = np.ctypeslib.as_array(
            (C.c_byte * (FRU_DTYPE.itemsize * self.nfru)).from_address(self.c.fru)
        ).view(FRU_DTYPE)
        self.fru[:] = frus

    def close(self):
        if self.c.fru:
            self.o.L.orc_cell_free(C.byref(self.c))

    def __del__(self):  # pragma: no cover - best effort
        try:
            self.close()
        except Exception:
            pass

    def distrib(self, t0, t1, dep0, depf):
        out = np.zeros(4)
        self.o.L.orc_distrib_simfru(C.byref(self.c), t0, t1,
                                    _p(np.ascontiguousarray(dep0, np.float64)),
                                    _p(np.ascontiguousarray(depf, np.float64)), _p(out))
        return out

    def flux(self, dep):
        out = np.zeros(4)
        self.o.L.orc_send_calc_fru_flux(C.byref(self.c), _p(np.ascontiguousarray(dep, np.float64)),
                                        _p(out))
        return out

    def avg(self, state):
        out = np.zeros(NOTHER)
        self.o.L.orc_send_calc_fru_avg(C.byref(self.c),
                                       _p(np.ascontiguousarray(state, np.float64)), _p(out))
        return out

    def save(self):
        self.o.L.orc_save_state(C.byref(self.c))

    def resume(self):
        self.o.L.orc_resume_state(C.byref(self.c))

    def rk54pd(self, t, tstep, oldh, state, cur, log_cap=0):
        log = np.zeros((max(log_cap, 1), 4))
        n = C.c_int(0)
        h = self.o.L.orc_rk54pd(C.byref(self.c), t, tstep, oldh, _p(state), _p(cur),
                                _p(log) if log_cap else None, log_cap, C.byref(n))
        return h, log[: n.value]

    def trace(self, cap):
        """Record the calls of the following rk54pd calls (cap records)."""
        self._ev = np.zeros(max(cap, 1), EVENT_DTYPE)
        self.c.ev = self._ev.ctypes.data if cap else None
        self.c.ev_cap = cap
        self.c.ev_n = 0

    def events(self):
        n = self.c.ev_n
        if n > self.c.ev_cap:
            raise OverflowError(f"{n} events for {self.c.ev_cap} records")
        return self._ev[:n].copy()


class Ref:
    """The reference's own objects (oracle/_ref/libgrw_ref.so).  Arrays use the
    reference's LP64 layout: MT words are ``unsigned long`` (8 bytes)."""

    def __init__(self, path: str = REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle` where the reference exists")
        L = C.CDLL(path)
        self.L = L
        L.ref_sgrnd.argtypes = [C.c_uint32, _DP, _DP]
        L.ref_mt_fast.argtypes = [_DP, _DP, C.c_int, _DP]
        L.ref_mt_one.argtypes = [_DP, _DP]
        L.ref_mt_one.restype = C.c_double
        L.ref_initialize_ran.argtypes = [C.c_int, _DP, _DP]
        L.ref_fru_rates.argtypes = [_DP] * 15
        L.ref_fru_rates.restype = C.c_double
        L.ref_fcn_fru.argtypes = [_DP] * 5
        L.ref_simfru.argtypes = [C.c_double, C.c_double] + [_DP] * 8
        L.ref_flux_sums.argtypes = [C.c_int, _DP, _DP, _DP, _DP]
        L.ref_avg_sums.argtypes = [C.c_int, _DP, _DP, _DP, _DP, _DP]
        L.ref_fcn.argtypes = [C.c_double, _DP, _DP, _DP, C.c_int] + [C.c_double] * 4
        L.ref_initialize_state.argtypes = [C.c_char_p, C.c_int, C.c_double] + [_DP] * 8
        L.ref_initialize_state.restype = C.c_int
        L.ref_mpi_init.restype = C.c_int
        L.ref_load_block.argtypes = [C.c_int, C.c_double] + [_DP] * 6
        L.ref_load_block.restype = C.c_int
        L.ref_get_block.argtypes = [C.c_int] + [_DP] * 6
        L.ref_get_block.restype = C.c_int
        L.ref_send_calc_fru_flux.argtypes = [_DP, _DP]
        L.ref_send_calc_fru_avg.argtypes = [_DP, _DP]
        L.ref_distrib_simfru.argtypes = [C.c_double, C.c_double, _DP, _DP, _DP]
        L.ref_set_errweight.argtypes = [_DP]
        L.ref_rk54pd.argtypes = [C.c_double, C.c_double, C.c_double, _DP, _DP]
        L.ref_rk54pd.restype = C.c_double
        L.ref_trace.argtypes = [_DP, C.c_int]
        L.ref_trace_n.restype = C.c_int
        L.ref_dynamic_fru.argtypes = [C.c_double] * 4 + [_DP] * 3
        L.ref_set_nfru.argtypes = [C.c_int, C.c_double]

    # FRU arrays in reference layout, for n FRUs
    @staticmethod
    def alloc(n):
        return dict(
            fru=np.zeros((n, 5)),
            ltype=np.zeros((n, 4, 2), np.int32),
            ryr=np.zeros((n, 4, 5), np.int32),
            ito2=np.zeros((n, 4), np.int32),
            mt=np.zeros((n, MTN + 1), np.uint64),
            mti=np.zeros(n, np.int32),
        )

    @staticmethod
    def from_orc(frus: np.ndarray):
        n = len(frus)
        a = Ref.alloc(n)
        a["fru"][:] = frus["conc"]
        a["ltype"][:, :, 0] = frus["lcc"]
        a["ltype"][:, :, 1] = frus["yinact"]
        a["ryr"][:] = frus["ryr"]
        a["ito2"][:] = frus["ito2"]
        a["mt"][:, :MTN] = frus["mt"]
        a["mti"][:] = frus["mti"]
        return a

    @staticmethod
    def to_orc(a) -> np.ndarray:
        n = len(a["fru"])
        f = np.zeros(n, FRU_DTYPE)
        f["conc"] = a["fru"]
        f["lcc"] = a["ltype"][:, :, 0]
        f["yinact"] = a["ltype"][:, :, 1]
        f["ryr"] = a["ryr"]
        f["ito2"] = a["ito2"]
        f["mt"] = a["mt"][:, :MTN].astype(np.uint32)
        f["mti"] = a["mti"]
        return f

    def sgrnd(self, seed):
        mt = np.zeros(MTN + 1, np.uint64)
        mti = np.zeros(1, np.int32)
        self.L.ref_sgrnd(seed & 0xFFFFFFFF, _p(mti), _p(mt))
        return mt, mti

    def initialize_ran(self, n):
        mt = np.zeros((n, MTN + 1), np.uint64)
        mti = np.zeros(n, np.int32)
        self.L.ref_initialize_ran(n, _p(mt), _p(mti))
        return mt, mti

    def rates(self, a, k, dep):
        """fru_rates_local on FRU k of array set a (mutates a)."""
        lrate = np.zeros((4, 4)); lidx = np.zeros((4, 4), np.int32); llen = np.zeros(4, np.int32)
        vrate = np.zeros(4); rrate = np.zeros((4, 5, 4)); ridx = np.zeros((4, 5, 4), np.int32)
        rlen = np.zeros((4, 5), np.int32); irate = np.zeros(4)
        dep = np.ascontiguousarray(dep, np.float64)
        row = lambda name: C.c_void_p(a[name].ctypes.data + k * a[name][0].nbytes)
        m = self.L.ref_fru_rates(row("ltype"), row("ryr"), row("ito2"), _p(dep), row("fru"),
                                 _p(lrate), _p(lidx), _p(llen), _p(vrate), _p(rrate), _p(ridx),
                                 _p(rlen), _p(irate), row("mti"), row("mt"))
        return m, dict(lcc_rate=lrate, lcc_dest=lidx, lcc_len=llen, vinact_rate=vrate,
                       ryr_rate=rrate, ryr_dest=ridx, ryr_len=rlen, ito2_rate=irate)

    def fcn_fru(self, conc, dep, lopen, nopen):
        conc = np.ascontiguousarray(conc, np.float64)
        dep = np.ascontiguousarray(dep, np.float64)
        lopen = np.ascontiguousarray(lopen, np.float64)
        nopen = np.ascontiguousarray(nopen, np.int32)
        d = np.zeros(5)
        self.L.ref_fcn_fru(_p(conc), _p(dep), _p(lopen), _p(nopen), _p(d))
        return d

    def simfru(self, t0, t1, dep0, depf, a, k):
        dep0 = np.array(dep0, np.float64)
        depf = np.array(depf, np.float64)
        row = lambda name: C.c_void_p(a[name].ctypes.data + k * a[name][0].nbytes)
        self.L.ref_simfru(t0, t1, _p(dep0), _p(depf), row("ltype"), row("ryr"), row("ito2"),
                          row("fru"), row("mti"), row("mt"))

    def flux_sums(self, a):
        out = np.zeros(5)
        self.L.ref_flux_sums(len(a["fru"]), _p(a["ltype"]), _p(a["ito2"]), _p(a["fru"]), _p(out))
        return out

    def avg_sums(self, a):
        out = np.zeros(11)
        self.L.ref_avg_sums(len(a["fru"]), _p(a["fru"]), _p(a["ltype"]), _p(a["ryr"]),
                            _p(a["ito2"]), _p(out))
        return out

    # -- the MPI master surface (mpi_master.c) over one rank
    def load_block(self, a, scale):
        """initialize_mpi_state (mpi_master.c:199) of array set a (<= 1000 FRUs)."""
        rc = self.L.ref_load_block(len(a["fru"]), scale, _p(a["fru"]), _p(a["ltype"]), _p(a["ryr"]),
                                   _p(a["ito2"]), _p(a["mti"]), _p(a["mt"]))
        if rc:
            raise RuntimeError(f"ref_load_block rc={rc}")

    def get_block(self, n):
        """parallel_get_FRUs (mpi_master.c:509) into a fresh array set."""
        a = self.alloc(n)
        rc = self.L.ref_get_block(n, _p(a["fru"]), _p(a["ltype"]), _p(a["ryr"]), _p(a["ito2"]),
                                  _p(a["mti"]), _p(a["mt"]))
        if rc:
            raise RuntimeError(f"ref_get_block rc={rc}")
        return a

    def send_calc_fru_flux(self, dep):
        out = np.zeros(4)
        self.L.ref_send_calc_fru_flux(_p(np.ascontiguousarray(dep, np.float64)), _p(out))
        return out

    def send_calc_fru_avg(self, state):
        out = np.zeros(NOTHER)
        self.L.ref_send_calc_fru_avg(_p(np.ascontiguousarray(state, np.float64)), _p(out))
        return out

    def distrib_simfru(self, t0, t1, dep0, depf):
        out = np.zeros(4)
        self.L.ref_distrib_simfru(t0, t1, _p(np.ascontiguousarray(dep0, np.float64)),
                                  _p(np.ascontiguousarray(depf, np.float64)), _p(out))
        return out

    def set_errweight(self, w):
        self._errw = np.ascontiguousarray(w, np.float64)
        self.L.ref_set_errweight(_p(self._errw))

    def trace(self, cap):
        self._ev = np.zeros(max(cap, 1), EVENT_DTYPE)
        self.L.ref_trace(_p(self._ev) if cap else None, cap)

    def events(self):
        n = self.L.ref_trace_n()
        if n > len(self._ev):
            raise OverflowError(f"{n} events for {len(self._ev)} records")
        return self._ev[:n].copy()

    def rk54pd(self, t, tstep, oldh, state, cur):
        """rk54pd (rk_pd54m.c:29) on the loaded block; state and cur updated in place."""
        return self.L.ref_rk54pd(t, tstep, oldh, _p(state), _p(cur))

    def dynamic_fru(self, t, V, pryr, plcc, phase, t_last):
        """dynamicFRU's phase decision (dynamicFRU.c:26-150), factor 1."""
        ph = np.array([phase], np.int32)
        tl = np.array([t_last], np.float64)
        tts = np.zeros(1)
        self.L.ref_dynamic_fru(t, V, pryr, plcc, _p(ph), _p(tl), _p(tts))
        return int(ph[0]), float(tl[0]), float(tts[0])

    def fcn(self, t, state, J4, keepc=1):
        st = np.array(state, np.float64)
        F = np.zeros(NGLOBAL)
        cur = np.zeros(NCUR)
        self.L.ref_fcn(t, _p(st), _p(F), _p(cur), keepc, *[float(x) for x in J4])
        return F, cur, st

    def initialize_state(self, ref_root, n, scale):
        a = self.alloc(n)
        state = np.zeros(NGLOBAL)
        errw = np.zeros(NGLOBAL)
        rc = self.L.ref_initialize_state(ref_root.encode(), n, scale, _p(state), _p(a["fru"]),
                                         _p(a["ltype"]), _p(a["ryr"]), _p(a["ito2"]), _p(a["mt"]),
                                         _p(a["mti"]), _p(errw))
        if rc:
            raise OSError(f"ref_initialize_state rc={rc}")
        return state, a, errw
