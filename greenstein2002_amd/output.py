"""The reference's output files and the driver loop that writes them.

Host-side file I/O around a ``Cell`` (formats byte for byte those of the
reference's C writers):

* ``states.N.txt``, ``currents.N.txt``, ``otherstates.N.txt`` with their header
  lines and row formats (data_io.c:22-140);
* the restart dump ``r_states.N.txt``, ``r_FRU.N.txt``, ``r_LCh.N.txt``,
  ``r_RyR.N.txt``, ``r_Ito2.N.txt`` and, optionally, the binary RNG-state
  records ``r_randomseeds.N.txt`` (``restart_data``, data_io.c:197-294);
* the per-FRU property files ``fruprops_j.N.txt``, ``fruprops_avg_j.N.txt``
  and ``fruprops_avg_tot.N.txt`` (write_fru_props.c:28-187);
* ``run_caf``: the output schedule of the reference's main loop
  (caf.c:127-276): a row of states and otherstates at the end of every
  time_step, the currents of each call's start one row later, a restart dump
  and a new file number every ``ndump`` steps, and the final currents from
  ``lastcall`` (fcn.c:1077-1092).

Reference quirks kept on purpose, each where it is implemented: the currents
row of the first call after a dump is not written (caf.c:171-180); the
per-FRU files index the LCC array with the cleft and field swapped
(write_fru_props.c:90-91).  One is not kept: the reference closes
``r_states`` twice (data_io.c:220,265), which glibc aborts on; here every file
is closed once.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import constants as K

# data_io.c:32-40, 56-60, 77-83 (the otherstates and currents headers end in
# a blank before the newline, the states header does not)
STATES_HEADER = ("Time V mNa hNa jNa Nai Ki Cai CaNSR xKs LTRPNCa HTRPNCa C0Kv43 C1Kv43 C2Kv43 "
                 "C3Kv43 OKv43 CI0Kv43 CI1Kv43 CI2Kv43 CI3Kv43 OIKv43 C0Kv14 C1Kv14 C2Kv14 "
                 "C3Kv14 OKv14 CI0Kv14 CI1Kv14 CI2Kv14 CI3Kv14 OIKv14 CaToT C1Herg C2Herg "
                 "C3Herg OHerg IHerg\n")
OTHER_HEADER = ("Time CaSSavg CaJSRavg JRyRtot PRyR_open PRyR_ready PNorm_mode PnotVinact "
                "PLType_open CaToT2 PIto2_open CaJSRtot CaSStot \n")
CURRENTS_HEADER = ("Time INa IKr IKs Ito1 IK1 IKp INaCa INaK IpCa ICab INab ICa JDHPR Jup Jtrpn "
                   "Jtr Jxfer IKv43 IKv14 IKv14_K IKv14_Na Ito2 Istim Itot \n")
# write_fru_props.c:148-151, 165-168, 181-184
FRUPROPS_HEADER = "Time JLType JRyR Jtr Jxfer CaSS CaJSR Ito2 NRyR_open\n"
FRUPROPS_AVG_HEADER = "Time JLType_tot JRyR_tot Jtr Jxfer_tot CaSS_avg CaJSR Ito2 PRyR_open\n"
FRUPROPS_AVGTOT_HEADER = ("Time JLType_totAll JRyR_totAll Jtr_totAll Jxfer_totAll CaSS_avgAll "
                          "CaJSR_avgAll Ito2 PRyR_openAll\n")

# parameters.h:177-184 (file name stems) and 186 (first file number)
STATES_FILE = "states"
CURRENTS_FILE = "currents"
OTHERSTATES_FILE = "otherstates"
FRUPROPS_FILE = "fruprops"
FRUPROPS_AVG_FILE = "fruprops_avg"
FRUPROPS_AVGTOT_FILE = "fruprops_avg_tot"
SEEDS_FILE = "r_randomseeds"


def _num(spec: str, x) -> str:
    """printf of one double as glibc prints it: Python's % formatting follows C
    for %g / %.8g except that glibc writes a NaN with its sign bit as "-nan"."""
    x = float(x)
    if x != x:
        return "-nan" if math.copysign(1.0, x) < 0 else "nan"
    return spec % x


def _tprint(time: float, ts_sec: bool) -> float:
    # data_io.c:96: output times in seconds when ts_sec (parameters.h:101)
    return time * 1.e-3 if ts_sec else time


def states_row(time: float, state, ts_sec: bool = True) -> str:
    """write_states (data_io.c:90-105): "%g " time, "%.8g " per state."""
    return _num("%g", _tprint(time, ts_sec)) + " " + "".join(_num("%.8g", v) + " " for v in state) + "\n"


def otherstates_row(time: float, other, ts_sec: bool = True) -> str:
    """write_otherstates (data_io.c:108-122): "%g " everywhere."""
    return _num("%g", _tprint(time, ts_sec)) + " " + "".join(_num("%g", v) + " " for v in other) + "\n"


def currents_row(time: float, current, ts_sec: bool = True) -> str:
    """write_currents (data_io.c:125-140): "%g " everywhere."""
    return _num("%g", _tprint(time, ts_sec)) + " " + "".join(_num("%g", v) + " " for v in current) + "\n"


def _path(output_dir: str, stem: str, filenumber: int) -> str:
    # sprintf("%s%s.%d.txt", output_dir, stem, filenumber): the directory is a
    # prefix, so it carries its own separator (parameters.h:177 "./")
    return "%s%s.%d.txt" % (output_dir, stem, filenumber)


def open_states(output_dir: str, filenumber: int):
    f = open(_path(output_dir, STATES_FILE, filenumber), "w")
    f.write(STATES_HEADER)
    return f


def open_otherstates(output_dir: str, filenumber: int):
    f = open(_path(output_dir, OTHERSTATES_FILE, filenumber), "w")
    f.write(OTHER_HEADER)
    return f


def open_currents(output_dir: str, filenumber: int):
    f = open(_path(output_dir, CURRENTS_FILE, filenumber), "w")
    f.write(CURRENTS_HEADER)
    return f


# ---------------------------------------------------------------- restart dump

def restart_data(output_dir: str, filenumber: int, state, fru, ltype, ryr, ito2,
                 mti=None, mt=None, save_seeds: bool = False, word_bytes: int = 8) -> list[str]:
    """restart_data (data_io.c:197-294) for the FRUs in the arrays.

    Text files: 37 lines "%g" of the global state; then, after a "%d" FRU count
    line, "%g " x5 per FRU, "%d %d" per cleft (LCC state, gate), "%d " x5 per
    cleft and "%d" per cleft.  With ``save_seeds`` the binary RNG records
    [iFRU, mti, mt[0..623]] in ``word_bytes``-byte little-endian words (8: the
    reference's LP64 ``unsigned long``; 4: the layout of the shipped
    ic/stable/ic_randomseeds_stable.txt).  Returns the paths written."""
    fru = np.asarray(fru, np.float64).reshape(-1, 5)
    n = fru.shape[0]
    ltype = np.asarray(ltype).reshape(n, 4, 2)
    ryr = np.asarray(ryr).reshape(n, 4, 5)
    ito2 = np.asarray(ito2).reshape(n, 4)
    paths = []
    p = "%sr_states.%d.txt" % (output_dir, filenumber)
    with open(p, "w") as f:
        f.write("".join(_num("%g", v) + "\n" for v in np.asarray(state).reshape(-1)))
    paths.append(p)
    names = [("r_FRU", []), ("r_LCh", []), ("r_RyR", []), ("r_Ito2", [])]
    parts = {k: v for k, v in names}
    for k in parts:
        parts[k].append("%d\n" % n)
    for i in range(n):
        parts["r_FRU"].append("".join(_num("%g", v) + " " for v in fru[i]) + "\n")
        for c in range(4):
            parts["r_LCh"].append("%d %d\n" % (int(ltype[i, c, 0]), int(ltype[i, c, 1])))
            parts["r_RyR"].append("".join("%d " % int(v) for v in ryr[i, c]) + "\n")
            parts["r_Ito2"].append("%d\n" % int(ito2[i, c]))
    for stem, lines in names:
        p = "%s%s.%d.txt" % (output_dir, stem, filenumber)
        with open(p, "w") as f:
            f.write("".join(lines))
        paths.append(p)
    if save_seeds:
        if mti is None or mt is None:
            raise ValueError("save_seeds needs mti and mt")
        p = _path(output_dir, SEEDS_FILE, filenumber)
        write_seeds(p, mti, mt, word_bytes)
        paths.append(p)
    return paths


def write_seeds(path: str, mti, mt, word_bytes: int = 8) -> None:
    """Binary RNG-state records (data_io.c:270-293): per FRU i the words
    i, mti[i], mt[i][0..623] in ``word_bytes``-byte little-endian words."""
    if word_bytes not in (4, 8):
        raise ValueError("word_bytes must be 4 or 8")
    mti = np.asarray(mti, np.int64).reshape(-1)
    mt = np.asarray(mt, np.uint64).reshape(len(mti), -1)[:, :624]
    n = len(mti)
    rec = np.empty((n, 626), np.uint64)
    rec[:, 0] = np.arange(n, dtype=np.uint64)
    rec[:, 1] = mti.astype(np.uint64)
    rec[:, 2:] = mt
    rec.astype("<u8" if word_bytes == 8 else "<u4").tofile(path)


# ------------------------------------------------------------ per-FRU properties

def fru_props_rows(time: float, V: float, Cai: float, CaNSR: float, fru, ltype, ryr, ito2,
                   fru_start: int = 0, fru_end: int = 9, ts_sec: bool = True):
    """write_fru_props (write_fru_props.c:28-134): returns (rows, avg_rows,
    avgtot_row); rows[j - fru_start] holds the four cleft lines of FRU j.

    The open-LCC test reads the LCC array of FRU j at (cleft, field) =
    (0, jcleft) and (1, jcleft) as the reference's
    LType_state[j][index_LCC_states][jcleft] and [j][index_LCC_Vinact][jcleft]
    do (write_fru_props.c:90-91): over the FRU's eight ints in cleft-major
    order, cleft jcleft's "state" is int jcleft and its "gate" int 2 + jcleft."""
    fru = np.asarray(fru, np.float64).reshape(-1, 5)
    nfru = fru.shape[0]
    if fru_end > nfru:  # write_fru_props.c:63-71
        raise ValueError("Inappropriate fru_end (%d) in write_fru_props" % fru_end)
    if fru_start < 0:
        raise ValueError("Inappropriate fru_start (%d) in write_fru_props" % fru_start)
    lflat = np.asarray(ltype).reshape(nfru, 8)
    ryr = np.asarray(ryr).reshape(nfru, 4, 5)
    ito2 = np.asarray(ito2).reshape(nfru, 4)
    tp = _tprint(time, ts_sec)
    VF_over_RT = V / K.RT_over_F
    VFsq_over_RT = (1000.0 * K.Faraday) * VF_over_RT
    JLType_totALL = JRyR_totALL = Jtr_totALL = Jxfer_totALL = 0.0
    CaSS_avgALL = CaJSR_avgALL = PRyR_openALL = Ito2_avgALL = 0.0
    rows, avg_rows = [], []
    for j in range(fru_start, fru_end + 1):
        JLType_tot1 = JRyR_tot1 = Jxfer_tot1 = CaSS_avg1 = PRyR_open1 = Ito2_tot1 = 0.0
        CaJSR = float(fru[j, 0])
        CaJSR_avgALL = CaJSR_avgALL + CaJSR
        Jtr = (CaNSR - CaJSR) / K.tautr * K.VJSR / K.Vmyo
        Jtr_totALL = Jtr_totALL + Jtr
        lines = []
        for c in range(4):
            CaSS = float(fru[j, c + 1])
            CaSS_avg1 += CaSS
            CaSS_avgALL += CaSS
            JL = 0.0
            s, g = int(lflat[j, c]), int(lflat[j, 2 + c])
            if (s == 6 or s == 12) and g == 2:
                JL = (K.PCa * 4.0 * VFsq_over_RT * (CaSS * math.exp(2.0 * VF_over_RT) - K.Cao * 0.341)
                      / (math.exp(2.0 * VF_over_RT) - 1.0) / (-2.0 * K.Vmyo * K.Faraday * 1000.0))
                JLType_tot1 += JL
                JLType_totALL += JL
            I2 = 0.0
            if int(ito2[j, c]) == 2:
                I2 = (K.PCl * VFsq_over_RT * (K.Cli * math.exp(-VF_over_RT) - K.Clo)
                      / (math.exp(-VF_over_RT) - 1.0))
                Ito2_tot1 += I2
                Ito2_avgALL += I2
            nopen = int(sum(1 for r in ryr[j, c] if r in (3, 4, 7)))
            PRyR_open1 = PRyR_open1 + float(nopen)
            PRyR_openALL = PRyR_openALL + float(nopen)
            JR = K.JRyRmax * float(nopen) * (CaJSR - CaSS) * K.VSS / K.Vmyo
            JRyR_tot1 = JRyR_tot1 + JR
            JRyR_totALL = JRyR_totALL + JR
            Jx = (CaSS - Cai) / K.tauxfer * K.VSS / K.Vmyo
            Jxfer_tot1 = Jxfer_tot1 + Jx
            Jxfer_totALL = Jxfer_totALL + Jx
            lines.append(" ".join(_num("%.8g", v) for v in (tp, JL, JR, Jtr, Jx, CaSS, CaJSR, I2))
                         + " %d\n" % nopen)
        rows.append(lines)
        CaSS_avg1 = CaSS_avg1 / 4
        PRyR_open1 = PRyR_open1 / float(5 * 4)
        avg_rows.append(" ".join(_num("%.8g", v) for v in (tp, JLType_tot1, JRyR_tot1, Jtr, Jxfer_tot1,
                                                           CaSS_avg1, CaJSR, Ito2_tot1, PRyR_open1))
                        + "\n")
    num_avg = float(fru_end - fru_start + 1)
    CaSS_avgALL = CaSS_avgALL / num_avg / 4
    CaJSR_avgALL = CaJSR_avgALL / num_avg
    PRyR_openALL = PRyR_openALL / (num_avg * float(5 * 4))
    avgtot = " ".join(_num("%.8g", v) for v in (tp, JLType_totALL, JRyR_totALL, Jtr_totALL, Jxfer_totALL,
                                                 CaSS_avgALL, CaJSR_avgALL, Ito2_avgALL, PRyR_openALL)) + "\n"
    return rows, avg_rows, avgtot


class FruPropsFiles:
    """open_fru_props / open_fru_props_avg / open_fru_props_avgtot
    (write_fru_props.c:137-187) and the rows of write_fru_props."""

    def __init__(self, output_dir: str, filenumber: int, fru_start: int = 0, fru_end: int = 9,
                 ts_sec: bool = True):
        self.fru_start, self.fru_end, self.ts_sec = fru_start, fru_end, ts_sec
        self.per, self.avg = [], []
        for j in range(fru_start, fru_end + 1):
            f = open("%s%s_%d.%d.txt" % (output_dir, FRUPROPS_FILE, j, filenumber), "w")
            f.write(FRUPROPS_HEADER)
            self.per.append(f)
        for j in range(fru_start, fru_end + 1):
            f = open("%s%s_%d.%d.txt" % (output_dir, FRUPROPS_AVG_FILE, j, filenumber), "w")
            f.write(FRUPROPS_AVG_HEADER)
            self.avg.append(f)
        self.tot = open(_path(output_dir, FRUPROPS_AVGTOT_FILE, filenumber), "w")
        self.tot.write(FRUPROPS_AVGTOT_HEADER)

    def write(self, time, V, Cai, CaNSR, fru, ltype, ryr, ito2) -> None:
        rows, avg_rows, avgtot = fru_props_rows(time, V, Cai, CaNSR, fru, ltype, ryr, ito2,
                                                self.fru_start, self.fru_end, self.ts_sec)
        for f, lines in zip(self.per, rows):
            f.write("".join(lines))
        for f, line in zip(self.avg, avg_rows):
            f.write(line)
        self.tot.write(avgtot)

    def close(self) -> None:
        for f in self.per + self.avg + [self.tot]:
            f.close()


# ------------------------------------------------------------------ main loop

@dataclass
class CafOutputs:
    filenumber: int       # the file number after the run (caf.c:225)
    n_steps: int          # calls of rk54pd
    dumps: list           # file numbers of the restart dumps written
    state: np.ndarray     # final global state


def run_caf(cell, time_end: float, output_dir: str = "./", time_start: float = 0.0,
            ndump: int = 10000, filenumber0: int = 0, ts_sec: bool = True,
            fru_props: bool = False, fru_start: int = 0, fru_end: int = 9,
            save_seeds: bool = False, driver: str = "device") -> CafOutputs:
    """The output schedule of caf.c:127-276 around ``cell`` (a ``Cell`` whose
    clock is at ``time_start``; time_step from its protocol).

    Per call k of rk54pd over [time_now, time_now + time_step]: the currents
    of the call's start go to currents.N under time_now except for the first
    call, and except for the first call after a restart dump, which reopens
    currents.N under the new file number instead (caf.c:171-180); states and
    otherstates go out under time_now + time_step (caf.c:184-185); with
    ``fru_props`` the per-FRU files (caf.c:187-194); every ``ndump`` calls the
    restart dump, after which states/otherstates (and the FRU files) continue
    under the next file number unless the run is within a step of its end
    (caf.c:196-228).  After the loop the currents of the final state
    (lastcall, fcn.c:1077-1092) go to currents.N under time_end (caf.c:254-255).
    Under the I-V clamp (protocol.ivclamp) every iv_period of simulated time
    the cell restarts from its initial condition and seeds with the next test
    potential (iv_n + 1), the clock running on (caf.c:232-244); a
    ``dynamic.DynamicCell`` then decides on its FRU count (caf.c:246-250)."""
    from .cell import fcn as host_fcn

    step = cell.protocol.time_step
    filenumber = filenumber0
    sf = open_states(output_dir, filenumber)
    cf = open_currents(output_dir, filenumber)
    of = open_otherstates(output_dir, filenumber)
    fp = FruPropsFiles(output_dir, filenumber, fru_start, fru_end, ts_sec) if fru_props else None
    n_restart = 0
    dumps = []
    calls = 0
    time_now = time_start
    iv_time = 0.0
    try:
        while time_now <= time_end - .0001 * step:
            n_restart = n_restart + 1
            tstep = time_now + step
            r = cell.run(1, log_cap=0, outputs=True, driver=driver)
            calls += 1
            state, current, other = r.states[0], r.currents[0], r.other[0]
            if time_now > time_start:
                if n_restart == 1:
                    cf.close()
                    cf = open_currents(output_dir, filenumber)
                else:
                    cf.write(currents_row(time_now, current, ts_sec))
            sf.write(states_row(tstep, state, ts_sec))
            of.write(otherstates_row(tstep, other, ts_sec))
            if fp is not None:
                g = cell.fru.parallel_get_FRUs()
                fp.write(tstep, float(state[0]), float(state[6]), float(state[7]), g["fru"], g["ltype"],
                         g["ryr"], g["ito2"])
            if n_restart >= ndump:
                n_restart = 0
                g = cell.fru.parallel_get_FRUs()
                restart_data(output_dir, filenumber, state, g["fru"], g["ltype"], g["ryr"], g["ito2"],
                             g["mti"], g["mt"], save_seeds=save_seeds)
                dumps.append(filenumber)
                filenumber = filenumber + 1
                sf.close()
                of.close()
                if fp is not None:
                    fp.close()
                sf = of = fp = None
                if time_now < time_end - 1.0001 * step:
                    sf = open_states(output_dir, filenumber)
                    of = open_otherstates(output_dir, filenumber)
                    if fru_props:
                        fp = FruPropsFiles(output_dir, filenumber, fru_start, fru_end, ts_sec)
            iv_time = iv_time + step
            if cell.protocol.ivclamp and iv_time >= cell.protocol.iv_period:  # caf.c:232-244
                iv_time = 0.0
                cell.protocol.iv_n = cell.protocol.iv_n + 1
                cell.restart()
            if hasattr(cell, "dynamic_step"):  # dynamicFRU after the step (caf.c:246-250)
                cell.dynamic_step(tstep, state, other)
            time_now += step
        # lastcall (fcn.c:1077-1092): fluxes of the final state, then fcn's currents
        st = np.array(cell.state, np.float64)
        J = cell.fru.send_calc_fru_flux(st[[0, 6, 7]])
        _, cur, _ = host_fcn(time_end, st, J, cell.protocol)
        cf.write(currents_row(time_end, cur, ts_sec))
    finally:
        for f in (cf, sf, of):
            if f is not None:
                f.close()
        if fp is not None:
            fp.close()
    return CafOutputs(filenumber=filenumber, n_steps=calls, dumps=dumps, state=np.array(cell.state))
