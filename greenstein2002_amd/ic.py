"""Initial conditions in the reference's formats.

``read_ic_dir`` parses the reference's ``ic/<name>/`` text files the way
``initialize_state.c:43-133`` does (one header count, then whitespace-separated
numbers; records are reused cyclically every min(NFRU, 1250) FRUs,
``initialize_state.c:38-41,62-65``).  ``load_stable`` returns the same data for
the shipped ``ic/stable`` set, packaged as ``data/ic_stable.npz``.
``total_ca`` is the cell-Ca bookkeeping of ``initialize_state.c:233-246`` and
``errweight`` the RK error weights of ``initialize_state.c:274-310``.
"""
from __future__ import annotations

import os
import re

import numpy as np

from . import constants as K

HERE = os.path.dirname(os.path.abspath(__file__))
STABLE_NPZ = os.path.join(HERE, "data", "ic_stable.npz")
REWIND = 1250

_NUM = re.compile(rb"[0-9eE+\-.]+")


def _tokens(path: str):
    with open(path, "rb") as f:
        return _NUM.findall(f.read())


def read_ic_dir(path: str, nfru: int, names=None) -> dict:
    """Read ic_{states,FRU,LCh,RyR,Ito2}_*.txt from ``path`` for ``nfru`` FRUs."""
    names = names or IC_NAMES
    rew = min(nfru, REWIND)

    def records(name, per, conv):
        tok = _tokens(os.path.join(path, name))
        nrec = int(tok[0])
        vals = np.array([conv(t) for t in tok[1:1 + per * min(nrec, rew)]])
        vals = vals.reshape(-1, per)
        idx = np.arange(nfru) % rew
        return vals[idx]

    out = dict(
        ryr=records(names["ryr"], 20, int).reshape(nfru, 4, 5).astype(np.int32),
        ltype=records(names["lch"], 8, int).reshape(nfru, 4, 2).astype(np.int32),
        ito2=records(names["ito2"], 4, int).reshape(nfru, 4).astype(np.int32),
        fru=records(names["fru"], 5, float).reshape(nfru, 5).astype(np.float64),
    )
    tok = _tokens(os.path.join(path, names["states"]))
    out["state"] = np.array([float(t) for t in tok[:37]])
    return out


IC_NAMES = dict(states="ic_states_stable.txt", fru="ic_FRU_stable.txt", lch="ic_LCh_stable.txt",
                ryr="ic_RyR_stable.txt", ito2="ic_Ito2_stable.txt")


def write_ic_dir(path: str, ic: dict | None = None, names=None) -> list[str]:
    """Write an initial-condition set in the reference's ic/ text formats (the
    files initialize_state.c:43-133 reads: the 37 global states, then a record
    count line and per FRU the 5 concentrations, 4 x "LCC gate", 4 x 5 RyR
    states and 4 Ito2 states).  Doubles are written with %.17g, so the
    reference's tokenizer (data_io.c:142-166, strtod) reads back the same
    values bit for bit.  ``ic`` defaults to the packaged ic/stable set."""
    if ic is None:
        d = np.load(STABLE_NPZ)
        ic = {k: d[k] for k in ("state", "fru", "ltype", "ryr", "ito2")}
    names = names or IC_NAMES
    os.makedirs(path, exist_ok=True)
    fru = np.asarray(ic["fru"], np.float64).reshape(-1, 5)
    n = len(fru)
    lt = np.asarray(ic["ltype"]).reshape(n, 4, 2)
    ryr = np.asarray(ic["ryr"]).reshape(n, 4, 5)
    ito2 = np.asarray(ic["ito2"]).reshape(n, 4)
    body = dict(
        states="".join("%.17g\n" % v for v in np.asarray(ic["state"], np.float64).reshape(-1)),
        fru="%d\n" % n + "".join(" ".join("%.17g" % v for v in row) + "\n" for row in fru),
        lch="%d\n" % n + "".join("%d %d\n" % (a, b) for a, b in lt.reshape(-1, 2)),
        ryr="%d\n" % n + "".join(" ".join("%d" % v for v in row) + "\n" for row in ryr.reshape(-1, 5)),
        ito2="%d\n" % n + "".join("%d\n" % v for v in ito2.reshape(-1)),
    )
    paths = []
    for key, text in body.items():
        p = os.path.join(path, names[key])
        with open(p, "w") as f:
            f.write(text)
        paths.append(p)
    return paths


def read_seeds(path: str, nfru: int, word_bytes: int = 8) -> dict:
    """Binary RNG-state records as initialize_state.c:135-160 reads them (and
    data_io.c:268-292 writes them): per FRU i the words i, mti, mt[0..623] in
    ``word_bytes``-byte little-endian words (8: an LP64 ``unsigned long``; 4:
    the shipped ic/stable/ic_randomseeds_stable.txt).  Records are read in
    order without rewinding.  Where the reference prints an error and goes on
    (a short record) or aborts (a record whose first word is not its FRU
    index), this raises ValueError.  Returns mti [nfru] and mt [nfru, 624]."""
    if word_bytes not in (4, 8):
        raise ValueError("word_bytes must be 4 or 8")
    rec = 626
    raw = np.fromfile(path, dtype="<u8" if word_bytes == 8 else "<u4")
    if raw.size < nfru * rec:
        raise ValueError("Unable to read seeds from file '%s': %d full records, %d wanted"
                         % (path, raw.size // rec, nfru))
    r = raw[: nfru * rec].reshape(nfru, rec).astype(np.uint64)
    bad = np.nonzero(r[:, 0] != np.arange(nfru, dtype=np.uint64))[0]
    if bad.size:
        i = int(bad[0])
        raise ValueError("RNG seed file fails consistency check: %d!=%d" % (int(r[i, 0]), i))
    mti = r[:, 1].astype(np.int64)
    if np.any(mti > 624):
        raise ValueError("RNG seed file holds mti > 624")
    return dict(mti=mti.astype(np.int32), mt=(r[:, 2:] & np.uint64(0xFFFFFFFF)).astype(np.uint32))


def load_stable(nfru: int, scale: float | None = None) -> dict:
    """ic/stable replicated to ``nfru`` FRUs, with state[CaTOT] recomputed for
    that FRU count and scale as initialize_state.c:233-246 does."""
    d = np.load(STABLE_NPZ)
    idx = np.arange(nfru) % min(nfru, REWIND)
    out = dict(fru=d["fru"][idx].copy(), ltype=d["ltype"][idx].copy(), ryr=d["ryr"][idx].copy(),
               ito2=d["ito2"][idx].copy(), state=d["state"].copy())
    if scale is None:
        scale = 12500.0 / nfru
    out["state"][31] = total_ca(out["fru"], out["state"], scale)
    return out


def total_ca(fru: np.ndarray, state: np.ndarray, scale: float) -> float:
    """initialize_state.c:233-246, sequential FRU order."""
    a1 = 0.0
    a2 = 0.0
    for k in range(len(fru)):
        cj = float(fru[k, 0])
        a1 = a1 + cj * (1.0 + K.CSQNtot / (K.KmCSQN + cj))
        for c in range(4):
            x = float(fru[k, c + 1])
            a2 = a2 + x * (1.0 + K.BSRtot / (K.KBSR + x) + K.BSLtot / (K.KBSL + x))
    Cai = float(state[6])
    return 1.e6 * ((a1 * K.VJSR + a2 * K.VSS) * scale + float(state[7]) * K.VNSR +
                   Cai * K.Vmyo * (1.0 + K.CMDNtot / (K.KmCMDN + Cai) + K.EGTAtot / (K.KmEGTA + Cai)) +
                   (float(state[9]) * K.LTRPNtot + float(state[10]) * K.HTRPNtot) * K.Vmyo)


def errweight() -> np.ndarray:
    """initialize_state.c:274-310."""
    w = np.ones(37)
    w[0] = 0.0
    w[4] = 0.1
    w[5] = 1 / 140.0
    w[6] = 1.e3
    w[7] = 2.5
    w[20] = 0.0
    w[30] = 0.0
    w[31] = 0.1
    w[36] = 0.0
    return w
