"""Ensemble fixtures: per-beat statistics of whole-cell runs for master seeds
1234..1249, reduced by greenstein2002_amd.beats.beat_stats from every accepted
RK step and the 1 ms samples, exactly as the device integrator reduces them.

    python tests/golden/gen_ensemble.py c1     # 250 FRUs, 2 beats  (~2 min per seed)
    python tests/golden/gen_ensemble.py c2     # 12,500 FRUs, 1 beat (~25 min per seed)

c1 runs the reference itself: its rk54pd (rk_pd54m.c) over its own MPI layer
(mpi_master.c, mpi_slave.c, one rank) and hot path, oracle/_ref/libgrw_ref.so,
called once per 1 ms output step as caf.c:150-169 does, with the accepted
steps read from the trace of the calls rk54pd makes.  c2 needs 12,500 FRUs,
beyond the reference's compiled array bound (MAX_LOCAL_NFRU = 1000,
parameters.h:50-51), so it runs the CPU oracle, which tests/test_oracle.py pins
to that reference driver call by call.

Writes tests/golden/ensemble_<name>.npz: seeds, nfru, ms, beats [seeds, nbeats, 12],
and for c1 the 1 ms traces of V, Cai, ICa (the FRU-flux ICa of each sample's
last accepted step) and JRyR (send_calc_fru_avg's JRyRtot, mpi_master.c:177).
"""
from __future__ import annotations

import os
import sys
import time
from multiprocessing import Pool

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

CONFIGS = {"c1": dict(nfru=250, ms=2011, nbeats=2), "c2": dict(nfru=12500, ms=1011, nbeats=1)}
SEEDS = list(range(1234, 1250))


def seeded_ic(n, seed):
    from greenstein2002_amd.ic import load_stable
    from oracle.pyoracle import FRU_DTYPE, Oracle

    o = Oracle()
    ic = load_stable(n)
    f = np.zeros(n, FRU_DTYPE)
    f["conc"] = ic["fru"]
    f["lcc"] = ic["ltype"][:, :, 0]
    f["yinact"] = ic["ltype"][:, :, 1]
    f["ryr"] = ic["ryr"]
    f["ito2"] = ic["ito2"]
    for j in range(n):
        m = o.mt_seed(o.seed_for_fru(seed, j))
        f["mt"][j] = m["mt"]
        f["mti"][j] = m["mti"]
    return ic, f


def accepted_steps(ev, state_after):
    """(t, V, Cai, ICa) of every accepted step in one rk54pd call's trace: a
    group of five sub-intervals not followed by a rollback.  t and ICa are those
    of the group's last sub-interval (rk_pd54m.c:208-213,248); V and Cai those of
    the accepted state, which the next flux call receives (:89-97), or the
    call's final state for its last step."""
    from oracle.pyoracle import EV_DISTRIB, EV_FLUX, EV_RESUME

    out = []
    kinds = ev["kind"]
    d = np.flatnonzero(kinds == EV_DISTRIB)
    for last in d[4::5]:
        nxt = kinds[last + 1] if last + 1 < len(ev) else None
        if nxt == EV_RESUME:
            continue
        if nxt == EV_FLUX:
            V, Cai = ev["dep"][last + 1][0], ev["dep"][last + 1][1]
        else:
            V, Cai = state_after[0], state_after[6]
        out.append((ev["t1"][last], V, Cai, ev["out"][last][2]))
    return np.array(out).reshape(-1, 4)


def run_seed(args):
    name, seed = args
    from greenstein2002_amd.beats import beat_stats

    cfg = CONFIGS[name]
    n, ms = cfg["nfru"], cfg["ms"]
    ic, f = seeded_ic(n, seed)
    state = np.array(ic["state"])
    cur = np.zeros(24)
    h = 0.1
    logs, states, other = [], np.zeros((ms, 37)), np.zeros((ms, 12))
    t0 = time.time()
    if name == "c1":
        from oracle.pyoracle import Oracle, Ref

        r = Ref()
        r.load_block(Ref.from_orc(f), 12500.0 / n)
        r.set_errweight(Oracle().errweight())
        for k in range(ms):
            r.trace(400000)
            h = r.rk54pd(float(k), float(k + 1), h, state, cur)
            logs.append(accepted_steps(r.events(), state))
            states[k] = state
            other[k] = r.send_calc_fru_avg(state)
    else:
        from oracle.pyoracle import Oracle, OracleCell

        cell = OracleCell(Oracle(), f, 12500.0 / n)
        for k in range(ms):
            h, log = cell.rk54pd(float(k), float(k + 1), h, state, cur, log_cap=200000)
            logs.append(log)
            states[k] = state
            other[k] = cell.avg(state)
        cell.close()
    log = np.concatenate(logs)
    rows_t = np.arange(1.0, ms + 1.0)
    beats = beat_stats(log, rows_t, states, other, 10.0, 1000.0)[: cfg["nbeats"]]
    # ICa of each 1 ms sample: that of the last accepted step ending at or before it
    idx = np.searchsorted(log[:, 0], rows_t, side="right") - 1
    ica = np.where(idx >= 0, log[np.maximum(idx, 0), 3], np.nan)
    print(f"{name} seed {seed}: {time.time() - t0:.0f} s, APD90 {beats[:, 4]}", flush=True)
    return seed, beats, states[:, 0].copy(), states[:, 6].copy(), ica, other[:, 2].copy()


def main():
    name = sys.argv[1]
    workers = int(sys.argv[2]) if len(sys.argv) > 2 else 6
    cfg = CONFIGS[name]
    with Pool(workers) as pool:
        res = sorted(pool.map(run_seed, [(name, s) for s in SEEDS], chunksize=1))
    out = dict(seeds=np.array(SEEDS), nfru=cfg["nfru"], ms=cfg["ms"],
               beats=np.array([r[1] for r in res]))
    if name == "c1":
        out.update(V=np.array([r[2] for r in res], np.float32), Cai=np.array([r[3] for r in res], np.float32),
                   ICa=np.array([r[4] for r in res], np.float32),
                   JRyR=np.array([r[5] for r in res], np.float32))
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), f"ensemble_{name}.npz"), **out)


if __name__ == "__main__":
    main()
