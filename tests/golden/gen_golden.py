"""Write the golden fixtures under tests/golden/ from the reference's own code.

Runs only where the reference tree exists (default /root/reference): it builds
oracle/_ref/libgrw_ref.so from the reference's unmodified sources, its MPI
layer and global integrator included (``make -C oracle``), and records inputs
and outputs of those compiled functions.  The fixtures are data (inputs and
expected outputs) and are committed; this script is how they were made.

    python tests/golden/gen_golden.py [--ref /root/reference] [--only NAME ...]
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle.pyoracle import MTN, Ref  # noqa: E402

A0_REF = 12350232  # initialize_ran.c:30


def gen_ic(r: Ref, ref_root: str):
    """ic/stable as read by the reference's initialize_state (1250 records)."""
    state, a, errw = r.initialize_state(ref_root, 1250, 10.0)
    np.savez_compressed(os.path.join(HERE, "ic_stable.npz"), state=state, fru=a["fru"],
                        ltype=a["ltype"], ryr=a["ryr"], ito2=a["ito2"], errweight=errw)
    return state, a


def gen_rng(r: Ref):
    mt, mti = r.initialize_ran(12500)
    seeds = mt[:, 0].astype(np.uint32)
    frus = np.array([0, 1, 721, 722, 1249, 12499])
    draws = np.zeros((len(frus), 1300), np.uint32)
    for k, j in enumerate(frus):
        m = mt[j].copy()
        i = mti[j:j + 1].copy()
        out = np.zeros(1300)
        import ctypes as C
        r.L.ref_mt_fast(i.ctypes.data_as(C.c_void_p), m.ctypes.data_as(C.c_void_p), 1300,
                        out.ctypes.data_as(C.c_void_p))
        draws[k] = np.round(out * 4294967296.0).astype(np.uint64).astype(np.uint32)
    np.savez_compressed(os.path.join(HERE, "rng_ref.npz"), a0=A0_REF, seeds=seeds, draw_frus=frus,
                        draws=draws)
    return mt, mti


def _fresh(n):
    return Ref.alloc(n)


def gen_rates(r: Ref, mt_all, rng):
    """fru_rates_local over random channel states, voltages and CaSS spanning the
    three RyR regimes (thresholds fru_rates_local.c:115-121)."""
    n = 400
    a = _fresh(n)
    a["ltype"][:, :, 0] = rng.integers(1, 13, (n, 4))
    a["ltype"][:, :, 1] = rng.integers(1, 3, (n, 4))
    a["ryr"][:] = rng.integers(1, 9, (n, 4, 5))
    a["ito2"][:] = rng.integers(1, 3, (n, 4))
    a["fru"][:, 0] = rng.uniform(0.2, 1.2, n)
    cass = np.choose(rng.integers(0, 5, (n, 4)),
                     [np.full((n, 4), 1e-5), np.full((n, 4), 1.2e-4), np.full((n, 4), 1e-3),
                      np.full((n, 4), 0.04), np.full((n, 4), 0.1)])
    cass *= rng.uniform(0.8, 1.25, (n, 4))
    a["fru"][:, 1:] = cass
    a["mt"][:] = mt_all[:n]
    a["mti"][:] = rng.integers(0, 625, n)
    V = rng.choice(np.arange(-100.0, 61.0, 10.0), n) + rng.uniform(-0.5, 0.5, n)
    inp = {k: v.copy() for k, v in a.items()}
    res = dict(max_rate=np.zeros(n), lcc_rate=np.zeros((n, 4, 4)), lcc_dest=np.zeros((n, 4, 4), np.int32),
               lcc_len=np.zeros((n, 4), np.int32), vinact_rate=np.zeros((n, 4)),
               ryr_rate=np.zeros((n, 4, 5, 4)), ryr_dest=np.zeros((n, 4, 5, 4), np.int32),
               ryr_len=np.zeros((n, 4, 5), np.int32), ito2_rate=np.zeros((n, 4)))
    for k in range(n):
        dep = np.array([V[k], 1.5e-4, 0.9])
        m, rr = r.rates(a, k, dep)
        res["max_rate"][k] = m
        for key, val in rr.items():
            res[key][k] = val
    np.savez_compressed(os.path.join(HERE, "rates_ref.npz"), V=V,
                        **{"in_" + k: v for k, v in inp.items() if k != "mt"},
                        in_seed_rows=np.arange(n), **res, out_ryr=a["ryr"], out_mti=a["mti"])


def gen_fcn_fru(r: Ref, rng):
    n = 400
    conc = np.column_stack([rng.uniform(0.1, 1.5, n), 10 ** rng.uniform(-5, -1, (n, 4))])
    V = rng.uniform(-100, 60, n)
    V[:10] = [0.0, 1e-7, -1e-7, 5e-7, 0.0, -47.13, 10.0, -90.0, 60.0, -100.0]
    dep = np.column_stack([V, 10 ** rng.uniform(-5, -3, n), rng.uniform(0.3, 1.5, n)])
    lopen = rng.integers(0, 2, (n, 4)).astype(np.float64)
    nopen = rng.integers(0, 6, (n, 4)).astype(np.int32)
    d = np.array([r.fcn_fru(conc[k], dep[k], lopen[k], nopen[k]) for k in range(n)])
    np.savez_compressed(os.path.join(HERE, "fcn_fru_ref.npz"), conc=conc, dep=dep, lopen=lopen,
                        nopen=nopen, d=d)


def gen_simfru(r: Ref, ic, mt_all, mti_all):
    """simfru_local on IC FRUs 0..7 under voltage clamp with frozen global state
    (V in {-80,-40,0,40} mV) and under a linear ramp, 5 ms in 0.05 ms sub-intervals.
    Snapshots every 10 sub-intervals; final MT rows."""
    nf = 8
    protos = [("clamp", -80.0, -80.0), ("clamp", -40.0, -40.0), ("clamp", 0.0, 0.0),
              ("clamp", 40.0, 40.0), ("ramp", -80.0, 40.0)]
    nsub, every = 100, 10
    P = len(protos)
    snaps = dict(fru=np.zeros((P, nf, nsub // every, 5)), ltype=np.zeros((P, nf, nsub // every, 4, 2), np.int32),
                 ryr=np.zeros((P, nf, nsub // every, 4, 5), np.int32),
                 ito2=np.zeros((P, nf, nsub // every, 4), np.int32), mti=np.zeros((P, nf, nsub // every), np.int32))
    mt_final = np.zeros((P, nf, MTN), np.uint32)
    for p, (_, v0, v1) in enumerate(protos):
        a = _fresh(nf)
        a["fru"][:] = ic["fru"][:nf]
        a["ltype"][:] = ic["ltype"][:nf]
        a["ryr"][:] = ic["ryr"][:nf]
        a["ito2"][:] = ic["ito2"][:nf]
        a["mt"][:] = mt_all[:nf]
        a["mti"][:] = mti_all[:nf]
        for s in range(nsub):
            t0, t1 = 0.05 * s, 0.05 * (s + 1)
            dep0 = [v0 + (v1 - v0) * s / nsub, 1.5e-4, 0.92]
            depf = [v0 + (v1 - v0) * (s + 1) / nsub, 1.6e-4, 0.92]
            for k in range(nf):
                r.simfru(t0, t1, dep0, depf, a, k)
            if (s + 1) % every == 0:
                q = (s + 1) // every - 1
                for key in ("fru", "ltype", "ryr", "ito2", "mti"):
                    snaps[key][p, :, q] = a[key]
        mt_final[p] = a["mt"][:, :MTN].astype(np.uint32)
    np.savez_compressed(os.path.join(HERE, "simfru_ref.npz"), nsub=nsub, every=every, dt_sub=0.05,
                        v0=np.array([p[1] for p in protos]), v1=np.array([p[2] for p in protos]),
                        cai0=1.5e-4, cai1=1.6e-4, nsr=0.92, mt_final=mt_final,
                        **{"snap_" + k: v for k, v in snaps.items()})


def gen_sums(r: Ref, ic, rng):
    a = _fresh(1250)
    for key in ("fru", "ltype", "ryr", "ito2"):
        a[key][:] = ic[key]
    flux = r.flux_sums(a)
    avg = r.avg_sums(a)
    # a scrambled variant that opens channels
    b = _fresh(1250)
    b["fru"][:] = ic["fru"] * rng.uniform(0.5, 2.0, (1250, 5))
    b["ltype"][:, :, 0] = rng.integers(1, 13, (1250, 4))
    b["ltype"][:, :, 1] = rng.integers(1, 3, (1250, 4))
    b["ryr"][:] = rng.integers(1, 9, (1250, 4, 5))
    b["ito2"][:] = rng.integers(1, 3, (1250, 4))
    flux_b = r.flux_sums(b)
    avg_b = r.avg_sums(b)
    np.savez_compressed(os.path.join(HERE, "sums_ref.npz"), flux=flux, avg=avg,
                        b_fru=b["fru"], b_ltype=b["ltype"], b_ryr=b["ryr"], b_ito2=b["ito2"],
                        flux_b=flux_b, avg_b=avg_b)


def seeds_o3(a0: int, n: int) -> np.ndarray:
    """initialize_ran.c:30-34 as gcc -O3 compiles it, for master seed a0 (the
    closed form the fixture rng_ref.npz pins for a0 = 12350232)."""
    j = np.arange(n, dtype=np.int64)
    return ((a0 + 23 * (j * (j + 1) // 2) + 4093 * j * j + 1) % 2 ** 32).astype(np.uint32)


def seeded_block(r: Ref, ic, n: int, a0: int):
    """ic/stable FRUs 0..n-1 with MT streams from sgrnd_local of the -O3 seeds."""
    a = _fresh(n)
    for key in ("fru", "ltype", "ryr", "ito2"):
        a[key][:] = ic[key][:n]
    for k, sd in enumerate(seeds_o3(a0, n)):
        mt, mti = r.sgrnd(int(sd))
        a["mt"][k] = mt
        a["mti"][k] = mti[0]
    return a


def gen_mpi_master(r: Ref, ic, state0, rng):
    """The master-side surface of mpi_master.c over one MPI rank: the flux
    conversion of send_calc_fru_flux (:41-113) and of distrib_simFRU (:401-468,
    zero-length and real sub-intervals), and send_calc_fru_avg's 12 outputs
    (:122-191), on the ic/stable block and a scrambled block of 1,000 FRUs (the
    reference's MAX_LOCAL_NFRU)."""
    n = 1000
    blocks = []
    a = _fresh(n)
    for key in ("fru", "ltype", "ryr", "ito2"):
        a[key][:] = ic[key][:n]
    blocks.append(a)
    b = _fresh(n)
    b["fru"][:] = ic["fru"][:n] * rng.uniform(0.5, 2.0, (n, 5))
    b["ltype"][:, :, 0] = rng.integers(1, 13, (n, 4))
    b["ltype"][:, :, 1] = rng.integers(1, 3, (n, 4))
    b["ryr"][:] = rng.integers(1, 9, (n, 4, 5))
    b["ito2"][:] = rng.integers(1, 3, (n, 4))
    blocks.append(b)
    # V on both sides of the |V| < 1e-6 Taylor branch (mpi_master.c:92,448)
    Vs = [-93.1, -40.0, -1.0e-6, -9.99e-7, -3.0e-7, 0.0, 1.0e-7, 9.99e-7, 1.0e-6, 1.01e-6,
          0.37, 20.0, 55.0]
    deps = np.array([[v, 1.4e-4 * (1 + 0.1 * k), 0.9 - 0.02 * k] for k, v in enumerate(Vs)])
    scales = [12.5, 1.0]
    flux = np.zeros((2, len(deps), 4))
    conv = np.zeros((2, len(deps), 4))
    for i, blk in enumerate(blocks):
        r.load_block(blk, scales[i])
        for k, d in enumerate(deps):
            flux[i, k] = r.send_calc_fru_flux(d)
            dep0 = d * [1.0, 0.9, 1.0] + [1.5, 0.0, 0.01]
            conv[i, k] = r.distrib_simfru(2.5, 2.5, dep0, d)
    states = state0 * rng.uniform(0.9, 1.1, (6, 37))
    avg = np.zeros((2, len(states), 12))
    for i, blk in enumerate(blocks):
        r.load_block(blk, scales[i])
        for k, st in enumerate(states):
            avg[i, k] = r.send_calc_fru_avg(st)
    # real sub-intervals through distrib_simFRU on 200 seeded FRUs
    m = 200
    c = seeded_block(r, ic, m, 1234)
    r.load_block(c, 12500.0 / m)
    sub = [(0.0, 0.05, -93.0, -92.0), (0.05, 0.2, -92.0, 1.0e-7), (0.2, 0.35, 1.0e-7, 30.0),
           (0.35, 0.5, 30.0, 10.0)]
    sub_out = np.zeros((len(sub), 4))
    for k, (t0, t1, v0, v1) in enumerate(sub):
        sub_out[k] = r.distrib_simfru(t0, t1, [v0, 1.5e-4, 0.92], [v1, 2.5e-4, 0.91])
    after = r.get_block(m)
    np.savez_compressed(
        os.path.join(HERE, "mpi_master_ref.npz"), scales=np.array(scales),
        b_fru=b["fru"], b_ltype=b["ltype"], b_ryr=b["ryr"], b_ito2=b["ito2"], deps=deps,
        flux=flux, conv=conv, states=states, avg=avg, sub=np.array(sub), sub_out=sub_out,
        sub_nfru=m, sub_a0=1234, after_fru=after["fru"], after_ltype=after["ltype"],
        after_ryr=after["ryr"], after_ito2=after["ito2"], after_mti=after["mti"])


def gen_rk54pd(r: Ref, ref_root: str, ms=25):
    """rk54pd (rk_pd54m.c:29-345) driving the reference's MPI layer over one
    rank: a 250-FRU cell (NFRU_scale 50) from ic/stable, master seed 1234, one
    call per 1 ms output step (caf.c:150-161) through the stimulus at 10 ms.
    Recorded: every call rk54pd makes (send_calc_fru_flux, send_save_state,
    distrib_simFRU with its sub-interval, send_resume_state) with its outputs,
    the returned step size, state and currents of each call, and the final FRU
    state."""
    n = 250
    state0, ic, errw = r.initialize_state(ref_root, n, 12500.0 / n)
    a = seeded_block(r, ic, n, 1234)
    r.load_block(a, 12500.0 / n)
    r.set_errweight(errw)
    state = np.array(state0, np.float64)
    cur = np.zeros(24)
    h = 0.1
    hs, states, curs, nev = [], [], [], []
    r.trace(200000)
    for k in range(ms):
        h = r.rk54pd(float(k), float(k + 1), h, state, cur)
        hs.append(h)
        states.append(state.copy())
        curs.append(cur.copy())
        nev.append(r.L.ref_trace_n())
    ev = r.events()
    r.trace(0)
    after = r.get_block(n)
    np.savez_compressed(
        os.path.join(HERE, "rk54pd_ref.npz"), nfru=n, a0=1234, ms=ms, h=np.array(hs),
        states=np.array(states), curs=np.array(curs), nev=np.array(nev), ev_kind=ev["kind"],
        ev_t=np.stack([ev["t0"], ev["t1"]], 1), ev_dep=ev["dep"], ev_out=ev["out"],
        after_fru=after["fru"], after_mti=after["mti"],
        after_chan=np.concatenate([after["ltype"].reshape(n, -1), after["ryr"].reshape(n, -1),
                                   after["ito2"]], 1))


def gen_fcn(r: Ref, state0, rng):
    n = 120
    st = state0 * rng.uniform(0.95, 1.05, (n, 37))
    st[:, 0] = rng.uniform(-100, 60, n)
    st[:6, 0] = [-47.13, 10.0, 0.0, 1e-7, -90.0, -95.0]
    t = rng.uniform(0, 3000, n)
    t[::7] = 10.0 + 1000.0 * (np.arange(len(t[::7])) % 3)
    J = rng.normal(size=(n, 4)) * [10, 0.1, 5, 0.5]
    F = np.zeros((n, 37))
    cur = np.zeros((n, 24))
    st_after = np.zeros((n, 37))
    for k in range(n):
        F[k], cur[k], st_after[k] = r.fcn(t[k], st[k], J[k])
    np.savez_compressed(os.path.join(HERE, "fcn_ref.npz"), t=t, state=st, J=J, F=F, cur=cur,
                        state_after=st_after)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    ap.add_argument("--only", nargs="*", default=None,
                    help="subset of: ic rng rates fcn_fru simfru sums fcn mpi_master rk54pd")
    args = ap.parse_args()
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), f"REF={args.ref}"], check=True)
    want = lambda name: args.only is None or name in args.only  # noqa: E731
    r = Ref()
    rng = np.random.default_rng(20021)
    state0, ic = gen_ic(r, args.ref)  # also the IC set the others start from
    if want("rng") or want("rates") or want("simfru"):
        mt_all, mti_all = gen_rng(r)
    if want("rates"):
        gen_rates(r, mt_all, rng)
    if want("fcn_fru"):
        gen_fcn_fru(r, rng)
    if want("simfru"):
        gen_simfru(r, ic, mt_all, mti_all)
    if want("sums"):
        gen_sums(r, ic, rng)
    if want("fcn"):
        gen_fcn(r, state0, rng)
    if want("mpi_master"):
        gen_mpi_master(r, ic, state0, np.random.default_rng(20022))
    if want("rk54pd"):
        gen_rk54pd(r, args.ref)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
