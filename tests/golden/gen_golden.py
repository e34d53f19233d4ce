"""Write the golden fixtures under tests/golden/ from the reference's own code.

Runs only where the reference tree exists (default /root/reference): it builds
oracle/_ref/libgrw_ref.so from the reference's unmodified hot-path sources
(``make -C oracle``) and records inputs and outputs of those compiled
functions.  The fixtures are data (inputs and expected outputs) and are
committed; this script is how they were made.

    python tests/golden/gen_golden.py [--ref /root/reference]
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
    deps = np.array([[-93.1, 1.4e-4, 0.9], [0.0, 1.4e-4, 0.9], [1e-7, 1.4e-4, 0.9], [20.0, 5e-4, 0.7]])
    conv = np.array([r.send_calc_fru_flux(b, d, 10.0) for d in deps])
    np.savez_compressed(os.path.join(HERE, "sums_ref.npz"), flux=flux, avg=avg,
                        b_fru=b["fru"], b_ltype=b["ltype"], b_ryr=b["ryr"], b_ito2=b["ito2"],
                        flux_b=flux_b, avg_b=avg_b, deps=deps, conv=conv, scale=10.0)


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
    args = ap.parse_args()
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), f"REF={args.ref}"], check=True)
    r = Ref()
    rng = np.random.default_rng(20021)
    state0, ic = gen_ic(r, args.ref)
    mt_all, mti_all = gen_rng(r)
    gen_rates(r, mt_all, rng)
    gen_fcn_fru(r, rng)
    gen_simfru(r, ic, mt_all, mti_all)
    gen_sums(r, ic, rng)
    gen_fcn(r, state0, rng)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
