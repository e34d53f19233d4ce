"""The benchmark configurations at their own shapes, each cut to a short window.

* voltage clamp of 12,500-FRU cells, mixed precision against FP64 (three test
  potentials x two seeds, hold -80 mV, step at 5 ms for 200 ms, t = 0..45 ms);
* mixed precision against FP64 on a 16-seed ensemble of whole 12,500-FRU beats
  (APD90 within 1%, peak Cai within 2%);
* 256 cells with FRU counts uniform in [1250, 12500] and cycle lengths uniform
  in [300, 2000] ms in one batch, a sample of them against one-cell batches;
* the parameter sweep: 1,024 cells x 12,500 FRUs in one mixed batch with
  k23 / gamma scales log-uniform in [0.5, 2] and SERCA uniform in [0.7, 1.3],
  and an FP64 rerun of the first cells.

The mixed form follows the FP64 path statistically, not event by event (its
FP32 step lengths move the micro-step grid), so mixed-vs-FP64 bounds are on
means over seeds; each bound is stated in the test.
"""
import sys
import time

import numpy as np
import pytest

from greenstein2002_amd.batch import Batch, CellSpec
from greenstein2002_amd.cell import Cell, Protocol

pytestmark = pytest.mark.gpu

NFRU = 12500


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_batch(specs, ms, precision, chunk=10):
    t0 = time.perf_counter()
    with Batch(specs, precision=precision) as b:
        done = 0
        while done < ms:
            m = min(chunk, ms - done)
            b.run(m)
            done += m
        state = b.state.copy()
        beats = [b.beats(k) for k in range(len(specs))]
        att, rej = b.attempts.copy(), b.rejects.copy()
    log(f"  {len(specs)} cells {precision}: {ms} ms in {time.perf_counter() - t0:.1f} s")
    return state, beats, att, rej


def clamp_specs(vts, seeds, shift=5.0, duration=200.0):
    specs = []
    for vt in vts:
        for sd in seeds:
            prot = Protocol(vclamp=True, vclamp_hold=-80.0, vclamp_set=float(vt), vclamp_shift=shift,
                            vclamp_duration=duration, vclamp_period=1000.0)
            specs.append(CellSpec(NFRU, seed=int(sd), protocol=prot))
    return specs


def test_c3_mixed_clamp_against_fp64():
    """12,500-FRU cells clamped from -80 mV to -30, 0 and +30 mV at 5 ms, two
    seeds each, 40 ms into the step (peak ICa, the LCC Ca entry and the RyR
    release of the step's first 40 ms).  Per test potential, the mixed means
    over the seeds against the FP64 means: Ca entry within 2%, Ca released and
    the EC-coupling gain (released / entered) within 4%, peak ICa within 15%
    (the peak of a few LCC openings at -30 mV, a noisy maximum); V is held
    exactly in both."""
    vts, seeds, ms = (-30.0, 0.0, 30.0), (1234, 1235), 45
    specs = clamp_specs(vts, seeds)
    res = {}
    for prec in ("fp64", "mixed"):
        state, beats, _, _ = run_batch(specs, ms, prec)
        assert np.all(np.isfinite(state))
        assert np.array_equal(state[:, 0], np.repeat(vts, len(seeds))), "V not held at the test potential"
        res[prec] = np.array([bt[0] for bt in beats]).reshape(len(vts), len(seeds), -1)
    for i, vt in enumerate(vts):
        f, m = res["fp64"][i], res["mixed"][i]
        assert np.all(f[:, 12] > 0) and np.all(f[:, 13] > 0) and np.all(f[:, 8] < 0), vt
        rel = lambda c: abs(m[:, c].mean() - f[:, c].mean()) / abs(f[:, c].mean())  # noqa: E731
        gain_f, gain_m = (f[:, 13] / f[:, 12]).mean(), (m[:, 13] / m[:, 12]).mean()
        errs = dict(entry=rel(12), release=rel(13), gain=abs(gain_m - gain_f) / gain_f, ica=rel(8))
        log(f"  Vt {vt:+.0f} mV: gain fp64 {gain_f:.3f} mixed {gain_m:.3f}; rel errors "
            + ", ".join(f"{k} {v:.4f}" for k, v in errs.items()))
        assert errs["entry"] < 0.02 and errs["release"] < 0.04 and errs["gain"] < 0.04, (vt, errs)
        assert errs["ica"] < 0.15, (vt, errs)


def test_c3_fp64_clamp_follows_oracle(oracle):
    """The FP64 clamp against the CPU oracle's whole cell under the same clamp
    (250 FRUs, -80 -> +10 mV at 5 ms, 15 ms): states within 1e-6 relative."""
    from greenstein2002_amd.ic import load_stable
    from oracle.pyoracle import FRU_DTYPE, OracleCell

    n, seed, ms = 250, 77, 15
    prot = Protocol(vclamp=True, vclamp_shift=5.0, vclamp_duration=200.0, vclamp_set=10.0,
                    vclamp_hold=-80.0, vclamp_period=1000.0)
    ic = load_stable(n)
    f = np.zeros(n, FRU_DTYPE)
    f["conc"] = ic["fru"]
    f["lcc"] = ic["ltype"][:, :, 0]
    f["yinact"] = ic["ltype"][:, :, 1]
    f["ryr"] = ic["ryr"]
    f["ito2"] = ic["ito2"]
    for j in range(n):
        m = oracle.mt_seed(oracle.seed_for_fru(seed, j))
        f["mt"][j] = m["mt"]
        f["mti"][j] = m["mti"]
    oprot = oracle.protocol(vclamp=1, vclamp_shift=5.0, vclamp_duration=200.0, vclamp_set=10.0,
                            vclamp_hold=-80.0, vclamp_period=1000.0)
    cell = OracleCell(oracle, f, 12500.0 / n, oprot)
    state = np.array(ic["state"], np.float64)
    state[0] = -80.0  # initialize_state.c:226-228 under voltage clamp
    cur = np.zeros(24)
    h = 0.1
    ref = np.zeros((ms, 37))
    for k in range(ms):
        h, _ = cell.rk54pd(float(k), float(k + 1), h, state, cur)
        ref[k] = state
    cell.close()
    with Cell(n, seed=seed, protocol=prot) as c:
        res = c.run(ms, driver="device")
    assert np.array_equal(res.states[:, 0], ref[:, 0])
    np.testing.assert_allclose(res.states, ref, rtol=1e-6, atol=1e-12)


def test_mixed_beat_statistics_against_fp64_ensemble():
    """16 seeds x 12,500 FRUs, one 1 Hz beat (stimulus at 10 ms, t = 0..420 ms),
    FP64 and mixed batches: ensemble-mean APD90 of the mixed form within 1% and
    mean peak Cai within 2% of FP64 (the tolerances the mixed form is held to)."""
    seeds = np.arange(1234, 1250)
    specs = [CellSpec(NFRU, seed=int(s)) for s in seeds]
    stats = {}
    for prec in ("fp64", "mixed"):
        state, beats, _, _ = run_batch(specs, 420, prec, chunk=60)
        assert np.all(np.isfinite(state))
        stats[prec] = np.array([bt[0] for bt in beats])
    f, m = stats["fp64"], stats["mixed"]
    assert np.all(np.isfinite(f[:, 4])) and np.all(np.isfinite(m[:, 4])), "a cell did not repolarise"
    apd = (m[:, 4].mean() - f[:, 4].mean()) / f[:, 4].mean()
    cai = (m[:, 5].mean() - f[:, 5].mean()) / f[:, 5].mean()
    log(f"  APD90 fp64 {f[:, 4].mean():.2f} mixed {m[:, 4].mean():.2f} ms ({apd:+.4%}); "
        f"peak Cai fp64 {f[:, 5].mean() * 1e3:.4f} mixed {m[:, 5].mean() * 1e3:.4f} uM ({cai:+.4%})")
    assert abs(apd) < 0.01
    assert abs(cai) < 0.02


def test_c4_256_uneven_cells_in_one_batch():
    """256 cells, FRU counts uniform in [1250, 12500] and cycle lengths uniform in
    [300, 2000] ms (host PRNG 1234), seeds 1234.., 14 ms through the stimulus at
    10 ms in one batch: every cell finite and fired, and a sample of eight cells
    bit-equal (state, attempts, rejections, beat records) to each run alone as a
    one-cell batch."""
    n, ms = 256, 14
    rng = np.random.default_rng(1234)
    nf = rng.integers(1250, 12501, n)
    cl = rng.uniform(300.0, 2000.0, n)
    specs = [CellSpec(int(nf[k]), seed=1234 + k, protocol=Protocol(period=float(cl[k]))) for k in range(n)]
    state, beats, att, rej = run_batch(specs, ms, "fp64", chunk=7)
    assert np.all(np.isfinite(state))
    assert all(bt[0][2] > 20.0 for bt in beats), "a cell did not fire"
    for k in np.linspace(0, n - 1, 8).astype(int):
        with Batch([specs[k]]) as b:
            b.run(7)
            b.run(7)
            assert np.array_equal(b.state[0], state[k]), f"cell {k} ({nf[k]} FRUs) differs alone"
            assert b.attempts[0] == att[k] and b.rejects[0] == rej[k], k
            assert np.array_equal(b.beats(0), beats[k], equal_nan=True), k


def test_c5_sweep_1024_cells_mixed_with_fp64_rerun():
    """1,024 cells x 12,500 FRUs (12.8 M FRUs, 33 GB of MT rows) in one mixed
    batch with the sweep's k23 / gamma / SERCA scales, 12 ms through the
    stimulus; then the first 16 cells again in FP64.  Every cell finite and
    fired; over the 16, the mean peak V of the mixed form within 1% of FP64 and
    the mean Cai at 12 ms within 5% (the upstroke's Ca entry)."""
    n, ms, m64 = 1024, 12, 16
    rng = np.random.default_rng(1234)
    k23 = np.exp(rng.uniform(np.log(0.5), np.log(2.0), n))
    gam = np.exp(rng.uniform(np.log(0.5), np.log(2.0), n))
    serca = rng.uniform(0.7, 1.3, n)
    specs = [CellSpec(NFRU, seed=1234 + k, k23_scale=float(k23[k]), gamma_scale=float(gam[k]),
                      protocol=Protocol(serca_scale=float(serca[k]))) for k in range(n)]
    state, beats, _, _ = run_batch(specs, ms, "mixed", chunk=6)
    assert np.all(np.isfinite(state))
    vpk = np.array([bt[0][2] for bt in beats])
    assert np.all(vpk > 20.0), "a cell did not fire"
    s64, b64, _, _ = run_batch(specs[:m64], ms, "fp64", chunk=6)
    v64 = np.array([bt[0][2] for bt in b64])
    dv = (vpk[:m64].mean() - v64.mean()) / abs(v64.mean())
    dc = (state[:m64, 6].mean() - s64[:, 6].mean()) / s64[:, 6].mean()
    log(f"  peak V mixed - fp64 {dv:+.4%}, Cai(12 ms) {dc:+.4%}")
    assert abs(dv) < 0.01 and abs(dc) < 0.05
