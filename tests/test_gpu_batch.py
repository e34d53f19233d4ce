"""Batches of independent cells (include/grw_batch.h) and the per-beat
statistics of the device integrator.

A cell in a batch runs the same kernels on the same block structure whatever
the other cells do (its FRUs start at a multiple of 64, its reductions span
only its own blocks, and its FRU order between attempts depends on its own
step counts alone), so its trajectory must equal the same cell run alone as a
one-cell batch bit for bit.  A single-cell context keeps its FRUs in index
order, so against it only the summation order of the flux sums differs.
"""
import numpy as np
import pytest

from greenstein2002_amd.batch import Batch, CellSpec
from greenstein2002_amd.beats import NSTAT, beat_stats
from greenstein2002_amd.cell import Cell, Protocol

pytestmark = pytest.mark.gpu


def single_run(sp: CellSpec, ms: int, precision="fp64"):
    with Cell(sp.nfru, seed=sp.seed, precision=precision, protocol=sp.protocol,
              k23_scale=sp.k23_scale, gamma_scale=sp.gamma_scale, nfru_scale=sp.nfru_scale) as c:
        r = c.run(ms, log_cap=0)
        return r.states[-1], r.stats["attempts"], r.stats["rejects"], c.beats()


def alone_run(sp: CellSpec, ms: int, precision="fp64"):
    with Batch([sp], precision=precision) as b:
        b.run(ms)
        return b.state[0].copy(), int(b.attempts[0]), int(b.rejects[0]), b.beats(0)


def test_batch_cells_equal_single_cell_runs():
    """Uneven cells with their own seeds, pacing, SERCA, RyR and LCC scales."""
    specs = [
        CellSpec(250, seed=11),
        CellSpec(300, seed=12, k23_scale=1.6, protocol=Protocol(period=300.0, shift=5.0)),
        CellSpec(64, seed=13, gamma_scale=0.6, protocol=Protocol(serca_scale=0.8)),
        CellSpec(513, seed=14, nfru_scale=30.0, protocol=Protocol(shift=2.0)),
    ]
    ms = 22
    with Batch(specs) as b:
        b.run(10)
        b.run([12, 12, 0, 12])   # cell 2 idles, then catches up alone
        b.run([0, 0, 12, 0])
        got = b.state.copy()
        att = b.attempts.copy()
        rej = b.rejects.copy()
        beats = [b.beats(k) for k in range(len(specs))]
    for k, sp in enumerate(specs):
        st, a, r, bt = alone_run(sp, ms)
        assert np.array_equal(got[k], st), f"cell {k} state differs"
        assert att[k] == a and rej[k] == r, f"cell {k} attempts {att[k]}/{a} rejects {rej[k]}/{r}"
        assert np.array_equal(beats[k], bt, equal_nan=True), f"cell {k} beat records differ"
        st1, a1, r1, bt1 = single_run(sp, ms)
        assert att[k] == a1 and rej[k] == r1, f"cell {k} vs single context: attempts {att[k]}/{a1}"
        np.testing.assert_allclose(got[k], st1, rtol=1e-8, atol=1e-13)
        np.testing.assert_allclose(beats[k], bt1, rtol=1e-7, atol=1e-12)


def test_mixed_batch_runs():
    specs = [CellSpec(200, seed=s) for s in range(3)]
    with Batch(specs, precision="mixed") as b:
        st = b.run(15)
    assert np.all(np.isfinite(st)) and np.all(st[:, 0] > 0.0)  # all three fired at 10 ms


def test_device_beat_records_match_host_reduction():
    """The online per-beat statistics equal beat_stats() of the same run's
    accepted-step log and 1 ms samples (three short beats, run in chunks)."""
    prot = Protocol(period=120.0, shift=10.0)
    with Cell(250, seed=5, protocol=prot) as c:
        parts = [c.run(n, log_cap=200000) for n in (50, 130, 200)]
        dev = c.beats()
    log = np.concatenate([p.log for p in parts])
    t = np.concatenate([p.t for p in parts])
    states = np.concatenate([p.states for p in parts])
    other = np.concatenate([p.other for p in parts])
    host = beat_stats(log, t, states, other, prot.shift, prot.period)
    assert dev.shape == (4, NSTAT)
    assert np.array_equal(dev[:, :13], host[:, :13], equal_nan=True)  # host log has no release flux
    # release flux: the peak over accepted steps bounds the 1 ms samples of the
    # same whole-cell JRyR (send_calc_fru_avg output 2) from above, and the
    # integral is close to the trapezoid over those samples
    for k in range(3):
        m = (t >= dev[k, 0]) & (t < dev[k, 0] + prot.period)
        assert dev[k, 14] >= other[m, 2].max() * (1.0 - 1e-9)
        approx = np.trapezoid(other[m, 2], t[m])
        assert abs(dev[k, 13] - approx) < 0.1 * abs(approx)
        assert dev[k, 12] > 0.0 and dev[k, 13] > dev[k, 12]  # gain above 1
    # three full beats with an action potential each; the first one, started from
    # the 1 Hz steady state, outlasts the 120 ms cycle, the paced ones repolarise
    assert np.all(dev[:3, 2] > 0.0)
    assert np.all((dev[1:3, 4] > 20.0) & (dev[1:3, 4] < 120.0))
    assert np.all(dev[:3, 8] < 0.0)  # inward ICa


def test_ensemble_matches_cpu_oracle():
    """16 seeds x 250 FRUs x 2 beats (1 Hz) in one batch against the reference's
    own runs on the same seeds (tests/golden/ensemble_c1.npz: its rk54pd over
    its MPI layer, reduced by the same beat_stats).  The FP64 path follows the oracle operation for operation
    (concentrations within an ulp of exp), so per seed the statistics agree
    closely; the ensemble means of APD90, peak V, peak Cai, peak ICa and the
    SR / JSR loads agree within 0.5%."""
    from pathlib import Path

    g = np.load(Path(__file__).parent / "golden" / "ensemble_c1.npz")
    seeds, ms = g["seeds"], int(g["ms"])
    ref = g["beats"]
    specs = [CellSpec(int(g["nfru"]), seed=int(s)) for s in seeds]
    with Batch(specs) as b:
        b.run(ms)
        got = np.array([b.beats(k)[: ref.shape[1], : ref.shape[2]] for k in range(len(seeds))])
    assert got.shape == ref.shape
    cols = [2, 4, 5, 8, 9, 10]  # V_peak, APD90, Cai_peak, ICa_peak, CaNSR, CaJSR
    close = np.isclose(got[:, :, cols], ref[:, :, cols], rtol=1e-6, atol=0)
    assert close.all(axis=(1, 2)).sum() >= len(seeds) - 2, close.all(axis=(1, 2))
    m_got, m_ref = got[:, :, cols].mean(axis=0), ref[:, :, cols].mean(axis=0)
    np.testing.assert_allclose(m_got, m_ref, rtol=5e-3)


def test_c1_output_traces_match_reference():
    """The 1 ms output traces of the 250-FRU shape: V, Cai, the FRU-flux ICa of
    each sample's last accepted step, and the whole-cell JRyR of
    send_calc_fru_avg (mpi_master.c:177), for master seeds 1234 and 1235 over
    the first beat (t = 1..300 ms), against the reference's own runs
    (tests/golden/ensemble_c1.npz, written by the reference's rk54pd over its
    MPI layer; stored as float32, so rtol 1e-5)."""
    from pathlib import Path

    g = np.load(Path(__file__).parent / "golden" / "ensemble_c1.npz")
    ms = 300
    for k in (0, 1):
        seed = int(g["seeds"][k])
        with Cell(int(g["nfru"]), seed=seed) as c:
            r = c.run(ms, log_cap=400000)
        idx = np.searchsorted(r.log[:, 0], r.t, side="right") - 1
        ica = r.log[np.maximum(idx, 0), 3]
        for name, got in (("V", r.states[:, 0]), ("Cai", r.states[:, 6]), ("ICa", ica),
                          ("JRyR", r.other[:, 2])):
            ref = g[name][k, :ms].astype(np.float64)
            np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-5 * np.abs(ref).max(),
                                       err_msg=f"seed {seed} {name}")


def test_full_size_beat_matches_cpu_oracle():
    """The benchmark configuration itself: 12,500 FRUs, one 1 Hz beat from
    ic/stable (t = 0..1011 ms), six master seeds in one batch, against the CPU
    oracle's runs on the same seeds (tests/golden/ensemble_c2.npz, made by
    gen_ensemble.py c2 and reduced by the same beat_stats).  The FP64 path
    consumes every seed's random streams exactly as the oracle does, so each
    beat's statistics agree to the round-off of exp and of the flux-sum order
    (within 2e-10 relative on all 16 seeds of the fixture); rtol 1e-7 here."""
    from pathlib import Path

    g = np.load(Path(__file__).parent / "golden" / "ensemble_c2.npz")
    seeds, ms = g["seeds"][:6], int(g["ms"])
    ref = g["beats"][:6]
    specs = [CellSpec(int(g["nfru"]), seed=int(s)) for s in seeds]
    with Batch(specs) as b:
        b.run(ms)
        got = np.array([b.beats(k)[: ref.shape[1], : ref.shape[2]] for k in range(len(seeds))])
    assert got.shape == ref.shape
    cols = [0, 1, 2, 4, 5, 7, 8, 9, 10]  # t_stim, V_rest, V_peak, APD90, Cai_peak, Cai_rest, ICa_peak, loads
    np.testing.assert_allclose(got[:, :, cols], ref[:, :, cols], rtol=1e-7, atol=0)
    assert np.array_equal(got[:, :, 11], ref[:, :, 11])  # the same number of accepted steps


def test_batch_calls_reject_a_single_cell_context():
    """grw_batch_load_ic / grw_batch_run on a context from grw_create fail with
    GRW_EINVAL and say why, instead of running with an empty protocol."""
    import ctypes as C

    from greenstein2002_amd import _lib
    from greenstein2002_amd.batch import _declare

    L = _lib.load()
    _declare(L)
    with Cell(64, seed=1) as c:
        ns = np.ones(1, np.int32)
        rc = L.grw_batch_run(c.fru._h, ns.ctypes.data_as(C.c_void_p), None, None)
        assert rc == -1 and b"batch context" in L.grw_last_error()
        z = np.zeros(64)
        rc = L.grw_batch_load_ic(c.fru._h, z.ctypes.data_as(C.c_void_p), z.ctypes.data_as(C.c_void_p), 1,
                                 z.ctypes.data_as(C.c_void_p), z.ctypes.data_as(C.c_void_p),
                                 z.ctypes.data_as(C.c_void_p), z.ctypes.data_as(C.c_void_p))
        assert rc == -1 and b"batch context" in L.grw_last_error()


def test_piecewise_run_equals_unbroken_run_with_fractional_time_step():
    """time_step = 0.1 ms: a run in pieces of 7, 13 and 20 output steps equals
    one unbroken run of 40 bit for bit (the clock adds time_step per step)."""
    prot = Protocol(time_step=0.1, shift=2.0)
    with Cell(128, seed=21, protocol=prot) as a:
        parts = [a.run(n, log_cap=0) for n in (7, 13, 20)]
        ta = a.t
    with Cell(128, seed=21, protocol=prot) as b:
        whole = b.run(40, log_cap=0)
        tb = b.t
    assert ta == tb
    assert np.array_equal(np.concatenate([p.t for p in parts]), whole.t)
    assert np.array_equal(np.concatenate([p.states for p in parts]), whole.states)

