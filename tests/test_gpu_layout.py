"""The two Monte Carlo kernels of the device integrator give the same numbers.

GRW_LAYOUT_WAVES (grw_set_layout) splits every micro-step of a FRU over the
four wavefronts of a block -- LCC, RyRs 0-1, RyRs 2-4, Heun step -- and groups
its flux sums into the partials of the four-lane kernel's blocks, so a run with
either kernel must give identical states, currents, averages, accepted-step
logs, attempt counts, beat statistics and FRU state, bit for bit (the MT
streams included).
"""
import ctypes as C

import numpy as np
import pytest

from greenstein2002_amd import _lib
from greenstein2002_amd.batch import Batch, CellSpec
from greenstein2002_amd.cell import Cell, Protocol

pytestmark = pytest.mark.gpu


def _run(n, seed, precision, layout, ms, graphs, pieces=1, prot=None):
    with Cell(n, seed=seed, precision=precision, protocol=prot) as cell:
        cell.set_layout(layout)
        cell.fru.set_graphs(graphs)
        parts = [cell.run(ms // pieces, driver="device") for _ in range(pieces)]
        fru = cell.fru.parallel_get_FRUs()
        beats = cell.beats()
    cat = lambda k: np.concatenate([getattr(p, k) for p in parts])
    stats = [p.stats for p in parts]
    return dict(states=cat("states"), currents=cat("currents"), other=cat("other"), log=cat("log"),
                stats=stats, fru=fru, beats=beats)


def _assert_same(a, b):
    for k in ("states", "currents", "other", "log"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    assert a["stats"] == b["stats"]
    for k in a["fru"]:
        np.testing.assert_array_equal(np.asarray(a["fru"][k]), np.asarray(b["fru"][k]), err_msg=k)
    np.testing.assert_array_equal(a["beats"], b["beats"])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_wave_layout_equals_quad_layout(precision):
    """512 FRUs over 30 ms (stimulus at 10 ms: upstroke and Ca release),
    graph replays on."""
    n, seed, ms = 512, 2024, 30
    print(f"layout parity {precision} n={n} seed={seed} ms={ms}", flush=True)
    q = _run(n, seed, precision, "quad", ms, graphs=True)
    w = _run(n, seed, precision, "waves", ms, graphs=True)
    assert q["states"][:, 0].max() > 0.0  # an action potential fired
    _assert_same(q, w)


def test_wave_layout_with_partial_blocks_and_pieces():
    """A FRU count that leaves the last 16-FRU and 64-FRU blocks part empty,
    graphs off, the run split in three pieces."""
    n, seed, ms = 500, 7, 24
    print(f"layout parity pieces n={n} seed={seed} ms={ms}", flush=True)
    q = _run(n, seed, "fp64", "quad", ms, graphs=False, pieces=3)
    w = _run(n, seed, "fp64", "waves", ms, graphs=False, pieces=3)
    _assert_same(q, w)


def test_wave_layout_under_voltage_clamp():
    """Clamp steps: large LCC and RyR activity from the first step."""
    prot = Protocol(vclamp=True, vclamp_set=0.0, vclamp_hold=-100.0, vclamp_shift=2.0, vclamp_duration=10.0,
                    vclamp_period=1000.0)
    n, seed, ms = 256, 99, 16
    print(f"layout parity vclamp n={n} seed={seed} ms={ms}", flush=True)
    q = _run(n, seed, "fp64", "quad", ms, graphs=True, prot=prot)
    w = _run(n, seed, "fp64", "waves", ms, graphs=True, prot=prot)
    _assert_same(q, w)


def test_switching_layout_mid_run_continues_the_same_trajectory():
    n, seed = 320, 5
    print(f"layout switch n={n} seed={seed}", flush=True)
    with Cell(n, seed=seed) as a, Cell(n, seed=seed) as b:
        a.fru.set_graphs(True)
        b.fru.set_graphs(True)
        ra = [a.run(10), a.run(10)]
        b.set_layout("waves")
        rb = [b.run(10)]
        b.set_layout("quad")
        rb.append(b.run(10))
    np.testing.assert_array_equal(np.concatenate([r.states for r in ra]), np.concatenate([r.states for r in rb]))


def test_wave_layout_is_refused_for_batches():
    L = _lib.load()
    with Batch([CellSpec(nfru=64, seed=1), CellSpec(nfru=64, seed=2)]) as bt:
        assert L.grw_set_layout(bt._h, 2) != 0
        assert L.grw_set_layout(bt._h, 1) == 0
        assert L.grw_set_layout(bt._h, 4) == 0
        assert L.grw_set_layout(bt._h, 5) != 0  # the queue layout is switched off
        assert L.grw_set_layout(bt._h, 8) != 0  # the polling layout runs single cells only
    with Cell(64) as cell:
        assert L.grw_set_layout(cell.fru._h, 12) != 0
        assert L.grw_set_layout(cell.fru._h, 0) != 0


QUEUE_OFF = pytest.mark.skip(reason="the queue layout is switched off: its one device run (384 FRUs, FP64, "
                                     "20 ms) did not finish and the cause is not known")


@pytest.mark.parametrize("layout,precision", [("quad3", "fp64"), ("quad4", "fp64"), ("quad3", "mixed"),
                                              ("wave1", "fp64"), ("wave1", "mixed"), ("wave1_3", "fp64"),
                                              ("wave1_3", "mixed"), ("quadp", "fp64"), ("quadp", "mixed"),
                                              ("wave1_4", "mixed"), ("refill", "fp64"), ("refill", "mixed"),
                                              ("refill_3", "fp64"), ("refill_3", "mixed"),
                                              pytest.param("queue", "fp64", marks=QUEUE_OFF),
                                              pytest.param("queue", "mixed", marks=QUEUE_OFF)])
def test_other_layouts_equal_quad(layout, precision):
    n, seed, ms = 384, 31, 20
    print(f"layout parity {layout} {precision} n={n} seed={seed} ms={ms}", flush=True)
    q = _run(n, seed, precision, "quad", ms, graphs=True)
    w = _run(n, seed, precision, layout, ms, graphs=True)
    _assert_same(q, w)


def _batch_specs():
    return [CellSpec(nfru=n, seed=100 + k, protocol=Protocol(period=cl))
            for k, (n, cl) in enumerate([(300, 400.0), (130, 1000.0), (517, 700.0), (64, 300.0)])]


def _batch_run(layout=None, tune=False, ms=14):
    with Batch(_batch_specs()) as bt:
        bt.set_graphs(True)
        info = None
        if layout:
            bt.set_layout(layout)
        if tune:
            info = bt.autotune(2)
        st = bt.run([ms] * 4)
        beats = [bt.beats(k) for k in range(4)]
    return st, beats, info


def test_batch_layouts_and_autotune_give_the_same_run():
    """A batch of four uneven cells: the register-capped layouts and a run
    after the autotuner (which times every candidate from a snapshot and
    restores it) end in the same states and beats as the default."""
    print("batch layouts + autotune", flush=True)
    ref, rb, _ = _batch_run()
    for lay in ("quad3", "quad4", "wave1", "wave1_3", "wave1_4", "refill", "refill_3"):
        st, b, _ = _batch_run(layout=lay)
        np.testing.assert_array_equal(st, ref, err_msg=lay)
        for x, y in zip(b, rb):
            np.testing.assert_array_equal(x, y)
    st, b, info = _batch_run(tune=True)
    assert info["layout"] in ("quad", "quad3", "quad4", "wave1", "wave1_3", "wave1_4", "refill", "refill_3")
    assert len(info["candidates"]) == 16 and all(c["ms"] > 0 for c in info["candidates"])
    np.testing.assert_array_equal(st, ref)
    for x, y in zip(b, rb):
        np.testing.assert_array_equal(x, y)


def test_cell_autotune_leaves_the_cell_as_it_was():
    n, seed = 512, 77
    print(f"cell autotune n={n} seed={seed}", flush=True)
    with Cell(n, seed=seed) as a, Cell(n, seed=seed) as b:
        ra0 = a.run(6)
        rb0 = b.run(6)
        info = b.autotune(trial_steps=4)
        assert info["layout"] in ("quad", "waves", "quad3", "wave1", "quadp")
        assert len(info["candidates"]) == 10 and all(c["ms"] > 0 for c in info["candidates"])
        assert b.layout == info["layout"]
        ra = a.run(14)
        rb = b.run(14)
        fa, fb = a.fru.parallel_get_FRUs(), b.fru.parallel_get_FRUs()
        ba, bb = a.beats(), b.beats()
    np.testing.assert_array_equal(ra0.states, rb0.states)
    for k in ("states", "currents", "other", "log"):
        np.testing.assert_array_equal(getattr(ra, k), getattr(rb, k), err_msg=k)
    assert ra.stats == rb.stats
    for k in fa:
        np.testing.assert_array_equal(np.asarray(fa[k]), np.asarray(fb[k]), err_msg=k)
    np.testing.assert_array_equal(ba, bb)
