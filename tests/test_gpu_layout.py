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
    with Cell(64) as cell:
        assert L.grw_set_layout(cell.fru._h, 3) != 0
