"""The device integrator with its RK attempts replayed from a captured HIP
graph (grw_set_graphs) against the same runs launched attempt by attempt.

The graph holds the same kernel launches with the same arguments, in the same
order on the same stream, so every output must be identical bit for bit.  Each
run is split into pieces, so the graph captured in the first piece is replayed
in later calls, with the host's own synchronisations and downloads between the
replays, and every output millisecond of every replay is compared.
"""
import numpy as np
import pytest

from greenstein2002_amd.batch import Batch, CellSpec
from greenstein2002_amd.cell import Cell, Protocol

pytestmark = pytest.mark.gpu


def cell_run(graphs: bool, precision: str, timing: bool):
    n, seed = 512, 2468
    with Cell(n, seed=seed, precision=precision) as c:
        c.fru.set_graphs(graphs)
        c.fru.set_timing(timing)
        parts = [c.run(ms, driver="device", log_cap=4000) for ms in (9, 7, 11)]
        beats = c.beats()
        st = c.fru.stats()
    return parts, beats, st


@pytest.mark.parametrize("precision,timing", [("fp64", True), ("fp64", False), ("mixed", True)])
def test_cell_graph_replays_equal_direct_launches(precision, timing):
    ref, bref, _ = cell_run(False, precision, timing)
    got, bgot, st = cell_run(True, precision, timing)
    for k, (a, b) in enumerate(zip(ref, got)):
        assert a.stats["attempts"] == b.stats["attempts"], f"piece {k}"
        assert a.stats["rejects"] == b.stats["rejects"], f"piece {k}"
        assert np.array_equal(a.states, b.states), f"piece {k}: states differ"
        assert np.array_equal(a.currents, b.currents), f"piece {k}: currents differ"
        assert np.array_equal(a.other, b.other), f"piece {k}: averages differ"
        assert np.array_equal(a.log, b.log), f"piece {k}: step logs differ"
    assert np.array_equal(bref, bgot, equal_nan=True)
    assert got[1].states[:, 0].max() > 0.0  # the stimulus at 10 ms fired in the second piece
    if timing:
        assert st["timed_launches"] > 0 and st["kernel_ms"] > 0.0
    # every FRU-step sits in a wavefront slot that ran as long as its slowest FRU
    assert 0 < st["fru_steps"] <= st["slot_steps"] <= 16 * st["fru_steps"]


def test_batch_graph_replays_equal_direct_launches():
    """Uneven cells: the regrouping and block-ordering kernels are in the graph."""
    specs = [
        CellSpec(250, seed=21),
        CellSpec(700, seed=22, k23_scale=1.4, protocol=Protocol(period=300.0, shift=5.0)),
        CellSpec(96, seed=23, gamma_scale=0.7),
    ]

    def run(graphs):
        with Batch(specs) as b:
            b.set_graphs(graphs)
            b.run(8)
            s1 = b.state.copy()
            b.run([6, 0, 6])
            b.run([0, 6, 0])
            return s1, b.state.copy(), b.attempts.copy(), b.rejects.copy(), \
                [b.beats(k) for k in range(len(specs))]

    ref, got = run(False), run(True)
    assert np.array_equal(ref[0], got[0])
    assert np.array_equal(ref[1], got[1])
    assert np.array_equal(ref[2], got[2]) and np.array_equal(ref[3], got[3])
    for a, b in zip(ref[4], got[4]):
        assert np.array_equal(a, b, equal_nan=True)
