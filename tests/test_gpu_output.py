"""run_caf on a real cell: the files it writes hold exactly the rows a twin
cell (same seed, same FRUs) returns from Cell.run, in the caf.c:150-255
schedule, the restart dump holds that twin's FRU state after the dump's step,
and the last currents row is lastcall's (fcn.c:1077-1092)."""
import os

import numpy as np
import pytest

from greenstein2002_amd import output as O
from greenstein2002_amd.cell import Cell, fcn

pytestmark = pytest.mark.gpu


def _lines(path):
    with open(path) as f:
        return f.readlines()


def test_run_caf_files_match_a_twin_cell(tmp_path):
    print("run_caf: 250 FRUs, 6 ms, Ndump 4", flush=True)
    d = str(tmp_path) + "/"
    with Cell(250, seed=1234) as a:
        res = O.run_caf(a, 6.0, output_dir=d, ndump=4, fru_props=True)
    assert res.n_steps == 6 and res.dumps == [0] and res.filenumber == 1
    with Cell(250, seed=1234) as b:
        r1 = b.run(4, log_cap=0)
        g = b.fru.parallel_get_FRUs()
        twin = tmp_path / "twin"
        twin.mkdir()
        O.restart_data(str(twin) + "/", 0, r1.states[3], g["fru"], g["ltype"], g["ryr"], g["ito2"])
        r2 = b.run(2, log_cap=0)
        st = np.array(b.state)
        J = b.fru.send_calc_fru_flux(st[[0, 6, 7]])
        _, cur_end, _ = fcn(6.0, st, J, b.protocol)
    S = np.concatenate([r1.states, r2.states])
    Ot = np.concatenate([r1.other, r2.other])
    Cu = np.concatenate([r1.currents, r2.currents])
    t = np.arange(1, 7, dtype=np.float64)
    assert _lines(d + "states.0.txt")[1:] == [O.states_row(t[k], S[k]) for k in range(4)]
    assert _lines(d + "states.1.txt")[1:] == [O.states_row(t[k], S[k]) for k in (4, 5)]
    assert _lines(d + "otherstates.0.txt")[1:] == [O.otherstates_row(t[k], Ot[k]) for k in range(4)]
    assert _lines(d + "otherstates.1.txt")[1:] == [O.otherstates_row(t[k], Ot[k]) for k in (4, 5)]
    # calls 2-4 write their start currents under time_now = 1..3; call 5, the
    # first after the dump, opens currents.1; call 6 writes under 5; lastcall at 6
    assert _lines(d + "currents.0.txt")[1:] == [O.currents_row(t[k] - 1.0, Cu[k]) for k in (1, 2, 3)]
    assert _lines(d + "currents.1.txt")[1:] == [O.currents_row(5.0, Cu[5]), O.currents_row(6.0, cur_end)]
    for stem in ("r_states", "r_FRU", "r_LCh", "r_RyR", "r_Ito2"):
        assert _lines(f"{d}{stem}.0.txt") == _lines(f"{twin}/{stem}.0.txt"), stem
    for j in range(10):
        assert len(_lines(f"{d}fruprops_{j}.0.txt")) == 1 + 4 * 4
        assert len(_lines(f"{d}fruprops_avg_{j}.1.txt")) == 1 + 2
    assert not os.path.exists(d + "r_states.1.txt")
