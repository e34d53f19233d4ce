"""dynamicFRU (dynamicFRU.c:26-152): the phase decision against the
reference's own compiled routine, and the redistribution rules -- CPU only.
The whole-cell runs with switching are in tests/test_gpu_cell.py."""
import os

import numpy as np
import pytest

from greenstein2002_amd.cell import Protocol
from greenstein2002_amd.dynamic import HIGH, LOW, average_down, phase_decision, replicate_up

REF_BUILT = os.path.exists(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                        "oracle", "_ref", "libgrw_ref.so"))


def _cases():
    rng = np.random.default_rng(17)
    cases = []
    for _ in range(3000):
        t = float(rng.choice([rng.uniform(0, 3000), 1000 * rng.integers(0, 3) + rng.choice(
            [9.0, 10.0, 10.5, 59.9, 60.0, 60.1, 909.9, 910.0, 910.1, 911.0])]))
        t_last = float(t - rng.choice([0.5, 1.0, 1.0000001, 2.0, rng.uniform(0, 50)]))
        V = float(rng.choice([-92.0, -80.0, -79.9, -80.1, 20.0]))
        pr = float(rng.choice([0.0, 1.9e-3, 2e-3, 5e-3, 6e-3, 6.1e-3]))
        pl = float(rng.choice([0.0, 0.9e-6, 1e-6, 1.5e-6, 2e-6, 2.1e-6]))
        cases.append((t, V, pr, pl, int(rng.integers(0, 2)), t_last))
    return cases


@pytest.mark.skipif(not REF_BUILT, reason="oracle/_ref not built (needs the reference tree)")
def test_phase_decision_matches_reference_dynamicFRU():
    """3,000 inputs around every threshold of dynamicFRU.c:49-105 (the 1 ms
    hold-off, 100 ms before and 50 ms after the stimulus, V -80 mV, the open
    fractions): the same phase, time of last change and time to stimulus as the
    reference's compiled dynamicFRU (factor 1, no FRU moved)."""
    from oracle.pyoracle import Ref

    r = Ref()
    switched = {LOW: 0, HIGH: 0}
    for t, V, pr, pl, ph, tl in _cases():
        ph_r, tl_r, tts_r = r.dynamic_fru(t, V, pr, pl, ph, tl)
        ph_o, tl_o, changed, tts = phase_decision(t, V, pr, pl, ph, tl, Protocol())
        assert (ph_o, tl_o) == (ph_r, tl_r), (t, V, pr, pl, ph, tl)
        if tts is not None:
            assert tts == tts_r
        switched[ph] += changed
    assert switched[LOW] > 100 and switched[HIGH] > 10  # both directions exercised


def test_redistribution_conserves_fru_calcium_and_streams():
    rng = np.random.default_rng(2)
    n_low, f = 50, 4
    n_high = n_low * f
    high = dict(fru=rng.uniform(1e-4, 1.0, (n_high, 5)), ltype=rng.integers(1, 13, (n_high, 4, 2)),
                ryr=rng.integers(1, 9, (n_high, 4, 5)), ito2=rng.integers(1, 3, (n_high, 4)),
                mti=rng.integers(0, 625, n_high), mt=rng.integers(0, 2**32, (n_high, 624), dtype=np.uint64))
    low_conc = average_down(high["fru"], n_low)
    # whole-cell FRU Ca: sum over FRUs x NFRU_scale (12500 / n) is unchanged
    np.testing.assert_allclose(low_conc.sum(0) * (12500 / n_low), high["fru"].sum(0) * (12500 / n_high),
                               rtol=1e-13)
    low = {k: v[:n_low].copy() for k, v in high.items()}
    low["fru"] = low_conc
    up = replicate_up(low, high, n_low)
    for i in (0, 7, n_low, n_low + 7, n_high - 1):
        j = i % n_low
        assert np.array_equal(up["fru"][i], low["fru"][j]) and np.array_equal(up["ryr"][i], low["ryr"][j])
        src = low if i < n_low else high
        assert np.array_equal(up["mt"][i], src["mt"][i]) and up["mti"][i] == src["mti"][i]
