"""GPU parity of the FRU hot path against the CPU oracle and the golden fixtures.

Every test drives the HIP library through its C ABI (FruCell) and compares with
oracle/libgrw_oracle.so (the restatement pinned to the reference) on the same
seeded inputs.  Bar: channel states, RNG words and positions bit-exact;
concentrations within 1e-12 relative (the device exp() may differ from glibc's
in the last ulp, everything else follows the reference's operation order).
"""
import numpy as np
import pytest

from greenstein2002_amd import FruCell
from greenstein2002_amd.ic import load_stable

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def orc_frus(o, ic, a0, nfru):
    from oracle.pyoracle import FRU_DTYPE

    f = np.zeros(nfru, FRU_DTYPE)
    f["conc"] = ic["fru"][:nfru]
    f["lcc"] = ic["ltype"][:nfru, :, 0]
    f["yinact"] = ic["ltype"][:nfru, :, 1]
    f["ryr"] = ic["ryr"][:nfru]
    f["ito2"] = ic["ito2"][:nfru]
    for j in range(nfru):
        m = o.mt_seed(o.seed_for_fru(a0, j))
        f["mt"][j] = m["mt"]
        f["mti"][j] = m["mti"]
    return f


def load_cell(cell, f):
    lt = np.stack([f["lcc"], f["yinact"]], axis=-1)
    cell.initialize_mpi_state(f["conc"], lt, f["ryr"], f["ito2"], f["mti"], f["mt"])


def assert_same(got, f, what=""):
    assert np.array_equal(got["ltype"][:, :, 0], f["lcc"]), f"LCC states differ {what}"
    assert np.array_equal(got["ltype"][:, :, 1], f["yinact"]), f"gate states differ {what}"
    assert np.array_equal(got["ryr"], f["ryr"]), f"RyR states differ {what}"
    assert np.array_equal(got["ito2"], f["ito2"]), f"Ito2 states differ {what}"
    assert np.array_equal(got["mti"], f["mti"]), f"RNG positions differ {what}"
    assert np.array_equal(got["mt"], f["mt"]), f"RNG words differ {what}"
    np.testing.assert_allclose(got["fru"], f["conc"], rtol=RTOL, atol=0, err_msg=what)


def test_device_seeding_matches_reference(golden, oracle):
    """grw_seed reproduces initialize_ran + sgrnd_local for 12,500 FRUs."""
    g = golden("rng_ref.npz")
    n = 12500
    with FruCell(n) as cell:
        cell.seed(int(g["a0"]))
        st = cell.parallel_get_FRUs()
    assert np.array_equal(st["mt"][:, 0], g["seeds"])
    assert np.all(st["mti"] == 624)
    for j in (0, 1, 721, 722, 12499):
        m = oracle.mt_seed(oracle.seed_for_fru(int(g["a0"]), j))
        assert np.array_equal(st["mt"][j], m["mt"])


@pytest.mark.parametrize("lanes", [4, 1])
def test_simfru_golden_clamp_and_ramp(golden, oracle, lanes):
    """Single-FRU event sequences under voltage clamp with frozen global state and
    under a ramp: the reference's own simfru_local outputs (fixture); both kernel
    layouts (4 lanes per FRU and 1 lane per FRU)."""
    g = golden("simfru_ref.npz")
    ic = load_stable(1250)
    nf = g["snap_fru"].shape[1]
    nsub, every, dts = int(g["nsub"]), int(g["every"]), float(g["dt_sub"])
    for p in range(len(g["v0"])):
        f = orc_frus(oracle, ic, 12350232, nf)
        with FruCell(nf, nfru_scale=1.0, lanes_per_fru=lanes) as cell:
            load_cell(cell, f)
            v0, v1 = float(g["v0"][p]), float(g["v1"][p])
            for s in range(nsub):
                dep0 = [v0 + (v1 - v0) * s / nsub, float(g["cai0"]), float(g["nsr"])]
                depf = [v0 + (v1 - v0) * (s + 1) / nsub, float(g["cai1"]), float(g["nsr"])]
                cell.simfru_local(dts * s, dts * (s + 1), dep0, depf)
                if (s + 1) % every == 0:
                    q = (s + 1) // every - 1
                    st = cell.parallel_get_FRUs()
                    assert np.array_equal(st["ltype"], g["snap_ltype"][p, :, q]), (p, q)
                    assert np.array_equal(st["ryr"], g["snap_ryr"][p, :, q]), (p, q)
                    assert np.array_equal(st["ito2"], g["snap_ito2"][p, :, q]), (p, q)
                    assert np.array_equal(st["mti"], g["snap_mti"][p, :, q]), (p, q)
                    np.testing.assert_allclose(st["fru"], g["snap_fru"][p, :, q], rtol=RTOL, atol=0)
            assert np.array_equal(cell.parallel_get_FRUs()["mt"], g["mt_final"][p])


@pytest.mark.parametrize("seq,lanes", [(False, 4), (True, 4), (False, 1), (True, 1)])
def test_distrib_matches_oracle(oracle, seq, lanes):
    """4,096 FRUs through an action-potential-like sequence of sub-intervals."""
    n = 4096
    ic = load_stable(n)
    rng = np.random.default_rng(7)
    f = orc_frus(oracle, ic, 1234, n)
    # spread the channel states so every LCC/RyR branch is exercised
    f["lcc"][: n // 2] = rng.integers(1, 13, (n // 2, 4))
    f["ryr"][: n // 2] = rng.integers(1, 9, (n // 2, 4, 5))
    f["conc"][: n // 4, 1:] *= 10 ** rng.uniform(0, 2.5, (n // 4, 4))
    f["mti"][:64] = rng.integers(0, 625, 64)
    volt = [-85.0, -60.0, 0.0, 30.0, 45.0, 20.0, -10.0, -40.0]
    with FruCell(n, sequential_sums=seq, lanes_per_fru=lanes) as cell:
        load_cell(cell, f)
        t = 0.0
        for k in range(len(volt) - 1):
            dep0 = np.array([volt[k], 1.5e-4 * (1 + k), 0.92])
            depf = np.array([volt[k + 1], 1.5e-4 * (2 + k), 0.91])
            out = cell.distrib_simFRU(t, t + 0.05, dep0, depf)
            for j in range(n):
                oracle.simfru(t, t + 0.05, dep0, depf, f, j)
            num = oracle.flux_sums(f)
            ref = oracle.flux_currents(num, depf, n, 12500.0 / n)
            if seq:
                np.testing.assert_allclose(out, ref, rtol=1e-14, atol=0)
            else:
                np.testing.assert_allclose(out, ref, rtol=1e-9, atol=1e-12)
            t += 0.05
        assert_same(cell.parallel_get_FRUs(), f, f"seq={seq}")
        sums = cell.calc_fru_flux_local()
        np.testing.assert_allclose(sums, oracle.flux_sums(f), rtol=2e-12 if not seq else 0, atol=0)
        avg = cell.calc_fru_avg_local()
        np.testing.assert_allclose(avg, oracle.avg_sums(f), rtol=1e-12 if not seq else 0, atol=1e-15)


@pytest.mark.parametrize("lanes", [4, 1])
def test_save_resume_replays_identically(oracle, lanes):
    """send_save_state / send_resume_state: rollback restores FRU and RNG state
    bit for bit, and the replayed Monte Carlo repeats the same events."""
    n = 1024
    ic = load_stable(n)
    f = orc_frus(oracle, ic, 99, n)
    dep0 = np.array([-20.0, 3e-4, 0.9])
    depf = np.array([30.0, 6e-4, 0.9])
    with FruCell(n, lanes_per_fru=lanes) as cell:
        load_cell(cell, f)
        cell.distrib_simFRU(0.0, 0.1, dep0, dep0)
        cell.send_save_state()
        saved = cell.parallel_get_FRUs()
        # long enough that every FRU regenerates more than 624 words
        for k in range(40):
            cell.distrib_simFRU(0.1 + 0.1 * k, 0.2 + 0.1 * k, dep0, depf)
        first = cell.parallel_get_FRUs()
        cell.send_resume_state()
        back = cell.parallel_get_FRUs()
        for key in saved:
            assert np.array_equal(back[key], saved[key]), key
        for k in range(40):
            cell.distrib_simFRU(0.1 + 0.1 * k, 0.2 + 0.1 * k, dep0, depf)
        again = cell.parallel_get_FRUs()
        for key in first:
            assert np.array_equal(again[key], first[key]), key
        # a short rollback inside one generation
        cell.send_save_state()
        s2 = cell.parallel_get_FRUs()
        cell.distrib_simFRU(5.0, 5.02, dep0, depf)
        cell.send_resume_state()
        b2 = cell.parallel_get_FRUs()
        for key in s2:
            assert np.array_equal(b2[key], s2[key]), key


@pytest.mark.parametrize("lanes", [4, 1])
def test_state_roundtrip_all_rng_positions(oracle, lanes):
    """load/get convert the reference's (mt, mti) to the device's lazy stream and
    back for every position 0..624, including after regeneration."""
    n = 625
    ic = load_stable(n)
    f = orc_frus(oracle, ic, 5, n)
    f["mti"] = np.arange(625)
    with FruCell(n, lanes_per_fru=lanes) as cell:
        load_cell(cell, f)
        st = cell.parallel_get_FRUs()
        assert np.array_equal(st["mti"], f["mti"])
        assert np.array_equal(st["mt"], f["mt"])
        dep = np.array([10.0, 2e-4, 0.9])
        # rollback from every checkpoint position, including positions inside a
        # loaded generation (no word regenerated yet at the checkpoint)
        cell.send_save_state()
        cell.simfru_local(0.0, 0.3, dep, dep)
        cell.send_resume_state()
        back = cell.parallel_get_FRUs()
        assert np.array_equal(back["mti"], f["mti"])
        assert np.array_equal(back["mt"], f["mt"])
        cell.simfru_local(0.0, 0.3, dep, dep)
        for j in range(n):
            oracle.simfru(0.0, 0.3, dep, dep, f, j)
        assert_same(cell.parallel_get_FRUs(), f, "after 0.3 ms")


def test_invalid_state_rejected():
    n = 64
    ic = load_stable(n)
    lt = ic["ltype"].copy()
    lt[3, 1, 0] = 13
    with FruCell(n) as cell:
        with pytest.raises(Exception, match="invalid"):
            cell.initialize_mpi_state(ic["fru"], lt, ic["ryr"], ic["ito2"], np.full(n, 624),
                                      np.zeros((n, 624), np.uint32))


def test_mixed_precision_runs_close(oracle):
    """The mixed form keeps FP64 state/time/thresholds but its FP32 step lengths
    shift the micro-step grid, so it follows the FP64 path only statistically:
    compare the ensemble sums after 2 ms at rest."""
    n = 4096
    ic = load_stable(n)
    f = orc_frus(oracle, ic, 1234, n)
    dep0 = np.array([-85.0, 1.5e-4, 0.92])
    with FruCell(n, precision="fp64") as a, FruCell(n, precision="mixed") as b:
        load_cell(a, f)
        load_cell(b, f)
        for k in range(20):
            oa = a.distrib_simFRU(0.1 * k, 0.1 * (k + 1), dep0, dep0)
            ob = b.distrib_simFRU(0.1 * k, 0.1 * (k + 1), dep0, dep0)
        sa, sb = a.calc_fru_flux_local(), b.calc_fru_flux_local()
        st = b.parallel_get_FRUs()
    assert np.all(np.isfinite(st["fru"]))
    np.testing.assert_allclose(sb[:2], sa[:2], rtol=2e-3)
    assert np.all(np.isfinite(ob))
