"""Output files (greenstein2002_amd/output.py) and RNG-state records (ic.py).

The row formats, the restart dump and the per-FRU property files are compared
byte for byte with the reference's own writers (data_io.c, write_fru_props.c,
compiled from the reference tree into oracle/_ref/libgrw_ref.so).  The
reference side runs in a child process: its restart_data closes r_states twice
(data_io.c:220,265), which glibc aborts on after r_states and r_FRU are
complete.  The other restart files are checked by reading them back with the
reference's initialize_state.  The caf.c output schedule is checked with a
stand-in cell on the CPU; tests/test_gpu_output.py runs it on a real cell."""
import os
import subprocess
import sys
import textwrap

import numpy as np
import pytest

from greenstein2002_amd import output as O
from greenstein2002_amd.ic import read_ic_dir, read_seeds

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libgrw_ref.so")
SEEDS = "/root/reference/ic/stable/ic_randomseeds_stable.txt"
needs_ref = pytest.mark.skipif(not os.path.exists(REF_LIB), reason="oracle/_ref not built (no reference tree)")

PRELUDE = textwrap.dedent(f"""
    import ctypes as C
    import numpy as np
    L = C.CDLL({REF_LIB!r})
    libc = C.CDLL(None)
    for name in ("open_states", "open_otherstates", "open_currents", "open_fru_props_avgtot"):
        getattr(L, name).restype = C.c_void_p
    libc.fclose.argtypes = [C.c_void_p]
    d = dict(np.load("in.npz"))
    vp = lambda a: a.ctypes.data_as(C.c_void_p)
""")


def run_ref(cwd, script, **arrays):
    np.savez(os.path.join(cwd, "in.npz"), **arrays)
    return subprocess.run([sys.executable, "-c", PRELUDE + textwrap.dedent(script)], cwd=cwd,
                          capture_output=True, text=True, timeout=120)


def tricky_values(rng, n):
    """Doubles over the whole range plus decimal rounding ties and specials."""
    v = rng.standard_normal(n) * 10.0 ** rng.uniform(-30, 30, n)
    special = [0.0, -0.0, 1.0, -1.0, 0.5, 2.5, 1e-5, 1e-4, 0.0001, 123456.5, 1234565.0, 1234575.0,
               123456785.0, 1e16, 5e-324, 1.7976931348623157e308, np.inf, -np.inf, np.nan,
               -np.nan, 2.0 ** -1074, 999999.5, 9999995.0, 0.30000000000000004]
    v[: len(special)] = special
    return v


@needs_ref
def test_rows_match_reference_writers(tmp_path):
    rng = np.random.default_rng(11)
    S = tricky_values(rng, 40 * 37).reshape(40, 37)
    Oth = tricky_values(rng, 40 * 12)[::-1].copy().reshape(40, 12)
    Cur = tricky_values(rng, 40 * 24).reshape(40, 24)
    times = np.concatenate([[0.0, 1.0, 1000.5, 12000.0, 0.1 + 0.2], rng.uniform(0, 1e5, 35)])
    ref = tmp_path / "ref"
    ref.mkdir()
    r = run_ref(ref, """
        for opener, writer, key in ((L.open_states, L.write_states, "S"),
                                    (L.open_otherstates, L.write_otherstates, "O"),
                                    (L.open_currents, L.write_currents, "Cu")):
            f = opener(7)
            for t, row in zip(d["t"], d[key]):
                row = np.ascontiguousarray(row)
                writer(C.c_void_p(f), C.c_double(t), vp(row))
            libc.fclose(f)
    """, S=S, O=Oth, Cu=Cur, t=times)
    assert r.returncode == 0, r.stderr
    ours = tmp_path / "ours"
    ours.mkdir()
    for opener, row, arr in ((O.open_states, O.states_row, S), (O.open_otherstates, O.otherstates_row, Oth),
                             (O.open_currents, O.currents_row, Cur)):
        with opener(str(ours) + "/", 7) as f:
            for t, x in zip(times, arr):
                f.write(row(t, x))
    for name in ("states.7.txt", "otherstates.7.txt", "currents.7.txt"):
        a, b = (ref / name).read_bytes(), (ours / name).read_bytes()
        assert a == b, name


def random_frus(rng, n):
    fru = rng.uniform(0.0, 1.5, (n, 5)) * 10.0 ** rng.integers(-4, 1, (n, 5))
    ltype = np.stack([rng.integers(1, 13, (n, 4)), rng.integers(1, 3, (n, 4))], axis=-1).astype(np.int32)
    ryr = rng.integers(1, 9, (n, 4, 5)).astype(np.int32)
    ito2 = rng.integers(1, 3, (n, 4)).astype(np.int32)
    mt = rng.integers(0, 2 ** 32, (n, 625), dtype=np.uint64)
    mti = rng.integers(0, 625, n).astype(np.int32)
    return fru, ltype, ryr, ito2, mt, mti


@needs_ref
def test_restart_dump_matches_reference(tmp_path):
    rng = np.random.default_rng(5)
    n = 23
    fru, ltype, ryr, ito2, mt, mti = random_frus(rng, n)
    state = tricky_values(rng, 37)
    state[16] = -np.nan
    ref = tmp_path / "ref"
    ref.mkdir()
    r = run_ref(ref, """
        n = len(d["fru"])
        L.ref_set_nfru(n, C.c_double(1.0))
        L.restart_data(3, vp(d["state"]), vp(d["fru"]), vp(d["lt"]), vp(d["ryr"]), vp(d["ito2"]),
                       vp(d["mt"]), vp(d["mti"]))
    """, state=state, fru=fru, lt=ltype, ryr=ryr, ito2=ito2, mt=mt, mti=mti)
    # the reference aborts in its second fclose of r_states (glibc double free)
    assert r.returncode in (0, -6), r.stderr
    ours = tmp_path / "ours"
    ours.mkdir()
    paths = O.restart_data(str(ours) + "/", 3, state, fru, ltype, ryr, ito2)
    assert [os.path.basename(p) for p in paths] == ["r_states.3.txt", "r_FRU.3.txt", "r_LCh.3.txt",
                                                    "r_RyR.3.txt", "r_Ito2.3.txt"]
    for name in ("r_states.3.txt", "r_FRU.3.txt"):
        assert (ref / name).read_bytes() == (ours / name).read_bytes(), name


@needs_ref
def test_restart_dump_reads_back_through_reference_reader(tmp_path):
    """The five restart files, renamed to ic/stable's names, read back by the
    reference's initialize_state give the dumped state at %g precision."""
    rng = np.random.default_rng(8)
    n = 40
    fru, ltype, ryr, ito2, mt, mti = random_frus(rng, n)
    fru[:, 1:] *= 1e-3
    state = np.array(read_stable_state(), np.float64)
    icdir = tmp_path / "ic" / "stable"
    icdir.mkdir(parents=True)
    O.restart_data(str(icdir) + "/", 0, state, fru, ltype, ryr, ito2)
    for stem in ("states", "FRU", "LCh", "RyR", "Ito2"):
        os.rename(icdir / f"r_{stem}.0.txt", icdir / f"ic_{stem}_stable.txt")
    r = run_ref(tmp_path, f"""
        n = {n}
        st = np.zeros(37); fru = np.zeros((n, 5)); lt = np.zeros((n, 4, 2), np.int32)
        ryr = np.zeros((n, 4, 5), np.int32); ito2 = np.zeros((n, 4), np.int32)
        mt = np.zeros((n, 625), np.uint64); mti = np.zeros(n, np.int32); ew = np.zeros(37)
        rc = L.ref_initialize_state({str(tmp_path)!r}.encode(), n, C.c_double(1.0), vp(st), vp(fru), vp(lt),
                                    vp(ryr), vp(ito2), vp(mt), vp(mti), vp(ew))
        assert rc == 0, rc
        np.savez("out.npz", st=st, fru=fru, lt=lt, ryr=ryr, ito2=ito2)
    """, dummy=np.zeros(1))
    assert r.returncode == 0, r.stderr
    got = np.load(tmp_path / "out.npz")
    g = np.vectorize(lambda x: float("%g" % x))
    np.testing.assert_array_equal(got["lt"], ltype)
    np.testing.assert_array_equal(got["ryr"], ryr)
    np.testing.assert_array_equal(got["ito2"], ito2)
    np.testing.assert_array_equal(got["fru"], g(fru))
    keep = np.arange(37) != 31  # CaTOT is recomputed on reading (initialize_state.c:233-246)
    np.testing.assert_array_equal(got["st"][keep], g(state)[keep])
    # and this package's reader agrees with the reference's
    mine = read_ic_dir(str(icdir), n)
    np.testing.assert_array_equal(mine["fru"], got["fru"])
    np.testing.assert_array_equal(mine["ltype"], got["lt"])
    np.testing.assert_array_equal(mine["ryr"], got["ryr"])
    np.testing.assert_array_equal(mine["state"][keep], got["st"][keep])


def read_stable_state():
    from greenstein2002_amd.ic import load_stable
    return load_stable(100)["state"]


@needs_ref
@pytest.mark.parametrize("V", [-80.0, -3.25, 0.4, 25.0])
def test_fru_props_match_reference(tmp_path, V):
    rng = np.random.default_rng(int(abs(V) * 100))
    n = 16
    fru, ltype, ryr, ito2, _, _ = random_frus(rng, n)
    ltype[:, :, 0] = rng.choice([6, 12, 6, 3], (n, 4))  # mostly open LCC states
    ref = tmp_path / "ref"
    ref.mkdir()
    r = run_ref(ref, f"""
        n = len(d["fru"])
        L.ref_set_nfru(n, C.c_double(1.0))
        per = (C.c_void_p * 200)(); avg = (C.c_void_p * 200)()
        L.open_fru_props(2, per); L.open_fru_props_avg(2, avg)
        tot = L.open_fru_props_avgtot(2)
        for t in (0.5, 17.0):
            L.write_fru_props(per, avg, C.c_void_p(tot), C.c_double(t), C.c_double({V!r}),
                              C.c_double(1.3e-4), C.c_double(0.9), vp(d["fru"]), vp(d["lt"]),
                              vp(d["ryr"]), vp(d["ito2"]))
        for j in range(10):
            libc.fclose(per[j]); libc.fclose(avg[j])
        libc.fclose(tot)
    """, fru=fru, lt=ltype, ryr=ryr, ito2=ito2)
    assert r.returncode == 0, r.stderr
    ours = tmp_path / "ours"
    ours.mkdir()
    fp = O.FruPropsFiles(str(ours) + "/", 2)
    for t in (0.5, 17.0):
        fp.write(t, V, 1.3e-4, 0.9, fru, ltype, ryr, ito2)
    fp.close()
    names = ([f"fruprops_{j}.2.txt" for j in range(10)] + [f"fruprops_avg_{j}.2.txt" for j in range(10)]
             + ["fruprops_avg_tot.2.txt"])
    for name in names:
        assert (ref / name).read_bytes() == (ours / name).read_bytes(), name


def test_fru_props_read_the_lcc_array_as_the_reference_indexes_it():
    """write_fru_props.c:90-91 reads LType_state[j][0][c] and [j][1][c]: over
    the FRU's cleft-major eight ints, cleft c's state is int c and its gate int
    2 + c.  Cleft 1 holds (6, 2) (ints 2, 3) and cleft 2 holds (2, 1) (ints 4,
    5), so the reference counts only cleft 2 as open (int 2 = 6, int 4 = 2),
    where the natural reading would count cleft 1."""
    fru = np.full((10, 5), 1e-4)
    fru[:, 0] = 1.0
    lt = np.ones((10, 4, 2), np.int32)
    lt[:, 1] = (6, 2)     # ints 2, 3
    lt[:, 2] = (2, 1)     # ints 4, 5
    ryr = np.ones((10, 4, 5), np.int32)
    ito2 = np.ones((10, 4), np.int32)
    rows, _, _ = O.fru_props_rows(1.0, 10.0, 1e-4, 1.0, fru, lt, ryr, ito2)
    jl = [float(line.split()[1]) for line in rows[0]]
    # cleft 2: state int 2 = 6, gate int 4 = 2 -> open; the others are closed
    assert jl[2] != 0.0 and jl[0] == jl[1] == jl[3] == 0.0
    # the natural reading (cleft 1 open) would have flagged cleft 1 instead
    assert jl[1] == 0.0


def test_seeds_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    mt = rng.integers(0, 2 ** 32, (9, 624), dtype=np.uint64)
    mti = rng.integers(0, 625, 9)
    for wb in (4, 8):
        p = str(tmp_path / f"s{wb}.bin")
        O.write_seeds(p, mti, mt, wb)
        assert os.path.getsize(p) == 9 * 626 * wb
        back = read_seeds(p, 9, wb)
        np.testing.assert_array_equal(back["mti"], mti)
        np.testing.assert_array_equal(back["mt"], mt.astype(np.uint32))
        with pytest.raises(ValueError, match="Unable to read seeds"):
            read_seeds(p, 10, wb)


@pytest.mark.skipif(not os.path.exists(SEEDS), reason="reference tree not present")
def test_shipped_seeds_file_reads_as_the_reference_reader_would():
    """ic/stable/ic_randomseeds_stable.txt holds 32-bit words: 125 records of
    626 words less 9 bytes.  No CR LF pair is left in it (a CR LF to LF
    conversion), and records 33, 34, 42, ... start one byte earlier each
    time, so a strict reader (initialize_state.c:150-158) accepts records 0-32
    and fails the index check at 33; with LP64 words it fails at record 0."""
    raw = open(SEEDS, "rb").read()
    assert len(raw) == 125 * 626 * 4 - 9 and b"\r\n" not in raw
    d = read_seeds(SEEDS, 33, 4)
    assert np.all(d["mti"] <= 624)
    with pytest.raises(ValueError, match="consistency check"):
        read_seeds(SEEDS, 34, 4)
    with pytest.raises(ValueError, match="consistency check"):
        read_seeds(SEEDS, 1, 8)


class _FakeFru:
    def __init__(self, cell):
        self.cell = cell

    def parallel_get_FRUs(self):
        n = 12
        k = self.cell.calls
        return dict(fru=np.full((n, 5), 1e-4 * (k + 1)), ltype=np.ones((n, 4, 2), np.int32),
                    ryr=np.ones((n, 4, 5), np.int32), ito2=np.ones((n, 4), np.int32),
                    mti=np.full(n, 624, np.int32), mt=np.zeros((n, 624), np.uint32))

    def send_calc_fru_flux(self, dep):
        return np.zeros(4)


class _FakeCell:
    """Stand-in for Cell: call k returns states k + 1, currents -(k + 1),
    other 100 (k + 1), so each written row names the call it came from."""

    def __init__(self):
        from greenstein2002_amd.cell import Protocol
        from greenstein2002_amd.ic import load_stable
        self.protocol = Protocol()
        assert self.protocol.time_step == 1.0
        self.calls = 0
        self.state = load_stable(12)["state"].copy()
        self.fru = _FakeFru(self)

    def run(self, n, log_cap=0, outputs=True, driver="device"):
        assert n == 1
        self.calls += 1
        k = self.calls

        class R:
            states = np.full((1, 37), float(k))
            currents = np.full((1, 24), -float(k))
            other = np.full((1, 12), 100.0 * k)
        return R


def _rows(path):
    with open(path) as f:
        return [line.split() for line in f.readlines()[1:]]


def test_caf_output_schedule(tmp_path):
    """caf.c:150-255 with time_end 7 and Ndump 3: dumps after calls 3 and 6
    (files 0 and 1), states / otherstates of calls 1-3, 4-6 and 7 in files
    0, 1, 2.  Currents: calls 2, 3 into currents.0 (call 1 is at time_start);
    call 4, the first after a dump, opens currents.1 instead of writing; calls
    5, 6 into currents.1; call 7 opens currents.2; lastcall's row at time_end
    goes to currents.2."""
    cell = _FakeCell()
    d = str(tmp_path) + "/"
    res = O.run_caf(cell, 7.0, output_dir=d, ndump=3, fru_props=True, fru_end=3)
    assert res.n_steps == 7 and res.dumps == [0, 1] and res.filenumber == 2
    for fn, calls in ((0, [1, 2, 3]), (1, [4, 5, 6]), (2, [7])):
        st = _rows(f"{d}states.{fn}.txt")
        assert [float(r[0]) for r in st] == [k * 1e-3 for k in calls]
        assert [float(r[1]) for r in st] == [float(k) for k in calls]
        ot = _rows(f"{d}otherstates.{fn}.txt")
        assert [float(r[1]) for r in ot] == [100.0 * k for k in calls]
        fp = _rows(f"{d}fruprops_avg_tot.{fn}.txt")
        assert [float(r[0]) for r in fp] == [k * 1e-3 for k in calls]
    files = sorted(os.listdir(tmp_path))
    assert "r_states.0.txt" in files and "r_states.1.txt" in files and "r_states.2.txt" not in files
    c2 = _rows(f"{d}currents.2.txt")
    # currents.2 opened at call 7 (no row), then lastcall's row at time_end 7
    assert [float(r[0]) for r in c2] == [7e-3]
    c1 = _rows(f"{d}currents.1.txt")
    assert [(float(r[0]), float(r[1])) for r in c1] == [(4e-3, -5.0), (5e-3, -6.0)]
    c0 = _rows(f"{d}currents.0.txt")
    assert [(float(r[0]), float(r[1])) for r in c0] == [(1e-3, -2.0), (2e-3, -3.0)]
