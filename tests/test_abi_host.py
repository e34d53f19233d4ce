"""CPU-side checks of the native library: it loads, exports every entry point
declared in include/*.h, rejects bad input without a GPU, and its host-side
global model (grw_fcn) equals the reference's fcn.c on the golden fixture."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    """grw_* functions of every header in include/ (grw_mpi.h declares the
    reference's own names, checked against libgrw_mpi.so below)."""
    import glob

    names = set()
    for path in sorted(glob.glob(os.path.join(ROOT, "include", "*.h"))):
        text = open(path).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(grw_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    from greenstein2002_amd import _lib

    L = _lib.load()
    names = declared_functions()
    assert len(names) >= 20
    for n in sorted(names):
        assert hasattr(L, n), f"{n} declared in include/ but not exported"
    assert set(_lib.EXPORTS) <= names


def test_create_without_gpu_fails_cleanly():
    import greenstein2002_amd as g

    if os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK):
        pytest.skip("a GPU is visible")
    with pytest.raises(g.GrwError):
        g.FruCell(64)


def test_bad_config_rejected():
    from greenstein2002_amd import _lib

    L = _lib.load()
    cfg = _lib.GrwConfig()
    L.grw_config_default(C.byref(cfg))
    h = C.c_void_p()
    cfg.nfru = 0
    assert L.grw_create(C.byref(cfg), C.byref(h)) == -1
    cfg.nfru = 64
    cfg.block = 96
    assert L.grw_create(C.byref(cfg), C.byref(h)) == -1
    assert b"block" in L.grw_last_error()


def test_context_switches_reject_a_null_context():
    """grw_set_timing and grw_set_graphs return GRW_EINVAL (-1) and set the
    error text for a null context, without touching a device."""
    from greenstein2002_amd import _lib

    L = _lib.load()
    for fn in (L.grw_set_timing, L.grw_set_graphs):
        assert fn(None, 1) == -1
        assert b"null context" in L.grw_last_error()


def test_host_fcn_matches_reference_fixture(golden):
    from greenstein2002_amd.cell import fcn

    g = golden("fcn_ref.npz")
    for k in range(len(g["t"])):
        F, cur, st = fcn(g["t"][k], g["state"][k], g["J"][k])
        assert np.array_equal(F, g["F"][k]), k
        assert np.array_equal(cur, g["cur"][k]), k
        assert np.array_equal(st, g["state_after"][k]), k


def test_host_fcn_matches_oracle_under_clamp_and_chf(oracle):
    from greenstein2002_amd.cell import Protocol, fcn

    rng = np.random.default_rng(5)
    st0 = np.load(os.path.join(ROOT, "tests", "golden", "ic_stable.npz"))["state"]
    for vclamp, chf, serca in ((1, 0, 1.0), (0, 1, 1.0), (1, 1, 0.8), (0, 0, 1.3)):
        prot = Protocol(vclamp=bool(vclamp), chf=bool(chf), serca_scale=serca, vclamp_set=12.0,
                        vclamp_shift=7.0)
        op = oracle.protocol(vclamp=vclamp, chf=chf, serca_scale=serca, vclamp_set=12.0,
                             vclamp_shift=7.0)
        for _ in range(20):
            s = st0 * rng.uniform(0.95, 1.05, 37)
            t = rng.uniform(0, 600)
            J = rng.normal(size=4)
            F1, c1, s1 = fcn(t, s, J, prot)
            F2, c2, s2 = oracle.fcn(t, s, J, prot=op)
            assert np.array_equal(F1, F2) and np.array_equal(c1, c2) and np.array_equal(s1, s2)


def _clamp_closed_form(kind, t, p):
    """The clamp potential as a piecewise function of the phase within the
    protocol's period (fcn.c:448-535 read as intervals), or None outside."""
    if kind == "vclamp":
        x = t - p["vclamp_shift"]
        if x < 0:
            return p["vclamp_hold"]
        ph = x - np.floor(x / p["vclamp_period"]) * p["vclamp_period"]
        return p["vclamp_set"] if ph < p["vclamp_duration"] else p["vclamp_hold"]
    if kind == "vpp":
        x = t - p["vpp_shift"]
        if x < 0:
            return p["vpp_hold"]
        ph = x - np.floor(x / p["vpp_period"]) * p["vpp_period"]
        if ph < p["vpp_ppduration"]:
            return p["vpp_ppset"]
        if ph < p["vpp_ppduration"] + p["vpp_duration"]:
            return p["vpp_set"]
        return p["vpp_hold"]
    k = np.floor(t / p["iv_period"])
    ph = t - k * p["iv_period"]
    on = ph >= p["iv_shift"] and ph < min(p["iv_period"], p["iv_shift"] + p["iv_duration"])
    return p["iv_set"] + p["iv_n"] * p["iv_step"] if on else p["iv_hold"]


@pytest.mark.parametrize("kind", ["vclamp", "vpp", "iv"])
def test_clamp_protocols_hold_the_closed_form_potential(oracle, kind):
    """The plain clamp, the clamp with a prepulse (fcn.c:485-516) and the I-V
    steps (fcn.c:521-535): at times on a grid through several periods (away
    from the switching instants by 1e-6 ms, and exactly on the integer ones),
    the host fcn and the oracle hold V at the closed-form potential, with
    dV = 0 (fcn.c:811-838); and the two agree bit for bit on every output."""
    from greenstein2002_amd.cell import Protocol, fcn

    params = dict(
        vclamp=dict(vclamp=1, vclamp_shift=7.0, vclamp_period=60.0, vclamp_duration=25.0,
                    vclamp_set=12.0, vclamp_hold=-80.0),
        vpp=dict(vppclamp=1, vpp_shift=5.0, vpp_period=90.0, vpp_ppduration=20.0, vpp_duration=30.0,
                 vpp_ppset=-40.0, vpp_set=15.0, vpp_hold=-85.0),
        iv=dict(ivclamp=1, iv_period=70.0, iv_shift=10.0, iv_duration=80.0, iv_set=-40.0,
                iv_hold=-100.0, iv_step=10.0, iv_n=3),
    )[kind]
    prot = Protocol(**{k: (bool(v) if k in ("vclamp", "vppclamp", "ivclamp") else v) for k, v in params.items()})
    op = oracle.protocol(**params)
    pc = {f: getattr(prot, f) for f in prot.__dataclass_fields__}
    st0 = np.load(os.path.join(ROOT, "tests", "golden", "ic_stable.npz"))["state"]
    rng = np.random.default_rng(3)
    ts = np.concatenate([np.arange(0.0, 400.0, 1.0), np.arange(0.5e-6, 400.0, 2.5) + 1e-6,
                         rng.uniform(0.0, 400.0, 200)])
    levels = set()
    for t in ts:
        s = st0 * rng.uniform(0.97, 1.03, 37)
        J = rng.normal(size=4)
        F1, c1, s1 = fcn(t, s, J, prot)
        F2, c2, s2 = oracle.fcn(t, s, J, prot=op)
        assert np.array_equal(F1, F2) and np.array_equal(c1, c2) and np.array_equal(s1, s2), t
        want = _clamp_closed_form(kind, t, pc)
        assert s1[0] == want, (t, s1[0], want)
        assert F1[0] == 0.0
        levels.add(want)
    assert len(levels) == (2 if kind != "vpp" else 3)  # every phase of the protocol was visited
    # the I-V test potential steps with iv_n
    if kind == "iv":
        prot.iv_n = 5
        _, _, s5 = fcn(25.0, st0, np.zeros(4), prot)
        assert s5[0] == -40.0 + 5 * 10.0
    # without a clamp V is the state's and the stimulus drives dV
    F0, _, s0 = fcn(10.5, st0, np.zeros(4), Protocol())
    assert s0[0] == st0[0] and F0[0] != 0.0


def test_ic_reader_and_packaged_ic():
    """The text reader (initialize_state.c:43-133 rules) on the reference's files
    gives the packaged ic/stable data; records repeat every 1250 FRUs."""
    from greenstein2002_amd.ic import load_stable, read_ic_dir

    d = load_stable(2600)
    assert np.array_equal(d["fru"][1250:2500], d["fru"][:1250])
    assert np.array_equal(d["ryr"][2500:], d["ryr"][:100])
    ref = "/root/reference/ic/stable"
    if not os.path.isdir(ref):
        pytest.skip("reference ic/stable not present")
    r = read_ic_dir(ref, 2600)
    for k in ("fru", "ltype", "ryr", "ito2"):
        assert np.array_equal(r[k], d[k]), k
    assert np.array_equal(r["state"][:31], d["state"][:31])


def test_total_ca_matches_reference_initialisation(golden):
    """initialize_state.c:233-246 recomputes state[CaTOT] for NFRU and scale."""
    from greenstein2002_amd.ic import load_stable

    g = golden("ic_stable.npz")
    d = load_stable(1250, 10.0)
    assert d["state"][31] == g["state"][31]


def test_beat_stats_closed_form():
    """beat_stats on a piecewise-linear action potential, ICa and release flux
    whose APD90, peaks, rest values and integrals are known in closed form."""
    from greenstein2002_amd.beats import NSTAT, beat_stats, ca_influx

    shift, period = 10.0, 100.0
    # rest -90 mV; upstroke 10 -> 12 ms to +30; linear repolarisation to -90 at 92 ms
    t = np.concatenate([np.arange(1.0, 10.01, 1.0), np.arange(10.5, 12.01, 0.5), np.arange(13.0, 92.01, 1.0),
                        np.arange(93.0, 140.01, 1.0)])

    def V(x):
        if x <= 10.0:
            return -90.0
        if x <= 12.0:
            return -90.0 + 60.0 * (x - 10.0)
        if x <= 92.0:
            return 30.0 - 1.5 * (x - 12.0)
        return -90.0

    def ICa(x):  # triangle of area 4 (uA/uF ms) between 10 and 14 ms, peak -2 at 12 ms
        return -(x - 10.0) if 10.0 <= x <= 12.0 else (-(14.0 - x) if 12.0 < x <= 14.0 else 0.0)

    Vs = np.array([V(x) for x in t])
    cai = 1e-4 + 1e-3 * np.exp(-((t - 40.0) / 10.0) ** 2)
    ica = np.array([ICa(x) for x in t])
    log = np.stack([t, Vs, cai, ica], axis=1)
    jrel = np.full(len(t), 2e-3)
    rows_t = np.arange(1.0, 141.0)
    states = np.zeros((len(rows_t), 37))
    other = np.zeros((len(rows_t), 12))
    states[:, 0] = [V(x) for x in rows_t]
    states[:, 6] = 1e-4 + 1e-3 * np.exp(-((rows_t - 40.0) / 10.0) ** 2)
    states[:, 7] = 0.7
    other[:, 1] = 0.65
    rec = beat_stats(log, rows_t, states, other, shift, period, jrel=jrel)
    assert rec.shape == (2, NSTAT)
    b = rec[0]
    assert b[0] == 10.0 and b[1] == -90.0 and b[2] == 30.0 and b[3] == 2.0
    # threshold 30 - 0.9 * 120 = -78 mV reached at 12 + 108/1.5 = 84 ms
    assert abs(b[4] - (84.0 - 10.0)) < 1e-12
    assert b[5] == cai.max() and b[6] == 40.0 - 10.0
    assert b[8] == -2.0 and b[9] == 0.7 and b[10] == 0.65
    assert b[11] == np.count_nonzero((t >= 10.0) & (t < 110.0))
    # trapezoids are exact on piecewise-linear data with kinks on the grid: Ca in
    # = -4 uA/uF ms converted, Ca released = 2e-3 mM/ms x (109 - 9) ms
    assert abs(b[12] - ca_influx(-4.0)) < 1e-15 and ca_influx(-4.0) > 0
    assert abs(b[13] - 2e-3 * 100.0) < 1e-15 and b[14] == 2e-3
    # the second (partial) beat has no action potential: no APD90
    assert rec[1][0] == 110.0 and np.isnan(rec[1][4])
    # without a release-flux column the release fields are NaN
    assert np.isnan(beat_stats(log, rows_t, states, other, shift, period)[0][13])


def test_mpi_shim_exports_the_reference_mpi_layer():
    """libgrw_mpi.so exports mpi_master.c's nine entry points and the slave loop
    caf.c names (include/grw_mpi.h), and nothing else."""
    import re
    import subprocess

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    hdr = open(os.path.join(root, "include", "grw_mpi.h")).read()
    declared = set(re.findall(r"^void (\w+)\(", hdr, re.M))
    assert declared == {"initialize_mpi", "end_mpi", "parallel", "initialize_mpi_state",
                        "parallel_get_FRUs", "distrib_simFRU", "send_calc_fru_flux",
                        "send_calc_fru_avg", "send_save_state", "send_resume_state"}
    so = os.path.join(root, "greenstein2002_amd", "libgrw_mpi.so")
    out = subprocess.run(["nm", "-D", "--defined-only", so], capture_output=True, text=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    assert exported == declared
    import ctypes as C
    L = C.CDLL(so)  # loads without the program's NFRU_sim / NFRU_scale (weak references)
    for name in declared:
        assert getattr(L, name)


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                                    "oracle", "_ref", "stoch_grw")),
                    reason="oracle/_ref/stoch_grw not built (needs the reference tree)")
def test_reference_driver_links_against_the_shim():
    """oracle/_ref/stoch_grw -- the reference's caf.c, rk_pd54m.c, fcn.c, IC and
    output code -- takes the MPI layer's entry points from libgrw_mpi.so, has
    no MPI library and none of the reference's FRU kernels."""
    import subprocess

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "oracle", "_ref", "stoch_grw")
    und = subprocess.run(["nm", "-D", "--undefined-only", exe], capture_output=True, text=True).stdout
    need = {ln.split()[-1] for ln in und.splitlines()}
    for name in ("initialize_mpi", "initialize_mpi_state", "distrib_simFRU", "send_calc_fru_flux",
                 "send_calc_fru_avg", "send_save_state", "send_resume_state", "parallel_get_FRUs",
                 "end_mpi", "parallel"):
        assert name in need, name
    syms = subprocess.run(["nm", exe], capture_output=True, text=True).stdout
    for kernel in ("simfru_local", "fru_rates_local", "fcn_fru", "calc_fru_flux_local", "MPI_Init"):
        assert kernel not in syms, kernel
    needed = lambda f: re.findall(r"NEEDED.*\[(.*)\]", subprocess.run(  # noqa: E731
        ["readelf", "-d", f], capture_output=True, text=True).stdout)
    assert "libgrw_mpi.so" in needed(exe) and not any("libmpi" in n for n in needed(exe))
    assert "libgrw_fru.so" in needed(os.path.join(root, "greenstein2002_amd", "libgrw_mpi.so"))

