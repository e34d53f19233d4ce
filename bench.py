"""Benchmark: simulated ms per wall-second of one 12,500-FRU canine myocyte.

Workload (the plain FP64 drop-in configuration): one cell of 12,500 FRUs
(50,000 LCCs, 250,000 RyRs, 50,000 Ito2 channels) started from ic/stable,
current-clamp paced at 1 Hz (stimulus -100 uA/uF for 1 ms at t = 10 + 1000 k ms),
MT19937 streams seeded from master seed 1234.  One "step" is one call of the
global RK5(4)6M integrator over time_step = 1 ms of simulated time, i.e. every
Monte Carlo sub-interval of all 12,500 FRUs in that millisecond.  The default
times steps [warmup, warmup + 1000) ms: a whole 1 Hz beat including the
stimulus, the Ca release and diastole.

    python bench.py [--gpus 1] [--steps 1000] [--warmup 5]

Prints one JSON line.  ``roofline`` is measured for the dominant kernel (the
FRU Monte Carlo sub-interval kernel with the integrator stage in its tail) from
HIP events recorded around every launch on the library's stream;
``cpu_baseline`` times the reference's own MPI program (oracle/_ref/stoch,
built from its unmodified sources) on this host's cores over the same simulated
window, at its compiled 1,000-FRU size scaled to 12,500 FRUs (the CPU oracle,
single-threaded, where that program is not built).  Two further measurements share the line:

* ``mixed``: the same beat with the mixed-precision form (FP32 rates and local
  Ca right-hand side), and its speed relative to FP64;
* ``batched``: 256 independent cells in one batch, FRU counts uniform in
  [1250, 12500] and pacing cycle lengths uniform in [300, 2000] ms (host PRNG
  seed 1234, master seeds 1234 + k), over the first --batch-ms simulated ms
  (each cell's first stimulus at 10 ms included); beats/s counts simulated
  cycle lengths completed per wall second, and a sample of the same cells run
  one after another through the single-cell driver gives the sequential rate.

``roofline.traffic`` is the kernel's measured HBM traffic per launch: bench.py
runs itself twice more under ``rocprofv3 --pmc`` (FETCH_SIZE, then WRITE_SIZE)
over the first 35 simulated ms of the same cell, and twice for a 256 MiB
device-to-device copy whose counter values give the bytes per counter unit
(``--no-traffic`` skips this; without rocprofv3 it is null).
"""
from __future__ import annotations

import argparse
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Algorithmic HBM bytes per FRU-step and per FRU sub-interval call:
#   32 uniforms x 12 B (4 B read of the word + 8 B amortised regeneration read/write)
#   and 120 B of FRU state loaded and stored per call.
BYTES_PER_STEP = 384.0
BYTES_PER_CALL = 120.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000, help="timed simulated ms")
    ap.add_argument("--warmup", type=int, default=5, help="untimed simulated ms before timing")
    ap.add_argument("--nfru", type=int, default=12500)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--lanes", type=int, default=4, choices=[1, 4], help="lanes per FRU")
    ap.add_argument("--driver", choices=["device", "host"], default="device",
                    help="integrator on the GPU (default) or on the host calling the FRU ABI")
    ap.add_argument("--cpu-ms", type=float, default=3.0,
                    help="0 disables the CPU baseline; otherwise the simulated ms of the CPU-oracle "
                         "sample used when the reference's MPI program is not built")
    ap.add_argument("--no-timing", action="store_true",
                    help="do not record per-launch HIP events (for profiler runs)")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision run")
    ap.add_argument("--graphs", type=int, default=0, choices=[0, 1],
                    help="1: the device integrator replays RK attempts from a captured HIP graph")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 counter runs that measure roofline.traffic")
    ap.add_argument("--batch-cells", type=int, default=256, help="cells of the batched run (0 skips it)")
    ap.add_argument("--batch-ms", type=int, default=30, help="simulated ms of the batched run")
    ap.add_argument("--batch-seq-sample", type=int, default=8,
                    help="cells of the batch also run one at a time for the sequential rate")
    ap.add_argument("--ensemble", type=int, default=0,
                    help="only run the seed ensemble: this many master seeds from 1234, one cell each, "
                         "FP64 and mixed batches, per-beat statistics vs each other and vs the CPU "
                         "reference fixture when one exists for the shape")
    ap.add_argument("--ensemble-ms", type=int, default=1011, help="simulated ms of each ensemble cell")
    ap.add_argument("--clamp", type=int, default=0,
                    help="only run the voltage-clamp set: this many seeds per test potential "
                         "(-40..+60 mV step 10, hold -80 mV), FP64 and mixed batches; peak ICa, "
                         "Ca entry, Ca release and EC-coupling gain of the step")
    ap.add_argument("--clamp-shift", type=float, default=100.0, help="ms held before the step")
    ap.add_argument("--clamp-duration", type=float, default=200.0, help="ms of the step")
    ap.add_argument("--sweep", type=int, default=0,
                    help="only run the parameter sweep: this many cells of --nfru FRUs (1,024 x 12,500 "
                         "= 12.8 M FRUs for the full sweep) with RyR-opening and LCC mode-switch scales "
                         "log-uniform in [0.5, 2] and SERCA scale uniform in [0.7, 1.3], mixed precision "
                         "in one batch, then the first --sweep-fp64 of them again in FP64")
    ap.add_argument("--sweep-ms", type=int, default=30, help="simulated ms of the sweep (stimulus at 10 ms)")
    ap.add_argument("--sweep-fp64", type=int, default=64, help="cells of the sweep rerun in FP64")
    ap.add_argument("--summarize", default="",
                    help="only summarise a rocprofv3 output folder (kernel trace and / or counter "
                         "CSVs) per kernel and grid size, as markdown")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="only run a device-to-device copy of this many MiB (5 times): the "
                         "reference kernel for scaling FETCH_SIZE / WRITE_SIZE counters")
    return ap.parse_args()


def calibrate_copy(mib: int) -> None:
    """A plain device-to-device copy of known size (hipMemcpy, five times) whose
    HBM counters calibrate the profiler's FETCH_SIZE / WRITE_SIZE."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    nbytes = mib * (1 << 20)
    a, b = C.c_void_p(), C.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    if hip.hipMemset(a, 1, C.c_size_t(nbytes)) != 0:
        raise RuntimeError("hipMemset failed")
    for _ in range(5):
        if hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"calibration": "copy", "bytes_read_per_launch": nbytes,
                      "bytes_written_per_launch": nbytes, "launches": 5}), flush=True)


def _counter_means(folder: str, counter: str) -> dict:
    """Mean value per dispatch of one counter, per (kernel, workgroups), from a
    rocprofv3 --output-format csv folder."""
    import csv
    import glob
    from collections import defaultdict

    acc = defaultdict(list)
    for path in glob.glob(os.path.join(folder, "**", "*counter_collection.csv"), recursive=True):
        with open(path) as f:
            for r in csv.DictReader(f):
                if r["Counter_Name"] != counter:
                    continue
                grid = int(r["Grid_Size"]) // max(int(r["Workgroup_Size"]), 1)
                acc[(r["Kernel_Name"], grid)].append(float(r["Counter_Value"]))
    return {k: (sum(v) / len(v), len(v)) for k, v in acc.items()}


def measure_traffic(args) -> dict | None:
    """HBM bytes per launch of k_dev_stage: FETCH_SIZE and WRITE_SIZE (KiB per
    dispatch) over t = 0..35 ms of the same cell (stimulus at 10 ms and the
    start of release), each in its own rocprofv3 run,
    scaled by the bytes per KiB-unit the same counter shows for a 256 MiB
    device-to-device copy (FETCH_SIZE reads half the copied bytes on MI355X).
    Returns None when rocprofv3 is missing or a run fails."""
    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(prof):
        return None
    env = dict(os.environ, TMPDIR=os.environ.get("TMPDIR", "/tmp"))
    win_ms = 35  # from t = 0 (no warmup, so that the launch count covers every dispatch)
    cell_cmd = [sys.executable, os.path.abspath(__file__), "--steps", str(win_ms), "--warmup", "0",
                "--nfru", str(args.nfru), "--seed", str(args.seed), "--no-mixed", "--batch-cells", "0",
                "--cpu-ms", "0", "--no-timing", "--no-traffic", "--graphs", str(args.graphs)]
    copy_cmd = [sys.executable, os.path.abspath(__file__), "--calibrate-copy-mib", "256"]
    out = {}
    with tempfile.TemporaryDirectory(prefix="grw_pmc_", dir=env["TMPDIR"]) as tmp:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            for tag, cmd in (("cell", cell_cmd), ("copy", copy_cmd)):
                d = os.path.join(tmp, f"{tag}_{counter}")
                try:
                    r = subprocess.run([prof, "--pmc", counter, "--output-format", "csv", "-d", d, "-o", tag, "--"] + cmd,
                                       capture_output=True, text=True, timeout=180, env=env, start_new_session=True)
                except subprocess.TimeoutExpired:
                    return None
                if tag == "cell":
                    lines = [ln for ln in r.stdout.splitlines() if ln.startswith("{")]
                    if not lines:
                        return None
                    out[f"cell_{counter}"] = json.loads(lines[-1])
                means = _counter_means(d, counter)
                if not means:
                    return None
                out[f"{tag}_{counter}_means"] = means
    nbytes = 256 * (1 << 20)
    res = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        copy = [v for (k, g), v in out[f"copy_{counter}_means"].items() if "copyBuffer" in k]
        stage = [(v, g) for (k, g), v in out[f"cell_{counter}_means"].items() if "k_dev_stage" in k]
        if not copy or not stage:
            return None
        (kib_copy, _n) = max(copy)
        (kib, ndisp), grid = max(stage, key=lambda x: x[0][1])
        res[counter] = dict(kib_per_dispatch=round(kib, 1), dispatches=ndisp, workgroups=grid,
                            bytes_per_kib=round(nbytes / (kib_copy * 1024.0), 4))
    cfg = out["cell_FETCH_SIZE"]["config"]
    launches = cfg.get("launches") or 0
    if not launches:
        return None
    f, w = res["FETCH_SIZE"], res["WRITE_SIZE"]
    read = f["kib_per_dispatch"] * 1024.0 * f["bytes_per_kib"] * f["dispatches"] / launches
    written = w["kib_per_dispatch"] * 1024.0 * w["bytes_per_kib"] * w["dispatches"] / launches
    alg = cfg.get("alg_bytes_per_launch")
    return dict(bytes_per_launch=round(read + written, 1), read_per_launch=round(read, 1),
                written_per_launch=round(written, 1),
                window="t = 0..%d ms of the same cell (%d launches)" % (win_ms, launches),
                alg_bytes_per_launch_window=alg,
                ratio_to_algorithmic=round((read + written) / alg, 3) if alg else None,
                counters=res)


def summarize(folder: str) -> None:
    """Per (kernel, grid size): dispatches and mean duration from
    *kernel_trace.csv, mean counter values per dispatch from
    *counter_collection.csv (rocprofv3 --output-format csv)."""
    import csv
    import glob
    from collections import defaultdict

    def short(name):
        name = name.replace("(anonymous namespace)::", "")
        return name.split("(")[0].replace("void ", "")

    for path in sorted(glob.glob(os.path.join(folder, "**", "*kernel_trace.csv"), recursive=True)):
        acc = defaultdict(list)
        with open(path) as f:
            for r in csv.DictReader(f):
                grid = int(r["Grid_Size_X"]) // max(int(r["Workgroup_Size_X"]), 1)
                acc[(short(r["Kernel_Name"]), grid)].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
        print(f"\n{os.path.relpath(path, folder)}\n")
        print("| kernel | workgroups | dispatches | mean us | min us | max us | total ms |")
        print("|---|---|---|---|---|---|---|")
        for (k, g), v in sorted(acc.items(), key=lambda kv: -sum(kv[1])):
            print(f"| {k} | {g} | {len(v)} | {sum(v) / len(v):.2f} | {min(v):.2f} | {max(v):.2f} | {sum(v) / 1e3:.2f} |")
    for path in sorted(glob.glob(os.path.join(folder, "**", "*counter_collection.csv"), recursive=True)):
        acc = defaultdict(lambda: defaultdict(list))
        with open(path) as f:
            for r in csv.DictReader(f):
                grid = int(r["Grid_Size"]) // max(int(r["Workgroup_Size"]), 1)
                acc[(short(r["Kernel_Name"]), grid)][r["Counter_Name"]].append(float(r["Counter_Value"]))
        print(f"\n{os.path.relpath(path, folder)}\n")
        names = sorted({c for v in acc.values() for c in v})
        print("| kernel | workgroups | dispatches | " + " | ".join(f"mean {c}" for c in names) + " |")
        print("|---|---|---|" + "---|" * len(names))
        for (k, g), v in sorted(acc.items()):
            n = max(len(x) for x in v.values())
            print(f"| {k} | {g} | {n} | " + " | ".join(
                f"{sum(v[c]) / len(v[c]):.1f}" if v.get(c) else "" for c in names) + " |")


def reference_cpu_baseline(t0: float, t1: float, ranks: int) -> dict | None:
    """The reference's own CPU path: its whole MPI program (caf.c + rk_pd54m.c +
    mpi_master.c / mpi_slave.c + the hot path, unmodified, built by
    oracle/Makefile into oracle/_ref/stoch) under MPICH's mpiexec with `ranks`
    processes on this host, started from ic/stable (written here from the
    package's copy, exact to the bit), timed between the progress lines it
    prints at simulated t0 and t1 (caf.c:152, one line per 1 ms output step).
    Its configuration is compiled in: 1,000 FRUs, NFRU_scale 12.5 (parameters.h:
    43-51), the same 1 Hz current clamp.  Returns None when the program or
    mpiexec is not there."""
    stoch = os.path.join(ROOT, "oracle", "_ref", "stoch")
    mpiexec = shutil.which("mpiexec", path=os.pathsep.join(["/opt/conda/bin", os.environ.get("PATH", "")]))
    stdbuf = shutil.which("stdbuf")
    if not (os.path.exists(stoch) and mpiexec and stdbuf):
        return None
    import signal

    from greenstein2002_amd.ic import write_ic_dir

    with tempfile.TemporaryDirectory(prefix="grw_ref_") as run:
        write_ic_dir(os.path.join(run, "ic", "stable"))
        p = subprocess.Popen([mpiexec, "-n", str(ranks), stdbuf, "-oL", stoch], cwd=run,
                             stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True,
                             start_new_session=True)
        marks = {}
        deadline = time.perf_counter() + 600.0
        try:
            for line in p.stdout:
                tok = line.split()
                if len(tok) == 4 and tok[2] == "1000":
                    try:
                        t = float(tok[0])
                    except ValueError:
                        continue
                    marks[t] = time.perf_counter()
                    if t >= t1 or time.perf_counter() > deadline:
                        break
        finally:
            try:
                os.killpg(p.pid, signal.SIGTERM)
            except ProcessLookupError:
                pass
            try:
                p.wait(timeout=30)
            except subprocess.TimeoutExpired:
                os.killpg(p.pid, signal.SIGKILL)
                p.wait()
    if t0 not in marks or t1 not in marks:
        return None
    wall = marks[t1] - marks[t0]
    rate = (t1 - t0) / wall
    return dict(value=round(rate * 1000.0 / 12500.0, 5), unit="simulated ms per wall-second", cores=ranks,
                kind="reference",
                sample=f"the reference's MPI program (oracle/_ref/stoch: caf.c, rk_pd54m.c, mpi_master.c, "
                       f"mpi_slave.c and the hot path, unmodified, gcc -O3, MPICH) on {ranks} ranks of this "
                       f"host, its compiled 1,000-FRU cell, 1 Hz current clamp from ic/stable, timed over "
                       f"t = {t0:g}..{t1:g} ms (the GPU line's window): {rate:.3f} simulated ms/s at 1,000 FRUs "
                       f"in {wall:.2f} s, scaled by 1,000/12,500 to the 12,500-FRU cell (its run time is "
                       f"per-FRU work: >= 95% in the hot path)",
                value_1000fru=round(rate, 4), wall_s=round(wall, 3))


def cpu_baseline(nfru: int, ms: float, seed: int):
    """The CPU oracle (oracle/libgrw_oracle.so, one thread) on the same cell
    for the first `ms` simulated ms from ic/stable; used when the reference's
    program is not built."""
    import numpy as np

    from greenstein2002_amd.ic import load_stable
    from oracle.pyoracle import FRU_DTYPE, Oracle, OracleCell

    o = Oracle()
    ic = load_stable(nfru)
    f = np.zeros(nfru, FRU_DTYPE)
    f["conc"] = ic["fru"]
    f["lcc"] = ic["ltype"][:, :, 0]
    f["yinact"] = ic["ltype"][:, :, 1]
    f["ryr"] = ic["ryr"]
    f["ito2"] = ic["ito2"]
    for j in range(nfru):
        m = o.mt_seed(o.seed_for_fru(seed, j))
        f["mt"][j] = m["mt"]
        f["mti"][j] = m["mti"]
    cell = OracleCell(o, f, 12500.0 / nfru)
    state = np.array(ic["state"])
    cur = np.zeros(24)
    h = 0.1
    n = int(round(ms))
    t0 = time.perf_counter()
    for k in range(n):
        h, _ = cell.rk54pd(float(k), float(k + 1), h, state, cur)
    dt = time.perf_counter() - t0
    steps = int(cell.c.n_fru_steps)
    cell.close()
    return dict(value=n / dt, unit="simulated ms per wall-second", cores=1, kind="port",
                sample=f"{nfru}-FRU cell, t = 0..{n} ms from ic/stable, CPU oracle single-threaded, "
                       f"{steps} FRU-steps in {dt:.1f} s = {dt / max(steps, 1) * 1e9:.0f} ns/FRU-step")


def host_cores() -> int:
    """Cores this process may use, capped at 16 (one GPU's share of the host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def run_cell(args, precision, timing):
    """The single-cell beat: warmup, then args.steps ms timed in 100 ms chunks."""
    from greenstein2002_amd.cell import Cell, Protocol

    prot = Protocol()
    cell = Cell(args.nfru, seed=args.seed, precision=precision, protocol=prot, block=args.block,
                lanes_per_fru=args.lanes)
    cell.fru.set_graphs(bool(args.graphs))
    if args.warmup:
        cell.run(args.warmup, outputs=False, log_cap=0, driver=args.driver)
    cell.fru.set_timing(timing)
    cell.fru.reset_stats()
    chunks = []
    t0 = time.perf_counter()
    done = 0
    while done < args.steps:
        n = min(100, args.steps - done)
        chunks.append(cell.run(n, log_cap=0, driver=args.driver))
        done += n
        print(f"[bench] {precision}: t = {args.warmup + done} ms simulated, {time.perf_counter() - t0:.1f} s wall",
              file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    st = cell.fru.stats()
    beats = cell.beats() if args.driver == "device" else None
    cell.close()
    return wall, chunks, st, beats


def batched(args):
    """256 uneven cells in one batch vs the same cells one at a time."""
    import numpy as np

    from greenstein2002_amd.batch import Batch, CellSpec
    from greenstein2002_amd.cell import Cell, Protocol

    n = args.batch_cells
    rng = np.random.default_rng(1234)
    nf = rng.integers(1250, 12501, n)
    cl = rng.uniform(300.0, 2000.0, n)
    specs = [CellSpec(int(nf[k]), seed=1234 + k, protocol=Protocol(period=float(cl[k]))) for k in range(n)]
    b = Batch(specs, precision=args.precision)
    b.set_graphs(bool(args.graphs))
    b.set_timing(True)
    b.reset_stats()
    t0 = time.perf_counter()
    done = 0
    while done < args.batch_ms:
        m = min(10, args.batch_ms - done)
        b.run(m)
        done += m
        print(f"[bench] batch of {n}: t = {done} ms simulated, {time.perf_counter() - t0:.1f} s wall",
              file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    st = b.stats()
    att = int(b.attempts.sum())
    b.close()
    alg = BYTES_PER_STEP * st["fru_steps"] + BYTES_PER_CALL * st["simfru_calls"]
    # the integrator times a systematic sample of its launches: algorithmic bytes
    # per launch over the mean duration of the timed ones
    avg_s = st["kernel_ms"] / 1e3 / max(st["timed_launches"], 1)
    achieved = alg / max(st["launches"], 1) / avg_s / 1e9 if st["kernel_ms"] else 0.0
    # the same cells one at a time (an evenly spaced sample of them)
    pick = np.unique(np.linspace(0, n - 1, max(1, min(args.batch_seq_sample, n))).astype(int))
    seq = []
    for k in pick:
        c = Cell(int(nf[k]), seed=1234 + int(k), precision=args.precision, protocol=Protocol(period=float(cl[k])))
        ts = time.perf_counter()
        c.run(args.batch_ms, outputs=False, log_cap=0)
        seq.append(time.perf_counter() - ts)
        c.close()
    seq_wall = float(np.mean(seq)) * n
    beats = float(np.sum(args.batch_ms / cl))
    return dict(
        workload=f"{n} cells, FRUs uniform in [1250, 12500] (total {int(nf.sum())}), cycle lengths uniform "
                 f"in [300, 2000] ms, t = 0..{args.batch_ms} ms from ic/stable, {args.precision}",
        cells=n, frus=int(nf.sum()), window_ms=args.batch_ms, wall_s=round(wall, 3),
        cell_ms_per_s=round(n * args.batch_ms / wall, 2), beats_per_s=round(beats / wall, 4),
        rk_attempts=att, fru_steps=st["fru_steps"],
        fru_steps_per_s=round(st["fru_steps"] / wall, 1),
        roofline=dict(bound="hbm", achieved=round(achieved, 3), peak=HBM_PEAK_GBS, unit="GB/s",
                      frac=round(achieved / HBM_PEAK_GBS, 6),
                      kernel="k_dev_stage<%s> over all cells" % ("double" if args.precision == "fp64" else "float"),
                      avg_launch_us=round(avg_s * 1e6, 3),
                      launches=st["launches"], timed_launches=st["timed_launches"]),
        sequential=dict(sample_cells=[int(k) for k in pick], sample_walls_s=[round(x, 3) for x in seq],
                        est_wall_s=round(seq_wall, 2), beats_per_s=round(beats / seq_wall, 4)),
        speedup_vs_sequential=round(seq_wall / wall, 2))


def ensemble(args) -> dict:
    """Seeds 1234.. as one batch per precision; beat 0 statistics of each cell,
    ensemble means, mixed - FP64 differences, and FP64 vs the CPU oracle's runs
    on the same seeds (tests/golden/ensemble_*.npz) when the shape matches."""
    import glob

    import numpy as np

    from greenstein2002_amd.batch import Batch, CellSpec
    from greenstein2002_amd.beats import FIELDS

    seeds = np.arange(1234, 1234 + args.ensemble)
    keys = ["APD90", "V_peak", "Cai_peak", "ICa_peak", "CaNSR_rest", "CaJSR_rest"]
    idx = [FIELDS.index(k) for k in keys]
    res, walls = {}, {}
    for prec in ("fp64", "mixed"):
        specs = [CellSpec(args.nfru, seed=int(s)) for s in seeds]
        b = Batch(specs, precision=prec)
        t0 = time.perf_counter()
        done = 0
        while done < args.ensemble_ms:
            m = min(100, args.ensemble_ms - done)
            b.run(m)
            done += m
            print(f"[bench] ensemble {prec}: t = {done} ms, {time.perf_counter() - t0:.1f} s wall",
                  file=sys.stderr, flush=True)
        walls[prec] = time.perf_counter() - t0
        res[prec] = np.array([b.beats(k)[0] for k in range(len(seeds))])
        b.close()
    out = dict(metric="beat statistics of a seed ensemble, FP64 vs mixed vs CPU reference",
               workload=f"{len(seeds)} cells x {args.nfru} FRUs, seeds {seeds[0]}..{seeds[-1]}, 1 Hz, "
                        f"t = 0..{args.ensemble_ms} ms from ic/stable, beat at 10 ms",
               wall_s={k: round(v, 2) for k, v in walls.items()}, stats={},
               beats={k: v.tolist() for k, v in res.items()}, fields=list(FIELDS))
    f64, mx = res["fp64"][:, idx], res["mixed"][:, idx]
    for j, k in enumerate(keys):
        a, m = f64[:, j], mx[:, j]
        out["stats"][k] = dict(fp64_mean=float(np.nanmean(a)), fp64_sd=float(np.nanstd(a)),
                               mixed_mean=float(np.nanmean(m)), mixed_sd=float(np.nanstd(m)),
                               mixed_rel_err=float((np.nanmean(m) - np.nanmean(a)) / abs(np.nanmean(a))),
                               sem_rel=float(np.nanstd(a) / np.sqrt(len(a)) / abs(np.nanmean(a))))
    for path in sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "ensemble_*.npz"))):
        g = np.load(path)
        if int(g["nfru"]) == args.nfru and int(g["ms"]) <= args.ensemble_ms and \
                np.array_equal(g["seeds"][: len(seeds)], seeds):
            ref = g["beats"][: len(seeds), 0][:, idx]
            same = np.isclose(f64, ref, rtol=1e-6, atol=0).all(axis=1)
            out["cpu_reference"] = dict(
                fixture=os.path.relpath(path, ROOT), seeds_equal_1e6=int(same.sum()),
                mean_rel_diff={k: float((np.nanmean(f64[:, j]) - np.nanmean(ref[:, j])) / abs(np.nanmean(ref[:, j])))
                               for j, k in enumerate(keys)})
            break
    return out


def clamp(args) -> dict:
    """Voltage clamp from -80 mV to Vt for --clamp-duration ms (fcn.c:448-483),
    seeds 1234.. per test potential, one batch per precision; statistics of the
    clamp period containing the step (beat records 8, 12, 13)."""
    import numpy as np

    from greenstein2002_amd.batch import Batch, CellSpec
    from greenstein2002_amd.cell import Protocol

    vts = np.arange(-40.0, 61.0, 10.0)
    seeds = np.arange(1234, 1234 + args.clamp)
    ms = int(args.clamp_shift + args.clamp_duration + 50.0)
    specs = []
    for vt in vts:
        for sd in seeds:
            prot = Protocol(vclamp=True, vclamp_hold=-80.0, vclamp_set=float(vt), vclamp_shift=args.clamp_shift,
                            vclamp_duration=args.clamp_duration, vclamp_period=1000.0)
            specs.append(CellSpec(args.nfru, seed=int(sd), protocol=prot))
    res, walls = {}, {}
    for prec in ("fp64", "mixed"):
        b = Batch(specs, precision=prec)
        t0 = time.perf_counter()
        done = 0
        while done < ms:
            m = min(50, ms - done)
            b.run(m)
            done += m
            print(f"[bench] clamp {prec}: t = {done} ms, {time.perf_counter() - t0:.1f} s wall", file=sys.stderr,
                  flush=True)
        walls[prec] = time.perf_counter() - t0
        res[prec] = np.array([b.beats(k)[0] for k in range(len(specs))]).reshape(len(vts), len(seeds), -1)
        b.close()
    out = dict(metric="voltage clamp, FP64 vs mixed",
               workload=f"{len(vts)} test potentials x {len(seeds)} seeds x {args.nfru} FRUs, hold -80 mV, "
                        f"step at {args.clamp_shift} ms for {args.clamp_duration} ms, t = 0..{ms} ms",
               wall_s={k: round(v, 2) for k, v in walls.items()}, rows=[])
    for i, vt in enumerate(vts):
        row = dict(Vt=float(vt))
        for name, col in (("ICa_peak", 8), ("Ca_in", 12), ("Ca_rel", 13)):
            a, m = res["fp64"][i, :, col], res["mixed"][i, :, col]
            row[name] = dict(fp64=float(a.mean()), sd=float(a.std()), mixed=float(m.mean()),
                             rel_err=float((m.mean() - a.mean()) / abs(a.mean())) if a.mean() else None)
        ga = res["fp64"][i, :, 13] / res["fp64"][i, :, 12]
        gm = res["mixed"][i, :, 13] / res["mixed"][i, :, 12]
        row["gain"] = dict(fp64=float(ga.mean()), sd=float(ga.std()), mixed=float(gm.mean()),
                           rel_err=float((gm.mean() - ga.mean()) / abs(ga.mean())))
        out["rows"].append(row)
    return out


def sweep(args) -> dict:
    """The many-cell parameter sweep in one batch (the streaming regime: the
    MT rows of all cells, 2.5 KB per FRU, far exceed L2 and the Infinity
    Cache).  Mixed precision for all cells, FP64 for the first few again with
    the same seeds and scales; beat statistics of the window (records 2, 5, 8:
    peak V, peak Cai, peak ICa) compared cell by cell."""
    import numpy as np

    from greenstein2002_amd.batch import Batch, CellSpec
    from greenstein2002_amd.cell import Protocol

    n = args.sweep
    rng = np.random.default_rng(1234)
    k23 = np.exp(rng.uniform(np.log(0.5), np.log(2.0), n))
    gam = np.exp(rng.uniform(np.log(0.5), np.log(2.0), n))
    serca = rng.uniform(0.7, 1.3, n)
    specs = [CellSpec(args.nfru, seed=1234 + k, k23_scale=float(k23[k]), gamma_scale=float(gam[k]),
                      protocol=Protocol(serca_scale=float(serca[k]))) for k in range(n)]

    def run(cells, prec):
        b = Batch(cells, precision=prec)
        b.set_graphs(bool(args.graphs))
        b.set_timing(True)
        b.reset_stats()
        t0 = time.perf_counter()
        done = 0
        while done < args.sweep_ms:
            m = min(10, args.sweep_ms - done)
            b.run(m)
            done += m
            print(f"[bench] sweep {prec} x {len(cells)}: t = {done} ms, {time.perf_counter() - t0:.1f} s wall",
                  file=sys.stderr, flush=True)
        wall = time.perf_counter() - t0
        st = b.stats()
        beats = np.array([b.beats(k, include_current=True)[0] for k in range(len(cells))])
        b.close()
        alg = BYTES_PER_STEP * st["fru_steps"] + BYTES_PER_CALL * st["simfru_calls"]
        avg_s = st["kernel_ms"] / 1e3 / max(st["timed_launches"], 1)
        achieved = alg / max(st["launches"], 1) / avg_s / 1e9 if st["kernel_ms"] else 0.0
        line = dict(precision=prec, cells=len(cells), wall_s=round(wall, 2),
                    cell_ms_per_s=round(len(cells) * args.sweep_ms / wall, 2),
                    fru_steps=st["fru_steps"], fru_steps_per_s=round(st["fru_steps"] / wall, 1),
                    roofline=dict(bound="hbm", achieved=round(achieved, 3), peak=HBM_PEAK_GBS, unit="GB/s",
                                  frac=round(achieved / HBM_PEAK_GBS, 6), avg_launch_us=round(avg_s * 1e6, 3),
                                  launches=st["launches"], timed_launches=st["timed_launches"],
                                  alg_bytes_per_launch=round(alg / max(st["launches"], 1), 1)))
        return line, beats

    mixed, bm = run(specs, "mixed")
    m64 = min(args.sweep_fp64, n)
    fp64, bf = run(specs[:m64], "fp64")
    cols = {"V_peak": 2, "Cai_peak": 5, "ICa_peak": 8}
    diff = {}
    for k, c in cols.items():
        a, m = bf[:, c], bm[:m64, c]
        rel = (m - a) / np.abs(a)
        diff[k] = dict(fp64_mean=float(a.mean()), mixed_mean=float(m.mean()),
                       mean_rel_err=float((m.mean() - a.mean()) / abs(a.mean())),
                       max_abs_rel_err_per_cell=float(np.abs(rel).max()))
    return dict(
        metric="parameter sweep of many 12,500-FRU cells in one batch (mixed), FP64 rerun of a subset",
        workload=f"{n} cells x {args.nfru} FRUs ({n * args.nfru:,} FRUs, {n * args.nfru * 2560 / 1e9:.1f} GB of MT "
                 f"rows), k23 and gamma scales log-uniform [0.5, 2], SERCA uniform [0.7, 1.3] (host PRNG "
                 f"1234), seeds 1234.., 1 Hz, t = 0..{args.sweep_ms} ms from ic/stable, stimulus at 10 ms",
        mixed=mixed, fp64_subset=fp64, mixed_vs_fp64_subset=diff,
        scales=dict(k23=[float(k23.min()), float(k23.max())], gamma=[float(gam.min()), float(gam.max())],
                    serca=[float(serca.min()), float(serca.max())]))


def main():
    args = parse()
    if args.sweep:
        print(json.dumps(sweep(args)), flush=True)
        return
    if args.clamp:
        print(json.dumps(clamp(args)), flush=True)
        return
    if args.ensemble:
        print(json.dumps(ensemble(args)), flush=True)
        return
    if args.summarize:
        summarize(args.summarize)
        return
    if args.calibrate_copy_mib:
        calibrate_copy(args.calibrate_copy_mib)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported on this "
              "single-GPU build (independent cells would shard across GPUs with no exchange)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from greenstein2002_amd.beats import as_dicts

    wall, chunks, st, beats = run_cell(args, args.precision, not args.no_timing)
    states = np.concatenate([c.states for c in chunks])
    attempts = sum(c.stats["attempts"] for c in chunks)
    rejects = sum(c.stats["rejects"] for c in chunks)

    value = args.steps / wall
    roof = None
    if not args.no_timing and st["launches"] and st["timed_launches"]:
        alg_bytes = BYTES_PER_STEP * st["fru_steps"] + BYTES_PER_CALL * st["simfru_calls"]
        # HIP events time a systematic sample of the launches in the timed region
        # (one RK attempt in eight); per-launch figures use every launch
        avg_s = st["kernel_ms"] / 1e3 / st["timed_launches"]
        achieved = alg_bytes / st["launches"] / avg_s / 1e9
        roof = dict(bound="hbm", achieved=round(achieved, 3), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 6), traffic=None,
                    kernel=(("k_dev_stage" if args.driver == "device" else
                             "k_simfru4" if args.lanes == 4 else "k_simfru") +
                            ("<double>" if args.precision == "fp64" else "<float>") +
                            " (FRU sub-interval + fused flux reduction" +
                            (" + integrator stage)" if args.driver == "device" else ")")),
                    avg_launch_us=round(avg_s * 1e6, 3),
                    launches=st["launches"], timed_launches=st["timed_launches"], fru_steps=st["fru_steps"],
                    critical_steps_per_launch=round(st["critical_steps"] / st["launches"], 2),
                    us_per_critical_step=round(avg_s * 1e6 * st["launches"] / max(st["critical_steps"], 1), 3),
                    serial_tail_us=round(st["tail_ms"] * 1e3 / st["launches"], 3),
                    alg_bytes_per_launch=round(alg_bytes / st["launches"], 1))
    V = states[:, 0]
    line = {
        "metric": "simulated ms per wall-second, 12,500-FRU cell",
        "value": round(value, 4),
        "unit": "simulated ms / s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32-rates/f64-state",
        "data": "synthetic: ic/stable initial state, seeded MT19937 (master seed %d)" % args.seed,
        "config": {"workload": "one canine myocyte, %d FRUs, 1 Hz current clamp, ic/stable, "
                               "%s, t = %d..%d ms" % (args.nfru, "FP64 drop-in" if args.precision == "fp64"
                                                      else "mixed FP32/FP64", args.warmup,
                                                      args.warmup + args.steps),
                   "nfru": args.nfru, "precision": args.precision, "lanes_per_fru": args.lanes,
                   "hip_graphs": bool(args.graphs),
                   "driver": args.driver,
                   "rk_attempts": attempts, "rk_rejects": rejects,
                   "fru_steps": st["fru_steps"], "launches": st["launches"],
                   "alg_bytes_per_launch": round((BYTES_PER_STEP * st["fru_steps"] + BYTES_PER_CALL *
                                                  st["simfru_calls"]) / max(st["launches"], 1), 1),
                   "peak_V_mV": round(float(V.max()), 3),
                   "peak_Cai_uM": round(float(states[:, 6].max()) * 1e3, 4)},
        "roofline": roof,
    }
    if beats is not None and len(beats):
        line["config"]["beat"] = {k: (round(v, 6) if v == v else None) for k, v in as_dicts(beats[:1])[0].items()}
    if not args.no_mixed and args.precision == "fp64":
        mw, mchunks, mst, mbeats = run_cell(args, "mixed", False)
        mbeat = as_dicts(mbeats[:1])[0] if mbeats is not None and len(mbeats) else {}
        fbeat = as_dicts(beats[:1])[0] if beats is not None and len(beats) else {}
        line["mixed"] = dict(value=round(args.steps / mw, 4), unit="simulated ms / s",
                             speedup_vs_fp64=round(wall / mw, 4),
                             fru_steps=mst["fru_steps"],
                             beat_vs_fp64={k: (round(mbeat[k] - fbeat[k], 6) if mbeat[k] == mbeat[k]
                                               and fbeat[k] == fbeat[k] else None)
                                           for k in ("APD90", "V_peak", "Cai_peak", "ICa_peak")}
                             if mbeat and fbeat else None)
    if args.batch_cells > 0:
        line["batched"] = batched(args)
    if roof is not None and not args.no_traffic and args.driver == "device" and args.precision == "fp64":
        tr = measure_traffic(args)
        if tr is not None:
            roof["traffic"] = tr["bytes_per_launch"]
            roof["traffic_detail"] = tr
    if args.cpu_ms > 0:
        ref = None
        if args.nfru == 12500 and args.precision == "fp64":
            ref = reference_cpu_baseline(float(args.warmup), float(args.warmup + args.steps), host_cores())
        if ref is None:
            ref = cpu_baseline(args.nfru, args.cpu_ms, args.seed)
            ref["value"] = round(ref["value"], 5)
        line["cpu_baseline"] = ref
        line["vs_cpu_baseline"] = round(value / ref["value"], 2)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
