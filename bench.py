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
``cpu_baseline`` times the CPU oracle (the reference's algorithm, bit-exact with
its rk54pd and MPI layer) on the same cell over the same simulated window, its
FRUs on this host's cores (up to 16); ``cpu_reference_mpi`` times the
reference's own MPI program (oracle/_ref/stoch, built from its unmodified
sources, where it is built) on the same cores over the same window, at its
compiled 1,000-FRU size scaled to 12,500 FRUs.  Two further measurements share the line:

* ``mixed``: the same beat with the mixed-precision form (FP32 rates and local
  Ca right-hand side), and its speed relative to FP64;
* ``batched``: 256 independent cells in one batch, FRU counts uniform in
  [1250, 12500] and pacing cycle lengths uniform in [300, 2000] ms (host PRNG
  seed 1234, master seeds 1234 + k), over the first --batch-ms simulated ms
  (each cell's first stimulus at 10 ms included); beats/s counts simulated
  cycle lengths completed per wall second, and a sample of the same cells run
  one after another through the single-cell driver gives the sequential rate.

``roofline.avg_launch_us`` is the mean duration of the kernel over the timed
window from a kernel trace: bench.py runs the same window again as a child
under ``rocprofv3 --kernel-trace`` (``--no-trace`` skips this and falls back to
the HIP events).  ``roofline.traffic`` is the kernel's measured traffic between
L2 and memory per launch: bench.py runs itself twice more under
``rocprofv3 --pmc`` (the EA read requests by size, then the write requests) over
the first 35 simulated ms of the same cell, and twice for a 256 MiB
device-to-device copy that checks the byte formulas (``--no-traffic`` skips
this; without rocprofv3 it is null).
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
    ap.add_argument("--cpu-ms", type=int, default=20,
                    help="simulated ms of the CPU baselines' sample, t = warmup .. warmup + this (the threaded "
                         "oracle, and the reference's MPI program where it is built); the GPU run times the "
                         "same window as its first chunk and reports its rate beside them; 0 disables them")
    ap.add_argument("--no-timing", action="store_true",
                    help="do not record per-launch HIP events (for profiler runs)")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision run")
    ap.add_argument("--layout", choices=["auto", "quad", "waves", "quad3", "quad4", "wave1", "wave1_3", "quadp", "wave1_4",
                                         "refill", "refill_3"], default="auto",
                    help="Monte Carlo kernel of the device integrator (grw_set_layout); auto: the autotuner "
                         "picks it and the graph replays from timed trials (grw_autotune_cell / _batch)")
    ap.add_argument("--tune-steps", type=int, default=10, help="simulated ms of each single-cell autotuner trial")
    ap.add_argument("--graphs", type=int, default=0, choices=[0, 1],
                    help="1: the device integrator replays RK attempts from a captured HIP graph")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 counter runs that measure roofline.traffic")
    ap.add_argument("--no-trace", action="store_true",
                    help="skip the rocprofv3 kernel-trace run that measures roofline.avg_launch_us")
    ap.add_argument("--batch-cells", type=int, default=256, help="cells of the batched run (0 skips it)")
    ap.add_argument("--batch-ms", type=int, default=30, help="simulated ms of the batched run")
    ap.add_argument("--batch-tune-steps", type=int, default=12,
                    help="simulated ms of each batch autotuner trial, from t = 0 (12: through the first "
                         "stimulus at 10 ms, so that the trials see the costly phase the timed window holds)")
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
    ap.add_argument("--tune-table", default="",
                    help="only run the autotuner on single cells of these comma-separated FRU counts, both "
                         "precisions, from t = --warmup (stimulus at 10 ms), trials of --tune-steps ms, and "
                         "print every candidate's time (one JSON line)")
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


def _profiler():
    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    return prof if os.path.exists(prof) else None


def _cell_cmd(args, steps, warmup):
    """This script on the same cell, FP64, with the layout the run used,
    without the other legs and without HIP events (the child a profiler
    watches).  Always direct launches: a graph replay submits the same
    kernels, and the profiler traces directly launched ones."""
    return [sys.executable, os.path.abspath(__file__), "--steps", str(steps), "--warmup", str(warmup),
            "--nfru", str(args.nfru), "--seed", str(args.seed), "--no-mixed", "--batch-cells", "0",
            "--cpu-ms", "0", "--no-timing", "--no-traffic", "--no-trace", "--graphs", "0",
            "--layout", TUNED["fp64"]["layout"] if "fp64" in TUNED else args.layout]


def _run_profiled(prof_args, cmd, folder, timeout=300):
    env = dict(os.environ, TMPDIR=os.environ.get("TMPDIR", "/tmp"))
    r = subprocess.run([_profiler(), *prof_args, "--output-format", "csv", "-d", folder, "-o", "p", "--"] + cmd,
                       capture_output=True, text=True, timeout=timeout, env=env, start_new_session=True)
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith("{")]
    return json.loads(lines[-1]) if lines else None


def kernel_trace(args) -> dict | None:
    """Mean duration of k_dev_stage (or of k_mc_wave + k_dev_tail, the
    single-wavefront layout's two launches per stage) over the timed window, from a kernel trace:
    this script's same window (--warmup .. --warmup + --steps) run again as a
    child under rocprofv3 --kernel-trace, without HIP events.  The window's
    dispatches are the last `launches` k_dev_stage dispatches that did work
    (longer than 3 us: the host enqueues attempts ahead, and those that find
    the run finished return at once, in about 1 us)."""
    import csv
    import glob

    if not _profiler():
        return None
    with tempfile.TemporaryDirectory(prefix="grw_kt_") as tmp:
        try:
            child = _run_profiled(["--kernel-trace"], _cell_cmd(args, args.steps, args.warmup), tmp)
        except subprocess.TimeoutExpired:
            return None
        if not child:
            return None
        disp, mc, tail = [], [], []
        for path in glob.glob(os.path.join(tmp, "**", "*kernel_trace.csv"), recursive=True):
            with open(path) as f:
                for r in csv.DictReader(f):
                    ab = (int(r["Start_Timestamp"]), int(r["End_Timestamp"]))
                    if "k_dev_stage" in r["Kernel_Name"]:
                        disp.append(ab)
                    elif "k_mc_wave" in r["Kernel_Name"]:
                        mc.append(ab)
                    elif "k_dev_tail" in r["Kernel_Name"]:
                        tail.append(ab)
    if mc and not disp:
        # single-wavefront layout: a stage is a Monte Carlo launch and its tail
        # launch; its duration is the sum of the two (the gap between them left out)
        mc.sort()
        tail.sort()
        if len(mc) != len(tail):
            return None
        disp = [(a, a + (b - a) + (d - c)) for (a, b), (c, d) in zip(mc, tail)]
    disp.sort()
    real = [(a, b) for a, b in disp if b - a > 3000]
    n = int(child["config"]["launches"])
    if len(real) < n or n == 0:
        return None
    w = real[-n:]
    dur = [(b - a) / 1e3 for a, b in w]
    return dict(avg_launch_us=sum(dur) / n, launches=n, span_ms=(w[-1][1] - w[0][0]) / 1e6,
                busy_ms=sum(dur) / 1e3, noop_dispatches=len(disp) - len(real),
                child_value=child["value"])


# Per-request-size read and write counters of the L2 -> memory interface (EA):
# bytes read = 32 RDREQ_32B + 64 RDREQ_64B + 128 RDREQ_128B, bytes written =
# 64 WRREQ_64B + 32 (WRREQ - WRREQ_64B) (WRITE_SIZE's own expression).  Four
# TCC counters per pass at most, so reads and writes take one pass each.
READ_COUNTERS = ("TCC_EA0_RDREQ_32B_sum", "TCC_EA0_RDREQ_64B_sum", "TCC_EA0_RDREQ_128B_sum", "TCC_EA0_RDREQ_sum")
WRITE_COUNTERS = ("TCC_EA0_WRREQ_sum", "TCC_EA0_WRREQ_64B_sum")


def _ea_bytes(folder: str, match: str):
    """Read and written bytes per dispatch of the kernels whose name contains
    `match` (the largest-grid dispatch group), and the dispatch count."""
    def means(counter):
        m = _counter_means(folder, counter)
        sel = [(v, g) for (k, g), v in m.items() if match in k]
        return max(sel, key=lambda x: (x[1], x[0][1])) if sel else None
    got = {c: means(c) for c in READ_COUNTERS + WRITE_COUNTERS}
    return got


def measure_traffic(args) -> dict | None:
    """Bytes per launch of k_dev_stage between L2 and memory (Infinity Cache /
    HBM), from the per-request-size EA counters over t = 0..35 ms of the same
    cell (stimulus at 10 ms and the start of release): one rocprofv3 --pmc pass
    for the read requests and one for the write requests.  The same two passes
    over a 256 MiB device-to-device copy check the byte formulas: the factor
    by which they miss the copied bytes scales the cell's figures (1.0 when the
    formulas are exact).  Returns None when rocprofv3 is missing or a run fails."""
    if not _profiler():
        return None
    win_ms = 35  # from t = 0 (no warmup, so that the launch count covers every dispatch)
    cell_cmd = _cell_cmd(args, win_ms, 0)
    copy_cmd = [sys.executable, os.path.abspath(__file__), "--calibrate-copy-mib", "256"]
    nbytes = 256 * (1 << 20)
    res = {}
    with tempfile.TemporaryDirectory(prefix="grw_pmc_") as tmp:
        for tag, cmd in (("cell", cell_cmd), ("copy", copy_cmd)):
            for name, counters in (("read", READ_COUNTERS), ("write", WRITE_COUNTERS)):
                d = os.path.join(tmp, f"{tag}_{name}")
                try:
                    child = _run_profiled(["--pmc", *counters], cmd, d, timeout=240)
                except subprocess.TimeoutExpired:
                    return None
                if tag == "cell":
                    if not child:
                        return None
                    res["cell_line"] = child
                vals = {}
                for c in counters:
                    m = _counter_means(d, c)
                    key = "k_dev_stage" if tag == "cell" else "copyBuffer"
                    sel = [(v, g) for (k, g), v in m.items() if key in k]
                    if not sel:
                        return None
                    (mean, ndisp), grid = max(sel, key=lambda x: (x[0][1], x[1]))
                    vals[c] = (mean, ndisp, grid)
                res[f"{tag}_{name}"] = vals
    def read_bytes(v):
        return (32.0 * v["TCC_EA0_RDREQ_32B_sum"][0] + 64.0 * v["TCC_EA0_RDREQ_64B_sum"][0] +
                128.0 * v["TCC_EA0_RDREQ_128B_sum"][0])

    def write_bytes(v):
        w64 = v["TCC_EA0_WRREQ_64B_sum"][0]
        return 64.0 * w64 + 32.0 * (v["TCC_EA0_WRREQ_sum"][0] - w64)

    f_read = nbytes / max(read_bytes(res["copy_read"]), 1.0)
    f_write = nbytes / max(write_bytes(res["copy_write"]), 1.0)
    cfg = res["cell_line"]["config"]
    launches = cfg.get("launches") or 0
    if not launches:
        return None
    ndisp = res["cell_read"]["TCC_EA0_RDREQ_sum"][1]
    rd = read_bytes(res["cell_read"]) * f_read * ndisp / launches
    wr = write_bytes(res["cell_write"]) * f_write * res["cell_write"]["TCC_EA0_WRREQ_sum"][1] / launches
    v = res["cell_read"]
    unsplit = v["TCC_EA0_RDREQ_sum"][0] - (v["TCC_EA0_RDREQ_32B_sum"][0] + v["TCC_EA0_RDREQ_64B_sum"][0] +
                                           v["TCC_EA0_RDREQ_128B_sum"][0])
    alg = cfg.get("alg_bytes_per_launch")
    return dict(bytes_per_launch=round(rd + wr, 1), read_per_launch=round(rd, 1), written_per_launch=round(wr, 1),
                window="t = 0..%d ms of the same cell (%d launches)" % (win_ms, launches),
                alg_bytes_per_launch_window=alg,
                ratio_to_algorithmic=round((rd + wr) / alg, 3) if alg else None,
                copy_factor=dict(read=round(f_read, 4), write=round(f_write, 4)),
                read_requests_per_dispatch={k.replace("TCC_EA0_", "").replace("_sum", ""): round(x[0], 1)
                                            for k, x in v.items()},
                read_requests_not_in_size_counts=round(unsplit, 1),
                dispatches=ndisp)


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
        ends = defaultdict(list)
        with open(path) as f:
            for r in csv.DictReader(f):
                grid = int(r["Grid_Size_X"]) // max(int(r["Workgroup_Size_X"]), 1)
                a, b = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
                acc[(short(r["Kernel_Name"]), grid)].append((b - a) / 1e3)
                ends[(short(r["Kernel_Name"]), grid)].append((a, b))
        print(f"\n{os.path.relpath(path, folder)}\n")
        # span: first start to last end of the group's dispatches
        print("| kernel | workgroups | dispatches | mean us | min us | max us | total ms | span ms |")
        print("|---|---|---|---|---|---|---|---|")
        for (k, g), v in sorted(acc.items(), key=lambda kv: -sum(kv[1])):
            e = ends[(k, g)]
            span = (max(b for _, b in e) - min(a for a, _ in e)) / 1e6
            print(f"| {k} | {g} | {len(v)} | {sum(v) / len(v):.2f} | {min(v):.2f} | {max(v):.2f} | "
                  f"{sum(v) / 1e3:.2f} | {span:.2f} |")
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
                    if t % 5.0 == 0.0:
                        print(f"[bench] reference MPI program: t = {t:g} ms simulated", file=sys.stderr, flush=True)
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
                       f"t = {t0:g}..{t1:g} ms (the CPU oracle's window): {rate:.3f} simulated ms/s at 1,000 FRUs "
                       f"in {wall:.2f} s, scaled by 1,000/12,500 to the 12,500-FRU cell (its run time is "
                       f"per-FRU work: >= 95% in the hot path)",
                value_1000fru=round(rate, 4), wall_s=round(wall, 3))


def cpu_baseline(nfru: int, t0: int, t1: int, seed: int, threads: int) -> dict:
    """The CPU oracle (oracle/libgrw_oracle.so, the reference's algorithm,
    bit-exact with it; its FRUs within a sub-interval on `threads` OpenMP
    threads) on the same cell, run from ic/stable and timed over simulated
    t0..t1 ms (a bounded sample of the GPU line's workload)."""
    import ctypes as C

    import numpy as np

    from greenstein2002_amd.ic import load_stable
    from oracle.pyoracle import FRU_DTYPE, Oracle, OracleCell

    o = Oracle()
    try:
        C.CDLL("libgomp.so.1").omp_set_num_threads(int(threads))
    except OSError:
        threads = 1
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
    for k in range(t0):
        h, _ = cell.rk54pd(float(k), float(k + 1), h, state, cur)
    steps0 = int(cell.c.n_fru_steps)
    ts = time.perf_counter()
    for k in range(t0, t1):
        h, _ = cell.rk54pd(float(k), float(k + 1), h, state, cur)
        if (k + 1 - t0) % 5 == 0:
            print(f"[bench] CPU oracle: t = {k + 1} ms simulated, {time.perf_counter() - ts:.1f} s wall",
                  file=sys.stderr, flush=True)
    dt = time.perf_counter() - ts
    steps = int(cell.c.n_fru_steps) - steps0
    cell.close()
    return dict(value=round((t1 - t0) / dt, 5), unit="simulated ms per wall-second", cores=int(threads),
                kind="port",
                sample=f"the CPU oracle (oracle/grw_oracle.c, the reference's algorithm, bit-exact with its "
                       f"rk54pd and MPI layer) on the same {nfru}-FRU cell and seed, FRUs of each "
                       f"sub-interval on {threads} OpenMP threads, timed over t = {t0}..{t1} ms (the stimulus "
                       f"at 10 ms, upstroke and start of release; the GPU's rate over the same window is "
                       f"gpu_value_same_window): {steps} FRU-steps in {dt:.2f} s",
                fru_steps=steps, wall_s=round(dt, 3))


def host_cores() -> int:
    """Cores this process may use, capped at 16 (one GPU's share of the host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


# autotuner results of this run, by leg ("fp64", "mixed", "batch")
TUNED: dict = {}
# (simulated ms, wall s) of each timed chunk of the single-cell runs, by precision
RUN_WALLS: dict = {}


def run_cell(args, precision, timing):
    """The single-cell beat: warmup, then args.steps ms timed in 100 ms chunks."""
    from greenstein2002_amd.cell import Cell, Protocol

    prot = Protocol()
    cell = Cell(args.nfru, seed=args.seed, precision=precision, protocol=prot, block=args.block,
                lanes_per_fru=args.lanes)
    cell.fru.set_graphs(bool(args.graphs))
    if args.driver == "device" and args.layout != "auto":
        cell.set_layout(args.layout)
    if args.warmup:
        cell.run(args.warmup, outputs=False, log_cap=0, driver=args.driver)
    if args.driver == "device" and args.layout == "auto":
        # timed trials from the current state, which the tuner restores
        TUNED[precision] = cell.autotune(args.tune_steps)
        print(f"[bench] {precision}: autotuner chose {TUNED[precision]['layout']}, graphs "
              f"{TUNED[precision]['graphs']}", file=sys.stderr, flush=True)
    cell.fru.set_timing(timing)
    cell.fru.reset_stats()
    chunks = []
    walls = []
    t0 = time.perf_counter()
    done = 0
    # the first chunk is the CPU baselines' window (t = warmup .. warmup + cpu_ms)
    first = args.cpu_ms if 0 < args.cpu_ms < args.steps else 0
    while done < args.steps:
        n = first if (done == 0 and first) else min(100, args.steps - done)
        tc = time.perf_counter()
        chunks.append(cell.run(n, log_cap=0, driver=args.driver))
        walls.append((n, time.perf_counter() - tc))
        done += n
        print(f"[bench] {precision}: t = {args.warmup + done} ms simulated, {time.perf_counter() - t0:.1f} s wall",
              file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    st = cell.fru.stats()
    beats = cell.beats() if args.driver == "device" else None
    cell.close()
    RUN_WALLS[precision] = walls
    return wall, chunks, st, beats


def batch_layout(b, args, tag: str) -> str:
    """The batch's Monte Carlo layout: the autotuner's choice from trials of
    --batch-tune-steps ms (--layout auto), else the requested one (a
    single-cell-only layout leaves the default)."""
    if args.layout == "auto":
        info = b.autotune(args.batch_tune_steps)
        print(f"[bench] {tag}: autotuner chose {info['layout']}, graphs {info['graphs']}", file=sys.stderr, flush=True)
        return info["layout"] + (" + graphs" if info["graphs"] else "")
    if args.layout in ("waves", "quadp"):
        return "quad"
    b.set_layout(args.layout)
    return args.layout


def stage_kernels(layout: str, prec: str) -> str:
    """The kernels of one RK stage under a layout (what a stage's timed span covers)."""
    t = "double" if prec == "fp64" else "float"
    if layout in ("wave1", "wave1_3", "wave1_4"):
        return "k_mc_wave%s<%s> + k_dev_tail<%s> (one RK stage)" % ("_occ" if layout != "wave1" else "", t, t)
    if layout in ("refill", "refill_3"):
        return "k_mc_refill%s<%s> + k_dev_tail<%s> (one RK stage)" % ("3" if layout == "refill_3" else "", t, t)
    if layout == "quadp":
        return "k_dev_stage_poll<%s>" % t
    return "k_dev_stage<%s>" % t


def batched(args, precision: str | None = None, sequential: bool = True):
    """256 uneven cells in one batch vs the same cells one at a time (the
    sequential sample only when `sequential`)."""
    import numpy as np

    from greenstein2002_amd.batch import Batch, CellSpec
    from greenstein2002_amd.cell import Cell, Protocol

    n = args.batch_cells
    rng = np.random.default_rng(1234)
    nf = rng.integers(1250, 12501, n)
    cl = rng.uniform(300.0, 2000.0, n)
    specs = [CellSpec(int(nf[k]), seed=1234 + k, protocol=Protocol(period=float(cl[k]))) for k in range(n)]
    prec = precision or args.precision
    key = "batch" if prec == args.precision else "batch_" + prec
    b = Batch(specs, precision=prec)
    b.set_graphs(bool(args.graphs))
    if args.layout == "auto":
        TUNED[key] = b.autotune(args.batch_tune_steps)
        print(f"[bench] batch ({prec}): autotuner chose {TUNED[key]['layout']}, graphs {TUNED[key]['graphs']}",
              file=sys.stderr, flush=True)
    elif args.layout not in ("waves", "quadp"):
        b.set_layout(args.layout)
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
    pick = np.unique(np.linspace(0, n - 1, max(1, min(args.batch_seq_sample, n))).astype(int)) if sequential else []
    seq = []
    for k in pick:
        c = Cell(int(nf[k]), seed=1234 + int(k), precision=prec, protocol=Protocol(period=float(cl[k])))
        if prec in TUNED:  # the single cell's tuned settings
            c.set_layout(TUNED[prec]["layout"])
            c.fru.set_graphs(TUNED[prec]["graphs"])
        else:
            c.fru.set_graphs(bool(args.graphs))
        ts = time.perf_counter()
        c.run(args.batch_ms, outputs=False, log_cap=0)
        seq.append(time.perf_counter() - ts)
        c.close()
        print(f"[bench] batch cell {k} alone: {seq[-1]:.2f} s wall", file=sys.stderr, flush=True)
    seq_wall = float(np.mean(seq)) * n if seq else None
    beats = float(np.sum(args.batch_ms / cl))
    # whole beats inside the window: beat k of a cell runs from its stimulus
    # at 10 + k CL ms to the next one
    done_beats = int(np.sum(np.floor(np.maximum(args.batch_ms - 10.0, 0.0) / cl)))
    return dict(
        workload=f"{n} cells, FRUs uniform in [1250, 12500] (total {int(nf.sum())}), cycle lengths uniform "
                 f"in [300, 2000] ms, t = 0..{args.batch_ms} ms from ic/stable, {prec}",
        cells=n, frus=int(nf.sum()), window_ms=args.batch_ms, wall_s=round(wall, 3),
        cell_ms_per_s=round(n * args.batch_ms / wall, 2),
        beats_per_s_est=round(beats / wall, 4),
        beats_completed=done_beats,
        beats_completed_per_s=round(done_beats / wall, 4),
        beats_per_s_basis=("estimate, not a count of finished beats: each cell's simulated ms over the window "
                           "divided by its cycle length, over the wall time; the window holds the stimulus, "
                           "upstroke and start of release, which cost more per simulated ms than the rest of "
                           "a beat, so whole beats run faster than this"),
        layout=TUNED[key]["layout"] if key in TUNED else ("quad" if args.layout in ("waves", "quadp") else args.layout),
        hip_graphs=bool(TUNED[key]["graphs"]) if key in TUNED else bool(args.graphs),
        autotune=TUNED.get(key),
        rk_attempts=att, fru_steps=st["fru_steps"],
        fru_steps_per_s=round(st["fru_steps"] / wall, 1),
        roofline=dict(bound="hbm", achieved=round(achieved, 3), peak=HBM_PEAK_GBS, unit="GB/s",
                      frac=round(achieved / HBM_PEAK_GBS, 6),
                      kernel=stage_kernels(TUNED[key]["layout"] if key in TUNED else args.layout, prec) +
                      " over all cells",
                      avg_launch_us=round(avg_s * 1e6, 3),
                      launches=st["launches"], timed_launches=st["timed_launches"],
                      lane_use=round(st["fru_steps"] / st["slot_steps"], 4) if st["slot_steps"] else None),
        sequential=dict(sample_cells=[int(k) for k in pick], sample_walls_s=[round(x, 3) for x in seq],
                        est_wall_s=round(seq_wall, 2), beats_per_s_est=round(beats / seq_wall, 4)) if seq else None,
        speedup_vs_sequential=round(seq_wall / wall, 2) if seq else None)


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
    res, walls, layouts = {}, {}, {}
    for prec in ("fp64", "mixed"):
        specs = [CellSpec(args.nfru, seed=int(s)) for s in seeds]
        b = Batch(specs, precision=prec)
        layouts[prec] = batch_layout(b, args, f"ensemble {prec}")
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
               wall_s={k: round(v, 2) for k, v in walls.items()}, layout=layouts, stats={},
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
    res, walls, layouts = {}, {}, {}
    for prec in ("fp64", "mixed"):
        b = Batch(specs, precision=prec)
        layouts[prec] = batch_layout(b, args, f"clamp {prec}")
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
               wall_s={k: round(v, 2) for k, v in walls.items()}, layout=layouts, rows=[])
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
        lay = batch_layout(b, args, f"sweep {prec}")
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
        line = dict(precision=prec, cells=len(cells), layout=lay, wall_s=round(wall, 2),
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


def tune_table(args) -> dict:
    """The autotuner's trials (grw_autotune_cell) on single cells of several
    FRU counts: every layout / graph candidate's wall time for the same
    --tune-steps simulated ms from t = --warmup, and the choice."""
    from greenstein2002_amd.cell import Cell

    rows = []
    for n in [int(x) for x in args.tune_table.split(",") if x]:
        for prec in ("fp64", "mixed"):
            with Cell(n, seed=args.seed, precision=prec) as c:
                if args.warmup:
                    c.run(args.warmup, outputs=False, log_cap=0)
                info = c.autotune(args.tune_steps)
            by = {(k["layout"], k["graphs"]): k["ms"] for k in info["candidates"]}
            quad = by[("quad", False)]
            best = min(by.values())
            rows.append(dict(nfru=n, precision=prec, chosen=info["layout"], graphs=info["graphs"],
                             candidates_ms={f"{a}{'+graphs' if g else ''}": v for (a, g), v in by.items()},
                             fastest_vs_quad=round(quad / best, 3)))
            print(f"[bench] tune table: {n} FRUs {prec}: {info['layout']} (x{quad / best:.3f} over quad)",
                  file=sys.stderr, flush=True)
    return dict(table="autotuner trials", window_ms=[args.warmup, args.warmup + args.tune_steps], rows=rows)


def main():
    args = parse()
    if args.tune_table:
        print(json.dumps(tune_table(args)), flush=True)
        return
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
        # HIP events (hipExtLaunchKernelGGL start / stop) on one RK attempt in
        # eight; an event on a kernel's completion makes that kernel end with a
        # system-scope release, so the sampled launches run longer than the
        # others.  The kernel trace of the same window, without events, gives
        # the duration the roofline uses.
        ev_s = st["kernel_ms"] / 1e3 / st["timed_launches"]
        kt = None
        if not args.no_trace and args.driver == "device":
            kt = kernel_trace(args)
        avg_s = kt["avg_launch_us"] / 1e6 if kt else ev_s
        achieved = alg_bytes / st["launches"] / avg_s / 1e9
        roof = dict(bound="hbm", achieved=round(achieved, 3), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 6), traffic=None,
                    kernel=((stage_kernels(TUNED["fp64"]["layout"] if "fp64" in TUNED else args.layout,
                                           args.precision) +
                             " (FRU sub-interval + flux reduction + integrator stage)")
                            if args.driver == "device" else
                            ("k_simfru4" if args.lanes == 4 else "k_simfru") +
                            ("<double>" if args.precision == "fp64" else "<float>") +
                            " (FRU sub-interval + fused flux reduction)"),
                    avg_launch_us=round(avg_s * 1e6, 3),
                    avg_launch_us_source=("rocprofv3 --kernel-trace of this window (child run, no events)"
                                          if kt else "HIP events, one RK attempt in eight"),
                    avg_launch_us_events=round(ev_s * 1e6, 3),
                    launches=st["launches"], timed_launches=st["timed_launches"], fru_steps=st["fru_steps"],
                    critical_steps_per_launch=round(st["critical_steps"] / st["launches"], 2),
                    us_per_critical_step=round(avg_s * 1e6 * st["launches"] / max(st["critical_steps"], 1), 3),
                    serial_tail_us=round(st["tail_ms"] * 1e3 / st["launches"], 3),
                    lane_use=round(st["fru_steps"] / st["slot_steps"], 4) if st["slot_steps"] else None,
                    alg_bytes_per_launch=round(alg_bytes / st["launches"], 1))
        if kt:
            roof["kernel_busy_ms"] = round(kt["busy_ms"], 3)
            roof["window_span_ms"] = round(kt["span_ms"], 3)
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
                   "hip_graphs": bool(TUNED["fp64"]["graphs"]) if "fp64" in TUNED else bool(args.graphs),
                   "layout": TUNED["fp64"]["layout"] if "fp64" in TUNED else args.layout,
                   "autotune": TUNED.get("fp64"),
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
                             layout=TUNED["mixed"]["layout"] if "mixed" in TUNED else args.layout,
                             hip_graphs=bool(TUNED["mixed"]["graphs"]) if "mixed" in TUNED else bool(args.graphs),
                             beat_vs_fp64={k: (round(mbeat[k] - fbeat[k], 6) if mbeat[k] == mbeat[k]
                                               and fbeat[k] == fbeat[k] else None)
                                           for k in ("APD90", "V_peak", "Cai_peak", "ICa_peak")}
                             if mbeat and fbeat else None)
    if args.batch_cells > 0:
        line["batched"] = batched(args)
        if args.precision == "fp64" and not args.no_mixed:
            bm = batched(args, "mixed", sequential=False)
            line["batched"]["mixed"] = dict(cell_ms_per_s=bm["cell_ms_per_s"], beats_per_s_est=bm["beats_per_s_est"],
                                            fru_steps_per_s=bm["fru_steps_per_s"], wall_s=bm["wall_s"],
                                            layout=bm["layout"], hip_graphs=bm["hip_graphs"],
                                            speedup_vs_fp64=round(bm["cell_ms_per_s"] /
                                                                  line["batched"]["cell_ms_per_s"], 4))
    if roof is not None and not args.no_traffic and args.driver == "device" and args.precision == "fp64":
        tr = measure_traffic(args)
        if tr is not None:
            roof["traffic"] = tr["bytes_per_launch"]
            roof["traffic_detail"] = tr
    if args.cpu_ms > 0:
        cores = host_cores()
        w0, w1 = args.warmup, args.warmup + min(args.cpu_ms, args.steps)
        line["cpu_baseline"] = cpu_baseline(args.nfru, w0, w1, args.seed, cores)
        # the GPU's rate over the same window: the timed run's first chunk, or
        # the whole timed run when the window covers it
        if args.cpu_ms < args.steps:
            n0, s0 = RUN_WALLS[args.precision][0]
            gpu_same = n0 / s0
        else:
            gpu_same = value
        line["cpu_baseline"]["gpu_value_same_window"] = round(gpu_same, 4)
        line["vs_cpu_baseline"] = round(gpu_same / line["cpu_baseline"]["value"], 2)
        if args.nfru == 12500 and args.precision == "fp64":
            runs = [reference_cpu_baseline(float(w0), float(w1), cores) for _ in range(3)]
            runs = [r for r in runs if r]
            if runs:
                ref = sorted(runs, key=lambda r: r["value"])[len(runs) // 2]
                ref["repeats"] = [r["value_1000fru"] for r in runs]
                ref["gpu_value_same_window"] = round(gpu_same, 4)
                line["cpu_reference_mpi"] = ref
                line["vs_cpu_reference_mpi"] = round(gpu_same / ref["value"], 2)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
