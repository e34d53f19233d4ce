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
FRU Monte Carlo sub-interval kernel) from HIP events recorded around every
launch on the library's stream; ``cpu_baseline`` times the CPU oracle (a
single-threaded restatement of the reference, bit-exact with it) on a bounded
sample of the same workload on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
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
                    help="simulated ms of the CPU-oracle sample (0 disables)")
    ap.add_argument("--no-timing", action="store_true",
                    help="do not record per-launch HIP events (for profiler runs)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="only run a device-to-device copy of this many MiB (5 times): the "
                         "reference kernel for scaling FETCH_SIZE / WRITE_SIZE counters")
    return ap.parse_args()


def calibrate_copy(mib: int) -> None:
    """A plain copy kernel of known size whose HBM counters calibrate the
    profiler's FETCH_SIZE / WRITE_SIZE (reads and writes mib MiB per launch)."""
    import torch

    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "copy", "bytes_read_per_launch": n * 4,
                      "bytes_written_per_launch": n * 4, "launches": 5}), flush=True)


def cpu_baseline(nfru: int, ms: float, seed: int):
    """The CPU oracle (oracle/libgrw_oracle.so, one thread) on the same cell
    for the first `ms` simulated ms from ic/stable."""
    import numpy as np

    from greenstein2002_amd.ic import errweight, load_stable
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
                sample=f"{nfru}-FRU cell, t = 0..{n} ms from ic/stable (diastole before the "
                       f"stimulus, the cheapest phase of the beat), CPU oracle single-threaded, "
                       f"{steps} FRU-steps in {dt:.1f} s = {dt / max(steps, 1) * 1e9:.0f} ns/FRU-step")


def main():
    args = parse()
    if args.calibrate_copy_mib:
        calibrate_copy(args.calibrate_copy_mib)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported on this "
              "single-GPU build (independent cells would shard across GPUs with no exchange)",
              file=sys.stderr)
        sys.exit(2)

    from greenstein2002_amd.cell import Cell, Protocol

    prot = Protocol()
    cell = Cell(args.nfru, seed=args.seed, precision=args.precision, protocol=prot, block=args.block,
                lanes_per_fru=args.lanes)
    if args.warmup:
        cell.run(args.warmup, outputs=False, log_cap=0, driver=args.driver)
    if not args.no_timing:
        cell.fru.set_timing(True)
    cell.fru.reset_stats()
    import numpy as np

    chunks = []
    t0 = time.perf_counter()
    done = 0
    while done < args.steps:
        n = min(100, args.steps - done)
        chunks.append(cell.run(n, log_cap=0, driver=args.driver))
        done += n
        print(f"[bench] t = {args.warmup + done} ms simulated, {time.perf_counter() - t0:.1f} s wall",
              file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    states = np.concatenate([c.states for c in chunks])
    attempts = sum(c.stats["attempts"] for c in chunks)
    rejects = sum(c.stats["rejects"] for c in chunks)
    st = cell.fru.stats()
    cell.close()

    value = args.steps * prot.time_step / wall
    roof = None
    if not args.no_timing and st["launches"]:
        alg_bytes = BYTES_PER_STEP * st["fru_steps"] + BYTES_PER_CALL * st["simfru_calls"]
        kernel_s = st["kernel_ms"] / 1e3
        achieved = alg_bytes / kernel_s / 1e9
        roof = dict(bound="hbm", achieved=round(achieved, 3), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 6), traffic=None,
                    kernel=(("k_dev_stage" if args.driver == "device" else
                             "k_simfru4" if args.lanes == 4 else "k_simfru") +
                            ("<double>" if args.precision == "fp64" else "<float>") +
                            " (FRU sub-interval + fused flux reduction" +
                            (" + integrator stage)" if args.driver == "device" else ")")),
                    avg_launch_us=round(st["kernel_ms"] / st["launches"] * 1e3, 3),
                    launches=st["launches"], fru_steps=st["fru_steps"],
                    critical_steps_per_launch=round(st["critical_steps"] / st["launches"], 2),
                    us_per_critical_step=round(st["kernel_ms"] * 1e3 / max(st["critical_steps"], 1), 3),
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
                   "driver": args.driver,
                   "rk_attempts": attempts, "rk_rejects": rejects,
                   "fru_steps": st["fru_steps"], "peak_V_mV": round(float(V.max()), 3),
                   "peak_Cai_uM": round(float(states[:, 6].max()) * 1e3, 4)},
        "roofline": roof,
    }
    if args.cpu_ms > 0:
        line["cpu_baseline"] = cpu_baseline(args.nfru, args.cpu_ms, args.seed)
        line["cpu_baseline"]["value"] = round(line["cpu_baseline"]["value"], 5)
        line["vs_cpu_baseline"] = round(value / line["cpu_baseline"]["value"], 2)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
