"""The drop-in itself: the reference's own driver program against the MI355X
library.

oracle/_ref/stoch_grw is the reference's caf.c, rk_pd54m.c, fcn.c, IC reader
and writers, linked against greenstein2002_amd/libgrw_mpi.so (mpi_master.c's
entry points over libgrw_fru.so) in place of its MPI layer and FRU kernels;
oracle/_ref/stoch is the reference's whole MPI program, unmodified, on the
host's cores.  Both start from ic/stable with the reference's own seeds
(initialize_ran.c, a = 12350232) and its compiled 1,000-FRU, 1 Hz protocol;
their output files must agree.  Both binaries are built by oracle/Makefile
where the reference tree exists.
"""
import os
import signal
import subprocess
import sys

import numpy as np
import pytest

from greenstein2002_amd.ic import write_ic_dir

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref")


def run_until(cmd, cwd, t_stop, env=None, timeout=300):
    """Run a reference driver program in cwd until its progress line for
    simulated time t_stop (caf.c:152), then stop it."""
    write_ic_dir(os.path.join(cwd, "ic", "stable"))
    p = subprocess.Popen(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                         env=env, start_new_session=True)
    seen = None
    try:
        for line in p.stdout:
            tok = line.split()
            if len(tok) == 4 and tok[2] == "1000":
                seen = float(tok[0])
                if seen >= t_stop:
                    break
    finally:
        try:
            os.killpg(p.pid, signal.SIGTERM)
        except ProcessLookupError:
            pass
        try:
            _, err = p.communicate(timeout=30)
        except subprocess.TimeoutExpired:
            os.killpg(p.pid, signal.SIGKILL)
            _, err = p.communicate()
    assert seen is not None and seen >= t_stop, (cmd, seen, err[-2000:])
    return seen


def rows(path):
    """Complete data rows of a reference output file (the tail of a file cut
    by stopping the program may be a partial row)."""
    with open(path) as f:
        lines = f.read().split("\n")[1:-1]  # header; the last piece has no newline
    return [np.array([float(x) for x in ln.split()]) for ln in lines]


def test_reference_driver_on_the_gpu_matches_reference_program(tmp_path):
    stoch, grw = os.path.join(REF, "stoch"), os.path.join(REF, "stoch_grw")
    if not (os.path.exists(stoch) and os.path.exists(grw)):
        pytest.skip("oracle/_ref binaries not built (they need the reference tree)")
    mpiexec = "/opt/conda/bin/mpiexec"
    t_stop = 80.0
    a, b = tmp_path / "ref", tmp_path / "gpu"
    a.mkdir()
    b.mkdir()
    print(f"[dropin] reference program, 4 ranks, to t = {t_stop} ms", file=sys.stderr, flush=True)
    run_until([mpiexec, "-n", "4", "stdbuf", "-oL", stoch], a, t_stop)
    print(f"[dropin] reference driver on libgrw_mpi.so, to t = {t_stop} ms", file=sys.stderr, flush=True)
    env = dict(os.environ, GRW_SEQUENTIAL_SUMS="1")
    run_until(["stdbuf", "-oL", grw], b, t_stop, env=env)
    for name, rtol in (("states.0.txt", 1e-6), ("otherstates.0.txt", 1e-5)):
        ra, rb = rows(a / name), rows(b / name)
        n = min(len(ra), len(rb))
        assert n >= 40, (name, len(ra), len(rb))
        A, B = np.array(ra[:n]), np.array(rb[:n])
        assert np.array_equal(A[:, 0], B[:, 0])
        np.testing.assert_allclose(B, A, rtol=rtol, atol=1e-12, err_msg=name)
        same = sum(np.array_equal(x, y) for x, y in zip(A, B))
        print(f"[dropin] {name}: {n} rows (t = {A[0, 0]:g}..{A[-1, 0]:g} s), {same} printed identically",
              file=sys.stderr, flush=True)
    # the action potential at 10 ms is in the compared rows
    V = np.array(rows(b / "states.0.txt"))[:, 1]
    assert V.max() > 20.0
