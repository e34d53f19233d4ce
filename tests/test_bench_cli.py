"""bench.py's command line without a GPU: the multi-GPU refusal and the
rocprofv3 summary tables (kernel trace and counters) from CSV files of the
profiler's format."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, BENCH, "--gpus", "2"], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stderr


def test_summarize_kernel_trace_and_counters(tmp_path):
    d = tmp_path / "run"
    d.mkdir()
    (d / "p_kernel_trace.csv").write_text(
        "Kernel_Name,Grid_Size_X,Workgroup_Size_X,Start_Timestamp,End_Timestamp\n"
        "void k_dev_stage<double>(SimArgs),50432,256,1000,91000\n"
        "void k_dev_stage<double>(SimArgs),50432,256,95000,185000\n"
        "void k_dev_begin<double>(SimArgs),50176,256,0,1000\n")
    (d / "p_counter_collection.csv").write_text(
        "Kernel_Name,Grid_Size,Workgroup_Size,Counter_Name,Counter_Value\n"
        "void k_dev_stage<double>(SimArgs),50432,256,SQ_INSTS_VALU,100\n"
        "void k_dev_stage<double>(SimArgs),50432,256,SQ_INSTS_VALU,300\n")
    r = subprocess.run([sys.executable, BENCH, "--summarize", str(d)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = r.stdout
    # two dispatches of 90 us over a span of 184 us, 197 workgroups
    assert "| k_dev_stage<double> | 197 | 2 | 90.00 | 90.00 | 90.00 | 0.18 | 0.18 |" in out
    assert "| k_dev_stage<double> | 197 | 2 | 200.0 |" in out
