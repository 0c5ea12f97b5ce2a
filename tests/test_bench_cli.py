"""bench.py's command-line contract on the CPU: a multi-GPU request is refused
with a message, and the CPU-side helpers give the documented figures."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_multi_gpu_request_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


def test_algorithmic_flops_of_the_benchmark_job():
    """The credited flops per pair launch: 62 per PRT pair, 11 per KDE match,
    9 per MCS entry, summed over the job's replicas (about 1.39e10 flops)."""
    sys.path.insert(0, ROOT)
    import bench
    import geauxdock_amd as g
    job = g.synth_complex(bench.P, bench.L, bench.NLIG, bench.NPRT, bench.NTMP, floor_temp=0.004,
                          ceiling_temp=0.04, seed=bench.SEED)
    flops, pairs = bench.algorithmic_flops_per_step(job)
    assert pairs == job.n_rep * bench.L * bench.P
    assert 62 * pairs < flops < 1.1 * 62 * pairs
    assert abs(flops - 1.386e10) / 1.386e10 < 0.01
