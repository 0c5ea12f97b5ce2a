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


def test_rates_of_a_timed_run():
    """2,000 MC steps of the 1,760-replica job in 0.259 s: 1.359e7
    replica-steps/s and 0.2 jobs in 0.259 s = 0.772 jobs/s (one job is the
    whole 10,000-step run)."""
    sys.path.insert(0, ROOT)
    import bench
    rs, jobs = bench.c2_rates(1760, 2000, 0.259)
    assert abs(rs - 1760 * 2000 / 0.259) < 1e-6 * rs
    assert abs(jobs - 0.2 / 0.259) < 1e-12
    rs, jobs = bench.c2_rates(1760, 10000, 1.3)   # the whole job
    assert abs(jobs - 1 / 1.3) < 1e-12


def test_workload_names_the_steps_timed():
    sys.path.insert(0, ROOT)
    import bench
    w = bench.c2_workload(2000)
    assert "2000 MC steps of the 10000-step job timed" in w
    assert "1760 replicas" in w and "P=3000" in w and "L=40" in w


def test_cpu_leg_describes_its_cores():
    sys.path.insert(0, ROOT)
    import bench
    logical, phys, phys_all = bench.host_cores()
    assert 1 <= phys <= logical and phys <= phys_all
    f = bench.cpu_fields(bench.cpu_threads())
    assert 1 <= f["cores"] <= min(16, phys)
    assert f["cpu_model"] and f["cpu_model"] != "unknown"


def test_batched_workload_shapes():
    """The batched legs' complexes: C1's shape (P=1200 x 5, L=24 x 5, one
    temperature: 25 replicas) for c1b, one 4,000-point target with 12-80-atom
    ligands x 440 replicas for c3, logU[1500, 9000] proteins for c4."""
    import numpy as np
    sys.path.insert(0, ROOT)
    import bench
    z = bench._sizes(np.random.default_rng(0), 4, "c1b")
    assert all((d["P"], d["L"], d["nlig"] * d["nprt"] * d["ntmp"]) == (1200, 24, 25) for d in z)
    z = bench._sizes(np.random.default_rng(0), 64, "c3")
    assert all(d["P"] == 4000 and 12 <= d["L"] <= 80 and d["nlig"] * d["nprt"] * d["ntmp"] == 440 for d in z)
    z = bench._sizes(np.random.default_rng(0), 64, "c4")
    assert all(1500 <= d["P"] <= 9000 and d["nlig"] * d["nprt"] * d["ntmp"] == 220 for d in z)
