"""bench.py's command-line contract on the CPU: a multi-GPU request is refused
with a message, and the CPU-side helpers give the documented figures."""
import os
import subprocess
import sys

import numpy as np

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


def _write_counter_csv(path, counter, rows):
    hdr = ("Correlation_Id,Dispatch_Id,Agent_Id,Queue_Id,Process_Id,Thread_Id,Grid_Size,Kernel_Id,Kernel_Name,"
           "Workgroup_Size,LDS_Block_Size,Scratch_Size,VGPR_Count,Accum_VGPR_Count,SGPR_Count,Counter_Name,"
           "Counter_Value,Start_Timestamp,End_Timestamp\n")
    with open(path, "w") as f:
        f.write(hdr)
        for i, (name, grid, val) in enumerate(rows):
            f.write('%d,%d,"Agent 1",1,1,1,%d,1,"%s",512,0,0,64,0,80,"%s",%f,0,1\n'
                    % (i + 1, i + 1, grid, name, counter, val))


def test_traffic_summary_per_step(tmp_path):
    """HBM bytes per MC step: two equal half-grid dispatches summed, each counter
    scaled by its reading for a 256 MiB copy (fetch reads half: x2)."""
    sys.path.insert(0, ROOT)
    import json
    import bench
    pair, copy = "void gd_pair_kernel<4>(GdJob const*, GdItem const*, int)", "gd_copy_kernel(float4)"
    dirs = {k: tmp_path / k for k in ("f", "w", "cf", "cw")}
    for d in dirs.values():
        d.mkdir()
    _write_counter_csv(dirs["f"] / "c2_counter_collection.csv", "FETCH_SIZE",
                       [(pair, 1000, 100.0), (pair, 1000, 100.0), (pair, 2000, 150.0)])
    _write_counter_csv(dirs["w"] / "c2_counter_collection.csv", "WRITE_SIZE",
                       [(pair, 1000, 50.0), (pair, 1000, 50.0), (pair, 2000, 100.0)])
    _write_counter_csv(dirs["cf"] / "copy_counter_collection.csv", "FETCH_SIZE", [(copy, 64, 131072.0)])
    _write_counter_csv(dirs["cw"] / "copy_counter_collection.csv", "WRITE_SIZE", [(copy, 64, 262144.0)])
    out = tmp_path / "t.json"
    r = bench.traffic_summary(str(dirs["f"]), str(dirs["w"]), str(dirs["cf"]), str(dirs["cw"]), str(out))
    assert abs(r["fetch_scale"] - 2.0) < 1e-12 and abs(r["write_scale"] - 1.0) < 1e-12
    half = (100.0 * 2.0 + 50.0) * 1024.0
    assert abs(r["hbm_bytes_per_launch"] - 2 * half) < 1e-6
    assert abs(r["hbm_bytes_whole_launch"] - (150.0 * 2.0 + 100.0) * 1024.0) < 1e-6
    assert json.load(open(out))["kernel"] == pair


def test_trace_summary_two_streams(tmp_path):
    """Step period from the first stream's half-launch starts; span of a step's
    two halves on different streams; whole-grid dispatches apart."""
    sys.path.insert(0, ROOT)
    import bench
    pair = "void gd_pair_kernel<4>(GdJob const*, GdItem const*, int)"
    rows = []
    cid = 0
    for k in range(5):                       # 5 two-stream steps, period 100, halves 60 long, offset 30
        t = 1000 + 100 * k
        for sid, t0 in ((1, t), (2, t + 30)):
            cid += 1
            rows.append((cid, sid, 500, t0, t0 + 60))
    cid += 1
    rows.append((cid, 1, 1000, 2000, 2090))  # a whole-grid swap pass
    path = tmp_path / "kt.csv"
    with open(path, "w") as f:
        f.write("Kernel_Name,Correlation_Id,Stream_Id,Grid_Size_X,Start_Timestamp,End_Timestamp\n")
        for c, sid, grid, a, b in rows:
            f.write('"%s",%d,%d,%d,%d,%d\n' % (pair, c, sid, grid, a * 1000, b * 1000))
    r = bench.trace_summary(str(path))["pair"]
    assert r["two_stream_steps"] == 5 and r["whole_job_dispatches"] == 1
    assert abs(r["two_stream_step_period_us"] - 100.0) < 1e-9
    assert abs(r["two_stream_step_span_mean_us"] - 90.0) < 1e-9
    assert abs(r["half_dispatch_mean_us"] - 60.0) < 1e-9 and abs(r["whole_job_mean_us"] - 90.0) < 1e-9


def test_fifo_latencies():
    """The plain server's latencies: one job at a time, in arrival order."""
    import bench
    np.testing.assert_allclose(bench.fifo_latencies([0, 0, 0], [1, 2, 3]), [1, 3, 6])    # a queue builds
    np.testing.assert_allclose(bench.fifo_latencies([0, 10], [1, 1]), [1, 1])            # idle in between
    np.testing.assert_allclose(bench.fifo_latencies([0, 0.5, 5], [2, 2, 1]), [2, 3.5, 1])
