"""bench.py's command line on the CPU: this build measures one GPU, and asking for more exits with a
message instead of running."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_workloads_are_listed():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0
    for w in ("vnoise8192", "spheres512", "batch256"):
        assert w in r.stdout


def test_step_counts_are_checked():
    for argv in (["--steps", "0"], ["--warmup", "-1"]):
        r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py")] + argv, capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 2, argv
        assert "--steps must be at least 1" in r.stderr


def test_hw_queue_count_is_bounded():
    """the HIP runtime's hardware queues are requested before it starts, 0..32 (0 keeps the environment's)"""
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--hw-queues", "40"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "--hw-queues must be between 0 and 32" in r.stderr


def _pmc_csv(path, counter, rows):
    """a rocprofv3 --pmc counter_collection.csv with the columns the summary reads"""
    os.makedirs(path, exist_ok=True)
    with open(os.path.join(path, "p_counter_collection.csv"), "w") as f:
        f.write('"Dispatch_Id","Grid_Size","Kernel_Name","Counter_Name","Counter_Value"\n')
        for did, name, grid, v in rows:
            f.write('%d,%d,"%s","%s",%f\n' % (did, grid, name, counter, v))


def test_traffic_summary(tmp_path):
    """bench.py --traffic: the kernel's FETCH_SIZE / WRITE_SIZE per dispatch (KiB, summed over a dispatch's rows),
    scaled by what the counters read for the plain copy of --calibrate-copy MiB, written as bytes per launch"""
    d = str(tmp_path)
    # the copy of 4 MiB: FETCH_SIZE reads half of its 4096 KiB (scale 2), WRITE_SIZE all of it (scale 1)
    _pmc_csv(os.path.join(d, "cf"), "FETCH_SIZE", [(1, "fill", 64, 3.0), (2, "__amd_rocclr_copyBuffer", 64, 2048.0)])
    _pmc_csv(os.path.join(d, "cw"), "WRITE_SIZE", [(1, "fill", 64, 3.0), (2, "__amd_rocclr_copyBuffer", 64, 4096.0)])
    # two launches of the kernel (one split over two rows, as per-XCD rows are), one other kernel to ignore
    _pmc_csv(os.path.join(d, "kf"), "FETCH_SIZE", [(5, "void k_final_pixels<0, false>(LevelArgs)", 512, 60.0),
                                                   (5, "void k_final_pixels<0, false>(LevelArgs)", 512, 40.0),
                                                   (6, "void k_final_pixels<0, false>(LevelArgs)", 256, 50.0),
                                                   (7, "k_fix(LevelArgs)", 256, 999.0)])
    _pmc_csv(os.path.join(d, "kw"), "WRITE_SIZE", [(5, "void k_final_pixels<0, false>(LevelArgs)", 512, 10.0),
                                                   (6, "void k_final_pixels<0, false>(LevelArgs)", 256, 30.0),
                                                   (7, "k_fix(LevelArgs)", 256, 999.0)])
    out = os.path.join(d, "t.json")
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--traffic", os.path.join(d, "kf"), os.path.join(d, "kw"),
                        "--copy", os.path.join(d, "cf"), os.path.join(d, "cw"), "--calibrate-copy", "4",
                        "--workload", "vnoise8192", "--roofline-kernel", "k_final_pixels", "--traffic-out", out],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    import json
    t = json.load(open(out))
    assert t["launches_counted"] == 2
    assert abs(t["fetch_scale"] - 2.0) < 1e-12 and abs(t["write_scale"] - 1.0) < 1e-12
    # fetch (100 + 50) KiB x 2, write (10 + 30) KiB x 1, over two launches
    assert t["fetch_bytes_per_launch"] == 150 * 1024
    assert t["write_bytes_per_launch"] == 20 * 1024
    assert t["bytes_per_launch"] == 170 * 1024
    assert "| k_final_pixels<0, false> | 2 |" in r.stdout
