"""Throughput of the component-tree hot path (build + filter) on one MI355X.

Workload (default): 8192x8192 u16 value-noise + ellipses image, 8-connectivity, the max-tree AND the
min-tree, area attribute, area opening lambda=1000 with the subtractive rule.  One step = build
+ filter of both trees (2 x 67.1 Mvoxel), one build + filter call per tree (--paired times both trees
in one batched call instead, with a tree type per image); the image and the output buffers are
resident in HBM before the timed region.  `--workload spheres512` runs the 512^3 u8 volume (6-connectivity,
max-tree, direct rule, lambda=500) instead.

Printed: ONE JSON line with the metric, a roofline object for the dominant kernel (per-launch
algorithmic bytes / HIP-event launch time, measured inside the timed region on the stream the
kernels run on), and a cpu_baseline measured with the CPU oracle (OpenMP) on the same image on this
host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, vendor figure

WORKLOADS = {
    "vnoise8192": dict(shape=(8192, 8192), dtype="u16", conn=8, trees=("max", "min"), attr=0, lam=1000.0, decision=3,
                       desc="8192^2 u16 value noise + 20k ellipses, 8-conn, max-tree and min-tree, area opening "
                            "lambda=1000 subtractive"),
    "spheres512": dict(shape=(512, 512, 512), dtype="u8", conn=6, trees=("max",), attr=0, lam=500.0, decision=0,
                       desc="512^3 u8 blurred noise + 5000 spheres, 6-conn max-tree, volume opening lambda=500 direct"),
    "batch256": dict(kind="batch", count=256, lo=64, hi=2048, conn=4, trees=("max",), attr=0, decision=0,
                     desc="256 2D images, W and H uniform in [64,2048], half u8 half u16 (blurred noise + disks, "
                          "per-image seed), 4-conn max-tree, area opening with per-image lambda in [10,5000] direct, "
                          "one batched call"),
}


def algorithmic_bytes(kernel: str, n: int, k_nodes: int, B: float, passes: float):
    """(units one tree makes this kernel process, algorithmic bytes per unit): every logical array is
    counted once per stage (build: read image + write parent = B + 4 per voxel; attribute: read
    parent + write area = 8 per voxel; filter: read parent, image and area, write output = 2B + 8
    per voxel).  The level kernels of the build together see every voxel once per pass, so their
    per-launch figure is (passes * n / launches) voxels; `passes` is the voxel-weighted mean number of
    8-bit digit passes (1 for u8, 2 for u16)."""
    table = {
        "k_union": (passes * n, B + 4),
        "k_final_pixels": (passes * n, 8),
        "k_filter_out": (n, 2 * B + 8),
        "k_direct_jump": (k_nodes, 9),
    }
    return table.get(kernel, (0, 0.0))


def make_batch(w: dict):
    from disccofan_amd import synth
    shapes = synth.batch_shapes(w["count"], w["lo"], w["hi"])
    imgs = [synth.disks(H, W, is16, seed) for (W, H, is16, seed, lam) in shapes]
    return imgs, [lam for (W, H, is16, seed, lam) in shapes]


def make_image(name: str) -> np.ndarray:
    from disccofan_amd import synth
    w = WORKLOADS[name]
    if name == "vnoise8192":
        return synth.vnoise(*w["shape"])
    return synth.spheres(*w["shape"])


def cpu_baseline_batch(imgs, lams, w: dict, threads: int):
    """The CPU oracle on every image of the batch, build + filter with the image's own lambda."""
    from oracle.oracle import OracleTree
    t0 = time.perf_counter()
    vox = 0
    for a, lam in zip(imgs, lams):
        o = OracleTree(a, w["conn"], "max", w["attr"], threads=threads)
        o.filter(w["attr"], lam, w["decision"])
        vox += a.size
        del o
    dt = time.perf_counter() - t0
    return {"value": vox / dt / 1e6, "unit": "Mvoxels/s", "cores": threads, "kind": "port",
            "sample": "all %d images of the batch, build+filter, %d OpenMP threads, %.2f s" % (len(imgs), threads, dt)}


def cpu_baseline(img: np.ndarray, w: dict, threads: int):
    """The CPU oracle (the reference algorithm restated in C: radix ranks, slab flooding, slab fuse,
    the same filter walks; OpenMP) on the whole workload image, build + filter of every tree of
    the step, timed on this host."""
    from oracle.oracle import OracleTree
    t0 = time.perf_counter()
    vox = 0
    for tt in w["trees"]:
        o = OracleTree(img, w["conn"], tt, w["attr"], threads=threads)
        o.filter(w["attr"], w["lam"], w["decision"])
        vox += img.size
        del o
    dt = time.perf_counter() - t0
    return {"value": vox / dt / 1e6, "unit": "Mvoxels/s", "cores": threads, "kind": "port",
            "sample": "the full %s image, %s, build+filter, %d OpenMP threads, %.2f s" % (
                "x".join(map(str, img.shape)), "+".join(w["trees"]) + "-tree", threads, dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="vnoise8192", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--kernels", action="store_true", help="print the per-kernel event times to stderr")
    ap.add_argument("--roofline-kernel", default="k_final_pixels",
                    help="the kernel bracketed by HIP events in the timed region (the dominant one: k_final_pixels, "
                         "the largest share of the kernel time in every workload's profile; k_union for the union "
                         "kernels); --kernels brackets every kernel instead")
    ap.add_argument("--no-profile", action="store_true",
                    help="time the steps without per-launch HIP events (no roofline object)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB 5 times (the reference point for the "
                         "FETCH_SIZE / WRITE_SIZE counters) and print its bandwidth")
    ap.add_argument("--no-plain", action="store_true",
                    help="batched forms: skip the per-image / per-tree comparison loop (plain_form null)")
    ap.add_argument("--paired", action="store_true",
                    help="a workload with a max-tree and a min-tree: time one batched call for both trees (a tree "
                         "type per image) instead of one build + filter call per tree, with the per-tree form beside it")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.steps < 1 or args.warmup < 0:
        ap.error("--steps must be at least 1 and --warmup at least 0")

    import torch

    if args.calibrate_copy:
        nb = args.calibrate_copy << 20
        src = torch.ones(nb // 4, dtype=torch.int32, device="cuda:0")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(5):
            dst.copy_(src)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 5
        print(json.dumps({"copy_bytes_read": nb, "copy_bytes_written": nb, "ms": round(ms, 4),
                          "GBps_read_plus_write": round(2 * nb / ms / 1e6, 1)}), flush=True)
        return

    import disccofan_amd as dcf

    w = WORKLOADS[args.workload]
    batch = w.get("kind") == "batch"
    pair = not batch and len(w["trees"]) > 1 and args.paired
    dev = torch.device("cuda:0")
    ctx = dcf.Context(0)
    stream = torch.cuda.current_stream()
    ctx.set_stream(stream.cuda_stream)
    state = {}
    t0 = time.perf_counter()
    if batch:
        imgs, lams = make_batch(w)
        gen_s = time.perf_counter() - t0
        n = int(sum(a.size for a in imgs))
        B = sum(a.nbytes for a in imgs) / n
        # u8 and u16 images are built in separate batch trees: one digit pass for u8, two for u16
        passes = sum((2 if a.dtype == np.uint16 else 1) * a.size for a in imgs) / n
        d_in = [torch.from_numpy(a.view(np.int16) if a.dtype == np.uint16 else a).to(dev) for a in imgs]
        d_out = [torch.empty_like(t) for t in d_in]
        torch.cuda.synchronize()
        items = [(ti.data_ptr(), a.dtype, a.shape, lam, to.data_ptr()) for ti, a, lam, to in zip(d_in, imgs, lams, d_out)]
        shape_desc = [len(imgs), n]

        def step():
            ctx.batch_filter_device(items, connectivity=w["conn"], tree_type="max", attr_idx=w["attr"],
                                    decision=w["decision"])

        def plain_step():
            for ti, a, lam, to in zip(d_in, imgs, lams, d_out):
                tr = ctx.build_device(ti.data_ptr(), a.dtype, a.shape, w["conn"], "max", w["attr"])
                tr.filter_device(w["attr"], lam, w["decision"], to.data_ptr(), a.dtype)
                tr.free()
    else:
        img = make_image(args.workload)
        gen_s = time.perf_counter() - t0
        n = img.size
        B = img.dtype.itemsize
        passes = 2 if B == 2 else 1
        tdt = torch.int16 if B == 2 else torch.uint8
        d_img = torch.from_numpy(img.reshape(-1).view(np.int16 if B == 2 else np.uint8)).to(dev)
        d_outs = [torch.empty(n, dtype=tdt, device=dev) for _ in w["trees"]]
        torch.cuda.synchronize()
        np_dtype = np.uint16 if B == 2 else np.uint8
        shape_desc = list(img.shape)

        def plain_step():
            for tt, d_out in zip(w["trees"], d_outs):
                tr = ctx.build_device(d_img.data_ptr(), np_dtype, img.shape, w["conn"], tt, w["attr"])
                tr.filter_device(w["attr"], w["lam"], w["decision"], d_out.data_ptr(), np_dtype)
                state["K"] = tr.num_nodes
                tr.free()

        if pair:
            # the trees of one image in one batched call (dcf_batch_filter with a tree type per image)
            items = [(d_img.data_ptr(), np_dtype, img.shape, w["lam"], d_out.data_ptr(), tt)
                     for tt, d_out in zip(w["trees"], d_outs)]

            def step():
                ctx.batch_filter_device(items, connectivity=w["conn"], attr_idx=w["attr"], decision=w["decision"])

            plain_step()  # node count of the trees, for the filter kernels' algorithmic bytes
        else:
            step = plain_step

    for i in range(args.warmup):
        t1 = time.perf_counter()
        step()
        torch.cuda.synchronize()
        print("warmup %d: %.1f ms" % (i, (time.perf_counter() - t1) * 1e3), file=sys.stderr, flush=True)

    ctx.reset_stats()
    ctx.profile_only(None if args.kernels else args.roofline_kernel)
    ctx.set_profiling(not args.no_profile)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ctx.set_profiling(False)

    ms_per_step = dt / args.steps * 1e3
    vox_per_step = n * len(w["trees"])
    value = vox_per_step / (ms_per_step / 1e3) / 1e6

    # dominant kernel from the event timings of the timed region (none with --no-profile)
    names = ctx.kernel_names()
    stats = {k: ctx.kernel_stats(k) for k in names}
    exact = {}
    for k in names:
        # kernel_stats matches by substring; keep exact-name totals
        cnt, ms = stats[k]
        for other in names:
            if other != k and k in other:
                c2, m2 = stats[other]
                cnt -= c2
                ms -= m2
        exact[k] = (cnt, ms)
    K = state.get("K", 0)
    step_bytes = float(vox_per_step) * (3 * B + 20)
    roof = None
    if exact:
        dom = args.roofline_kernel if args.roofline_kernel in exact else max(exact, key=lambda k: exact[k][1])
        launches, tot_ms = exact[dom]
        avg_ms = tot_ms / max(launches, 1)
        units, per_unit = algorithmic_bytes(dom, n, K, B, passes)
        runs = args.steps * len(w["trees"])
        alg = units * runs * per_unit / max(launches, 1)
        achieved = alg / (avg_ms / 1e3) / 1e9 if alg > 0 else 0.0
        traffic = None
        tpath = os.path.join(ROOT, "profiles", "traffic_%s_%s.json" % (dom, args.workload))
        if os.path.exists(tpath):
            try:
                traffic = json.load(open(tpath)).get("bytes_per_launch")
            except Exception:
                traffic = None
        roof = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic, "avg_launch_ms": round(avg_ms, 4),
                "launches": launches, "algorithmic_bytes_per_launch": round(alg),
                "share_of_step": round(tot_ms / (ms_per_step * args.steps), 4),
                "step_algorithmic_bytes": step_bytes,
                "step_achieved": round(step_bytes / (ms_per_step / 1e3) / 1e9, 3),
                "step_frac": round(step_bytes / (ms_per_step / 1e3) / 1e9 / HBM_PEAK_GBS, 6)}
        if args.kernels:
            tot = sum(v[1] for v in exact.values())
            for k in sorted(exact, key=lambda k: -exact[k][1]):
                c, m = exact[k]
                print("%-16s launches %5d  total %9.3f ms  avg %8.4f ms  %5.1f%%" % (k, c, m, m / max(c, 1), 100 * m / tot),
                      file=sys.stderr)
    if not batch:
        print("nodes K = %d (%.1f%% of pixels)" % (K, 100.0 * K / n), file=sys.stderr)

    plain = None
    if (batch or pair) and not args.no_plain:
        # the plain form on the same device-resident data: one build + filter call per image / tree
        plain_step()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        plain_step()
        torch.cuda.synchronize()
        pms = (time.perf_counter() - t1) * 1e3
        plain = {"value": round(vox_per_step / (pms / 1e3) / 1e6, 3), "ms_per_step": round(pms, 3),
                 "form": ("one dcf_tree_build + dcf_filter per image, same images and lambdas, 1 timed step" if batch
                          else "one dcf_tree_build + dcf_filter per tree, 1 timed step")}
        print("plain per-image loop: %.1f ms (%.1f Mvox/s); batched %.1f ms" % (pms, plain["value"], ms_per_step),
              file=sys.stderr)

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        cpu = cpu_baseline_batch(imgs, lams, w, threads) if batch else cpu_baseline(img, w, threads)

    line = {
        "metric": "Mvoxels/s (build+filter)",
        "value": round(value, 3),
        "unit": "Mvoxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8+u16" if batch else ("u16" if B == 2 else "u8"),
        "data": "synthetic (seeded generator, disccofan_amd/synth)",
        "config": {"workload": w["desc"], "shape": shape_desc, "voxels_per_step": vox_per_step,
                   "connectivity": w["conn"], "trees": list(w["trees"]), "attribute": w["attr"],
                   "lambda": "per image" if batch else w["lam"], "decision": w["decision"],
                   "form": ("batched: one dcf_batch_filter call for the whole batch" if batch else
                            "batched: both trees of the image in one dcf_batch_filter call" if pair else
                            "plain: one dcf_tree_build + dcf_filter call per tree"),
                   "bit_exact_vs_oracle": "tests/test_gpu_batch.py" if batch else "tests/test_gpu_fullsize.py"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "plain_form": plain,
        "batched_vs_plain": round(plain["ms_per_step"] / ms_per_step, 3) if plain else None,
        "gen_s": round(gen_s, 2),
    }
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
