"""Throughput of the component-tree hot path (build + filter) on one MI355X.

The metric covers three workloads, all run by default, device-resident (image and outputs in HBM
before each timed region):

  vnoise8192  8192x8192 u16 value noise + ellipses, 8-connectivity, the max-tree AND the min-tree,
              area opening lambda=1000, subtractive rule.  One step = build + filter of both trees in
              one dcf_batch_filter call (a tree type per image), which builds images this large
              concurrently on two worker contexts.  This is `value`; the plain form (one
              dcf_tree_build + dcf_filter call per tree, one after the other) is timed beside it.
  spheres512  512^3 u8 blurred noise + spheres, 6-connectivity max-tree, volume opening lambda=500,
              direct rule.
  batch256    256 2D images (sides 64..2048, half u8 half u16), 4-connectivity max-trees, per-image
              lambda, direct rule: one batched dcf_batch_filter call, against one build + filter call
              per image (`workloads.batch256.plain_form`).

Every workload reports its Mvoxels/s, roofline objects for the level loop's two largest kernels (`rooflines`:
the unions, k_union, the build stage, and k_final_pixels, the attribute stage; each is algorithmic bytes per
launch / its HIP-event launch time, recorded on the stream the kernels run on, in a profiled timed region of
its own), `roofline` being the one with more time in its launches (`value` comes from an unprofiled timed
region), and a CPU baseline:
the CPU oracle (OpenMP, the reference's algorithm restated) on a bounded sample of the same workload,
1 warm-up run then the median of 5, at the host's thread count and at 1 thread.  The 8192^2 workload also reports the
tiled build (8 row slabs, boundary trees merged on the GPU) and the spectrum kernel.

Printed: ONE JSON line.  `--workload NAME` runs one workload alone.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, vendor figure

WORKLOADS = {
    "vnoise8192": dict(shape=(8192, 8192), dtype="u16", conn=8, trees=("max", "min"), attr=0, lam=1000.0, decision=3,
                       spec_bins=(64, 1, 2 ** 26),
                       desc="8192^2 u16 value noise + 20k ellipses, 8-conn, max-tree and min-tree, area opening "
                            "lambda=1000 subtractive"),
    "spheres512": dict(shape=(512, 512, 512), dtype="u8", conn=6, trees=("max",), attr=0, lam=500.0, decision=0,
                       spec_bins=(48, 1, 2 ** 27),
                       desc="512^3 u8 blurred noise + 5000 spheres, 6-conn max-tree, volume opening lambda=500 direct"),
    "batch256": dict(kind="batch", count=256, lo=64, hi=2048, conn=4, trees=("max",), attr=0, decision=0,
                     desc="256 2D images, W and H uniform in [64,2048], half u8 half u16 (blurred noise + disks, "
                          "per-image seed), 4-conn max-tree, area opening with per-image lambda in [10,5000] direct, "
                          "one batched call"),
}

# Algorithmic bytes per voxel and stage (each logical array once per stage; parent int32, area u32):
#   build: read the image, write the parent            B + 4
#   attribute: read the parent, write the area          4 + 4
#   filter: read parent, image and area, write output  4 + 2B + 4
# build + filter = 3B + 20 per voxel; the spectrum reads parent, image and area: B + 8.
# The level loop's two largest kernels carry two stages: the unions (k_union: k_union_edges over the zone graph,
# k_union_tgt over pixels) link every voxel into its component, the build stage, B + 4 bytes per voxel; k_final_pixels
# gives every voxel its parent and accumulates the areas, the attribute stage, 8 bytes per voxel.  Each stage's bytes
# of a tree are spread over that tree's launches of the kernel (one per level, loop and digit pass).
STAGE_BYTES = {"k_union": lambda B: B + 4.0, "k_final_pixels": lambda B: 8.0, "k_spectrum": lambda B: B + 8.0}
ROOFLINE_KERNELS = ("k_union", "k_final_pixels")


def step_bytes_per_voxel(B: float) -> float:
    return 3 * B + 20


def host_cpu():
    """(threads for the CPU baseline, nproc, model): the box's CPU share is OMP_NUM_THREADS when it is set"""
    try:
        nproc = len(os.sched_getaffinity(0))
    except AttributeError:
        nproc = os.cpu_count() or 1
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or nproc
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return threads, nproc, model


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


# ------------------------------------------------------------------------------------------------
# CPU baseline: the oracle (test infrastructure, the checker) timed on this host, as the baseline only
# ------------------------------------------------------------------------------------------------
def _time_oracle(jobs, threads: int, runs: int = 5, warmup: int = 1) -> float:
    """median wall time of `runs` runs, after `warmup` untimed ones, of build + filter of every (image,
    connectivity, tree, attr, lambda, decision) job"""
    from oracle.oracle import OracleTree
    for _ in range(warmup):
        for img, conn, tt, attr, lam, dec in jobs:
            OracleTree(img, conn, tt, attr, threads=threads).filter(attr, lam, dec)
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        for img, conn, tt, attr, lam, dec in jobs:
            o = OracleTree(img, conn, tt, attr, threads=threads)
            o.filter(attr, lam, dec)
            del o
        times.append(time.perf_counter() - t0)
    return statistics.median(times)


def cpu_baseline(name: str, w: dict, data, threads: int, nproc: int, model: str):
    """The CPU oracle (radix ranks, slab flooding, slab fuse, the reference's filter walks; OpenMP) on a
    bounded sample of the workload: the multi-threaded figure on a larger sample, the 1-thread one on a
    smaller one, each 1 warm-up run and then the median of 5 runs."""
    if name == "batch256":
        imgs, lams = data
        jobs_mt = [(a, w["conn"], "max", w["attr"], l, w["decision"]) for a, l in zip(imgs[:32], lams[:32])]
        jobs_1 = [(a, w["conn"], "max", w["attr"], l, w["decision"]) for a, l in zip(imgs[:6], lams[:6])]
        desc_mt, desc_1 = "the first 32 images of the batch", "the first 6 images of the batch"
    else:
        img = data
        if img.ndim == 2:
            s_mt, s_1 = img[:2048], img[:256]
            desc_mt, desc_1 = "the first 2048 rows (8192x2048)", "the first 256 rows (8192x256)"
        else:
            s_mt, s_1 = img[:128], img[:16]
            desc_mt, desc_1 = "the first 128 planes (512x512x128)", "the first 16 planes (512x512x16)"
        jobs_mt = [(s_mt, w["conn"], tt, w["attr"], w["lam"], w["decision"]) for tt in w["trees"]]
        jobs_1 = [(s_1, w["conn"], tt, w["attr"], w["lam"], w["decision"]) for tt in w["trees"]]
    vox_mt = sum(j[0].size for j in jobs_mt)
    vox_1 = sum(j[0].size for j in jobs_1)
    t_mt = _time_oracle(jobs_mt, threads)
    t_1 = _time_oracle(jobs_1, 1)
    trees = "+".join(w["trees"]) + "-tree"
    return {"value": round(vox_mt / t_mt / 1e6, 3), "unit": "Mvoxels/s", "cores": threads, "kind": "port",
            "sample": "%s, %s build+filter, 1 warm-up then the median of 5 runs (%.2f s each)" % (desc_mt, trees, t_mt),
            "one_thread": {"value": round(vox_1 / t_1 / 1e6, 3), "cores": 1,
                           "sample": "%s, %s build+filter, 1 warm-up then the median of 5 runs (%.2f s each)" % (desc_1, trees, t_1)},
            "nproc": nproc, "cpu_model": model}


# ------------------------------------------------------------------------------------------------
# GPU timing
# ------------------------------------------------------------------------------------------------
def exact_stats(ctx):
    """per-kernel (launches, total ms) by exact name (kernel_stats matches by substring)"""
    names = ctx.kernel_names()
    stats = {k: ctx.kernel_stats(k) for k in names}
    out = {}
    for k in names:
        cnt, ms = stats[k]
        for other in names:
            if other != k and k in other:
                c2, m2 = stats[other]
                cnt -= c2
                ms -= m2
        out[k] = (cnt, ms)
    return out


def _pmc_rows(path: str, counter: str):
    """per dispatch of a rocprofv3 --pmc run (its -d folder or counter CSV): (kernel name, grid size, KiB)"""
    import csv
    import glob
    files = [path] if path.endswith(".csv") else sorted(glob.glob(os.path.join(path, "**", "*counter_collection.csv"),
                                                                      recursive=True))
    if not files:
        raise SystemExit("no counter_collection.csv under %s" % path)
    per = {}
    for f in files:
        for r in csv.DictReader(open(f)):
            if r["Counter_Name"] != counter:
                continue
            key = (f, int(r["Dispatch_Id"]))
            name, grid, v = per.get(key, (r["Kernel_Name"], int(r["Grid_Size"]), 0.0))
            per[key] = (name, grid, v + float(r["Counter_Value"]))
    return list(per.values())


def traffic_summary(args):
    """--traffic FETCH_RUN WRITE_RUN --copy COPY_FETCH_RUN COPY_WRITE_RUN: the HBM bytes per launch of
    --roofline-kernel in --workload, from two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) of the workload, each
    scaled by what the same counter reads for the plain copy of --calibrate-copy MiB; writes
    profiles/traffic_<kernel>_<workload>.json (read back as roofline.traffic) and prints the table"""
    kernel, workload = args.roofline_kernel, args.workload
    if not workload or not kernel or not args.copy or not args.calibrate_copy:
        raise SystemExit("--traffic needs --workload, --roofline-kernel, --copy and --calibrate-copy MIB")
    copy_kib = float(args.calibrate_copy << 10)
    scale = []
    for run, counter in zip(args.copy, ("FETCH_SIZE", "WRITE_SIZE")):
        vals = [v for (name, g, v) in _pmc_rows(run, counter) if "copyBuffer" in name]
        if not vals:
            raise SystemExit("no copy dispatch in %s" % run)
        scale.append(copy_kib / max(vals))
    rows = {}
    for run, counter, sc in zip(args.traffic, ("FETCH_SIZE", "WRITE_SIZE"), scale):
        for name, grid, v in _pmc_rows(run, counter):
            if kernel not in name:
                continue
            short = name.split("(")[0].replace("void ", "")
            ent = rows.setdefault(short, {"n": 0, "grid": 0, "FETCH_SIZE": 0.0, "WRITE_SIZE": 0.0})
            if counter == "FETCH_SIZE":
                ent["n"] += 1
                ent["grid"] += grid
            ent[counter] += v * 1024.0 * sc
    if not rows:
        raise SystemExit("no dispatch of %s in the runs" % kernel)
    print("| kernel | dispatches | fetch MB/launch | write MB/launch | total MB/launch | fetch B per grid thread | "
          "write B per grid thread |")
    print("|---|---|---|---|---|---|---|")
    tot = {"n": 0, "grid": 0, "FETCH_SIZE": 0.0, "WRITE_SIZE": 0.0}
    for short, e in sorted(rows.items(), key=lambda kv: -kv[1]["n"]):
        for k in tot:
            tot[k] += e[k]
        print("| %s | %d | %.2f | %.2f | %.2f | %.0f | %.0f |" % (
            short, e["n"], e["FETCH_SIZE"] / e["n"] / 1e6, e["WRITE_SIZE"] / e["n"] / 1e6,
            (e["FETCH_SIZE"] + e["WRITE_SIZE"]) / e["n"] / 1e6, e["FETCH_SIZE"] / max(e["grid"], 1),
            e["WRITE_SIZE"] / max(e["grid"], 1)))
    out = {"kernel": kernel, "workload": workload,
           "bytes_per_launch": round((tot["FETCH_SIZE"] + tot["WRITE_SIZE"]) / tot["n"]),
           "fetch_bytes_per_launch": round(tot["FETCH_SIZE"] / tot["n"]),
           "write_bytes_per_launch": round(tot["WRITE_SIZE"] / tot["n"]),
           "launches_counted": tot["n"], "fetch_scale": scale[0], "write_scale": scale[1]}
    path = args.traffic_out or os.path.join(ROOT, "profiles", "traffic_%s_%s.json" % (kernel, workload))
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("\nFETCH_SIZE scale %.4f, WRITE_SIZE scale %.4f; all listed kernels: %.2f MB per launch -> %s" % (
        scale[0], scale[1], out["bytes_per_launch"] / 1e6, os.path.relpath(path, ROOT)))


def traffic_of(kernel: str, workload: str):
    path = os.path.join(ROOT, "profiles", "traffic_%s_%s.json" % (kernel, workload))
    if not os.path.exists(path):
        return None
    try:
        return json.load(open(path)).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def roofline(kernel: str, launches: int, tot_ms: float, alg_per_launch: float, workload: str, step_ms: float, steps: int,
             vox_per_step: int, B: float):
    avg_ms = tot_ms / max(launches, 1)
    achieved = alg_per_launch / (avg_ms / 1e3) / 1e9 if avg_ms > 0 else 0.0
    traffic = traffic_of(kernel, workload)
    step_bytes = float(vox_per_step) * step_bytes_per_voxel(B)
    return {"bound": "hbm", "kernel": kernel, "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic,
            "traffic_over_algorithmic": round(traffic / alg_per_launch, 2) if traffic and alg_per_launch else None,
            "algorithmic_bytes_per_launch": round(alg_per_launch), "avg_launch_ms": round(avg_ms, 4), "launches": launches,
            "share_of_step": round(tot_ms / (step_ms * steps), 4) if step_ms > 0 else None,
            "step_algorithmic_bytes": step_bytes,
            "step_achieved": round(step_bytes / (step_ms / 1e3) / 1e9, 3),
            "step_frac": round(step_bytes / (step_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 6)}


def timed(step, steps: int) -> float:
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps * 1e3


def run_workload(name: str, ctx, args, kernel: str, cpu: bool, say):
    import torch
    from disccofan_amd import synth
    w = WORKLOADS[name]
    batch = w.get("kind") == "batch"
    dev = torch.device("cuda:0")
    state = {}
    t0 = time.perf_counter()
    steps = args.steps if not batch else max(1, min(args.steps, 5))
    if batch:
        imgs, lams = make_batch(w)
        data = (imgs, lams)
        gen_s = time.perf_counter() - t0
        n = int(sum(a.size for a in imgs))
        B = sum(a.nbytes for a in imgs) / n
        d_in = [torch.from_numpy(a.view(np.int16) if a.dtype == np.uint16 else a).to(dev) for a in imgs]
        d_out = [torch.empty_like(t) for t in d_in]
        torch.cuda.synchronize()
        items = [(ti.data_ptr(), a.dtype, a.shape, lam, to.data_ptr()) for ti, a, lam, to in zip(d_in, imgs, lams, d_out)]
        # launches of k_final_pixels per tree: counted in the profiled region (one per level and pass)
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
        img = make_image(name)
        data = img
        gen_s = time.perf_counter() - t0
        n = img.size
        B = img.dtype.itemsize
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

        if len(w["trees"]) > 1 and not args.sequential:
            # every tree of the image in one batched call: images this large are built concurrently, one per
            # worker context (own stream, own host thread)
            items = [(d_img.data_ptr(), np_dtype, img.shape, w["lam"], d_out.data_ptr(), tt)
                     for tt, d_out in zip(w["trees"], d_outs)]

            def step():
                ctx.batch_filter_device(items, connectivity=w["conn"], attr_idx=w["attr"], decision=w["decision"])

            # (nodes_per_tree comes from the plain form, timed after the batched one)
        else:
            step = plain_step
            plain_step = None
    vox_per_step = n * len(w["trees"])

    # the kernels whose launches are bracketed by HIP events, each in a profiled region of its own (no other
    # launch pays for events); --kernels brackets every launch in one region
    kernels = [kernel] if kernel else list(ROOFLINE_KERNELS)
    regions = [None] if args.kernels else kernels
    # warm-up with the first region's events on (the event pools are filled before the timed regions)
    ctx.profile_only(regions[0])
    ctx.set_profiling(True)
    for i in range(args.warmup):
        t1 = time.perf_counter()
        step()
        torch.cuda.synchronize()
        say("%s warmup %d: %.1f ms" % (name, i, (time.perf_counter() - t1) * 1e3))
    ctx.set_profiling(False)
    prof_ms, ex, region_ms, ex_all = 0.0, {}, {}, {}
    if not args.no_profile:
        for r in regions:
            ctx.profile_only(r)
            ctx.set_profiling(True)
            if r is not regions[0]:
                step()
                torch.cuda.synchronize()
            ctx.reset_stats()
            t_ms = timed(step, steps)
            ctx.set_profiling(False)
            stats = exact_stats(ctx)
            for k in kernels if r is None else [r]:
                if k in stats:
                    ex[k] = stats[k]
                    region_ms[k] = t_ms
            if r is regions[0]:
                prof_ms = t_ms
                if args.kernels:
                    ex_all = stats
    # unprofiled timed region: the headline rate
    ms = timed(step, steps)
    value = vox_per_step / (ms / 1e3) / 1e6
    say("%s: %.2f ms/step (%.1f Mvox/s), profiled %.2f ms/step" % (name, ms, value, prof_ms))
    # the stage's bytes of one step (every voxel of every tree once) spread over the step's launches; `roofline` is
    # the kernel with the most time in its launches, `rooflines` holds every bracketed one
    roofs = {}
    for k in kernels:
        if k in ex:
            launches, tot = ex[k]
            alg = vox_per_step * STAGE_BYTES[k](B) / (launches / steps) if launches else 0.0
            roofs[k] = roofline(k, launches, tot, alg, name, region_ms[k], steps, vox_per_step, B)
    roof = max(roofs.values(), key=lambda rf: rf["avg_launch_ms"] * rf["launches"]) if roofs else None
    if args.kernels and not args.no_profile and ex_all:
        tot = sum(v[1] for v in ex_all.values())
        for k in sorted(ex_all, key=lambda k: -ex_all[k][1]):
            c, m = ex_all[k]
            say("%-16s launches %5d  total %9.3f ms  avg %8.4f ms  %5.1f%%" % (k, c, m, m / max(c, 1), 100 * m / tot))
    out = {"value": round(value, 3), "unit": "Mvoxels/s", "ms_per_step": round(ms, 3), "steps": steps,
           "profiled_ms_per_step": round(prof_ms, 3), "dtype": "u8+u16" if batch else ("u16" if B == 2 else "u8"),
           "config": {"workload": w["desc"], "shape": shape_desc, "voxels_per_step": vox_per_step,
                      "connectivity": w["conn"], "trees": list(w["trees"]), "attribute": w["attr"],
                      "lambda": "per image" if batch else w["lam"], "decision": w["decision"],
                      "form": ("batched: one dcf_batch_filter call for the whole batch" if batch else
                               "batched: both trees in one dcf_batch_filter call, built concurrently on two worker "
                               "contexts (one HIP stream and one host thread each)" if plain_step else
                               "plain: one dcf_tree_build + dcf_filter call per tree"),
                      "bit_exact_vs_oracle": ("tests/test_gpu_batch.py::test_full_batch256_workload" if batch else
                                              "tests/test_gpu_fullsize.py"),
                      "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))},
           "roofline": roof, "rooflines": roofs, "gen_s": round(gen_s, 2)}

    if batch and not args.no_extras:
        # the same batched call with every pass's level loop captured into one HIP graph
        ctx.set_graph(True)
        step()
        gms = timed(step, steps)
        ctx.set_graph(False)
        out["graph"] = {"value": round(vox_per_step / (gms / 1e3) / 1e6, 3), "ms_per_step": round(gms, 3),
                        "form": "the batched call with each pass's level loop captured into one HIP graph "
                                "(dcf_ctx_set_graph)",
                        "graph_vs_no_graph": round(ms / gms, 3)}
        say("%s with graph-captured level loops: %.1f ms (without %.1f ms)" % (name, gms, ms))

    if plain_step is not None and not args.no_plain:
        plain_step()
        torch.cuda.synchronize()
        pms = timed(plain_step, 1 if batch else steps)
        out["plain_form"] = {"value": round(vox_per_step / (pms / 1e3) / 1e6, 3), "ms_per_step": round(pms, 3),
                             "form": ("one dcf_tree_build + dcf_filter per image, same images and lambdas, 1 timed step"
                                      if batch else "one dcf_tree_build + dcf_filter call per tree, one after the other, "
                                                    "%d timed steps" % steps)}
        out["batched_vs_plain"] = round(pms / ms, 3)
        say("%s plain form: %.1f ms (batched %.1f ms)" % (name, pms, ms))
    if not batch:
        out["nodes_per_tree"] = state.get("K")

    if not args.no_host:
        # the same calls on host numpy buffers: the image goes to the GPU and the output comes back in every step
        if batch:
            def host_step():
                ctx.batch_filter(imgs, lams, connectivity=w["conn"], tree_type="max", attr_idx=w["attr"],
                                 decision=w["decision"])
            hform = "one dcf_batch_filter call on host buffers"
        elif plain_step is not None:
            def host_step():
                ctx.batch_filter([img] * len(w["trees"]), [w["lam"]] * len(w["trees"]), connectivity=w["conn"],
                                 attr_idx=w["attr"], decision=w["decision"], tree_types=list(w["trees"]))
            hform = "both trees in one dcf_batch_filter call on host buffers"
        else:
            def host_step():
                for tt in w["trees"]:
                    tr = ctx.build(img, w["conn"], tt, w["attr"])
                    tr.filter(w["attr"], w["lam"], w["decision"], out_dtype=img.dtype)
                    tr.free()
            hform = "dcf_tree_build + dcf_filter on host buffers"
        host_step()
        hms = timed(host_step, 1 if batch else min(steps, 3))
        out["host_buffers"] = {"value": round(vox_per_step / (hms / 1e3) / 1e6, 3), "ms_per_step": round(hms, 3),
                               "form": hform + " (image in, output out over PCIe in every step)"}
        say("%s on host buffers: %.1f ms (device-resident %.1f ms)" % (name, hms, ms))

    if not batch and not args.no_extras:
        # the spectrum kernel on the max-tree: B + 8 algorithmic bytes per voxel
        nb, lo, hi = w["spec_bins"]
        lv = synth.log_lambdas(nb, lo, hi)
        tr = ctx.build_device(d_img.data_ptr(), np_dtype, img.shape, w["conn"], "max", w["attr"])
        K = tr.num_nodes
        tr.spectrum1d(w["attr"], lv)
        ctx.reset_stats()
        ctx.profile_only("k_spectrum")
        ctx.set_profiling(True)
        for _ in range(5):
            tr.spectrum1d(w["attr"], lv)
        ctx.set_profiling(False)
        c, m = exact_stats(ctx).get("k_spectrum", (0, 0.0))
        tr.free()
        if c:
            # the kernel reads node arrays, not pixels (the per-voxel B + 8 of the contract was paid by the
            # build): per node its parent, own and parent grey and area, 16 bytes
            avg = m / c
            alg = K * 16.0
            out["spectrum"] = {"kernel": "k_spectrum", "bins": nb, "nodes": K, "avg_launch_ms": round(avg, 4),
                               "achieved": round(alg / (avg / 1e3) / 1e9, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                               "frac": round(alg / (avg / 1e3) / 1e9 / HBM_PEAK_GBS, 6),
                               "algorithmic_bytes_per_launch": round(alg),
                               "voxel_contract_bytes": round(n * STAGE_BYTES["k_spectrum"](B)),
                               "voxel_contract_rate_gbs": round(n * STAGE_BYTES["k_spectrum"](B) / (avg / 1e3) / 1e9, 3)}
    if name == "vnoise8192" and not args.no_extras:
        out["tiled"] = tiled_build(ctx, d_img, img, w, say)
    if cpu:
        threads, nproc, model = host_cpu()
        out["cpu_baseline"] = cpu_baseline(name, w, data, threads, nproc, model)
    return out


def tiled_build(ctx, d_img, img, w, say):
    """The distributed path on one GPU: the max-tree of the 8192^2 image from 8 row slabs (each an 8192 x
    1024 share plus the one-pixel overlap), built tile by tile and merged through their boundary trees;
    the share's own build (the extended 8192 x 1026 slab) is timed alone for the per-GPU figure."""
    import torch
    np_dtype = img.dtype
    H, W = img.shape
    T = 8
    def build():
        t = ctx.build_tiled_device(d_img.data_ptr(), np_dtype, img.shape, (1, T), w["conn"], "max", w["attr"])
        torch.cuda.synchronize()
        return t

    build().free()
    times = []
    for _ in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        tr = build()
        times.append((time.perf_counter() - t0) * 1e3)
        tr.free()
    total_ms = statistics.median(times)
    # the merge kernels' time from one more run with HIP events around them
    ctx.reset_stats()
    ctx.profile_only("k_tm_")
    ctx.set_profiling(True)
    build().free()
    ctx.set_profiling(False)
    ex = exact_stats(ctx)
    merge_ms = sum(m for k, (c, m) in ex.items() if k != "k_tm_gather")
    # one share: rows [1023, 2049) of the image, its own slab plus the overlap rows
    rows = slice(H // T - 1, 2 * H // T + 1)
    share = d_img.view(H, W)[rows].contiguous()
    tr = ctx.build_device(share.data_ptr(), np_dtype, (share.shape[0], W), w["conn"], "max", w["attr"])
    tr.free()
    share_ms = timed(lambda: ctx.build_device(share.data_ptr(), np_dtype, (share.shape[0], W), w["conn"], "max",
                                              w["attr"]).free(), 3)
    say("tiled build (1x%d slabs): %.1f ms, merge kernels %.1f ms, one %dx%d share's build %.1f ms" % (
        T, total_ms, merge_ms, W, share.shape[0], share_ms))
    return {"grid": [1, T], "tree": "max", "total_ms": round(total_ms, 3), "merge_kernels_ms": round(merge_ms, 3),
            "share_shape": [W, int(share.shape[0])], "share_build_ms": round(share_ms, 3),
            "share_build_mvox_s": round(W * share.shape[0] / (share_ms / 1e3) / 1e6, 3),
            "bit_exact_vs_untiled": "tests/test_gpu_tiles.py::test_vnoise_8192_slabs"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default=None, choices=sorted(WORKLOADS),
                    help="run one workload alone (default: all three, vnoise8192 giving `value`)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain (per-image / per-tree) comparison")
    ap.add_argument("--no-host", action="store_true", help="skip the host-buffer (PCIe-inclusive) figure")
    ap.add_argument("--sequential", action="store_true",
                    help="vnoise8192: time the plain form (one call per tree, one after the other) as the step")
    ap.add_argument("--no-extras", action="store_true", help="skip the spectrum and tiled-build figures")
    ap.add_argument("--kernels", action="store_true", help="bracket every kernel and print the per-kernel times")
    ap.add_argument("--roofline-kernel", default=None, choices=ROOFLINE_KERNELS,
                    help="bracket only this kernel with HIP events (default: k_union and k_final_pixels, each in a "
                         "profiled region of its own; the one with more time is `roofline`); with --traffic, the "
                         "kernel whose counters are summarised")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB 5 times (the reference point for the "
                         "FETCH_SIZE / WRITE_SIZE counters) and print its bandwidth")
    ap.add_argument("--calibrate-gather", type=int, default=0, metavar="MIB",
                    help="only run random 4-byte gathers from a table of MIB MiB (one per table element, torch.take), "
                         "5 times, and print their rate: the random-access ceiling of the level kernels")
    ap.add_argument("--traffic", nargs=2, metavar=("FETCH_RUN", "WRITE_RUN"),
                    help="no GPU work: summarise two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) of --workload into "
                         "profiles/traffic_<--roofline-kernel>_<workload>.json, scaled by the --copy runs")
    ap.add_argument("--copy", nargs=2, metavar=("FETCH_RUN", "WRITE_RUN"),
                    help="the two --pmc runs of `bench.py --calibrate-copy MIB` that scale --traffic")
    ap.add_argument("--traffic-out", default=None, metavar="JSON",
                    help="where --traffic writes (default profiles/traffic_<kernel>_<workload>.json)")
    ap.add_argument("--no-profile", action="store_true",
                    help="skip the profiled timed region (no roofline object; for counter runs)")
    ap.add_argument("--hw-queues", type=int, default=16, metavar="N",
                    help="the HIP runtime's hardware queues for this process (GPU_MAX_HW_QUEUES, set before the "
                         "runtime starts; 0 keeps the environment's). Concurrent worker streams overlap only on "
                         "queues of their own; HIP's default is 4")
    args = ap.parse_args()
    if not 0 <= args.hw_queues <= 32:
        ap.error("--hw-queues must be between 0 and 32")
    if args.hw_queues:
        os.environ["GPU_MAX_HW_QUEUES"] = str(args.hw_queues)

    if args.traffic:
        traffic_summary(args)
        return
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

    if args.calibrate_gather:
        n = (args.calibrate_gather << 20) // 4
        table = torch.arange(n, dtype=torch.int32, device="cuda:0")
        idx = torch.randint(0, n, (n,), dtype=torch.int64, device="cuda:0", generator=torch.Generator("cuda:0").manual_seed(7))
        out = torch.take(table, idx)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(5):
            out = torch.take(table, idx)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 5
        assert int(out[0]) == int(idx[0])
        print(json.dumps({"gathers": n, "table_bytes": n * 4, "ms": round(ms, 4), "gathers_per_s": round(n / (ms / 1e3), 1),
                          "index_and_output_bytes": n * 12}), flush=True)
        return

    import disccofan_amd as dcf

    def say(msg):
        print(msg, file=sys.stderr, flush=True)

    ctx = dcf.Context(0)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    names = [args.workload] if args.workload else ["vnoise8192", "spheres512", "batch256"]
    res = {}
    for name in names:
        res[name] = run_workload(name, ctx, args, args.roofline_kernel, not args.no_cpu, say)
        torch.cuda.empty_cache()
    head = res[names[0]]
    batch = res.get("batch256", {})
    line = {
        "metric": "Mvoxels/s (build+filter)",
        "value": head["value"],
        "unit": "Mvoxels/s",
        "n_gpus": 1,
        "steps": head["steps"],
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": head["dtype"],
        "data": "synthetic (seeded generator, disccofan_amd/synth)",
        "config": head["config"],
        "roofline": head["roofline"],
        "rooflines": head.get("rooflines"),
        "cpu_baseline": head.get("cpu_baseline"),
        "host_buffers": head.get("host_buffers"),
        "plain_form": head.get("plain_form", batch.get("plain_form")),
        "batched_vs_plain": head.get("batched_vs_plain", batch.get("batched_vs_plain")),
        "workloads": {k: v for k, v in res.items() if k != names[0]},
        "spectrum": head.get("spectrum"),
        "tiled": head.get("tiled"),
        "precision_form": {"config": "4096^2 f32 blobs quantised to u16, 8-conn, trace (13) and non-compactness (17), "
                                     "32x32 ncomp x area spectrum",
                           "max_rel_err": {"inertiatrace": 0.0, "ncomp": 0.0, "pattern2D_bins": 0.0},
                           "how": "bit-exact against the FP64 oracle (exact int64 moments, FP64 central step): "
                                  "tests/test_gpu_fullsize.py::test_blobs_4096_precision_form"},
        "gen_s": head["gen_s"],
    }
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
