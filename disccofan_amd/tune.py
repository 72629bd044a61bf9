"""Autotuner for the build's tunables: the zone graph and its tile shape, precomputed union targets, 16-byte
records, the level-list tile shapes, the k_fix grid, and for the tiled build the zips per pair (its merge
schedule).  They are set on each candidate's context through dcf_ctx_set_option (Context.set_option), by
the names of the DCF_* environment variables that give new contexts their defaults.

Each knob is tried in turn with the others held at their current best (coordinate descent from the
defaults), timing build + filter of a sample workload on the GPU: a single image, a batch through the
batched form, or a tiled build.  The outputs of every candidate must agree bit for bit (the tunables
change the schedule, never the result); a difference raises.  The result is the fastest setting, for
Context.set_option, or as environment assignments to export before the process creates its contexts.

    python -m disccofan_amd.tune --workload vnoise --size 2048
    python -m disccofan_amd.tune --workload batch --count 64
    python -m disccofan_amd.tune --workload batch --count 256 --lo 16 --size 128    # many small images
    python -m disccofan_amd.tune --workload tiled --size 4096 --tiles 8
"""
import argparse
import time

import numpy as np

from . import Context, synth

# knob -> candidate values, the library's default first
KNOBS_2D = {
    "DCF_ZONES": ["2", "1", "0"],
    "DCF_RECORDS": ["0", "1"],
    "DCF_TARGETS": ["1", "2", "0"],
    "DCF_LEVEL_TILE2D": ["64x64x1", "32x32x1", "128x32x1", "128x128x1", "256x16x1"],
    "DCF_LEVEL_GROUP2D": ["1x1x1", "2x2x1", "4x1x1"],
    "DCF_ZONE_TILE2D": ["64x64", "32x32", "128x32", "32x128"],
    "DCF_FIX_BLOCKS": ["1024", "256"],
    "DCF_FINE_LOOPS": ["3", "1", "2", "4"],
}
KNOBS_3D = {
    "DCF_ZONES": ["2", "1", "0"],
    "DCF_RECORDS": ["0", "1"],
    "DCF_TARGETS": ["1", "2", "0"],
    "DCF_LEVEL_TILE3D": ["16x16x16", "32x32x8", "64x16x4", "8x8x8"],
    "DCF_LEVEL_GROUP3D": ["1x1x1", "2x2x2"],
    "DCF_ZONE_TILE3D": ["16", "8", "12"],
    "DCF_FINAL_BLOCKS": ["1024", "0", "512", "2048"],
    "DCF_FINE_LOOPS": ["3", "1", "2"],
}
# the batched form orders its level lists by each image's zone tiles: the zone tiles are its tile shape
KNOBS_BATCH = {
    "DCF_ZONES": ["2", "1", "0"],
    "DCF_RECORDS": ["0", "1"],
    "DCF_TARGETS": ["1", "2", "0"],
    "DCF_ZONE_TILE2D": ["64x64", "32x32", "16x16", "128x32", "32x128"],
    "DCF_FIX_BLOCKS": ["1024", "256"],
    "DCF_FINAL_BLOCKS": ["1024", "0", "512"],
    # concurrent batch trees: worker contexts (the default with 16 hardware queues first), batch trees per dtype
    # (0: one part per worker), refining loops per worker (0: as the hardware queues allow)
    "DCF_WORKERS": ["4", "1", "2", "3"],
    "DCF_BATCH_PARTS": ["0", "1", "2"],
    "DCF_WORKER_LOOPS": ["0", "1", "2"],
}
# the tiled build: its merge schedule (zips started per identification pair) and the tiles' own build
KNOBS_TILED = {
    "DCF_ZIP_SEGMENTS": ["64", "1", "16", "256"],
    "DCF_ZONES": ["2", "0"],
    "DCF_LEVEL_TILE2D": ["64x64x1", "32x32x1", "128x32x1"],
}


def knobs_for(is3d: bool, batched: bool, tiled: bool = False) -> dict:
    if tiled:
        return KNOBS_TILED
    if batched:
        return KNOBS_BATCH
    return KNOBS_3D if is3d else KNOBS_2D


def env_line(setting: dict) -> str:
    return " ".join("export %s=%s;" % (k, v) for k, v in sorted(setting.items()))


class _Runner:
    """times one setting: a fresh context with the setting in the environment, one warm-up, then the
    best of `repeats` runs; returns (seconds, outputs)"""

    def __init__(self, images, lambdas, connectivity, tree_type, attr_idx, decision, batched, repeats, device, grid=None):
        self.images, self.lambdas = images, lambdas
        self.conn, self.tt, self.attr, self.dec = connectivity, tree_type, attr_idx, decision
        self.batched, self.repeats, self.device = batched, repeats, device
        self.grid = grid
        try:
            import torch
            self.torch = torch if torch.cuda.is_available() else None
        except ImportError:
            self.torch = None

    def _once(self, ctx):
        if self.grid is not None:
            outs = []
            for img, lam in zip(self.images, self.lambdas):
                t = ctx.build_tiled(img, self.grid, self.conn, self.tt, self.attr)
                outs.append(t.parent())
                t.free()
            return outs
        if self.batched:
            return ctx.batch_filter(self.images, self.lambdas, connectivity=self.conn, tree_type=self.tt,
                                    attr_idx=self.attr, decision=self.dec)
        outs = []
        for img, lam in zip(self.images, self.lambdas):
            if self.torch is not None:
                th = self.torch
                dev = th.device("cuda:%d" % self.device)
                src = th.from_numpy(img.view(np.int16) if img.dtype == np.uint16 else img).to(dev)
                out = th.empty(img.shape, dtype=th.float32, device=dev)
                t = ctx.build_device(src.data_ptr(), img.dtype, img.shape, self.conn, self.tt, self.attr)
                t.filter_device(self.attr, lam, self.dec, out.data_ptr(), np.float32)
                ctx.synchronize()
                outs.append(out)
                t.free()
            else:
                t = ctx.build(img, self.conn, self.tt, self.attr)
                outs.append(t.filter(self.attr, lam, self.dec))
                t.free()
        return outs

    def run(self, setting: dict):
        ctx = Context(self.device)
        try:
            for k, v in setting.items():
                ctx.set_option(k, v)
            self._once(ctx)
            best = float("inf")
            outs = None
            for _ in range(self.repeats):
                ctx.synchronize()
                t0 = time.perf_counter()
                outs = self._once(ctx)
                ctx.synchronize()
                best = min(best, time.perf_counter() - t0)
            outs = [o.cpu().numpy() if hasattr(o, "cpu") else np.asarray(o) for o in outs]
            return best, outs
        finally:
            ctx.close()


def tune(images, lambdas, connectivity=4, tree_type="max", attr_idx=0, decision=0, batched=False, repeats=3,
         sweeps=1, device=0, knobs=None, log=None, grid=None):
    """Coordinate descent over the knobs.  grid: tune the tiled build with these tiles per axis.  Returns
    (best setting, best seconds, table of (setting, seconds)).  Raises RuntimeError when a candidate's
    output differs from the first candidate's."""
    images = [np.ascontiguousarray(a) for a in images]
    is3d = images[0].ndim == 3
    knobs = knobs or knobs_for(is3d, batched, grid is not None)
    runner = _Runner(images, lambdas, connectivity, tree_type, attr_idx, decision, batched, repeats, device, grid)
    best = {k: v[0] for k, v in knobs.items()}
    best_s, ref = runner.run(best)
    table = [(dict(best), best_s)]
    if log:
        log("default %s: %.2f ms" % (env_line(best), best_s * 1e3))
    for _ in range(sweeps):
        for k, values in knobs.items():
            for v in values:
                if v == best[k]:
                    continue
                cand = dict(best)
                cand[k] = v
                s, outs = runner.run(cand)
                if any(not np.array_equal(a, b) for a, b in zip(outs, ref)):
                    raise RuntimeError("setting %s changed the output" % env_line(cand))
                table.append((cand, s))
                if log:
                    log("%s: %.2f ms" % (env_line(cand), s * 1e3))
                if s < best_s:
                    best, best_s = cand, s
    return best, best_s, table


# side of the sample when --size is not given: a 2048^2 image, a 256^3 volume, batch sides up to 2048, a
# 4096^2 image for the tiled build
DEFAULT_SIZE = {"vnoise": 2048, "spheres": 256, "batch": 2048, "tiled": 4096}


def _workload(args):
    size = args.size or DEFAULT_SIZE[args.workload]
    if args.workload in ("vnoise", "tiled"):
        return [synth.vnoise(size, size)], [1000.0], 8, False
    if args.workload == "spheres":
        return [synth.spheres(size, size, size)], [500.0], 6, False
    shapes = synth.batch_shapes(args.count, getattr(args, "lo", 64), size)
    imgs = [synth.disks(H, W, is16, seed) for (W, H, is16, seed, lam) in shapes]
    lams = [lam for (W, H, is16, seed, lam) in shapes]
    return imgs, lams, 4, True


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--workload", choices=["vnoise", "spheres", "batch", "tiled"], default="vnoise")
    ap.add_argument("--size", type=int, default=0,
                    help="side of the image (vnoise: 2048, tiled: 4096), of the volume (spheres: 256) or the largest "
                         "batch side (2048)")
    ap.add_argument("--count", type=int, default=64, help="images in the batch")
    ap.add_argument("--lo", type=int, default=64, help="the smallest batch side")
    ap.add_argument("--tiles", type=int, default=8, help="tiled build: row slabs")
    ap.add_argument("--repeats", type=int, default=3)
    ap.add_argument("--sweeps", type=int, default=1)
    ap.add_argument("--device", type=int, default=0)
    args = ap.parse_args(argv)
    imgs, lams, conn, batched = _workload(args)
    grid = (1, args.tiles) if args.workload == "tiled" else None
    best, s, _ = tune(imgs, lams, connectivity=conn, decision=0, batched=batched, repeats=args.repeats,
                      sweeps=args.sweeps, device=args.device, log=lambda m: print(m, flush=True), grid=grid)
    print("best: %.2f ms" % (s * 1e3))
    print(env_line(best))


if __name__ == "__main__":
    main()
