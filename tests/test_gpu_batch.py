"""The batched form (dcf_batch_filter): every image of an uneven batch is filtered with its own
tree and lambda; each output must equal the oracle's output for that image alone, bit for bit."""
import sys

import numpy as np
import pytest

import disccofan_amd as dcf
from disccofan_amd import synth
from oracle.oracle import OracleTree

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    # test_device_pointers uses torch: import it before the library loads (its HIP runtime first)
    import torch  # noqa: F401
    c = dcf.Context(0)
    yield c
    c.close()


def random_batch(rng, count, is3d=False, u16_share=0.5, maxside=120):
    imgs, lams = [], []
    for _ in range(count):
        if is3d:
            shape = tuple(int(v) for v in rng.integers(2, 18, size=3))
        else:
            shape = (int(rng.integers(1, maxside)), int(rng.integers(1, maxside)))
        if rng.random() < u16_share:
            img = rng.integers(0, int(rng.choice([3, 50, 4000, 65536])), size=shape).astype(np.uint16)
        else:
            img = rng.integers(0, int(rng.choice([2, 9, 256])), size=shape).astype(np.uint8)
        imgs.append(img)
        lams.append(float(rng.choice([1.0, 3.0, 7.5, 40.0, 1e9])))
    return imgs, lams


def check_batch(ctx, imgs, lams, conn, tree_type, decision=0, extract=False, attr=0, label=""):
    say("batch", label, len(imgs), "images conn", conn, tree_type, "decision", decision, "extract", extract)
    outs = ctx.batch_filter(imgs, lams, connectivity=conn, tree_type=tree_type, attr_idx=attr, decision=decision,
                            extract=extract)
    for k, (img, lam, out) in enumerate(zip(imgs, lams, outs)):
        o = OracleTree(img, conn, tree_type, attr)
        want = o.filter(attr, lam, decision, extract=extract).astype(img.dtype)
        bad = np.nonzero(out.reshape(-1) != want.reshape(-1))[0]
        assert len(bad) == 0, "%s image %d %s %s lam %g: %d pixels differ, first %s" % (
            label, k, img.shape, img.dtype, lam, len(bad), bad[:5])


@pytest.mark.parametrize("seed", range(3))
@pytest.mark.parametrize("conn", [4, 8])
@pytest.mark.parametrize("tree_type", ["max", "min"])
def test_mixed_batch(ctx, seed, conn, tree_type):
    rng = np.random.default_rng(500 + seed)
    imgs, lams = random_batch(rng, 14)
    check_batch(ctx, imgs, lams, conn, tree_type, decision=0, label="mixed s%d" % seed)


def test_rules_and_extract(ctx):
    rng = np.random.default_rng(77)
    imgs, lams = random_batch(rng, 9)
    for dec in (1, 2, 3):
        check_batch(ctx, imgs, lams, 4, "max", decision=dec, label="rule")
    check_batch(ctx, imgs, lams, 8, "min", decision=0, extract=True, label="extract")


def test_all_u8_batch(ctx):
    rng = np.random.default_rng(91)
    imgs, lams = random_batch(rng, 20, u16_share=0.0)
    check_batch(ctx, imgs, lams, 4, "max", label="u8 only")


def test_3d_batch(ctx):
    rng = np.random.default_rng(93)
    imgs, lams = random_batch(rng, 8, is3d=True)
    check_batch(ctx, imgs, lams, 6, "max", label="3d")
    check_batch(ctx, imgs, lams, 26, "min", label="3d")


def test_non_area_attribute(ctx):
    """the per-node lambda goes through the general attribute evaluation (here non-compactness)"""
    rng = np.random.default_rng(95)
    imgs, _ = random_batch(rng, 6, maxside=60)
    lams = [0.2, 0.3, 0.5, 0.25, 1.0, 0.18]
    check_batch(ctx, imgs, lams, 4, "max", decision=0, attr=17, label="ncomp")


def test_max_and_min_rules_non_increasing(ctx):
    """the max and min rules of non-compactness with per-image lambdas (the reference's raster-scan result
    per image); the images' greys start at 1, so a failing max-tree root spreads zeros"""
    rng = np.random.default_rng(96)
    imgs, _ = random_batch(rng, 8, maxside=50)
    imgs = [np.maximum(a, 1) for a in imgs]
    lams = [0.2, 0.3, 0.5, 0.25, 1.0, 0.18, 0.4, 0.22]
    for dec in (1, 2):
        check_batch(ctx, imgs, lams, 8, "max", decision=dec, attr=17, label="rule %d ncomp" % dec)
        check_batch(ctx, imgs, lams, 4, "min", decision=dec, attr=17, extract=True, label="rule %d ncomp min-tree" % dec)


def test_full_batch256_workload(ctx):
    """the batched workload at its real size: all 256 images (sides 64..2048, u8 and u16, 289 Mpixel) in
    one call with their own lambdas, every output against the oracle's tree of that image alone"""
    shapes = synth.batch_shapes(256, 64, 2048)
    imgs = [synth.disks(H, W, is16, seed) for (W, H, is16, seed, lam) in shapes]
    lams = [lam for (W, H, is16, seed, lam) in shapes]
    say("batch256: %d images, %d pixels" % (len(imgs), sum(a.size for a in imgs)))
    runs = {"context defaults": ctx.batch_filter(imgs, lams, connectivity=4, tree_type="max", attr_idx=0, decision=0)}
    # four workers (two batch trees per key width) with two refining loops each
    c4 = dcf.Context(0)
    c4.set_option("DCF_WORKERS", "4")
    c4.set_option("DCF_WORKER_LOOPS", "2")
    runs["4 workers, 2 loops"] = c4.batch_filter(imgs, lams, connectivity=4, tree_type="max", attr_idx=0, decision=0)
    c4.close()
    for k, (img, lam) in enumerate(zip(imgs, lams)):
        want = OracleTree(img, 4, "max", 0, threads=16).filter(0, lam, 0).astype(img.dtype)
        for how, outs in runs.items():
            bad = np.nonzero(outs[k].reshape(-1) != want.reshape(-1))[0]
            assert len(bad) == 0, "batch256 (%s) image %d %s %s lam %g: %d pixels differ, first %s" % (
                how, k, img.shape, img.dtype, lam, len(bad), bad[:5])
    say("batch256: 256 images bit-exact in both runs")


def test_uneven_workload_batch(ctx):
    """the uneven-size batch generator of the batched workload, first 24 images, sides 64..512"""
    shapes = synth.batch_shapes(24, 64, 512)
    imgs = [synth.disks(H, W, is16, seed) for (W, H, is16, seed, lam) in shapes]
    lams = [lam for (W, H, is16, seed, lam) in shapes]
    check_batch(ctx, imgs, lams, 4, "max", label="workload")


@pytest.mark.parametrize("conn", [4, 8])
def test_per_image_tree_types(ctx, conn):
    """every image with its own tree type (DCF_IMAGE_MAX_TREE / DCF_IMAGE_MIN_TREE) in one call,
    u8 and u16 mixed, against the oracle's tree of that type for the image alone"""
    rng = np.random.default_rng(310 + conn)
    imgs, lams = random_batch(rng, 12)
    tts = [("min" if rng.random() < 0.5 else "max") for _ in imgs]
    for dec in (0, 3):
        say("per-image tree types", conn, "decision", dec, tts)
        outs = ctx.batch_filter(imgs, lams, connectivity=conn, tree_type="max", decision=dec, tree_types=tts)
        for k, (img, lam, tt, out) in enumerate(zip(imgs, lams, tts, outs)):
            want = OracleTree(img, conn, tt, 0).filter(0, lam, dec).astype(img.dtype)
            bad = np.nonzero(out.reshape(-1) != want.reshape(-1))[0]
            assert len(bad) == 0, "image %d %s %s: %d pixels differ" % (k, img.shape, tt, len(bad))


def test_max_and_min_tree_of_one_image(ctx):
    """the two trees of one u16 image in one call (the pair the 8192^2 workload builds), subtractive rule"""
    img = synth.vnoise(300, 420)
    outs = ctx.batch_filter([img, img], [1000.0, 1000.0], connectivity=8, decision=3, tree_types=["max", "min"])
    for tt, out in zip(("max", "min"), outs):
        want = OracleTree(img, 8, tt, 0).filter(0, 1000.0, 3).astype(np.uint16)
        assert np.array_equal(out, want), tt


def test_min_rule_when_roots_fail(ctx):
    """the min rule with per-image lambdas above some images' areas, every image's root grey above 0"""
    rng = np.random.default_rng(321)
    imgs = [(rng.integers(5, 300, size=tuple(int(v) for v in rng.integers(2, 60, size=2)))).astype(np.uint16)
            for _ in range(6)] + [rng.integers(9, 200, size=(int(rng.integers(2, 50)), 33)).astype(np.uint8) for _ in range(4)]
    lams = [float(a.size + 1) if k % 2 == 0 else float(a.size // 2) for k, a in enumerate(imgs)]
    tts = [("max" if k % 3 else "min") for k in range(len(imgs))]
    outs = ctx.batch_filter(imgs, lams, connectivity=4, decision=1, tree_types=tts)
    for k, (img, lam, tt, out) in enumerate(zip(imgs, lams, tts, outs)):
        want = OracleTree(img, 4, tt, 0).filter(0, lam, 1).astype(img.dtype)
        assert np.array_equal(out, want), (k, img.shape, tt, lam)


def test_device_pointers(ctx):
    import torch
    rng = np.random.default_rng(97)
    imgs, lams = random_batch(rng, 6)
    dev = torch.device("cuda:0")
    tin = [torch.from_numpy(a.view(np.int16) if a.dtype == np.uint16 else a).to(dev) for a in imgs]
    tout = [torch.empty_like(t) for t in tin]
    torch.cuda.synchronize()
    ctx.batch_filter_device([(ti.data_ptr(), a.dtype, a.shape, lam, to.data_ptr())
                             for ti, a, lam, to in zip(tin, imgs, lams, tout)], connectivity=8)
    host = ctx.batch_filter(imgs, lams, connectivity=8)
    for to, h, a in zip(tout, host, imgs):
        got = to.cpu().numpy().view(a.dtype)
        assert np.array_equal(got, h)


def test_batch_errors(ctx):
    a2 = np.zeros((5, 6), np.uint8)
    a3 = np.zeros((3, 4, 5), np.uint8)
    with pytest.raises(dcf.DcfError) as e:
        ctx.batch_filter([a2, a3], [1, 1])
    assert e.value.code == -1
    with pytest.raises(dcf.DcfError) as e:
        ctx.batch_filter([a2.astype(np.float32)], [1])
    assert e.value.code == -1
    with pytest.raises(dcf.DcfError) as e:
        ctx.batch_filter([], [])
    assert e.value.code == -1
    # an image's flags may name one tree type, not both
    items = (dcf.BatchImage * 1)()
    out = np.empty_like(a2)
    items[0].img = a2.ctypes.data
    items[0].dtype = 0
    items[0].dims = (dcf.C.c_uint64 * 3)(6, 5, 1)
    items[0].lam = 1.0
    items[0].out = out.ctypes.data
    items[0].flags = dcf.DCF_IMAGE_MAX_TREE | dcf.DCF_IMAGE_MIN_TREE
    assert dcf.lib().dcf_batch_filter(ctx._h, items, 1, 4, 0, 0, 0, 0, 0) == -1


def test_graph_captured_level_loops(ctx):
    """dcf_ctx_set_graph: every pass's level loop captured into one HIP graph; three calls in a row, with
    eager calls between them, give the eager results (batch against the oracle, single trees against
    the eager build)"""
    shapes = synth.batch_shapes(20, 64, 400)
    imgs = [synth.disks(H, W, is16, seed) for (W, H, is16, seed, lam) in shapes]
    lams = [lam for (W, H, is16, seed, lam) in shapes]
    vol = synth.spheres(40, 48, 56)
    u16 = synth.vnoise(300, 410)
    eager = [OracleTree(img, 4, "max", 0).filter(0, lam, 0).astype(img.dtype) for img, lam in zip(imgs, lams)]
    e_vol = ctx.build(vol, 6, "max", 0).parent()
    e_u16 = ctx.build(u16, 8, "min", 17).filter(17, 0.5, 3)
    try:
        for rep in range(3):
            ctx.set_graph(True)
            say("graph batch replay", rep)
            outs = ctx.batch_filter(imgs, lams, connectivity=4, tree_type="max")
            g_vol = ctx.build(vol, 6, "max", 0).parent()
            g_u16 = ctx.build(u16, 8, "min", 17).filter(17, 0.5, 3)
            ctx.set_graph(False)
            for k, (out, want) in enumerate(zip(outs, eager)):
                assert np.array_equal(out, want), "graph batch image %d (replay %d)" % (k, rep)
            assert np.array_equal(g_vol, e_vol), "graph volume parent (replay %d)" % rep
            assert np.array_equal(g_u16, e_u16), "graph u16 min-tree filter (replay %d)" % rep
            # eager work between the replays
            assert np.array_equal(ctx.build(vol, 6, "max", 0).parent(), e_vol)
    finally:
        ctx.set_graph(False)


@pytest.mark.parametrize("workers", [1, 2, 3])
def test_large_images_on_worker_contexts(workers):
    """images of at least DCF_LARGE_PIXELS pixels are built one by one, the others in the u8 and u16 batch trees,
    these parts spread over worker contexts (own stream and host thread each) and run concurrently; one worker
    runs them in turn on the context's own stream.  Every output still equals the oracle's.  The threshold is
    lowered so that small images take the large-image path; u8 and u16, max- and min-trees, extract on"""
    c = dcf.Context(0)
    c.set_option("DCF_LARGE_PIXELS", "5000")
    c.set_option("DCF_WORKERS", str(workers))
    rng = np.random.default_rng(40 + workers)
    imgs, lams = random_batch(rng, 14, maxside=140)
    imgs += [synth.vnoise(130, 150), synth.disks(120, 90), synth.vnoise(100, 101)]
    lams += [300.0, 80.0, 55.0]
    tts = [("min" if rng.random() < 0.5 else "max") for _ in imgs]
    n_large = sum(a.size >= 5000 for a in imgs)
    say("worker contexts", workers, "large images", n_large, "of", len(imgs))
    assert n_large >= 4
    for dec in (0, 3):
        outs = c.batch_filter(imgs, lams, connectivity=8, tree_type="max", decision=dec, extract=dec == 3, tree_types=tts)
        for k, (img, lam, tt, out) in enumerate(zip(imgs, lams, tts, outs)):
            want = OracleTree(img, 8, tt, 0).filter(0, lam, dec, extract=dec == 3).astype(img.dtype)
            bad = np.nonzero(out.reshape(-1) != want.reshape(-1))[0]
            assert len(bad) == 0, "image %d %s %s decision %d: %d pixels differ" % (k, img.shape, tt, dec, len(bad))
    c.close()


def _small_images(rng, is3d):
    """images on both sides of the LDS build's 4096-pixel limit (from one pixel up), flat ones, u8 and u16, few
    and many grey levels"""
    if is3d:
        shapes = [(2, 1, 1), (2, 2, 2), (16, 16, 16), (16, 16, 17), (4, 64, 64), (8, 8, 37), (3, 5, 7)]
    else:
        shapes = [(1, 1), (1, 37), (37, 1), (64, 64), (64, 65), (128, 128), (17, 240), (100, 3), (90, 90)]
    imgs = []
    for k, shape in enumerate(shapes):
        levels = [2, 7, 256, 4000, 65536][k % 5]
        dt = np.uint16 if levels > 256 or k % 3 == 0 else np.uint8
        img = rng.integers(0, levels, size=shape).astype(dt)
        if k % 4 == 3:
            img[:] = img.reshape(-1)[0]
        imgs.append(img)
    imgs.append(synth.disks(96, 120, True, 7) if not is3d else synth.spheres(12, 20, 24))
    return imgs


@pytest.mark.parametrize("is3d,conn", [(False, 4), (False, 8), (True, 6), (True, 26)])
def test_small_images_built_in_lds(is3d, conn):
    """images of at most 4096 pixels are built whole in one workgroup's LDS (k_small_build): sorted by key,
    flooded with a union-find in that order, levelroots by pointer jumping; the larger ones take the level loop
    in the same call.  Every output equals the oracle's, for the max- and the min-tree of each image (one call,
    a tree type per image) and every rule, and equals the level loop's (DCF_SMALL_PIXELS=0)"""
    rng = np.random.default_rng(90 + conn)
    imgs = _small_images(rng, is3d)
    imgs = imgs + imgs
    tts = ["max"] * (len(imgs) // 2) + ["min"] * (len(imgs) // 2)
    lams = [float(rng.choice([1.0, 3.0, 40.0, 300.0, 1e9])) for _ in imgs]
    lds, loop = dcf.Context(0), dcf.Context(0)
    lds.set_option("DCF_SMALL_PIXELS", "4096")
    loop.set_option("DCF_SMALL_PIXELS", "0")
    for dec in range(4):
        say("LDS build", "3D" if is3d else "2D", "conn", conn, "decision", dec, len(imgs), "images")
        a = lds.batch_filter(imgs, lams, connectivity=conn, decision=dec, extract=dec == 1, tree_types=tts)
        b = loop.batch_filter(imgs, lams, connectivity=conn, decision=dec, extract=dec == 1, tree_types=tts)
        for k, (img, lam, tt) in enumerate(zip(imgs, lams, tts)):
            want = OracleTree(img, conn, tt, 0).filter(0, lam, dec, extract=dec == 1).astype(img.dtype)
            assert np.array_equal(a[k], want), "image %d %s %s %s lam %g decision %d" % (k, img.shape, img.dtype, tt, lam, dec)
            assert np.array_equal(b[k], want)
    lds.close()
    loop.close()


def test_small_images_with_other_attributes():
    """the LDS build accumulates the area only: with another attribute the small images take the level loop"""
    rng = np.random.default_rng(77)
    imgs = _small_images(rng, False)[:6]
    lams = [5.0] * len(imgs)
    c = dcf.Context(0)
    c.set_option("DCF_SMALL_PIXELS", "4096")
    for attr in (1, 13):
        check_batch(c, imgs, lams, 8, "max", decision=0, attr=attr, label="attr %d" % attr)
    c.close()


def test_many_small_images_queue():
    """the LDS build with more images than workgroups: every workgroup takes image after image from the queue"""
    rng = np.random.default_rng(5)
    imgs, lams = [], []
    for k in range(1500):
        side = int(rng.integers(1, 40)), int(rng.integers(1, 40))
        imgs.append(rng.integers(0, [3, 256, 65536][k % 3], size=side).astype(np.uint16 if k % 3 == 2 else np.uint8))
        lams.append(float(rng.choice([1.0, 4.0, 20.0])))
    c = dcf.Context(0)
    c.set_option("DCF_SMALL_PIXELS", "4096")
    outs = c.batch_filter(imgs, lams, connectivity=4, decision=0)
    for k in range(0, 1500, 7):
        want = OracleTree(imgs[k], 4, "max", 0).filter(0, lams[k], 0).astype(imgs[k].dtype)
        assert np.array_equal(outs[k], want), "image %d %s" % (k, imgs[k].shape)
    c.close()


def test_batch_tree_refining_loops():
    """one worker: the batch trees run in turn on the context itself, and a u16 batch tree of at least 2^22 pixels
    refines in three independent level loops (coarse regions dealt by node id); every output equals the oracle's"""
    rng = np.random.default_rng(31)
    imgs = [synth.vnoise(1400, 1300), synth.disks(1100, 1300, True, 5), synth.vnoise(900, 1500),
            rng.integers(0, 65536, size=(600, 700)).astype(np.uint16), synth.disks(500, 400, False, 8)]
    assert sum(a.size for a in imgs if a.dtype == np.uint16) >= 1 << 22
    lams = [300.0, 50.0, 1000.0, 20.0, 80.0]
    tts = ["max", "min", "max", "min", "max"]
    c = dcf.Context(0)
    c.set_option("DCF_WORKERS", "1")
    for dec in (0, 3):
        say("batch trees on the context itself, decision", dec)
        outs = c.batch_filter(imgs, lams, connectivity=8, decision=dec, tree_types=tts)
        for k, (img, lam, tt) in enumerate(zip(imgs, lams, tts)):
            want = OracleTree(img, 8, tt, 0, threads=16).filter(0, lam, dec).astype(img.dtype)
            assert np.array_equal(outs[k], want), "image %d %s %s decision %d" % (k, img.shape, tt, dec)
    c.close()
