"""The benchmark workloads at (or near) full size on the GPU, against the CPU oracle (OpenMP slab
build) bit for bit, plus size-independent properties where the oracle would take too long.
"""
import sys

import numpy as np
import pytest

import disccofan_amd as dcf
from disccofan_amd import synth
from oracle.oracle import OracleTree

pytestmark = pytest.mark.gpu
THREADS = 16


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = dcf.Context(0)
    yield c
    c.close()


def same(a, b, what):
    a = np.asarray(a).reshape(-1)
    b = np.asarray(b).reshape(-1)
    bad = np.nonzero(a != b)[0]
    assert len(bad) == 0, "%s differs at %d positions, first %s: gpu %r oracle %r" % (what, len(bad), bad[:5], a[bad[:5]],
                                                                                    b[bad[:5]])


@pytest.fixture(scope="module")
def vnoise8192():
    return synth.vnoise(8192, 8192)


@pytest.mark.parametrize("tree_type", ["max", "min"])
def test_vnoise_8192_u16(ctx, vnoise8192, tree_type):
    """8192^2 u16, 8-conn, area opening lambda=1000 subtractive, 64-bin area spectrum."""
    img = vnoise8192
    say("vnoise 8192^2 u16", tree_type, "sha", synth.sha256(img)[:16])
    t = ctx.build(img, 8, tree_type, 0)
    o = OracleTree(img, 8, tree_type, 0, threads=THREADS)
    same(t.parent(), o.parent().astype(np.int32), "parent")
    f = t.filter(0, 1000, 3, out_dtype=np.uint16)
    want = o.filter(0, 1000, 3).astype(np.uint16)
    _VNOISE_FILTERED[tree_type] = want
    same(f, want, "subtractive filter")
    lv = synth.log_lambdas(64, 1, 2 ** 26)
    same(t.spectrum1d(0, lv), o.spectrum1d(0, lv), "spectrum")


_VNOISE_FILTERED = {}


@pytest.mark.parametrize("loops", [2, 3])
def test_vnoise_8192_pair_on_workers(vnoise8192, loops):
    """the bench's step: the max-tree and the min-tree of the 8192^2 image in one batched call, the two trees on
    two worker contexts at once, each refining pass split into `loops` level loops on streams of their own"""
    img = vnoise8192
    for tt in ("max", "min"):
        if tt not in _VNOISE_FILTERED:
            _VNOISE_FILTERED[tt] = OracleTree(img, 8, tt, 0, threads=THREADS).filter(0, 1000, 3).astype(np.uint16)
    c = dcf.Context(0)
    c.set_option("DCF_WORKERS", "2")
    c.set_option("DCF_WORKER_LOOPS", str(loops))
    say("vnoise 8192^2 pair on two workers,", loops, "loops each")
    outs = c.batch_filter([img, img], [1000.0, 1000.0], connectivity=8, decision=3, tree_types=["max", "min"])
    c.close()
    for tt, out in zip(("max", "min"), outs):
        same(out, _VNOISE_FILTERED[tt], "%s-tree subtractive filter" % tt)


def test_spheres_256_u8(ctx):
    img = synth.spheres(256, 256, 256)
    say("spheres 256^3 u8 6-conn")
    t = ctx.build(img, 6, "max", 0)
    o = OracleTree(img, 6, "max", 0, threads=THREADS)
    same(t.parent(), o.parent().astype(np.int32), "parent")
    same(t.filter(0, 500, 0, out_dtype=np.uint8), o.filter(0, 500, 0).astype(np.uint8), "direct filter")
    lv = synth.log_lambdas(48, 1, 2 ** 27)
    same(t.spectrum1d(0, lv), o.spectrum1d(0, lv), "spectrum")


def test_spheres_512_u8(ctx):
    """512^3 u8, 6-conn max-tree at full size against the oracle: parent array, the volume opening
    lambda=500 with the direct rule and the 48-bin volume spectrum"""
    img = synth.spheres(512, 512, 512)
    say("spheres 512^3 u8 6-conn vs oracle, sha", synth.sha256(img)[:16])
    t = ctx.build(img, 6, "max", 0)
    o = OracleTree(img, 6, "max", 0, threads=THREADS)
    same(t.parent(), o.parent().astype(np.int32), "parent")
    same(t.filter(0, 500, 0, out_dtype=np.uint8), o.filter(0, 500, 0).astype(np.uint8), "direct filter")
    lv = synth.log_lambdas(48, 1, 2 ** 27)
    same(t.spectrum1d(0, lv), o.spectrum1d(0, lv), "spectrum")
    t.free()


def test_spheres_512_u8_properties(ctx):
    """512^3 u8: the full volume against closed-form properties."""
    img = synth.spheres(512, 512, 512)
    say("spheres 512^3 u8 6-conn properties")
    t = ctx.build(img, 6, "max", 0)
    # lambda = 1 keeps every node: the filter is the identity
    same(t.filter(0, 1, 0, out_dtype=np.uint8), img, "identity filter")
    # the pattern spectrum of a max-tree with background sums to the total grey, up to the
    # rounding of each node's float product dh * area (pattern.c:91 multiplies in float)
    lv = synth.log_lambdas(48, 1, 2 ** 27)
    spec = t.spectrum1d(0, lv)
    total = float(img.astype(np.int64).sum())
    assert abs(spec.sum() - total) <= 2.0 ** -23 * total, (spec.sum(), total)
    # the root is the one node whose area is the whole volume
    area = t.check(0)
    assert area.max() == np.float32(img.size)
    par = t.parent()
    assert (par < 0).sum() == 1


def test_blobs_4096_precision_form(ctx):
    """4096^2 f32 blobs quantised to u16, 8-conn, trace / ncomp and the 32x32 ncomp x area spectrum."""
    img = synth.blobs(4096, 4096)
    say("blobs 4096^2 u16 inertia")
    t = ctx.build(img, 8, "max", 13)
    o = OracleTree(img, 8, "max", 13, threads=THREADS)
    same(t.parent(), o.parent().astype(np.int32), "parent")
    same(t.check(13), o.check(13), "trace")
    same(t.check(17), o.check(17), "ncomp")
    l1 = synth.log_lambdas(32, 0.16, 100.0)
    l2 = synth.log_lambdas(32, 1, 2 ** 24)
    same(t.spectrum2d(17, l1, 0, l2), o.spectrum2d(17, l1, 0, l2), "pattern2D")


def test_float_four_passes(ctx):
    """4608x4096 f32 with ~17.8M distinct values (more than 2^24, so the build takes four 8-bit digit
    passes over the dense ranks): the ranks of a smooth field with ties broken at random, stored as
    float32 (exact below 2^24, every second integer above), so the tree is deep and nearly every
    pixel is its own node.  Parent array, direct and subtractive area filters, area spectrum."""
    H, W = 4608, 4096
    base = synth.vnoise(H, W).astype(np.float64)
    rng = np.random.default_rng(4)
    order = np.argsort(base.reshape(-1) + rng.random(H * W) * 0.5, kind="stable")
    img = np.empty(H * W, np.float32)
    img[order] = np.arange(H * W, dtype=np.float64).astype(np.float32)
    img = img.reshape(H, W)
    say("ranked float 4608x4096, distinct", len(np.unique(img)))
    for tt in ("max", "min"):
        say("ranked float", tt)
        t = ctx.build(img, 8, tt, 0)
        o = OracleTree(img, 8, tt, 0, threads=THREADS)
        same(t.parent(), o.parent().astype(np.int32), "parent " + tt)
        same(t.filter(0, 1000, 0), o.filter(0, 1000, 0), "direct filter " + tt)
        same(t.filter(0, 1000, 3), o.filter(0, 1000, 3), "subtractive filter " + tt)
        lv = synth.log_lambdas(32, 1, 2 ** 25)
        gs, os_ = t.spectrum1d(0, lv), o.spectrum1d(0, lv)
        assert np.allclose(gs, os_, rtol=1e-12, atol=0), (gs, os_)
        t.free()


@pytest.mark.parametrize("tree_type", ["max", "min"])
def test_vnoise_4096_rules_of_ncomp(ctx, tree_type):
    """4096^2 u16 value noise, 8-conn, non-compactness: all four rules, the min and max rules being the
    reference's raster-scan results (millions of nodes: the preorder range minima and doubling at size)"""
    img = synth.vnoise(4096, 4096)
    say("vnoise 4096 rules", tree_type)
    t = ctx.build(img, 8, tree_type, 17)
    o = OracleTree(img, 8, tree_type, 17, threads=THREADS)
    for dec in (0, 1, 2, 3):
        for lam in (0.3, 2.0):
            say("vnoise 4096 rule", tree_type, dec, lam)
            same(t.filter(17, lam, dec), o.filter(17, lam, dec), "rule %d lambda %g %s" % (dec, lam, tree_type))
    t.free()


@pytest.mark.parametrize("tree_type", ["max", "min"])
def test_vnoise_2048_u16_4conn(ctx, tree_type):
    """2048^2 u16 value noise at 4-connectivity (the default of the reference's 2D runs)"""
    img = synth.vnoise(2048, 2048)
    say("vnoise 2048^2 u16 4-conn", tree_type)
    t = ctx.build(img, 4, tree_type, 0)
    o = OracleTree(img, 4, tree_type, 0, threads=THREADS)
    same(t.parent(), o.parent().astype(np.int32), "parent")
    same(t.filter(0, 500, 3, out_dtype=np.uint16), o.filter(0, 500, 3).astype(np.uint16), "subtractive filter")
    same(t.filter(0, 500, 0, out_dtype=np.uint16), o.filter(0, 500, 0).astype(np.uint16), "direct filter")


def test_u16_volume_26conn(ctx):
    """160^3 u16 volume (blurred noise + spheres scaled to 16 bits) at 26-connectivity: two digit passes in 3D"""
    base = synth.spheres(160, 160, 160).astype(np.uint32)
    rng = np.random.default_rng(26)
    img = (base * 256 + rng.integers(0, 256, size=base.shape)).astype(np.uint16)
    say("u16 160^3 26-conn")
    for tt in ("max", "min"):
        t = ctx.build(img, 26, tt, 0)
        o = OracleTree(img, 26, tt, 0, threads=THREADS)
        same(t.parent(), o.parent().astype(np.int32), "parent " + tt)
        same(t.filter(0, 200, 0, out_dtype=np.uint16), o.filter(0, 200, 0).astype(np.uint16), "direct filter " + tt)
        lv = synth.log_lambdas(32, 1, 2 ** 22)
        same(t.spectrum1d(0, lv), o.spectrum1d(0, lv), "spectrum " + tt)
        t.free()
