"""GPU (libdcf.so, HIP on gfx950) against the CPU oracle, bit for bit.

Integer images: canonical parent arrays, check images, filtered images and spectra must be
identical.  Every case prints its name/shape before running so a failure names it.
"""
import os
import sys

import numpy as np
import pytest

import disccofan_amd as dcf
from disccofan_amd import synth
from oracle.oracle import OracleTree, lvec_read

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = dcf.Context(0)
    yield c
    c.close()


def assert_same(a, b, what):
    a = np.asarray(a)
    b = np.asarray(b)
    if not np.array_equal(a, b):
        bad = np.nonzero(a.reshape(-1) != b.reshape(-1))[0]
        raise AssertionError("%s differs at %d of %d positions, first %s: gpu %r oracle %r" % (
            what, len(bad), a.size, bad[:5], a.reshape(-1)[bad[:5]], b.reshape(-1)[bad[:5]]))


# ------------------------------------------------------------------------------------------------
# reference golden files
# ------------------------------------------------------------------------------------------------
def test_lena_golden(ctx, lena):
    img = lena["img"]
    say("lena 512x512 u8 4-conn")
    t = ctx.build(img, 4, "max", 0)
    assert_same(t.check(0), lena["check_area4"].astype(np.float32), "check area 4-conn")
    assert_same(t.filter(0, 200, 0), lena["filter_area200"].astype(np.float32), "filter area 200")
    assert t.num_nodes == len(lena["tree_index"])
    o = OracleTree(img, 4, "max", 0)
    assert_same(t.parent(), o.parent().astype(np.int32), "parent")
    t.accumulate(17)
    assert_same(t.check(17), lena["check_ncomp4"], "check ncomp")
    assert_same(t.filter(17, 0.2, 0), lena["filter_ncomp02"].astype(np.float32), "filter ncomp 0.2")
    say("lena 512x512 u8 8-conn")
    t8 = ctx.build(img, 8, "max", 0)
    assert_same(t8.check(0), lena["check_area8"].astype(np.float32), "check area 8-conn")


def test_lena_operations_driver(ctx, lena):
    say("run_operations on lena")
    lv = os.path.join(GOLDEN, "lvec.txt")
    ops = [("filter", 0, 200), ("filter", 17, 0.2), ("check", 17), ("pattern", 0, lv, 1)]
    outs = ctx.run_operations(lena["img"], ops, connectivity=4)
    assert_same(outs[0], lena["filter_area200"].astype(np.float32), "op filter area")
    assert_same(outs[1], lena["filter_ncomp02"].astype(np.float32), "op filter ncomp")
    assert_same(outs[2], lena["check_ncomp4"], "op check ncomp")
    o = OracleTree(lena["img"], 4, "max", 0)
    assert_same(outs[3], o.spectrum1d(0, lvec_read(lv)), "op spectrum")


# ------------------------------------------------------------------------------------------------
# seeded random images, every connectivity and both tree types
# ------------------------------------------------------------------------------------------------
def _compare_all(ctx, img, conn, tree_type, lam_area=20, lam_nc=0.25, label=""):
    say("case", label, img.shape, img.dtype, conn, tree_type)
    o = OracleTree(img, conn, tree_type, 0)
    t = ctx.build(img, conn, tree_type, 0)
    assert_same(t.parent(), o.parent().astype(np.int32), "parent " + label)
    assert_same(t.check(0), o.check(0), "area " + label)
    for dec in (0, 3):
        assert_same(t.filter(0, lam_area, dec), o.filter(0, lam_area, dec), "filter area d%d %s" % (dec, label))
        assert_same(t.filter(0, lam_area, dec, all_rules=True), o.filter(0, lam_area, dec), "general d%d %s" % (dec, label))
    for dec in (1, 2):
        assert_same(t.filter(0, lam_area, dec), o.filter(0, lam_area, dec), "filter area d%d %s" % (dec, label))
    assert_same(t.filter(0, lam_area, 2, all_rules=True), o.filter(0, lam_area, 2), "general max " + label)
    assert_same(t.filter(0, lam_area, 1, all_rules=True), o.filter(0, lam_area, 1), "general min " + label)
    assert_same(t.filter(0, lam_area, 0, extract=True), o.filter(0, lam_area, 0, extract=True), "extract " + label)
    lv = synth.log_lambdas(16, 1, 2 ** 14)
    assert_same(t.spectrum1d(0, lv), o.spectrum1d(0, lv), "spectrum " + label)
    o.reaccumulate(17)
    t.accumulate(17)
    assert_same(t.check(17), o.check(17), "ncomp " + label)
    assert_same(t.check(13), o.check(13), "trace " + label)
    for dec in (0, 1, 2, 3):
        assert_same(t.filter(17, lam_nc, dec), o.filter(17, lam_nc, dec), "filter ncomp d%d %s" % (dec, label))
    l1 = synth.log_lambdas(8, 0.05, 2.0)
    l2 = synth.log_lambdas(8, 1, 4096)
    assert_same(t.spectrum2d(17, l1, 0, l2), o.spectrum2d(17, l1, 0, l2), "spectrum2d " + label)


@pytest.mark.parametrize("seed", range(6))
def test_random_2d(ctx, seed):
    rng = np.random.default_rng(100 + seed)
    H, W = int(rng.integers(1, 150)), int(rng.integers(1, 150))
    levels = [2, 7, 256, 65536][seed % 4]
    dt = np.uint8 if levels <= 256 else np.uint16
    img = rng.integers(0, levels, size=(H, W)).astype(dt)
    for conn in (4, 8):
        for tt in ("max", "min"):
            _compare_all(ctx, img, conn, tt, label="rand2d s%d" % seed)


@pytest.mark.parametrize("seed", range(4))
def test_random_3d(ctx, seed):
    rng = np.random.default_rng(200 + seed)
    D, H, W = int(rng.integers(2, 20)), int(rng.integers(1, 40)), int(rng.integers(1, 40))
    levels = [3, 16, 256, 1000][seed % 4]
    dt = np.uint8 if levels <= 256 else np.uint16
    img = rng.integers(0, levels, size=(D, H, W)).astype(dt)
    for conn in (6, 26):
        for tt in ("max", "min"):
            _compare_all(ctx, img, conn, tt, label="rand3d s%d" % seed)


@pytest.mark.parametrize("tile2d,tile3d", [("32x32", "8"), ("128x32", "12"), ("16x128", "16")])
def test_zone_tiles(monkeypatch, tile2d, tile3d):
    """other zone tiles (DCF_ZONE_TILE2D / DCF_ZONE_TILE3D, non-square included), with the zone graph forced
    on: single images, and a batch, whose level lists and per-image tile tables follow the same tiles"""
    monkeypatch.setenv("DCF_ZONES", "1")
    monkeypatch.setenv("DCF_ZONE_TILE2D", tile2d)
    monkeypatch.setenv("DCF_ZONE_TILE3D", tile3d)
    c = dcf.Context(0)
    try:
        rng = np.random.default_rng(900 + len(tile2d))
        cases = [(synth.vnoise(150, 203), 8), (rng.integers(0, 9, size=(97, 61)).astype(np.uint8), 4),
                 (rng.integers(0, 4, size=(19, 23, 29)).astype(np.uint8), 6),
                 (rng.integers(0, 300, size=(13, 17, 21)).astype(np.uint16), 26)]
        for img, conn in cases:
            for tt in ("max", "min"):
                label = "zone tiles %s / %s %s %s" % (tile2d, tile3d, img.dtype, img.shape)
                say(label, conn, tt)
                o = OracleTree(img, conn, tt, 0)
                t = c.build(img, conn, tt, 0)
                assert_same(t.parent(), o.parent().astype(np.int32), "parent " + label)
                assert_same(t.filter(0, 15, 3), o.filter(0, 15, 3), "subtractive " + label)
                t.free()
        imgs = [synth.vnoise(int(rng.integers(20, 200)), int(rng.integers(20, 200))) for _ in range(5)]
        imgs += [rng.integers(0, 6, size=(int(rng.integers(5, 150)), int(rng.integers(5, 150)))).astype(np.uint8) for _ in range(5)]
        lams = [float(rng.choice([3, 20, 100])) for _ in imgs]
        say("zone tiles batch", tile2d)
        outs = c.batch_filter(imgs, lams, connectivity=4, decision=0)
        for img, lam, out in zip(imgs, lams, outs):
            want = OracleTree(img, 4, "max", 0).filter(0, lam, 0).astype(img.dtype)
            assert_same(out, want, "batch zone tiles %s %s" % (tile2d, img.shape))
    finally:
        c.close()


@pytest.mark.parametrize("zones,targets,records", [("0", "1", "1"), ("1", "0", "1"), ("1", "1", "0"), ("0", "0", "0")])
def test_build_paths(monkeypatch, zones, targets, records):
    """the build's optional paths on single images, each on and off: the zone graph of the first pass
    (DCF_ZONES, 1 forces it also for 3D u8 and zone-rich images), the union targets precomputed by the
    level scatter (DCF_TARGETS) and the 16-byte union-find records of refining passes (DCF_RECORDS); the
    library reads them when a context is created"""
    monkeypatch.setenv("DCF_ZONES", zones)
    monkeypatch.setenv("DCF_TARGETS", targets)
    monkeypatch.setenv("DCF_RECORDS", records)
    c = dcf.Context(0)
    try:
        rng = np.random.default_rng(700 + 4 * int(zones) + 2 * int(targets) + int(records))
        cases = [
            (rng.integers(0, 65536, size=(67, 91)).astype(np.uint16), 8),
            (synth.vnoise(96, 128), 4),
            (rng.integers(0, 12, size=(73, 59)).astype(np.uint8), 8),
            (rng.integers(0, 5, size=(11, 17, 19)).astype(np.uint8), 6),
            (rng.integers(0, 3000, size=(9, 13, 15)).astype(np.uint16), 26),
            (rng.normal(size=(41, 37)).astype(np.float32), 8),
        ]
        for img, conn in cases:
            for tt in ("max", "min"):
                label = "zones=%s targets=%s records=%s %s %s" % (zones, targets, records, img.dtype, img.shape)
                say("build path", label, conn, tt)
                o = OracleTree(img, conn, tt, 0)
                t = c.build(img, conn, tt, 0)
                assert_same(t.parent(), o.parent().astype(np.int32), "parent " + label)
                assert_same(t.check(0), o.check(0), "area " + label)
                assert_same(t.filter(0, 15, 3), o.filter(0, 15, 3), "subtractive " + label)
                t.accumulate(17)
                o.reaccumulate(17)
                assert_same(t.check(17), o.check(17), "ncomp " + label)
                t.free()
    finally:
        c.close()


@pytest.mark.parametrize("tree_type", ["max", "min"])
def test_min_rule_when_the_root_fails(ctx, tree_type):
    """the min rule with lambda above the image's area, on images whose root grey is not 0: the reference
    keeps the root's grey above the root's level (filtering.c:64-80), the direct rule gives 0"""
    rng = np.random.default_rng(81)
    for img, conn in ((rng.integers(100, 4000, size=(37, 29)).astype(np.uint16), 8),
                      (rng.integers(7, 40, size=(23, 31)).astype(np.uint8), 4),
                      (rng.integers(3, 20, size=(6, 9, 11)).astype(np.uint8), 26)):
        o = OracleTree(img, conn, tree_type, 0)
        t = ctx.build(img, conn, tree_type, 0)
        for lam in (img.size + 1, 1e9, img.size, img.size // 3):
            say("min rule", img.dtype, img.shape, conn, tree_type, lam)
            assert_same(t.filter(0, lam, 1), o.filter(0, lam, 1), "min rule lambda %g" % lam)
            assert_same(t.filter(0, lam, 1, extract=True), o.filter(0, lam, 1, extract=True), "min extract %g" % lam)
        t.free()


def _node_values(o, aidx, exact=True):
    """the distinct attribute values of the nodes; exact=False gives the midpoints between them, for an
    attribute whose GPU and CPU values may differ in the last bit (3D non-compactness divides by
    pow(A, 5/3), whose last bit differs between math libraries)"""
    par, g = o.parent(), o.grey()
    a, _ = o.attributes(aidx)
    vals = np.unique(a[(par < 0) | (g[np.maximum(par, 0)] != g)])
    return vals if exact or len(vals) < 2 else (vals[:-1] + vals[1:]) / 2


@pytest.mark.parametrize("tree_type", ["max", "min"])
def test_max_rule_of_non_increasing_attributes(ctx, tree_type):
    """the max rule of non-increasing attributes follows the reference's raster scan (filtering.c:156-208),
    including the stale values of pixels visited before their node is kept; lambdas sit exactly on node
    values (between them for 3D non-compactness)"""
    rng = np.random.default_rng(91 if tree_type == "max" else 92)
    cases = [(rng.integers(0, 12, size=(37, 29)).astype(np.uint8), 4),
             (rng.integers(0, 9, size=(41, 23)).astype(np.uint8), 8),
             (rng.integers(100, 140, size=(33, 35)).astype(np.uint16), 8),
             (rng.integers(0, 6, size=(7, 9, 11)).astype(np.uint8), 6),
             (rng.integers(0, 5, size=(6, 8, 9)).astype(np.uint8), 26)]
    for img, conn in cases:
        for aidx in (17, 13, 2, 6):
            o = OracleTree(img, conn, tree_type, aidx)
            t = ctx.build(img, conn, tree_type, aidx)
            vals = _node_values(o, aidx, exact=not (img.ndim == 3 and aidx == 17))
            for lam in (vals[len(vals) // 2], vals[len(vals) // 4], vals[-1] + 1, rng.choice(vals)):
                say("max rule", img.dtype, img.shape, conn, tree_type, aidx, lam)
                assert_same(t.filter(aidx, lam, 2), o.filter(aidx, lam, 2), "max rule a%d lambda %r" % (aidx, lam))
            lam = vals[len(vals) // 3]
            assert_same(t.filter(aidx, lam, 2, extract=True), o.filter(aidx, lam, 2, extract=True), "max extract a%d" % aidx)
            t.free()


@pytest.mark.parametrize("tree_type", ["max", "min"])
def test_min_rule_of_non_increasing_attributes(ctx, tree_type):
    """the min rule of non-increasing attributes follows the reference's raster scan (filtering.c:55-100),
    including the zeros a failing root of nonzero grey spreads; half the lambdas make the root fail"""
    rng = np.random.default_rng(94 if tree_type == "max" else 95)
    cases = [(rng.integers(1, 12, size=(37, 29)).astype(np.uint8), 4),
             (rng.integers(2, 9, size=(41, 23)).astype(np.uint8), 8),
             (rng.integers(100, 140, size=(33, 35)).astype(np.uint16), 8),
             (rng.integers(1, 6, size=(7, 9, 11)).astype(np.uint8), 6),
             (rng.integers(1, 5, size=(6, 8, 9)).astype(np.uint8), 26)]
    for img, conn in cases:
        for aidx in (17, 13, 2, 6):
            o = OracleTree(img, conn, tree_type, aidx)
            t = ctx.build(img, conn, tree_type, aidx)
            vals = _node_values(o, aidx, exact=not (img.ndim == 3 and aidx == 17))
            par = o.parent()
            a, _ = o.attributes(aidx)
            above = vals[vals > a[par < 0][0]]
            lams = [vals[len(vals) // 2], vals[len(vals) // 4], vals[-1] + 1]
            if len(above):
                lams += [above[0], above[len(above) // 2]]
            for lam in lams:
                say("min rule", img.dtype, img.shape, conn, tree_type, aidx, lam)
                assert_same(t.filter(aidx, lam, 1), o.filter(aidx, lam, 1), "min rule a%d lambda %r" % (aidx, lam))
            lam = lams[-1]
            assert_same(t.filter(aidx, lam, 1, extract=True), o.filter(aidx, lam, 1, extract=True), "min extract a%d" % aidx)
            t.free()


@pytest.mark.parametrize("tree_type", ["max", "min"])
def test_max_and_min_rules_vnoise_1024(ctx, tree_type):
    """the max and min rules of non-compactness on a 1024^2 u16 value-noise image (deep tree, many nodes)"""
    img = synth.vnoise(1024, 1024)
    o = OracleTree(img, 8, tree_type, 17, threads=8)
    t = ctx.build(img, 8, tree_type, 17)
    for dec in (1, 2):
        for lam in (0.2, 0.5, 3.0):
            say("rule vnoise 1024", tree_type, dec, lam)
            assert_same(t.filter(17, lam, dec), o.filter(17, lam, dec), "rule %d vnoise lambda %g" % (dec, lam))
    t.free()


def test_min_and_max_rules_float(ctx):
    """float greys: positive noise as max- and min-tree; negatives with zeros as a min-tree"""
    rng = np.random.default_rng(93)
    pos = (rng.random((31, 37)) * 50 + 1).astype(np.float32)
    neg = rng.integers(-3, 4, size=(29, 33)).astype(np.float32)
    for img, tt in ((pos, "max"), (pos, "min"), (neg, "min")):
        o = OracleTree(img, 4, tt, 17)
        t = ctx.build(img, 4, tt, 17)
        vals = _node_values(o, 17)
        for dec in (1, 2):
            for lam in (vals[len(vals) // 2], vals[len(vals) // 5], vals[-1]):
                say("rule float", tt, dec, lam)
                assert_same(t.filter(17, lam, dec), o.filter(17, lam, dec), "rule %d float %s lambda %r" % (dec, tt, lam))
        t.free()


def test_random_sweep_all_rules(ctx):
    """300 seeded random cases: 2D / 3D shapes from 1 pixel up, u8 / u16 / f32 greys of few to many levels,
    every connectivity and tree type, several attribute groups, the four rules at the median check value
    and above every node's value; parent arrays and outputs bit for bit"""
    rng = np.random.default_rng(2026)
    for it in range(300):
        is3d = rng.random() < 0.35
        shape = tuple(int(v) for v in (rng.integers(1, 24, size=3) if is3d else rng.integers(1, 160, size=2)))
        kind = str(rng.choice(["u8", "u16", "f32"]))
        levels = int(rng.choice([2, 5, 40, 256, 4000, 65536]))
        if kind == "u8":
            img = rng.integers(0, min(levels, 256), size=shape).astype(np.uint8)
        elif kind == "u16":
            img = rng.integers(0, levels, size=shape).astype(np.uint16)
        else:
            img = (rng.integers(0, levels, size=shape) * rng.choice([0.25, 1.5, -0.75])).astype(np.float32)
            img[img == 0] = 0.5  # no zero grey below a negative root (refused for the scan rules)
        conn = int(rng.choice([6, 26] if is3d else [4, 8]))
        tt = str(rng.choice(["max", "min"]))
        attr = int(rng.choice([0, 17, 13, 1, 6, 9] if kind != "f32" else [0, 17, 1, 6]))
        label = "sweep %d %s %s conn %d %s attr %d" % (it, kind, shape, conn, tt, attr)
        o = OracleTree(img, conn, tt, attr)
        t = ctx.build(img, conn, tt, attr)
        assert_same(t.parent(), o.parent().astype(np.int32), "parent " + label)
        chk = o.check(attr)
        for lam in (float(np.median(chk)), float(np.max(chk)) * 1.01 + 1e-9):
            for dec in (0, 1, 2, 3):
                assert_same(t.filter(attr, lam, dec), o.filter(attr, lam, dec), "%s rule %d lambda %r" % (label, dec, lam))
        t.free()


@pytest.mark.parametrize("shape,conn", [((57, 83), 4), ((61, 47), 8), ((9, 23, 31), 6), ((7, 19, 17), 26)])
def test_every_attribute(ctx, shape, conn):
    """check() of all 18 attributes (every group of AttribsArray) against the oracle; the eigenvalue
    attributes 14-16 come from closed-form 2x2 / 3x3 solvers on both sides and may differ in the last
    float32 bit, every other attribute is bit-exact."""
    rng = np.random.default_rng(300 + conn)
    img = rng.integers(0, 40, size=shape).astype(np.uint8)
    say("every attribute", shape, conn)
    t = ctx.build(img, conn, "max", 0)
    o = OracleTree(img, conn, "max", 0)
    for a in range(18):
        t.accumulate(a)
        o.reaccumulate(a)
        g, w = t.check(a), o.check(a)
        if a in (14, 15, 16):
            ulp = np.abs(g.view(np.int32).astype(np.int64) - w.view(np.int32).astype(np.int64))
            assert ulp.max() <= 1, "attribute %d differs by %d ulp" % (a, ulp.max())
        else:
            assert_same(g, w, "check attr %d" % a)
        lam = float(np.median(w))
        if a not in (14, 15, 16):
            assert_same(t.filter(a, lam, 0), o.filter(a, lam, 0), "direct attr %d" % a)


@pytest.mark.parametrize("case", ["few_values", "all_distinct", "3d", "negatives", "ramp", "many_distinct"])
def test_float_images(ctx, case):
    """f32 images (the reference's own grey type, transform_float keys flood.c:362-371): tree, areas and
    every filter rule bit-exact, the subtractive rule with the reference's float rounding order.
    The build runs on dense ranks of the float keys in 8-bit digit passes: up to 256 distinct values
    take one pass, up to 65536 two, "ramp" and "many_distinct" (~10^5 distinct values) three; the
    ramp's tree is one chain, so the third pass's coarse tree is ~350 nodes deep."""
    rng = np.random.default_rng(600 + len(case))
    if case == "few_values":
        img = (rng.integers(0, 9, size=(61, 77)) * 0.25).astype(np.float32)
    elif case == "all_distinct":
        img = rng.random((53, 67)).astype(np.float32) * 1000
    elif case == "3d":
        img = rng.normal(size=(9, 21, 23)).astype(np.float32)
    elif case == "ramp":
        img = (np.arange(300 * 301, dtype=np.float64).reshape(300, 301) * 0.37 - 9000).astype(np.float32)
    elif case == "many_distinct":
        img = rng.normal(size=(317, 333)).astype(np.float32)
    else:
        img = (rng.normal(size=(40, 50)) * 100).astype(np.float32)
    for conn in ((6, 26) if img.ndim == 3 else (4, 8)):
        for tt in ("max", "min"):
            say("float", case, conn, tt)
            o = OracleTree(img, conn, tt, 0)
            t = ctx.build(img, conn, tt, 0)
            assert_same(t.parent(), o.parent().astype(np.int32), "parent f32 %s" % case)
            assert_same(t.check(0), o.check(0), "area f32 %s" % case)
            for dec in (0, 3):
                assert_same(t.filter(0, 12, dec), o.filter(0, 12, dec), "filter f32 %s d%d" % (case, dec))
            for dec in (1, 2):
                assert_same(t.filter(0, 12, dec), o.filter(0, 12, dec), "filter f32 %s d%d" % (case, dec))
            # float spectra are double sums of float products; their order of summation differs (the
            # reference's own result depends on its thread count): relative 1e-12
            lv = synth.log_lambdas(12, 1, 4096)
            gs, os_ = t.spectrum1d(0, lv), o.spectrum1d(0, lv)
            assert np.allclose(gs, os_, rtol=1e-12, atol=0), ("spectrum f32", case, gs, os_)


EDGE = {
    "1x1": np.array([[7]], np.uint8),
    "1xN": (np.arange(300) * 7 % 251).astype(np.uint8).reshape(1, 300),
    "Nx1": (np.arange(300) * 13 % 253).astype(np.uint8).reshape(300, 1),
    "constant": np.full((97, 131), 42, np.uint8),
    "ramp": (np.arange(130 * 140) % 65536).astype(np.uint16).reshape(130, 140),
    "checker": ((np.indices((66, 67)).sum(0) % 2) * 200).astype(np.uint8),
    "tile_edges": np.random.default_rng(5).integers(0, 5, size=(129, 65)).astype(np.uint8),
    "zeros_and_max": np.where(np.random.default_rng(6).random((70, 70)) < 0.5, 0, 255).astype(np.uint8),
}


@pytest.mark.parametrize("name", sorted(EDGE))
def test_edge_cases(ctx, name):
    img = EDGE[name]
    for conn in (4, 8):
        _compare_all(ctx, img, conn, "max", lam_area=3, label=name)
    _compare_all(ctx, img, 4, "min", lam_area=3, label=name + " min")


def test_edge_3d_thin(ctx):
    rng = np.random.default_rng(9)
    for shape in [(3, 1, 1), (17, 1, 33), (2, 40, 1), (33, 17, 18)]:
        img = rng.integers(0, 9, size=shape).astype(np.uint8)
        for conn in (6, 26):
            _compare_all(ctx, img, conn, "max", lam_area=4, label="thin3d %s" % (shape,))


# ------------------------------------------------------------------------------------------------
# the benchmark workloads (reduced sizes here; full sizes in test_gpu_fullsize.py)
# ------------------------------------------------------------------------------------------------
def test_workload_disks_1024(ctx):
    img = synth.disks(1024, 1024)
    say("disks 1024^2 u8 4-conn max, sha", synth.sha256(img)[:16])
    o = OracleTree(img, 4, "max", 0)
    t = ctx.build(img, 4, "max", 0)
    assert_same(t.parent(), o.parent().astype(np.int32), "parent")
    assert_same(t.filter(0, 100, 0), o.filter(0, 100, 0), "filter")
    lv = np.zeros(32, np.float32)
    lv[1:] = 2.0 ** np.arange(31)
    assert_same(t.spectrum1d(0, lv), o.spectrum1d(0, lv), "spectrum")


@pytest.mark.parametrize("tree_type", ["max", "min"])
def test_workload_vnoise_1024(ctx, tree_type):
    img = synth.vnoise(1024, 1024)
    say("vnoise 1024^2 u16 8-conn", tree_type)
    o = OracleTree(img, 8, tree_type, 0, threads=8)
    t = ctx.build(img, 8, tree_type, 0)
    assert_same(t.parent(), o.parent().astype(np.int32), "parent")
    assert_same(t.filter(0, 1000, 3, out_dtype=np.uint16), o.filter(0, 1000, 3).astype(np.uint16), "filter u16 out")
    assert_same(t.filter(0, 1000, 3, all_rules=True), o.filter(0, 1000, 3), "filter general")
    lv = synth.log_lambdas(64, 1, 2 ** 26)
    assert_same(t.spectrum1d(0, lv), o.spectrum1d(0, lv), "spectrum")


def test_workload_spheres_64(ctx):
    img = synth.spheres(64, 64, 64)
    say("spheres 64^3 u8 6-conn")
    o = OracleTree(img, 6, "max", 0, threads=8)
    t = ctx.build(img, 6, "max", 0)
    assert_same(t.parent(), o.parent().astype(np.int32), "parent")
    assert_same(t.filter(0, 500, 0), o.filter(0, 500, 0), "filter")
    lv = synth.log_lambdas(48, 1, 2 ** 27)
    assert_same(t.spectrum1d(0, lv), o.spectrum1d(0, lv), "spectrum")


def test_workload_blobs_512(ctx):
    img = synth.blobs(512, 512)
    say("blobs 512^2 u16 8-conn inertia")
    o = OracleTree(img, 8, "max", 13, threads=8)
    t = ctx.build(img, 8, "max", 13)
    assert_same(t.parent(), o.parent().astype(np.int32), "parent")
    assert_same(t.check(13), o.check(13), "check trace")
    assert_same(t.check(17), o.check(17), "check ncomp")
    l1 = synth.log_lambdas(32, 0.16, 100.0)
    l2 = synth.log_lambdas(32, 1, 2 ** 24)
    assert_same(t.spectrum2d(17, l1, 0, l2), o.spectrum2d(17, l1, 0, l2), "pattern2D")


def test_tree_csv_matches_gt_tree(ctx, lena, tmp_path):
    """("tree",15) on lena, 4-connectivity: the CSV of write_tree (io.c:1431-1573) against the
    reference's GT_tree.txt, rebuilt line by line from its parsed columns with the reference's printf
    formats.  Every column is compared as text; elong / flat / spars come from closed-form
    eigenvalues (not EISPACK rs) and may differ by one unit of their last printed digit."""
    t = ctx.build(lena["img"], 4, "max", 15)
    path = tmp_path / "lena_tree.txt"
    t.write_csv(str(path), 15)
    lines = path.read_text().splitlines()
    assert lines[0] == ("index,x,y,z,gval,parent,flux,area,x_mean,y_mean,z_mean,x_wmean,y_wmean,z_wmean,"
                        "elong,flat,spars,ncomp")
    order = np.argsort(lena["tree_index"])
    cols = {k[5:]: lena[k][order] for k in lena if k.startswith("tree_")}
    n = len(order)
    assert len(lines) == n + 1
    fmt_int = ("index", "x", "y", "z")
    eig = {"elong": 14, "flat": 15, "spars": 16}
    bad = []
    for r in range(n):
        got = lines[r + 1].split(",")
        assert lines[r + 1].endswith(" "), "rows end with a space after ncomp"
        want = ["%d" % cols[c][r] for c in fmt_int]
        want += ["%f" % cols["gval"][r], "%d" % cols["parent"][r], "%f" % cols["flux"][r], "%d" % cols["area"][r]]
        want += ["%.2f" % cols[c][r] for c in ("x_mean", "y_mean", "z_mean", "x_wmean", "y_wmean", "z_wmean")]
        for j, w in enumerate(want):
            if got[j] != w:
                bad.append((r, j, got[j], w))
        for j, c in enumerate(("elong", "flat", "spars")):
            if abs(float(got[14 + j]) - cols[c][r]) > 0.0011:
                bad.append((r, 14 + j, got[14 + j], cols[c][r]))
        if got[17] != "%.4f " % cols["ncomp"][r]:
            bad.append((r, 17, got[17], "%.4f " % cols["ncomp"][r]))
    assert not bad, "%d cells differ, first %s" % (len(bad), bad[:5])


def test_tree_csv_area_group_3d(ctx, tmp_path):
    """group 0 (area) CSV of a 3D min-tree against the oracle's tree: one row per levelroot in raster
    order with index, coordinates, grey (inverted space), parent levelroot, flux and area"""
    rng = np.random.default_rng(47)
    img = rng.integers(0, 12, size=(6, 9, 11)).astype(np.uint8)
    t = ctx.build(img, 6, "min", 0)
    o = OracleTree(img, 6, "min", 0)
    path = tmp_path / "t.txt"
    t.write_csv(str(path), 0)
    lines = path.read_text().splitlines()
    assert lines[0] == "index,x,y,z,gval,parent,flux,area,"
    par = o.parent().reshape(-1)
    g = o.grey().reshape(-1).astype(np.float32)
    lr = np.nonzero((par < 0) | (g != g[np.maximum(par, 0)]))[0]
    area = o.check(0).reshape(-1)
    assert len(lines) == len(lr) + 1
    W, H = img.shape[2], img.shape[1]
    for r, i in enumerate(lr):
        p = par[i]
        dh = np.float32(g[i] - g[p]) if p >= 0 else g[i]
        sz = int(area[i])
        want = "%d,%d,%d,%d,%f,%d,%f,%d," % (i, i % (W * H) % W, i % (W * H) // W, i // (W * H), g[i], p,
                                             np.float32(dh * np.float32(sz)), sz)
        assert lines[r + 1] == want, (r, lines[r + 1], want)


@pytest.mark.parametrize("tree_type", ["max", "min"])
def test_dmp(ctx, tree_type):
    """the differential profile: g - filter(lambda_i) per lambda after lambda_0 (operation.c:63-77);
    a min-tree's profile stays in the inverted grey space"""
    rng = np.random.default_rng(41)
    img = rng.integers(0, 60, size=(70, 90)).astype(np.uint8)
    lv = synth.log_lambdas(6, 2, 500)
    t = ctx.build(img, 8, tree_type, 0)
    o = OracleTree(img, 8, tree_type, 0)
    prof = t.dmp(0, lv)
    assert prof.shape == (len(lv) - 1,) + img.shape
    for i in range(1, len(lv)):
        ext = o.filter(0, float(lv[i]), 0, extract=True)
        want = ext if tree_type == "max" else np.float32(o.gmax) - ext
        assert_same(prof[i - 1], want, "dmp %s lambda %g" % (tree_type, lv[i]))


def test_dmp_operation(ctx, tmp_path):
    rng = np.random.default_rng(43)
    img = rng.integers(0, 30, size=(40, 50)).astype(np.uint8)
    lvf = tmp_path / "lv.txt"
    lvf.write_text("#LambdaVector file. Keep this exact line in all similar lambda vector files\n3 10 40\n")
    outs = ctx.run_operations(img, [("dmp", 0, str(lvf))], connectivity=4)
    t = ctx.build(img, 4, "max", 0)
    assert_same(outs[0], t.dmp(0, lvec_read(str(lvf))), "dmp operation")


# ------------------------------------------------------------------------------------------------
# error behaviour (status codes, no device fault)
# ------------------------------------------------------------------------------------------------
def test_errors(ctx, tmp_path):
    img = np.zeros((8, 8), np.uint8)
    with pytest.raises(dcf.DcfError) as e:
        ctx.build(img, 5)
    assert e.value.code == -1
    with pytest.raises(dcf.DcfError) as e:
        ctx.build(img, 4, "max", 18)
    assert e.value.code == -1
    t = ctx.build(img, 4, "max", 0)
    with pytest.raises(dcf.DcfError) as e:
        t.check(17)
    assert e.value.code == -6
    with pytest.raises(dcf.DcfError) as e:
        t.filter(0, 5, 7)
    assert e.value.code == -1
    with pytest.raises(dcf.DcfError) as e:
        t.spectrum2d(0, [0, 1], 17, [0, 1])
    assert e.value.code == -1
    with pytest.raises(dcf.DcfError) as e:
        ctx.run_operations(img, [("pattern", 0, str(tmp_path / "none.txt"))])
    assert e.value.code == -5
    with pytest.raises(dcf.DcfError) as e:
        ctx.run_operations(img, [("csl", 0, "x")])
    assert e.value.code == -4
    # the min and max rules of a non-increasing attribute on a float max-tree with a node of grey 0
    # above its negative root: refused
    fimg = np.array([[-2.0, 0.0, 3.0], [0.0, 1.5, -2.0]], np.float32)
    tf = ctx.build(fimg, 4, "max", 17)
    for dec in (1, 2):
        with pytest.raises(dcf.DcfError) as e:
            tf.filter(17, 0.3, dec)
        assert e.value.code == -4


def test_dimension_errors(ctx):
    """dimensions whose product overflows 64 bits (2^22 x 2^22 x 2^20 wraps to 0) or passes 2^31 are
    refused before any copy or launch, in the single and the batched form; NULL dimensions are refused
    by the operator driver.  The arguments are raw C-ABI calls: the Python wrapper takes its
    dimensions from numpy shapes and cannot express these."""
    import ctypes as C
    L = dcf.lib()
    img = np.zeros(64, np.uint8)
    out = np.zeros(64, np.uint8)
    for dims in ((1 << 22, 1 << 22, 1 << 20), (1 << 30, 1 << 30, 1 << 30), (1 << 16, 1 << 15, 1),
                 (0x7FFFFFFF, 0x7FFFFFFF, 0x7FFFFFFF), (1 << 32, 1, 1)):
        say("dims", dims)
        d = (C.c_uint64 * 3)(*dims)
        h = C.c_void_p()
        rc = L.dcf_tree_build(ctx._h, img.ctypes.data, dcf.DCF_U8, d, 4, 0, 0, None, 0, C.byref(h))
        assert rc == -1 and not h.value, (dims, rc)
        b = (dcf.BatchImage * 1)()
        b[0].img, b[0].dtype, b[0].lam, b[0].out, b[0].flags = img.ctypes.data, dcf.DCF_U8, 1.0, out.ctypes.data, 0
        b[0].dims[:] = list(dims)
        assert L.dcf_batch_filter(ctx._h, b, 1, 4, 0, 0, 0, 0, 0) == -1, dims
    assert L.dcf_run_operations(ctx._h, img.ctypes.data, dcf.DCF_U8, None, 4, 0, None, b'("filter",0,5)', None,
                                None) == -1
    # spectra of more than DCF_MAX_SPECTRUM_BINS (2^28) bins are refused before anything is read
    say("spectrum bin limits")
    t = ctx.build(np.zeros((4, 4), np.uint8), 4, "max", 0)
    l1, l2, sp = np.zeros(1 << 15, np.float32), np.zeros(1 << 14, np.float32), np.zeros(4, np.float64)
    assert L.dcf_spectrum2d(t._h, 0, l1.ctypes.data, 1 << 15, 0, l2.ctypes.data, 1 << 14, sp.ctypes.data) == -1
    assert L.dcf_spectrum1d(t._h, 0, l1.ctypes.data, (1 << 28) + 1, sp.ctypes.data) == -1
    assert np.array_equal(t.spectrum1d(0, [0, 1, 4]), t.spectrum1d(0, [0, 1, 4]))


def test_c_caller_on_gpu(tmp_path):
    """a plain C program (tests/c/c_caller.c) drives build + filter + spectrum through the C ABI;
    its outputs agree with the oracle"""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    libdir = os.path.join(root, "disccofan_amd", "lib")
    exe = str(tmp_path / "c_caller")
    subprocess.run(["gcc", "-std=c99", "-O1", "-I", os.path.join(root, "include"), os.path.join(root, "tests", "c", "c_caller.c"),
                    "-L", libdir, "-ldcf", "-Wl,-rpath," + libdir, "-o", exe], check=True)
    say("c_caller run 300 200")
    r = subprocess.run([exe, "run", "300", "200"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    got = dict(l.split() for l in r.stdout.splitlines())
    # the same synthetic image as c_caller.c (64-bit LCG, top 5 bits)
    s = 12345
    vals = np.empty(300 * 200, np.uint8)
    for i in range(vals.size):
        s = (s * 6364136223846793005 + 1442695040888963407) % (1 << 64)
        vals[i] = s >> 59
    img = vals.reshape(200, 300)
    o = OracleTree(img, 4, "max", 0)
    assert int(got["nodes"]) == o.stats()["nodes"]
    assert int(got["filter_sum"]) == int(o.filter(0, 25.0, 0).astype(np.int64).sum())
    assert float(got["spectrum_sum"]) == float(got["grey_sum"]) == float(img.astype(np.int64).sum())


def test_context_options():
    """dcf_ctx_set_option: the build's tunables by their environment names; unknown names and values out of
    range are DCF_EINVAL; a changed schedule gives the same tree"""
    ctx = dcf.Context(0)
    img = synth.vnoise(200, 260)
    ref = ctx.build(img, 8, "max", 0).parent()
    for name, value in (("DCF_ZONES", "0"), ("DCF_ZONE_TILE2D", "32x32"), ("DCF_LEVEL_TILE2D", "32x32x1"),
                        ("DCF_FIX_BLOCKS", "64"), ("DCF_TARGETS", "0"), ("DCF_ZIP_SEGMENTS", "4")):
        ctx.set_option(name, value)
        assert np.array_equal(ctx.build(img, 8, "max", 0).parent(), ref), (name, value)
    for name, value in (("DCF_NOPE", "1"), ("DCF_ZONES", "7"), ("DCF_ZONE_TILE2D", "100x100"), ("DCF_ZONE_TILE2D", "64"),
                        ("DCF_LEVEL_TILE3D", "0x1x1"), ("DCF_FIX_BLOCKS", "12x"), ("DCF_ZIP_SEGMENTS", "0")):
        with pytest.raises(dcf.DcfError) as e:
            ctx.set_option(name, value)
        assert e.value.code == -1, (name, value)
    ctx.close()


@pytest.mark.parametrize("tree_type", ["max", "min"])
def test_refining_pass_loops(tree_type):
    """the refining pass of an image of at least 2^22 pixels runs as DCF_FINE_LOOPS independent level loops (coarse
    regions dealt by node id, each loop on its own stream): 1 to 4 loops give the oracle's parent array, areas
    and subtractive filter; out-of-range loop counts are refused"""
    img = synth.vnoise(2048, 2048)
    o = OracleTree(img, 8, tree_type, 0, threads=16)
    want_par = o.parent().astype(np.int32)
    want_out = o.filter(0, 300.0, 3)
    ctx = dcf.Context(0)
    for loops in (1, 2, 3, 4):
        ctx.set_option("DCF_FINE_LOOPS", str(loops))
        say("refining pass in", loops, "loops,", tree_type, "tree")
        t = ctx.build(img, 8, tree_type, 0)
        assert np.array_equal(t.parent(), want_par), loops
        assert np.array_equal(t.filter(0, 300.0, 3), want_out), loops
        t.free()
    for bad in ("0", "5"):
        with pytest.raises(dcf.DcfError):
            ctx.set_option("DCF_FINE_LOOPS", bad)
    ctx.close()
