"""The tiled build (dcf_tree_build_tiled): the image cut into tiles with the reference's one-pixel
overlap, every tile's tree built on its own and the boundary trees merged on the GPU, must give the
tree of the untiled build bit for bit -- canonical parent array, attributes, filters and spectra --
as the reference's own tests require of its distributed runs (tests/test_mpi.sh:143-172 compares the
outputs of 1, 2, 4 ... processes)."""
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
    c = dcf.Context(0)
    yield c
    c.close()


def same(a, b, what):
    a = np.asarray(a).reshape(-1)
    b = np.asarray(b).reshape(-1)
    bad = np.nonzero(a != b)[0]
    assert len(bad) == 0, "%s differs at %d positions, first %s: tiled %r untiled %r" % (what, len(bad), bad[:5],
                                                                                       a[bad[:5]], b[bad[:5]])


def compare(ctx, img, grid, conn, tt, attr, lams, spec_lv=None, label=""):
    say("tiles", label, img.shape, img.dtype, "grid", grid, "conn", conn, tt, "attr", attr)
    t = ctx.build(img, conn, tt, attr)
    s = ctx.build_tiled(img, grid, conn, tt, attr)
    same(s.parent(), t.parent(), "parent " + label)
    assert s.num_nodes == t.num_nodes
    same(s.check(attr), t.check(attr), "check %d %s" % (attr, label))
    for lam in lams:
        for dec in (0, 3):
            same(s.filter(attr, lam, dec), t.filter(attr, lam, dec), "filter %d lambda %g %s" % (dec, lam, label))
    if spec_lv is not None:
        a, b = s.spectrum1d(attr, spec_lv), t.spectrum1d(attr, spec_lv)
        if img.dtype == np.float32:
            # float greys: double bins summed by atomics in a varying order
            assert np.allclose(a, b, rtol=1e-12, atol=0), (label, a, b)
        else:
            same(a, b, "spectrum " + label)
    s.free()
    t.free()


GRIDS_2D = [(1, 2), (2, 1), (2, 2), (3, 2), (1, 5), (4, 3)]
GRIDS_3D = [(1, 1, 2), (2, 2, 2), (1, 3, 2), (3, 1, 1)]


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("conn", [4, 8])
def test_random_2d(ctx, seed, conn):
    rng = np.random.default_rng(700 + seed * 10 + conn)
    for k in range(6):
        H, W = int(rng.integers(5, 70)), int(rng.integers(5, 70))
        kind = int(rng.integers(0, 3))
        if kind == 0:
            img = rng.integers(0, int(rng.choice([2, 5, 40, 256])), size=(H, W)).astype(np.uint8)
        elif kind == 1:
            img = rng.integers(0, int(rng.choice([7, 3000, 65536])), size=(H, W)).astype(np.uint16)
        else:
            img = (rng.integers(-20, 21, size=(H, W)) * 0.25).astype(np.float32)
        grid = GRIDS_2D[int(rng.integers(0, len(GRIDS_2D)))]
        grid = (min(grid[0], W), min(grid[1], H))
        tt = "min" if rng.random() < 0.5 else "max"
        attr = int(rng.choice([0, 1, 3, 6, 9, 13, 17])) if img.dtype != np.float32 else int(rng.choice([0, 2, 6, 13, 17]))
        lv = np.array([0, 1, 2, 4, 8, 16, 64, 256, 1024], np.float32)
        lams = [3.0, 50.0] if attr == 0 else [float(rng.uniform(0.1, 30.0))]
        compare(ctx, img, grid, conn, tt, attr, lams, lv if attr in (0, 2, 3, 6, 13, 17) else None,
                label="2d seed %d case %d" % (seed, k))


@pytest.mark.parametrize("seed", range(3))
@pytest.mark.parametrize("conn", [6, 26])
def test_random_3d(ctx, seed, conn):
    rng = np.random.default_rng(900 + seed * 10 + conn)
    for k in range(4):
        D, H, W = (int(v) for v in rng.integers(3, 24, size=3))
        if rng.random() < 0.5:
            img = rng.integers(0, int(rng.choice([3, 20, 256])), size=(D, H, W)).astype(np.uint8)
        else:
            img = rng.integers(0, int(rng.choice([9, 65536])), size=(D, H, W)).astype(np.uint16)
        grid = GRIDS_3D[int(rng.integers(0, len(GRIDS_3D)))]
        grid = (min(grid[0], W), min(grid[1], H), min(grid[2], D))
        tt = "min" if rng.random() < 0.5 else "max"
        attr = int(rng.choice([0, 1, 5, 6, 9, 13, 17]))
        lams = [4.0, 40.0] if attr == 0 else [float(rng.uniform(0.1, 30.0))]
        compare(ctx, img, grid, conn, tt, attr, lams, np.array([0, 1, 3, 9, 27, 81, 243], np.float32),
                label="3d seed %d case %d" % (seed, k))


def test_small_tiles_against_oracle(ctx):
    """single-pixel-wide owned ranges (grid = image size along an axis) and the oracle's tree"""
    img = synth.disks(40, 33)
    for grid in ((33, 1), (1, 40), (11, 8)):
        say("tiles against oracle, grid", grid)
        s = ctx.build_tiled(img, grid, 4, "max", 0)
        o = OracleTree(img, 4, "max", 0)
        same(s.parent(), o.parent().astype(np.int32), "parent grid %s" % (grid,))
        same(s.filter(0, 30, 0), o.filter(0, 30, 0), "direct filter grid %s" % (grid,))
        s.free()


def test_refusals(ctx):
    img = synth.disks(64, 64)
    for grid in ((65, 1), (0, 2), (2, 2, 2)):
        with pytest.raises(dcf.DcfError) as e:
            ctx.build_tiled(img, grid, 4, "max", 0)
        assert e.value.code == -1, grid


@pytest.mark.parametrize("attr", [1, 2, 3, 4, 5])
def test_extent_attributes(ctx, attr):
    """the extent group's minima and maxima merged across tiles (range queries over the class tour)"""
    img = synth.spheres(40, 36, 44)
    t = ctx.build(img, 26, "max", attr)
    s = ctx.build_tiled(img, (2, 3, 2), 26, "max", attr)
    same(s.parent(), t.parent(), "parent")
    same(s.check(attr), t.check(attr), "check %d" % attr)
    o = OracleTree(img, 26, "max", attr)
    same(s.check(attr), o.check(attr), "check %d vs oracle" % attr)
    s.free()
    t.free()


@pytest.fixture(scope="module")
def vnoise8192():
    return synth.vnoise(8192, 8192)


@pytest.mark.parametrize("tree_type", ["max", "min"])
def test_vnoise_8192_slabs(ctx, vnoise8192, tree_type):
    """the 8192^2 u16 workload in 2, 4 and 8 row slabs (8-connectivity): parent, areas, the lambda=1000
    subtractive filter and the 64-bin area spectrum equal the untiled build's"""
    img = vnoise8192
    t = ctx.build(img, 8, tree_type, 0)
    par, area = t.parent(), t.check(0)
    filt = t.filter(0, 1000, 3, out_dtype=np.uint16)
    lv = synth.log_lambdas(64, 1, 2 ** 26)
    spec = t.spectrum1d(0, lv)
    t.free()
    for T in (2, 4, 8):
        say("vnoise 8192 slabs", T, tree_type)
        s = ctx.build_tiled(img, (1, T), 8, tree_type, 0)
        same(s.parent(), par, "parent T=%d" % T)
        same(s.check(0), area, "area T=%d" % T)
        same(s.filter(0, 1000, 3, out_dtype=np.uint16), filt, "subtractive filter T=%d" % T)
        same(s.spectrum1d(0, lv), spec, "spectrum T=%d" % T)
        s.free()


@pytest.mark.parametrize("conn", [6, 26])
def test_spheres_256_blocks(ctx, conn):
    """256^3 u8 volume in 2x2x2 blocks: parent, volumes, the lambda=500 direct filter and the 48-bin
    spectrum equal the untiled build's"""
    img = synth.spheres(256, 256, 256)
    say("spheres 256^3 2x2x2 blocks conn", conn)
    t = ctx.build(img, conn, "max", 0)
    s = ctx.build_tiled(img, (2, 2, 2), conn, "max", 0)
    same(s.parent(), t.parent(), "parent")
    same(s.check(0), t.check(0), "volume")
    same(s.filter(0, 500, 0, out_dtype=np.uint8), t.filter(0, 500, 0, out_dtype=np.uint8), "direct filter")
    lv = synth.log_lambdas(48, 1, 2 ** 27)
    same(s.spectrum1d(0, lv), t.spectrum1d(0, lv), "spectrum")
    s.free()
    t.free()
