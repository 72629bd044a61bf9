"""Pins the oracle's attributes 1-13 and 17 and its pattern spectra by brute force, from each node's
pixel set and nothing the oracle computes.

A node's pixel set comes from the brute-force tree of ``oracle.brute_tree`` (connected components of
every upper level set, scipy.ndimage.label); the expected values are then taken from those pixels:

* extent attributes 1-5 (attributes.c:114-269): bounding boxes of the pixel sets, max exclusive;
* mean attributes 6-8 (attributes.c:271-350) and grey-weighted means / flux 9-12 (attributes.c:352-469):
  sums of the pixel coordinates and of the node-space greys (gmax - g for a min-tree);
* trace (13) and non-compactness (17) (attributes.c:705-723, 851-874): the trace of the central
  second-moment matrix of the pixel coordinates with A/12 on its diagonal, taken by numpy on the
  matrix, 2D (x, y) for a node whose z sum is 0 and 3D otherwise, divided by A^2 or A^(5/3);
* 1D and 2D pattern spectra (pattern.c:80-96, 170-191): every node adds (g - g_parent) * area to the
  bin numpy's searchsorted picks for its own attribute value among the thresholds (lambda_0 = 0
  first), a 2D bin being s1 + s2 * n1.

Pixel sizes are powers of two, so every product here is exact and the comparisons are bit for bit
except for the moment formulas (1e-9 relative: numpy's matrix route rounds differently).
"""
import numpy as np
import pytest

from oracle.oracle import OracleTree, brute_tree


def rand_img(rng, shape, levels):
    return rng.integers(0, levels, size=shape).astype(np.uint8)


def dims_of(shape):
    return (shape[-1], shape[-2], shape[0] if len(shape) == 3 else 1)


def brute_nodes(img, conn, tree_type):
    """levelroots of the brute-force tree, node-space grey of every pixel, parent of every levelroot and
    every node's pixel set (indices)"""
    g = img.reshape(-1).astype(np.int64)
    if tree_type == "min":
        g = g.max() - g
    work = g.reshape(img.shape)
    par, _ = brute_tree(work, conn)
    n = g.size
    is_lr = (par < 0) | (g[np.maximum(par, 0)] != g)
    members = {int(x): [] for x in np.nonzero(is_lr)[0]}
    for v in range(n):
        y = v if is_lr[v] else int(par[v])
        while y >= 0:
            members[y].append(v)
            y = int(par[y])
    return g, par, members


def coords(shape, idx):
    W, H, _ = dims_of(shape)
    idx = np.asarray(idx)
    return np.stack([idx % W, (idx // W) % H, idx // (W * H)], axis=1).astype(np.int64)


CASES = [((9, 13), 4), ((16, 5), 8), ((4, 5, 6), 6), ((3, 4, 7), 26)]


@pytest.mark.parametrize("shape,conn", CASES)
@pytest.mark.parametrize("tree_type", ["max", "min"])
@pytest.mark.parametrize("pixdim", [(1.0, 1.0, 1.0), (0.5, 2.0, 4.0)])
def test_attributes_1_to_12_against_pixel_sets(shape, conn, tree_type, pixdim):
    rng = np.random.default_rng(sum(shape) * 7 + conn + (tree_type == "min"))
    img = rand_img(rng, shape, 6)
    g, par, members = brute_nodes(img, conn, tree_type)
    pd = np.asarray(pixdim, dtype=np.float32)
    vol = float(pd[0] * pd[1] * pd[2])
    want = {a: {} for a in range(1, 13)}
    for x, pix in members.items():
        c = coords(shape, pix)
        lo, hi = c.min(axis=0), c.max(axis=0) + 1
        ext = (hi - lo).astype(np.float64)
        A = float(len(pix))
        want[1][x] = ext[0] * ext[1] * ext[2] * vol
        want[2][x] = float(sum((float(pd[i]) * hi[i] - float(pd[i]) * lo[i]) ** 2 for i in range(3)))
        for i in range(3):
            want[3 + i][x] = float(pd[i]) * ext[i]
        s = c.sum(axis=0).astype(np.float64)
        for i in range(3):
            want[6 + i][x] = float(pd[i]) * s[i] / A
        gv = g[pix].astype(np.float64)
        sg = gv.sum()
        sgx = (gv[:, None] * c).sum(axis=0).astype(np.float64)
        with np.errstate(divide="ignore", invalid="ignore"):
            for i in range(3):
                want[9 + i][x] = float(pd[i]) * sgx[i] / sg
        want[12][x] = sg
    o = OracleTree(img, conn, tree_type, 1, pixdim=pixdim, threads=1)
    assert np.array_equal(o.parent().astype(np.int64), par)
    for a in range(1, 13):
        if a in (6, 9, 12):
            o.reaccumulate(a)
        got = o.attributes(a)[0]
        for x, w in want[a].items():
            assert got[x] == w or (np.isnan(w) and np.isnan(got[x])), (a, x, got[x], w)


def moment_trace(c, A):
    """trace of the central second-moment matrix (with the pixel-size term A/12) and whether the node
    takes the 2D formulas (z sum 0)"""
    c = c.astype(np.float64)
    s1 = c.sum(axis=0)
    M = c.T @ c - np.outer(s1, s1) / A + np.eye(3) * (A / 12)
    flat = s1[2] == 0
    return (np.trace(M[:2, :2]) if flat else np.trace(M)), flat


@pytest.mark.parametrize("shape,conn", [((9, 13), 4), ((11, 9), 8), ((4, 5, 6), 6), ((5, 4, 7), 26), ((3, 6, 6), 6)])
@pytest.mark.parametrize("tree_type", ["max", "min"])
def test_trace_and_noncompactness_against_moment_matrix(shape, conn, tree_type):
    rng = np.random.default_rng(sum(shape) * 3 + conn + (tree_type == "min"))
    img = rand_img(rng, shape, 20)
    _, par, members = brute_nodes(img, conn, tree_type)
    o = OracleTree(img, conn, tree_type, 13, threads=1)
    assert np.array_equal(o.parent().astype(np.int64), par)
    tr, nc = o.attributes(13)[0], o.attributes(17)[0]
    branches = set()
    for x, pix in members.items():
        A = float(len(pix))
        t, flat = moment_trace(coords(shape, pix), A)
        branches.add(flat)
        ncomp = t / A ** 2 if flat else t / A ** (5.0 / 3.0)
        assert abs(tr[x] - t) <= 1e-9 * abs(t), (x, tr[x], t, flat)
        assert abs(nc[x] - ncomp) <= 1e-9 * abs(ncomp), (x, nc[x], ncomp, flat)
    if len(shape) == 3:
        # a volume has nodes on both branches: some lie in the plane z = 0 (hazard H7)
        assert branches == {True, False}, branches


def brute_attr(shape, pix, g_node, attr):
    c = coords(shape, pix)
    A = float(len(pix))
    if attr == 0:
        return A
    if attr == 3:
        return float(c[:, 0].max() + 1 - c[:, 0].min())
    if attr == 6:
        return float(c[:, 0].sum()) / A
    if attr == 17:
        t, flat = moment_trace(c, A)
        return t / A ** 2 if flat else t / A ** (5.0 / 3.0)
    raise ValueError(attr)


def scale_of(lv, a):
    """largest k with a >= lambda_k (bin 0 below lambda_0), by numpy's searchsorted"""
    return max(int(np.searchsorted(np.asarray(lv, dtype=np.float64), a, side="right")) - 1, 0)


def brute_spectrum(img, conn, tree_type, attr1, l1, attr2=None, l2=None):
    g, par, members = brute_nodes(img, conn, tree_type)
    out = np.zeros(len(l1) * (len(l2) if l2 is not None else 1), dtype=np.float64)
    for x, pix in members.items():
        dh = g[x] - (g[par[x]] if par[x] >= 0 else 0)
        if dh == 0:
            continue
        b = scale_of(l1, brute_attr(img.shape, pix, g, attr1))
        if l2 is not None:
            b += scale_of(l2, brute_attr(img.shape, pix, g, attr2)) * len(l1)
        out[b] += float(dh) * len(pix)
    return out


@pytest.mark.parametrize("shape,conn", [((13, 17), 4), ((12, 10), 8), ((5, 6, 7), 6), ((4, 6, 5), 26)])
@pytest.mark.parametrize("tree_type", ["max", "min"])
def test_spectra_against_brute_binning(shape, conn, tree_type):
    """each node lands in the bin of its own attribute value: binning by another node's value (or by the
    area instead of the attribute) moves weight between bins and fails"""
    rng = np.random.default_rng(sum(shape) * 5 + conn + (tree_type == "min"))
    img = rand_img(rng, shape, 7)
    W = dims_of(shape)[0]
    lam = {0: np.array([0, 1, 2, 3, 5, 8, 13, 21, 34, 55, 89], np.float32),
           3: np.array([0, 1, 2, 3, 4, 6, 9], np.float32),
           6: np.linspace(0, W, 9).astype(np.float32),
           17: np.array([0, 0.1, 0.15, 0.2, 0.3, 0.5, 1.0], np.float32)}
    grp = {0: 0, 3: 1, 6: 2, 17: 4}
    o = OracleTree(img, conn, tree_type, 0, threads=1)
    for a in (0, 3, 6, 17):
        if grp[a] != 0:
            o.reaccumulate(a)
        want = brute_spectrum(img, conn, tree_type, a, lam[a])
        got = o.spectrum1d(a, lam[a])
        assert np.array_equal(got, want), (a, got, want)
        if a != 0:
            # the 2D spectrum of (attribute, area): attr2 = 0 means the area of attr1's group (operation.c:97-98)
            want2 = brute_spectrum(img, conn, tree_type, a, lam[a], 0, lam[0])
            got2 = o.spectrum2d(a, lam[a], 0, lam[0])
            assert np.array_equal(got2, want2), (a, got2, want2)
    # the bins are not degenerate: several of them hold weight
    assert (brute_spectrum(img, conn, tree_type, 6, lam[6]) > 0).sum() >= 3
