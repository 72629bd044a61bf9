"""Pins the CPU oracle where the reference's golden vectors leave it open, with checks that do not
restate the oracle's own code:

* the canonical parent array and node areas against connected components of every upper level set
  (scipy.ndimage.label), in 2D and 3D, for every connectivity, for max-trees and (by inversion)
  min-trees, of integer and of float images (negative and fractional greys);
* the filter rules: for an increasing attribute (area) the direct, min, max and subtractive rules
  coincide (a node passes iff all its ancestors pass), so a wrong sign, index or ancestor step in
  any one of them breaks the equality; the direct rule is also checked against its definition
  walked on the brute-force tree;
* the 2D pattern spectrum against the 1D spectra of its two axes (marginals), which fails for a
  transposed bin index;
* find_scale against numpy's searchsorted;
* the max rule of non-increasing attributes, whose reference result depends on the raster scan
  (stale values), against a closed form of that scan derived per node (first pixel of each subtree,
  first pixel of its passing subtrees); the cases are checked to include stale values, where the
  textbook max rule differs;
* the min rule of any attribute against a closed form of its raster scan: the textbook min rule with
  the root left out, except for the zeros a failing root spreads, found top-down from write times; the
  cases are checked to include such zeros;
* the subtractive rule of any attribute, on integer and float images, against its per-node recurrence
  evaluated top-down in float32 in the reference's operation order (the oracle follows the reference's
  path stack instead);
* the eigenvalue attributes (elongation, flatness, sparseness), whose closed-form 2x2 / 3x3 solvers
  replace EISPACK rs(), against LAPACK's symmetric eigenvalues of each node's moment matrix, built from
  the node's pixel coordinates.
"""
import numpy as np
import pytest

from oracle.oracle import OracleTree, brute_tree, find_scale


def rand_img(rng, shape, levels):
    return rng.integers(0, levels, size=shape).astype(np.uint8)


CASES_2D = [((9, 13), 4), ((9, 13), 8), ((16, 5), 4), ((7, 7), 8)]
CASES_3D = [((4, 5, 6), 6), ((4, 5, 6), 26), ((3, 3, 7), 6)]


@pytest.mark.parametrize("shape,conn", CASES_2D + CASES_3D)
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_parent_and_area_match_brute_force(shape, conn, seed):
    rng = np.random.default_rng(seed * 101 + conn)
    img = rand_img(rng, shape, 6)
    bp, barea = brute_tree(img, conn)
    o = OracleTree(img, conn, "max", 0, threads=1)
    assert np.array_equal(o.parent().reshape(-1).astype(np.int64), bp)
    assert np.array_equal(o.check(0).reshape(-1), barea.astype(np.float32))


@pytest.mark.parametrize("shape,conn", CASES_2D + CASES_3D)
def test_float_images_match_brute_force(shape, conn):
    """f32 images with negative and fractional greys: the oracle orders them through the reference's
    float keys (flood.c:362-371), the brute force through numpy's float comparison; they agree for
    every value but -0.0 / +0.0, which the keys tell apart and which the image avoids.  The min-tree
    is the max-tree of gmax - g computed in float32 (preprocessing.c:84-91)."""
    rng = np.random.default_rng(40 + conn + len(shape))
    img = (rng.integers(-4, 5, size=shape) * 0.75 + rng.integers(0, 2, size=shape) * 0.125).astype(np.float32)
    img[img == 0] = 0.0625
    bp, barea = brute_tree(img, conn)
    o = OracleTree(img, conn, "max", 0, threads=1)
    assert np.array_equal(o.parent().reshape(-1).astype(np.int64), bp)
    assert np.array_equal(o.check(0).reshape(-1), barea.astype(np.float32))
    inv = (np.float32(img.max()) - img).astype(np.float32)
    bp, barea = brute_tree(inv, conn)
    o = OracleTree(img, conn, "min", 0, threads=1)
    assert np.array_equal(o.parent().reshape(-1).astype(np.int64), bp)
    assert np.array_equal(o.check(0).reshape(-1), barea.astype(np.float32))


@pytest.mark.parametrize("shape,conn", CASES_2D[:2] + CASES_3D[:2])
def test_min_tree_is_max_tree_of_the_inverse(shape, conn):
    rng = np.random.default_rng(7 + conn)
    img = rand_img(rng, shape, 5)
    inv = (img.max() - img).astype(np.uint8)
    bp, barea = brute_tree(inv, conn)
    o = OracleTree(img, conn, "min", 0, threads=1)
    assert np.array_equal(o.parent().reshape(-1).astype(np.int64), bp)
    assert np.array_equal(o.check(0).reshape(-1), barea.astype(np.float32))
    # filtering a min-tree returns gmax - (the max-tree filter of the inverse)
    om = OracleTree(inv, conn, "max", 0, threads=1)
    for lam in (2, 5, 11):
        assert np.array_equal(o.filter(0, lam, 0), img.max() - om.filter(0, lam, 0))


def brute_direct(img, parent, area, lam):
    """direct rule by definition: the grey of the nearest ancestor-or-self node whose area >= lam
    (each pixel's node is its levelroot), 0 when there is none"""
    flat = img.reshape(-1)
    out = np.zeros(flat.size, dtype=np.float32)
    for p in range(flat.size):
        v = p if parent[p] < 0 or flat[parent[p]] != flat[p] else parent[p]  # levelroot of p
        while v >= 0 and area[v] < lam:
            v = parent[v]
        out[p] = flat[v] if v >= 0 else 0
    return out.reshape(img.shape)


@pytest.mark.parametrize("shape,conn", CASES_2D + CASES_3D)
def test_rules_against_definition(shape, conn):
    rng = np.random.default_rng(99 + conn)
    img = rand_img(rng, shape, 7)
    bp, barea = brute_tree(img, conn)
    o = OracleTree(img, conn, "max", 0, threads=1)
    for lam in (1, 2, 3, 8, 25, img.size, img.size + 1):
        d = o.filter(0, lam, 0)
        assert np.array_equal(d, brute_direct(img, bp, barea, lam)), lam
        for rule in (1, 2, 3):
            assert np.array_equal(o.filter(0, lam, rule), d), (lam, rule)


@pytest.mark.parametrize("seed", [0, 1])
def test_rules_coincide_for_area_on_larger_images(seed):
    rng = np.random.default_rng(seed)
    img = rng.integers(0, 40, size=(48, 64)).astype(np.uint8)
    o = OracleTree(img, 8, "max", 0, threads=1)
    for lam in (3, 17, 200):
        d = o.filter(0, lam, 0)
        for rule in (1, 2, 3):
            assert np.array_equal(o.filter(0, lam, rule), d)


def test_spectrum2d_marginals_are_the_1d_spectra():
    rng = np.random.default_rng(5)
    img = rng.integers(0, 30, size=(40, 56)).astype(np.uint8)
    l1 = np.array([0.0, 0.16, 0.2, 0.3, 0.5, 1.0, 2.0], dtype=np.float32)  # ncomp axis
    l2 = np.array([0.0, 1, 2, 4, 8, 16, 64, 256, 1024], dtype=np.float32)  # area axis
    o4 = OracleTree(img, 4, "max", 17, threads=1)
    s2 = o4.spectrum2d(17, l1, 0, l2).reshape(len(l2), len(l1))  # bin s1 + s2 * n1
    assert np.array_equal(s2.sum(axis=0), o4.spectrum1d(17, l1))
    o0 = OracleTree(img, 4, "max", 0, threads=1)
    assert np.array_equal(s2.sum(axis=1), o0.spectrum1d(0, l2))
    # every bin sums to the total grey of the image (background = 1)
    assert s2.sum() == img.astype(np.int64).sum()


def test_find_scale_is_searchsorted():
    lv = np.array([0.0, 1.0, 1.5, 2.25, 3.375, 10.0, 100.0], dtype=np.float32)
    vals = np.concatenate([lv.astype(np.float64), lv.astype(np.float64) - 1e-9, [0.0, 5.0, 1e9, 99.99]])
    for a in vals:
        want = max(int(np.searchsorted(lv.astype(np.float64), a, side="right")) - 1, 0)
        assert find_scale(lv, a) == want, a


@pytest.mark.parametrize("shape,conn,root_first", [((9, 13), 4, False), ((16, 5), 8, False), ((4, 5, 6), 6, False),
                                                    ((9, 13), 4, True), ((12, 11), 8, True), ((3, 1, 1), 6, True)])
def test_min_rule_when_the_root_fails(shape, conn, root_first):
    """tree_filter_min (filtering.c:55-100) walks the pixels in raster order from their levelroot up to the
    first node an earlier walk reached, and never tests the root, which only the walk of a root-level pixel
    reaches (and sets to 0).  With lambda above the image's area every node fails, and the output has a
    closed form over the root's children, the connected components of {g > gmin}: a component whose first
    pixel precedes the first root-level pixel keeps gmin, any other 0, and the root-level pixels 0 (the
    direct rule gives 0 everywhere).  One lambda below, the root passes and the two rules agree."""
    from scipy import ndimage
    from oracle.oracle import _structure
    rng = np.random.default_rng(60 + conn + 7 * root_first)
    img = (rng.integers(0, 6, size=shape) + 3).astype(np.uint8)
    if root_first:
        img.reshape(-1)[0] = 3  # a root-level pixel first in raster order: every component gets 0
    elif len(shape) == 2:
        img[shape[0] // 2, :] = 3  # a band of the root's level: the components below it start after c_root
        img[: shape[0] // 2, :] = np.maximum(img[: shape[0] // 2, :], 4)
    m = img.min()
    flat = img.reshape(-1)
    lab, nl = ndimage.label(img > m, structure=_structure(img.ndim, conn))
    lab = lab.reshape(-1)
    first = np.full(nl + 1, flat.size, dtype=np.int64)
    np.minimum.at(first, lab, np.arange(flat.size))
    c_root = int(np.nonzero(flat == m)[0][0])
    want = np.where(flat == m, 0, np.where(first[lab] < c_root, m, 0)).astype(np.float32).reshape(img.shape)
    o = OracleTree(img, conn, "max", 0, threads=1)
    assert np.array_equal(o.filter(0, img.size + 1, 1), want)
    assert np.array_equal(o.filter(0, img.size + 1, 0), np.zeros(img.shape, np.float32))
    assert np.array_equal(o.filter(0, img.size, 1), o.filter(0, img.size, 0))


NONE = 1 << 62


def max_rule_closed_form(parent, g, attr, lam):
    """tree_filter_max (filtering.c:156-208) over one raster scan, solved per node.  first(x) is the
    first pixel of x's subtree, Q(x) the smallest first(w) over passing nodes w under x (when x is
    fixed, i.e. set to its own grey by a passing descendant's walk).  The walk of first(x) writes
    V(x) = g(a) for the parent a of x if a passes or Q(a) < first(x), else V(a); a failing root gives
    0.  A levelroot pixel ends with g when its node is ever fixed; another pixel v of node y reads
    g(y) when Q(y) < v, else V(y)."""
    n = len(parent)
    is_lr = (parent < 0) | (g[np.maximum(parent, 0)] != g)
    node = np.where(is_lr, np.arange(n), parent)
    lrs = np.nonzero(is_lr)[0]
    depth = {}
    for x in lrs:
        d, y = 0, x
        while parent[y] >= 0:
            y, d = parent[y], d + 1
        depth[x] = d
    order = sorted(lrs, key=lambda x: depth[x])
    passes = {x: attr[x] >= lam for x in lrs}
    first = {x: x for x in lrs}
    qc = {x: NONE for x in lrs}
    Q = {}
    for x in reversed(order):
        Q[x] = first[x] if passes[x] else qc[x]
        a = parent[x]
        if a >= 0:
            first[a] = min(first[a], first[x])
            qc[a] = min(qc[a], Q[x])
    V = {}
    for x in order:
        a = parent[x]
        V[x] = 0.0 if a < 0 else (g[a] if (passes[a] or Q[a] < first[x]) else V[a])
    out = np.empty(n, np.float32)
    for v in range(n):
        y = node[v]
        kept = Q[y] < NONE if v == y else Q[y] < v
        out[v] = g[y] if kept else V[y]
    return out


def textbook_max_rule(parent, g, attr, lam):
    """keep every node with a passing node in its subtree; a pixel takes its nearest kept node's grey"""
    n = len(parent)
    is_lr = (parent < 0) | (g[np.maximum(parent, 0)] != g)
    kept = np.zeros(n, bool)
    for x in np.nonzero(is_lr & (attr >= lam))[0]:
        while x >= 0 and not kept[x]:
            kept[x] = True
            x = parent[x]
    out = np.zeros(n, np.float32)
    for v in range(n):
        y = v if is_lr[v] else parent[v]
        while y >= 0 and not kept[y]:
            y = parent[y]
        out[v] = g[y] if y >= 0 else 0
    return out


def test_max_rule_of_non_increasing_attributes():
    rng = np.random.default_rng(77)
    stale = 0
    for it in range(120):
        if it % 3 == 2:
            shape, conn = tuple(int(s) for s in rng.integers(2, 7, size=3)), int(rng.choice([6, 26]))
        else:
            shape, conn = tuple(int(s) for s in rng.integers(2, 14, size=2)), int(rng.choice([4, 8]))
        img = rand_img(rng, shape, int(rng.integers(3, 10)))
        tt = "max" if it % 2 == 0 else "min"
        aidx = int(rng.choice([17, 13, 6, 14, 2]))
        o = OracleTree(img, conn, tt, aidx, threads=1)
        par, g = o.parent(), o.grey()
        a, _ = o.attributes(aidx)
        lr = (par < 0) | (g[np.maximum(par, 0)] != g)
        lam = float(rng.choice(np.unique(a[lr])))
        got = o.filter(aidx, lam, 2).reshape(-1)
        want = max_rule_closed_form(par, g, a, lam)
        text = textbook_max_rule(par, g, a, lam)
        if tt == "min":
            want, text = o.gmax - want, o.gmax - text
        assert np.array_equal(got, want), (it, shape, conn, tt, aidx, lam)
        stale += not np.array_equal(got, text)
    assert stale >= 20, stale  # the scan's stale values are exercised, not only the textbook rule


def min_rule_closed_form(parent, g, attr, lam):
    """tree_filter_min (filtering.c:55-100) over one raster scan, solved per node, top-down.  Without
    its root test the min rule gives g(parent(h)) for the highest failing node h below the root on a
    node's path (its own grey when there is none; the root: g or 0).  A failing root of nonzero grey is
    written 0 at its levelroot pixel R, and later walks stopping at it, or at a node already zeroed, copy
    the 0.  t(x) is when x is written, tau(x) the earliest t of its strict ancestors, W(x) whether its
    first-reached ancestor is zeroed; a node below a failing node is written at first(x) and zeroed when
    W and tau < first(x); any other node is zeroed when W and tau < its levelroot, and else written at its
    levelroot."""
    n = len(parent)
    is_lr = (parent < 0) | (g[np.maximum(parent, 0)] != g)
    node = np.where(is_lr, np.arange(n), parent)
    lrs = np.nonzero(is_lr)[0]
    depth = {}
    for x in lrs:
        d, y = 0, x
        while parent[y] >= 0:
            y, d = parent[y], d + 1
        depth[x] = d
    order = sorted(lrs, key=lambda x: depth[x])
    first = {x: x for x in lrs}
    for x in reversed(order):
        if parent[x] >= 0:
            first[parent[x]] = min(first[parent[x]], first[x])
    fail = {x: attr[x] < lam for x in lrs}
    t, tau, W, Z, h, val = {}, {}, {}, {}, {}, {}
    for x in order:
        p = parent[x]
        if p < 0:
            t[x], tau[x], W[x], h[x] = x, NONE, False, -1
            Z[x] = bool(fail[x] and g[x] != 0)
            val[x] = 0.0 if fail[x] else g[x]
            continue
        tau[x] = min(t[p], tau[p])
        W[x] = Z[p] if t[p] <= tau[p] else W[p]
        h[x] = h[p] if h[p] >= 0 else (x if fail[x] else -1)
        if h[x] >= 0:
            Z[x] = W[x] and tau[x] < first[x]
            t[x] = first[x]
        else:
            Z[x] = W[x] and tau[x] < x
            t[x] = max(first[x], tau[x] + 1) if Z[x] else x
        val[x] = 0.0 if Z[x] else (g[parent[h[x]]] if h[x] >= 0 else g[x])
    return np.array([val[node[v]] for v in range(n)], np.float32)


def textbook_min_rule_without_root(parent, g, attr, lam):
    n = len(parent)
    is_lr = (parent < 0) | (g[np.maximum(parent, 0)] != g)
    out = np.empty(n, np.float32)
    for v in range(n):
        y = v if is_lr[v] else parent[v]
        h, x = -1, y
        while parent[x] >= 0:
            if attr[x] < lam:
                h = x
            x = parent[x]
        out[v] = g[parent[h]] if h >= 0 else (g[y] if (y != x or attr[y] >= lam) else 0)
    return out


def test_min_rule_of_any_attribute():
    rng = np.random.default_rng(78)
    zeros = 0
    for it in range(150):
        if it % 3 == 2:
            shape, conn = tuple(int(s) for s in rng.integers(2, 7, size=3)), int(rng.choice([6, 26]))
        else:
            shape, conn = tuple(int(s) for s in rng.integers(2, 14, size=2)), int(rng.choice([4, 8]))
        tt = "max" if it % 4 else "min"
        img = rand_img(rng, shape, int(rng.integers(3, 10))) + np.uint8(tt == "max")  # a max-tree root above 0
        aidx = int(rng.choice([17, 13, 6, 2, 0]))
        o = OracleTree(img, conn, tt, aidx, threads=1)
        par, g = o.parent(), o.grey()
        a, _ = o.attributes(aidx)
        lr = (par < 0) | (g[np.maximum(par, 0)] != g)
        vals = np.unique(a[lr])
        above_root = vals[vals > a[par < 0][0]]
        # half the lambdas make the root fail
        lam = float(rng.choice(above_root if len(above_root) and it % 2 else vals))
        got = o.filter(aidx, lam, 1).reshape(-1)
        want = min_rule_closed_form(par, g, a, lam)
        text = textbook_min_rule_without_root(par, g, a, lam)
        if tt == "min":
            want, text = o.gmax - want, o.gmax - text
        assert np.array_equal(got, want), (it, shape, conn, tt, aidx, lam)
        zeros += not np.array_equal(got, text)
    assert zeros >= 10, zeros  # the zeros of a failing root are exercised


def subtractive_recurrence(parent, g, attr, lam):
    """tree_filter_subtractive (filtering.c:211-261) per node, top-down: the root keeps g when it passes
    (0 otherwise); a passing node u below p takes (g(u) + f(p)) - g(p), a failing one f(p), each sum
    rounded to float32 as the reference's float arithmetic does"""
    n = len(parent)
    is_lr = (parent < 0) | (g[np.maximum(parent, 0)] != g)
    lrs = np.nonzero(is_lr)[0]
    depth = {}
    for x in lrs:
        d, y = 0, x
        while parent[y] >= 0:
            y, d = parent[y], d + 1
        depth[x] = d
    f = {}
    for x in sorted(lrs, key=lambda x: depth[x]):
        p = parent[x]
        if p < 0:
            f[x] = g[x] if attr[x] >= lam else np.float32(0)
        elif attr[x] >= lam:
            f[x] = np.float32(np.float32(g[x] + f[p]) - g[p])
        else:
            f[x] = f[p]
    node = np.where(is_lr, np.arange(n), parent)
    return np.array([f[node[v]] for v in range(n)], np.float32)


def test_subtractive_rule_of_any_attribute():
    rng = np.random.default_rng(79)
    differs_from_direct = 0
    for it in range(120):
        shape = tuple(int(s) for s in rng.integers(2, 14, size=2))
        if it % 3 == 2:
            img = (rng.random(shape) * 9 - 3).astype(np.float32)
        else:
            img = rand_img(rng, shape, int(rng.integers(3, 12)))
        tt = "max" if it % 2 else "min"
        aidx = int(rng.choice([17, 13, 6, 2]))
        o = OracleTree(img, int(rng.choice([4, 8])), tt, aidx, threads=1)
        par, g = o.parent(), o.grey()
        a, _ = o.attributes(aidx)
        lr = (par < 0) | (g[np.maximum(par, 0)] != g)
        lam = float(rng.choice(np.unique(a[lr])))
        got = o.filter(aidx, lam, 3).reshape(-1)
        want = subtractive_recurrence(par, g, a, lam)
        if tt == "min":
            want = np.float32(o.gmax) - want
        assert np.array_equal(got, want), (it, shape, tt, aidx, lam)
        differs_from_direct += not np.array_equal(got, o.filter(aidx, lam, 0).reshape(-1))
    assert differs_from_direct >= 20, differs_from_direct  # non-increasing attributes: not the direct rule


@pytest.mark.parametrize("shape,conn", [((5, 6, 7), 6), ((4, 7, 6), 26), ((1, 9, 11), 6), ((13, 17), 4), ((11, 9), 8)])
def test_eigen_attributes_against_lapack(shape, conn):
    """inertia_elon / flat / spar (attributes.c:726-849): eigenvalues of the central moment matrix with
    the pixel-size correction A/12 on the diagonal; a node whose z sum is 0 uses the 2D formulas
    (flatness -1), as the reference tests inertia[3] == 0"""
    rng = np.random.default_rng(sum(shape) + conn)
    img = rand_img(rng, shape, 5)
    o = OracleTree(img, conn, "max", 14, threads=1)
    par, g = o.parent(), o.grey()
    n = len(par)
    is_lr = (par < 0) | (g[np.maximum(par, 0)] != g)
    dims = (shape[-1], shape[-2], shape[0] if len(shape) == 3 else 1)
    idx = np.arange(n)
    xyz = np.stack([idx % dims[0], (idx // dims[0]) % dims[1], idx // (dims[0] * dims[1])], axis=1).astype(np.float64)
    members = {x: [] for x in np.nonzero(is_lr)[0]}
    for v in range(n):
        y = v if is_lr[v] else par[v]
        while y >= 0:
            members[y].append(v)
            y = par[y]
    got = {a: o.attributes(a)[0] for a in (14, 15, 16)}
    for x, pix in members.items():
        c = xyz[pix]
        A = float(len(pix))
        s1 = c.sum(axis=0)
        M = c.T @ c - np.outer(s1, s1) / A + np.eye(3) * (A / 12)
        if s1[2] == 0:
            ev = np.linalg.eigvalsh(M[:2, :2])
            want = (ev[1] / ev[0], -1.0, np.pi * np.sqrt(4 * ev[1] / A) * np.sqrt(4 * ev[0] / A) / A)
        else:
            ev = np.linalg.eigvalsh(M)
            want = (ev[2] / ev[1], ev[1] / ev[0],
                    np.pi * np.sqrt(20 * ev[2] / A) * np.sqrt(20 * ev[1] / A) * np.sqrt(20 * ev[0] / A) / (6 * A))
        for a, w in zip((14, 15, 16), want):
            assert abs(got[a][x] - w) <= 1e-9 * abs(w), (x, a, got[a][x], w)
