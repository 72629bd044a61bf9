"""ctypes view of the CPU oracle (oracle/dcf_oracle.c) and a brute-force labelling oracle.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the cpu_baseline leg
of bench.py, as the checker.  The product path (disccofan_amd) never imports it.

The C oracle restates the reference's flooding build, attributes, filter rules and spectra;
``brute_tree`` re-derives the canonical max-tree from first principles (connected components of
every upper level set, scipy.ndimage.label) so the flooding oracle itself is pinned on small
inputs independently of the reference's algorithm.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None


def build() -> str:
    """Compile the oracle with its Makefile (gcc, no reference sources involved), always from the source:
    an existing library, however recent, is rebuilt."""
    subprocess.run(["make", "-s", "-B", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        vp, i64p = C.c_void_p, C.POINTER(C.c_int64)
        L.orc_build.restype = vp
        L.orc_build.argtypes = [vp, i64p, C.c_int, C.c_int, C.c_int, vp, C.c_int]
        L.orc_free.argtypes = [vp]
        L.orc_reaccumulate.argtypes = [vp, C.c_int]
        L.orc_size.restype = C.c_int64
        L.orc_size.argtypes = [vp]
        L.orc_connectivity.argtypes = [vp]
        L.orc_gmax.restype = C.c_float
        L.orc_gmax.argtypes = [vp]
        L.orc_parent.argtypes = [vp, vp]
        L.orc_grey.argtypes = [vp, vp]
        L.orc_attr_all.argtypes = [vp, C.c_int, vp, vp]
        L.orc_check.argtypes = [vp, C.c_int, vp]
        L.orc_filter.argtypes = [vp, C.c_int, C.c_int, C.c_double, C.c_int, vp]
        L.orc_find_scale.argtypes = [vp, C.c_int, C.c_double]
        L.orc_spectrum1d.argtypes = [vp, C.c_int, vp, C.c_int, vp]
        L.orc_spectrum2d.argtypes = [vp, C.c_int, vp, C.c_int, C.c_int, vp, C.c_int, vp]
        L.orc_lvec_read.argtypes = [C.c_char_p, C.c_double, vp, C.c_int]
        L.orc_tree_stats.argtypes = [vp, i64p, i64p, i64p]
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class OracleTree:
    """A canonical component tree built by the CPU oracle on a float copy of the image."""

    def __init__(self, img: np.ndarray, connectivity: int = 4, tree_type: str = "max", attr_idx: int = 0,
                 pixdim=(1.0, 1.0, 1.0), threads: int = 1):
        img = np.asarray(img)
        if img.ndim == 2:
            dims = (img.shape[1], img.shape[0], 1)
        elif img.ndim == 3:
            dims = (img.shape[2], img.shape[1], img.shape[0])
        else:
            raise ValueError("2D or 3D image expected")
        self.shape = img.shape
        self.dims = dims
        g = np.ascontiguousarray(img, dtype=np.float32)
        d = (C.c_int64 * 3)(*dims)
        pd = np.asarray(pixdim, dtype=np.float32)
        L = lib()
        self._t = L.orc_build(_ptr(g), d, int(connectivity), 1 if tree_type == "min" else 0, int(attr_idx),
                              _ptr(pd), int(threads))
        if not self._t:
            raise ValueError("oracle build failed (bad arguments or out of memory)")
        self.n = int(L.orc_size(self._t))
        self.connectivity = int(L.orc_connectivity(self._t))

    def __del__(self):
        t = getattr(self, "_t", None)
        if t:
            lib().orc_free(t)
            self._t = None

    @property
    def gmax(self) -> float:
        return float(lib().orc_gmax(self._t))

    def reaccumulate(self, attr_idx: int) -> None:
        if lib().orc_reaccumulate(self._t, int(attr_idx)) != 0:
            raise RuntimeError("re-accumulation failed")

    def parent(self) -> np.ndarray:
        out = np.empty(self.n, dtype=np.int64)
        lib().orc_parent(self._t, _ptr(out))
        return out

    def grey(self) -> np.ndarray:
        out = np.empty(self.n, dtype=np.float32)
        lib().orc_grey(self._t, _ptr(out))
        return out

    def attributes(self, attr_idx: int):
        """(attribute, area-function) of every pixel's record; meaningful at levelroots."""
        a = np.empty(self.n, dtype=np.float64)
        ar = np.empty(self.n, dtype=np.float64)
        if lib().orc_attr_all(self._t, int(attr_idx), _ptr(a), _ptr(ar)) != 0:
            raise ValueError("attribute not in the tree's accumulated group")
        return a, ar

    def check(self, attr_idx: int) -> np.ndarray:
        out = np.empty(self.n, dtype=np.float32)
        if lib().orc_check(self._t, int(attr_idx), _ptr(out)) != 0:
            raise ValueError("attribute not in the tree's accumulated group")
        return out.reshape(self.shape)

    def filter(self, attr_idx: int, lam: float, decision: int = 0, extract: bool = False) -> np.ndarray:
        out = np.empty(self.n, dtype=np.float32)
        if lib().orc_filter(self._t, int(attr_idx), int(decision), float(lam), 1 if extract else 0, _ptr(out)) != 0:
            raise ValueError("bad filter arguments")
        return out.reshape(self.shape)

    def spectrum1d(self, attr_idx: int, lambdas) -> np.ndarray:
        lv = np.ascontiguousarray(lambdas, dtype=np.float32)
        out = np.empty(len(lv), dtype=np.float64)
        if lib().orc_spectrum1d(self._t, int(attr_idx), _ptr(lv), len(lv), _ptr(out)) != 0:
            raise ValueError("bad spectrum arguments")
        return out

    def spectrum2d(self, attr1: int, l1, attr2: int, l2) -> np.ndarray:
        a = np.ascontiguousarray(l1, dtype=np.float32)
        b = np.ascontiguousarray(l2, dtype=np.float32)
        out = np.empty(len(a) * len(b), dtype=np.float64)
        rc = lib().orc_spectrum2d(self._t, int(attr1), _ptr(a), len(a), int(attr2), _ptr(b), len(b), _ptr(out))
        if rc != 0:
            raise ValueError("bad 2D spectrum arguments (%d)" % rc)
        return out

    def stats(self):
        k, d, l = C.c_int64(), C.c_int64(), C.c_int64()
        lib().orc_tree_stats(self._t, C.byref(k), C.byref(d), C.byref(l))
        return {"nodes": k.value, "max_depth": d.value, "leaves": l.value}


def find_scale(lambdas, a: float) -> int:
    lv = np.ascontiguousarray(lambdas, dtype=np.float32)
    return int(lib().orc_find_scale(_ptr(lv), len(lv), float(a)))


def lvec_read(path: str, imscale: float = 1.0) -> np.ndarray:
    n = lib().orc_lvec_read(path.encode(), float(imscale), None, 0)
    if n < 0:
        raise FileNotFoundError(path)
    out = np.zeros(n, dtype=np.float32)
    lib().orc_lvec_read(path.encode(), float(imscale), _ptr(out), n)
    return out


# ------------------------------------------------------------------------------------------
# brute force: canonical max-tree from connected components of every upper level set
# ------------------------------------------------------------------------------------------
def _structure(ndim: int, connectivity: int):
    from scipy import ndimage
    if ndim == 2:
        return ndimage.generate_binary_structure(2, 1 if connectivity in (4, 6) else 2)
    return ndimage.generate_binary_structure(3, 1 if connectivity in (4, 6) else 3)


def brute_tree(img: np.ndarray, connectivity: int = 4):
    """Canonical parent array (flat, levelroot = min index of the node's level-g pixels,
    root -> -1) and per-pixel node area of the max-tree, by labelling {img >= g} for every g."""
    from scipy import ndimage
    img = np.asarray(img)
    flat = img.reshape(-1)
    n = flat.size
    st = _structure(img.ndim, connectivity)
    levels = np.unique(flat)
    # comp_rep[level index][pixel] = levelroot of the pixel's component at that level (or -1)
    lr_at = {}
    area_at = {}
    for g in levels:
        lab, nl = ndimage.label(img >= g, structure=st)
        lab = lab.reshape(-1)
        rep = np.full(nl + 1, np.iinfo(np.int64).max, dtype=np.int64)
        idx = np.nonzero((flat == g) & (lab > 0))[0]
        np.minimum.at(rep, lab[idx], idx)
        cnt = np.bincount(lab, minlength=nl + 1)
        lr_at[g] = (lab, rep)
        area_at[g] = cnt
    parent = np.full(n, -1, dtype=np.int64)
    area = np.zeros(n, dtype=np.int64)
    lv = list(levels)
    for p in range(n):
        g = flat[p]
        lab, rep = lr_at[g]
        r = int(rep[lab[p]])
        area[p] = area_at[g][lab[p]]
        if r != p:
            parent[p] = r
            continue
        # levelroot: the next lower level whose component containing p has pixels at that level
        k = lv.index(g) - 1
        par = -1
        while k >= 0:
            g2 = lv[k]
            lab2, rep2 = lr_at[g2]
            r2 = rep2[lab2[p]]
            if r2 != np.iinfo(np.int64).max:
                par = int(r2)
                break
            k -= 1
        parent[p] = par
    return parent, area
