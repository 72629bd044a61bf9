"""disccofan_amd -- MI355X-native component trees (max-tree / min-tree) behind disccofan's operator surface.

Python host-side mirror of the reference's operator interface, over the C ABI of libdcf.so
(include/dcf.h).  Names and argument meanings follow the reference:

    tree_type       "max" | "min"                      (config.c tree_type, preprocessing.c:84-91)
    connectivity    4, 8 (2D) / 6, 26 (3D)             (io.c:64-79 mapping)
    attribute index 0..17 of AttribsArray              (attributes.c:1224-1248)
    decision        0 Direct, 1 Min, 2 Max, 3 Subtractive (filtering.c:271-277)
    operations      ("filter", attr, lambda[, decision]), ("extract", ...), ("check", attr),
                    ("pattern", attr, "lvec"[, imscale]),
                    ("pattern2D", attr1, "lvec1", s1, attr2, "lvec2", s2)   (config.c:297-433)

Every call runs on the GPU through libdcf.so; there is no CPU fallback: if the library is missing
or no GPU is visible, the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
import sys
import weakref

import numpy as np



def _ask_for_hw_queues(n: int = 16) -> None:
    """The HIP runtime gives a process 4 hardware queues unless GPU_MAX_HW_QUEUES says otherwise.  A
    stream takes its queue when it is created (once all are taken, the one with the fewest streams), and
    busy streams that share a queue run one after the other.  The library runs the parts of a batched
    call and the tiles of a tiled build on worker streams (up to 12 streams with the null stream), so it
    asks for 16 queues before the runtime starts (torch starts it on its first GPU call).  A value
    already set is kept.  The library reads the variable when a context is created to choose how many
    workers and level loops to run."""
    if "GPU_MAX_HW_QUEUES" in os.environ:
        return
    t = sys.modules.get("torch")
    if t is not None and t.cuda.is_initialized():
        return
    os.environ["GPU_MAX_HW_QUEUES"] = str(n)


_ask_for_hw_queues()

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libdcf.so")
SYNTH_PATH = os.path.join(_HERE, "lib", "libdcf_synth.so")

DCF_U8, DCF_U16, DCF_F32 = 0, 1, 2
DCF_DEVICE_PTR = 0x1
DCF_ALL_RULES = 0x2

STATUS = {0: "DCF_OK", -1: "DCF_EINVAL", -2: "DCF_ENOMEM", -3: "DCF_EDEVICE", -4: "DCF_EUNSUPPORTED",
          -5: "DCF_EIO", -6: "DCF_EGROUP"}

ATTRIBUTES = ["area", "rectarea", "rectdiag", "xext", "yext", "zext", "xmean", "ymean", "zmean",
              "xwmean", "ywmean", "zwmean", "flux", "inertiatrace", "elong", "flat", "spars", "ncomp"]
DECISIONS = ["Direct", "Min", "Max", "Subtractive"]

# exported symbols of libdcf.so (kept in sync with include/dcf.h)
EXPORTS = [
    "dcf_ctx_create", "dcf_ctx_destroy", "dcf_ctx_set_stream", "dcf_ctx_set_graph", "dcf_ctx_set_option",
    "dcf_ctx_stream", "dcf_last_error",
    "dcf_ctx_synchronize", "dcf_ctx_set_profiling", "dcf_ctx_profile_only", "dcf_batch_filter", "dcf_dmp", "dcf_tree_write_csv", "dcf_kernel_stats", "dcf_kernel_stats_reset",
    "dcf_kernel_name", "dcf_tree_build", "dcf_tree_build_tiled", "dcf_tree_free", "dcf_tree_accumulate", "dcf_tree_group",
    "dcf_tree_parent", "dcf_tree_num_nodes", "dcf_filter", "dcf_check", "dcf_spectrum1d", "dcf_spectrum2d",
    "dcf_spectrum1d_noalloc",
    "dcf_lvec_read", "dcf_lvec_free", "dcf_attr_name", "dcf_attr_short_name", "dcf_attr_group",
    "dcf_decision_name", "dcf_run_operations",
]


class DcfError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__("%s (%d): %s" % (STATUS.get(code, "?"), code, msg))
        self.code = code


class BatchImage(C.Structure):
    """dcf_batch_image of include/dcf.h"""
    _fields_ = [("img", C.c_void_p), ("dtype", C.c_int), ("dims", C.c_uint64 * 3), ("lam", C.c_double),
                ("out", C.c_void_p), ("flags", C.c_uint)]


DCF_IMAGE_MAX_TREE = 0x1
DCF_IMAGE_MIN_TREE = 0x2


def _image_flags(tree_type) -> int:
    """per-image tree type of a batch item: None = the call's"""
    if tree_type is None:
        return 0
    _tree_type_code(tree_type)
    return DCF_IMAGE_MIN_TREE if tree_type == "min" else DCF_IMAGE_MAX_TREE


_lib = None


def lib():
    """Load libdcf.so (built in-tree by disccofan_amd.build); raises when it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libdcf.so not built: run `python -m disccofan_amd.build` (%s missing)" % LIB_PATH)
    # PyTorch ships its own HIP runtime: when both live in one process, torch's must be initialised
    # first (a runtime loaded before it hides the GPU from torch), so do that when torch is imported
    torch = sys.modules.get("torch")
    if torch is not None:
        try:
            if torch.cuda.is_available():
                torch.cuda.init()
        except Exception:
            pass
    L = C.CDLL(LIB_PATH)
    vp, i, u, d = C.c_void_p, C.c_int, C.c_uint, C.c_double
    u64p = C.POINTER(C.c_uint64)
    L.dcf_ctx_create.argtypes = [i, u, C.POINTER(vp)]
    L.dcf_ctx_destroy.argtypes = [vp]
    L.dcf_ctx_set_stream.argtypes = [vp, vp]
    L.dcf_ctx_set_graph.argtypes = [vp, i]
    L.dcf_ctx_set_option.argtypes = [vp, C.c_char_p, C.c_char_p]
    L.dcf_ctx_stream.restype = vp
    L.dcf_ctx_stream.argtypes = [vp]
    L.dcf_last_error.restype = C.c_char_p
    L.dcf_last_error.argtypes = [vp]
    L.dcf_ctx_synchronize.argtypes = [vp]
    L.dcf_ctx_set_profiling.argtypes = [vp, i]
    L.dcf_ctx_profile_only.argtypes = [vp, C.c_char_p]
    L.dcf_batch_filter.argtypes = [vp, C.POINTER(BatchImage), i, i, i, i, i, i, u]
    L.dcf_kernel_stats.argtypes = [vp, C.c_char_p, C.POINTER(C.c_int64), C.POINTER(C.c_double)]
    L.dcf_kernel_stats_reset.argtypes = [vp]
    L.dcf_kernel_name.restype = C.c_char_p
    L.dcf_kernel_name.argtypes = [vp, i]
    L.dcf_tree_build.argtypes = [vp, vp, i, u64p, i, i, i, vp, u, C.POINTER(vp)]
    L.dcf_tree_build_tiled.argtypes = [vp, vp, i, u64p, i, i, i, vp, C.POINTER(C.c_int), u, C.POINTER(vp)]
    L.dcf_tree_free.argtypes = [vp]
    L.dcf_tree_accumulate.argtypes = [vp, i]
    L.dcf_tree_group.argtypes = [vp]
    L.dcf_tree_parent.argtypes = [vp, vp, u]
    L.dcf_tree_num_nodes.argtypes = [vp, C.POINTER(C.c_int64)]
    L.dcf_filter.argtypes = [vp, i, i, d, i, vp, i, u]
    L.dcf_check.argtypes = [vp, i, vp, u]
    L.dcf_spectrum1d_noalloc.argtypes = [vp, i, vp, i, vp]
    L.dcf_dmp.argtypes = [vp, i, i, vp, i, vp, i, u]
    L.dcf_tree_write_csv.argtypes = [vp, i, C.c_char_p]
    L.dcf_spectrum1d.argtypes = [vp, i, vp, i, vp]
    L.dcf_spectrum2d.argtypes = [vp, i, vp, i, i, vp, i, vp]
    L.dcf_lvec_read.argtypes = [C.c_char_p, d, C.POINTER(C.POINTER(C.c_float)), C.POINTER(i)]
    L.dcf_lvec_free.argtypes = [C.POINTER(C.c_float)]
    L.dcf_attr_name.restype = C.c_char_p
    L.dcf_attr_name.argtypes = [i]
    L.dcf_attr_short_name.restype = C.c_char_p
    L.dcf_attr_short_name.argtypes = [i]
    L.dcf_attr_group.argtypes = [i]
    L.dcf_decision_name.restype = C.c_char_p
    L.dcf_decision_name.argtypes = [i]
    L.dcf_run_operations.argtypes = [vp, vp, i, u64p, i, i, vp, C.c_char_p, vp, vp]
    _lib = L
    return L


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _dtype_code(dt) -> int:
    dt = np.dtype(dt)
    if dt == np.uint8:
        return DCF_U8
    if dt == np.uint16:
        return DCF_U16
    if dt == np.float32:
        return DCF_F32
    raise TypeError("image dtype must be uint8, uint16 or float32, got %s" % dt)


def _dims_of(shape):
    if len(shape) == 2:
        return (shape[1], shape[0], 1)
    if len(shape) == 3:
        return (shape[2], shape[1], shape[0])
    raise ValueError("2D (H,W) or 3D (D,H,W) image expected")


def _pixdim3(pixdim) -> np.ndarray:
    """data_properties.pixdim (x, y, z) as the three floats the library reads"""
    pd = np.ascontiguousarray(pixdim, dtype=np.float32).reshape(-1)
    if pd.size != 3:
        raise ValueError("pixdim must hold 3 values (x, y, z), got %d" % pd.size)
    return pd


def _tree_type_code(tree_type) -> int:
    if tree_type not in ("max", "min"):
        raise ValueError("tree_type must be 'max' or 'min', got %r" % (tree_type,))
    return 1 if tree_type == "min" else 0


def _batch_args(count: int, lambdas, tree_types):
    """one lambda and one optional tree type per image"""
    lams = [float(x) for x in lambdas]
    if len(lams) != count:
        raise ValueError("%d images but %d lambdas" % (count, len(lams)))
    tts = list(tree_types) if tree_types is not None else [None] * count
    if len(tts) != count:
        raise ValueError("%d images but %d tree types" % (count, len(tts)))
    return lams, tts


def lvec_read(path: str, imscale: float = 1.0) -> np.ndarray:
    """lambda_vector_read (lambdavec.c:35-84): lambda_0 = 0 prepended, values divided by imscale."""
    L = lib()
    p = C.POINTER(C.c_float)()
    n = C.c_int()
    rc = L.dcf_lvec_read(path.encode(), float(imscale), C.byref(p), C.byref(n))
    if rc != 0:
        raise DcfError(rc, "couldn't read the lvec file: %s" % path)
    out = np.ctypeslib.as_array(p, shape=(n.value,)).copy()
    L.dcf_lvec_free(p)
    return out


class Context:
    """A GPU context (one HIP stream, a caching allocator and its trees)."""

    def __init__(self, device: int = 0):
        L = lib()
        h = C.c_void_p()
        rc = L.dcf_ctx_create(int(device), 0, C.byref(h))
        if rc != 0:
            raise DcfError(rc, "dcf_ctx_create(device=%d) failed (no GPU visible?)" % device)
        self._h = h
        self._trees = weakref.WeakSet()

    def close(self):
        if getattr(self, "_h", None):
            for t in list(self._trees):
                t.free()
            lib().dcf_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int):
        if rc != 0:
            raise DcfError(rc, lib().dcf_last_error(self._h).decode())

    def set_stream(self, stream_handle: int | None):
        self._check(lib().dcf_ctx_set_stream(self._h, C.c_void_p(stream_handle) if stream_handle else None))

    def set_option(self, name: str, value):
        """one tunable of the build by its environment name (dcf_ctx_set_option), e.g. ("DCF_ZONE_TILE2D", "32x32")"""
        self._check(lib().dcf_ctx_set_option(self._h, name.encode(), str(value).encode()))

    def set_graph(self, on: bool):
        """capture every build pass's level loop into one HIP graph (dcf_ctx_set_graph)"""
        self._check(lib().dcf_ctx_set_graph(self._h, 1 if on else 0))

    def synchronize(self):
        self._check(lib().dcf_ctx_synchronize(self._h))

    def set_profiling(self, on: bool):
        self._check(lib().dcf_ctx_set_profiling(self._h, 1 if on else 0))

    def profile_only(self, name: str | None):
        """Record events only around kernels whose name contains `name` (None: every kernel)."""
        self._check(lib().dcf_ctx_profile_only(self._h, name.encode() if name else None))

    def kernel_stats(self, name: str):
        n = C.c_int64()
        ms = C.c_double()
        self._check(lib().dcf_kernel_stats(self._h, name.encode(), C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def kernel_names(self):
        out = []
        i = 0
        while True:
            s = lib().dcf_kernel_name(self._h, i)
            if not s:
                return out
            out.append(s.decode())
            i += 1

    def reset_stats(self):
        self._check(lib().dcf_kernel_stats_reset(self._h))

    def build(self, img, connectivity: int = 4, tree_type: str = "max", attr_idx: int = 0, pixdim=(1.0, 1.0, 1.0)):
        """build_tree_with_attributes (flood.c:175) on a host numpy image."""
        a = np.ascontiguousarray(img)
        return self._build(_ptr(a), _dtype_code(a.dtype), _dims_of(a.shape), a.shape, connectivity, tree_type,
                           attr_idx, pixdim, 0, keep=a)

    def build_device(self, ptr: int, dtype, shape, connectivity: int = 4, tree_type: str = "max", attr_idx: int = 0,
                     pixdim=(1.0, 1.0, 1.0)):
        """Same on an image already resident in HBM (e.g. a torch tensor's data_ptr()); the image must
        stay unchanged until the tree is freed."""
        return self._build(C.c_void_p(ptr), _dtype_code(dtype), _dims_of(shape), tuple(shape), connectivity, tree_type,
                           attr_idx, pixdim, DCF_DEVICE_PTR, keep=None)

    def build_tiled(self, img, grid, connectivity: int = 4, tree_type: str = "max", attr_idx: int = 0,
                    pixdim=(1.0, 1.0, 1.0)):
        """The reference's distributed build on one GPU (dcf_tree_build_tiled): the image cut into
        grid = (tiles along x, along y[, along z]) tiles with a one-pixel overlap, every tile built on its
        own and the tiles' boundary trees merged; the result equals build()'s tree."""
        a = np.ascontiguousarray(img)
        g = (C.c_int * 3)(*(tuple(int(v) for v in grid) + (1, 1, 1))[:3])
        return self._build(_ptr(a), _dtype_code(a.dtype), _dims_of(a.shape), a.shape, connectivity, tree_type,
                           attr_idx, pixdim, 0, keep=a, grid=g)

    def build_tiled_device(self, ptr: int, dtype, shape, grid, connectivity: int = 4, tree_type: str = "max",
                           attr_idx: int = 0, pixdim=(1.0, 1.0, 1.0)):
        """build_tiled on an image already resident in HBM."""
        g = (C.c_int * 3)(*(tuple(int(v) for v in grid) + (1, 1, 1))[:3])
        return self._build(C.c_void_p(ptr), _dtype_code(dtype), _dims_of(shape), tuple(shape), connectivity, tree_type,
                           attr_idx, pixdim, DCF_DEVICE_PTR, keep=None, grid=g)

    def _build(self, p, code, dims, shape, connectivity, tree_type, attr_idx, pixdim, flags, keep, grid=None):
        tt = _tree_type_code(tree_type)
        d = (C.c_uint64 * 3)(*dims)
        pd = _pixdim3(pixdim)
        h = C.c_void_p()
        if grid is not None:
            rc = lib().dcf_tree_build_tiled(self._h, p, code, d, int(connectivity), tt, int(attr_idx), _ptr(pd), grid,
                                            flags, C.byref(h))
        else:
            rc = lib().dcf_tree_build(self._h, p, code, d, int(connectivity), tt,
                                      int(attr_idx), _ptr(pd), flags, C.byref(h))
        self._check(rc)
        t = Tree(self, h, shape, code, keep)
        self._trees.add(t)
        return t

    def batch_filter(self, images, lambdas, connectivity: int = 4, tree_type: str = "max", attr_idx: int = 0,
                     decision: int = 0, extract: bool = False, all_rules: bool = False, tree_types=None):
        """The batched form: every host image (u8 or u16, all 2D or all 3D) gets its own tree and is
        filtered with its own lambda, all in one build over the batch.  tree_types: an optional
        per-image 'max' / 'min' (None entries take tree_type).  Returns the outputs."""
        arrs = [np.ascontiguousarray(a) for a in images]
        lams, tts = _batch_args(len(arrs), lambdas, tree_types)
        outs = [np.empty_like(a) for a in arrs]
        items = (BatchImage * len(arrs))()
        for k, (a, o, lam) in enumerate(zip(arrs, outs, lams)):
            items[k].img = a.ctypes.data
            items[k].dtype = _dtype_code(a.dtype)
            items[k].dims = (C.c_uint64 * 3)(*_dims_of(a.shape))
            items[k].lam = lam
            items[k].out = o.ctypes.data
            items[k].flags = _image_flags(tts[k])
        self._batch(items, len(arrs), connectivity, tree_type, attr_idx, decision, extract, all_rules, 0)
        return outs

    def batch_filter_device(self, items, connectivity: int = 4, tree_type: str = "max", attr_idx: int = 0,
                            decision: int = 0, extract: bool = False, all_rules: bool = False):
        """Same on device buffers: items = [(img_ptr, dtype, shape, lam, out_ptr[, tree_type]), ...]."""
        arr = (BatchImage * len(items))()
        for k, it in enumerate(items):
            ip, dt, shape, lam, op = it[:5]
            arr[k].img = ip
            arr[k].dtype = _dtype_code(dt)
            arr[k].dims = (C.c_uint64 * 3)(*_dims_of(shape))
            arr[k].lam = float(lam)
            arr[k].out = op
            arr[k].flags = _image_flags(it[5] if len(it) > 5 else None)
        self._batch(arr, len(items), connectivity, tree_type, attr_idx, decision, extract, all_rules, DCF_DEVICE_PTR)

    def _batch(self, arr, count, connectivity, tree_type, attr_idx, decision, extract, all_rules, flags):
        tt = _tree_type_code(tree_type)
        fl = flags | (DCF_ALL_RULES if all_rules else 0)
        self._check(lib().dcf_batch_filter(self._h, arr, int(count), int(connectivity), tt,
                                           int(attr_idx), int(decision), 1 if extract else 0, fl))

    def run_operations(self, img, operations, connectivity: int = 4, tree_type: str = "max", pixdim=(1.0, 1.0, 1.0)):
        """loop_operations (disccofan.c:137-205): run the reference's operation tuples in order.
        Returns one output per operation (image, check image, spectrum, or a differential profile
        stacked along a leading axis)."""
        a = np.ascontiguousarray(img)
        if not isinstance(operations, str):
            operations = ",".join(_tuple_text(t) for t in operations)
        outs = {}
        shape = a.shape

        @C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int64)
        def cb(user, idx, kind, data, count):
            if kind in (0, 1):
                arr = np.ctypeslib.as_array(C.cast(data, C.POINTER(C.c_float)), shape=(count,)).copy()
                outs[idx] = arr.reshape(shape)
            elif kind == 5:
                outs[idx] = C.string_at(data, count).decode()
            elif kind == 4:
                arr = np.ctypeslib.as_array(C.cast(data, C.POINTER(C.c_float)), shape=(count,)).copy()
                outs[idx] = arr.reshape((-1,) + tuple(shape))
            else:
                outs[idx] = np.ctypeslib.as_array(C.cast(data, C.POINTER(C.c_double)), shape=(count,)).copy()

        d = (C.c_uint64 * 3)(*_dims_of(a.shape))
        pd = _pixdim3(pixdim)
        rc = lib().dcf_run_operations(self._h, _ptr(a), _dtype_code(a.dtype), d, int(connectivity),
                                      _tree_type_code(tree_type), _ptr(pd), operations.encode(),
                                      C.cast(cb, C.c_void_p), None)
        self._check(rc)
        return [outs[i] for i in sorted(outs)]


def _tuple_text(t) -> str:
    parts = []
    for v in t:
        parts.append('"%s"' % v if isinstance(v, str) else repr(v))
    return "(" + ",".join(parts) + ")"


class Tree:
    """A component tree resident on the GPU (dcf_tree)."""

    def __init__(self, ctx: Context, h, shape, dtype_code, keep):
        self.ctx = ctx
        self._h = h
        self.shape = tuple(shape)
        self.n = int(np.prod(self.shape))
        self.dtype_code = dtype_code
        self._keep = keep

    def free(self):
        if getattr(self, "_h", None):
            if getattr(self.ctx, "_h", None):  # a closed context already freed its trees
                lib().dcf_tree_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @property
    def num_nodes(self) -> int:
        k = C.c_int64()
        self.ctx._check(lib().dcf_tree_num_nodes(self._h, C.byref(k)))
        return k.value

    @property
    def group(self) -> int:
        return int(lib().dcf_tree_group(self._h))

    def accumulate(self, attr_idx: int):
        """flood_attributes (flood.c:184) when attr_idx is in another group."""
        self.ctx._check(lib().dcf_tree_accumulate(self._h, int(attr_idx)))

    def parent(self) -> np.ndarray:
        out = np.empty(self.n, dtype=np.int32)
        self.ctx._check(lib().dcf_tree_parent(self._h, _ptr(out), 0))
        return out

    def filter(self, attr_idx: int, lam: float, decision: int = 0, extract: bool = False, out_dtype=np.float32,
               all_rules: bool = False) -> np.ndarray:
        out = np.empty(self.shape, dtype=out_dtype)
        flags = DCF_ALL_RULES if all_rules else 0
        self.ctx._check(lib().dcf_filter(self._h, int(attr_idx), int(decision), float(lam), 1 if extract else 0,
                                         _ptr(out), _dtype_code(out_dtype), flags))
        return out

    def filter_device(self, attr_idx: int, lam: float, decision: int, out_ptr: int, out_dtype, extract: bool = False,
                      all_rules: bool = False):
        flags = DCF_DEVICE_PTR | (DCF_ALL_RULES if all_rules else 0)
        self.ctx._check(lib().dcf_filter(self._h, int(attr_idx), int(decision), float(lam), 1 if extract else 0,
                                         C.c_void_p(out_ptr), _dtype_code(out_dtype), flags))

    def dmp(self, attr_idx: int, lambdas, decision: int = 0, out_dtype=np.float32) -> np.ndarray:
        """perform_dmp_operation (operation.c:63-77): g - filter(lambda_i) for every lambda after
        lambda_0, stacked along a leading axis (min-trees stay in the inverted grey space)."""
        lv = np.ascontiguousarray(lambdas, dtype=np.float32).reshape(-1)
        out = np.empty((max(len(lv) - 1, 0),) + tuple(self.shape), dtype=out_dtype)
        self.ctx._check(lib().dcf_dmp(self._h, int(attr_idx), int(decision), _ptr(lv), len(lv), _ptr(out),
                                      _dtype_code(out_dtype), 0))
        return out

    def write_csv(self, path: str, attr_idx: int):
        """write_tree (io.c:1431-1573): the tree as the reference's CSV"""
        self.ctx._check(lib().dcf_tree_write_csv(self._h, int(attr_idx), os.fsencode(path)))

    def check(self, attr_idx: int) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.float32)
        self.ctx._check(lib().dcf_check(self._h, int(attr_idx), _ptr(out), 0))
        return out

    def spectrum1d(self, attr_idx: int, lambdas) -> np.ndarray:
        lv = np.ascontiguousarray(lambdas, dtype=np.float32).reshape(-1)
        out = np.empty(len(lv), dtype=np.float64)
        self.ctx._check(lib().dcf_spectrum1d(self._h, int(attr_idx), _ptr(lv), len(lv), _ptr(out)))
        return out

    def spectrum2d(self, attr1: int, l1, attr2: int, l2) -> np.ndarray:
        a = np.ascontiguousarray(l1, dtype=np.float32).reshape(-1)
        b = np.ascontiguousarray(l2, dtype=np.float32).reshape(-1)
        out = np.empty(len(a) * len(b), dtype=np.float64)
        self.ctx._check(lib().dcf_spectrum2d(self._h, int(attr1), _ptr(a), len(a), int(attr2), _ptr(b), len(b),
                                             _ptr(out)))
        return out
