"""dcf_spectrum1d_noalloc (the allocation-free 1D spectrum, scratch in a fixed device array) against
dcf_spectrum1d on the same trees, through the C ABI: two calls in a row, the second with a smaller
tree and fewer bins, and a request beyond the maximum bin count."""
import ctypes as C

import numpy as np
import pytest

import disccofan_amd as dcf
from disccofan_amd import synth

pytestmark = pytest.mark.gpu


def noalloc(tree, attr, lv):
    lv = np.ascontiguousarray(lv, dtype=np.float32)
    out = np.full(len(lv), -1.0, dtype=np.float64)
    rc = dcf.lib().dcf_spectrum1d_noalloc(tree._h, int(attr), lv.ctypes.data_as(C.c_void_p), len(lv),
                                          out.ctypes.data_as(C.c_void_p))
    return rc, out


def test_noalloc_spectrum_matches_the_allocating_one():
    ctx = dcf.Context(0)
    cases = [
        # (image, connectivity, tree type, attribute, thresholds): integer greys give exact u64 bins
        (synth.disks(700, 900), 4, "max", 0, synth.log_lambdas(200, 1, 2 ** 30)),
        (synth.vnoise(130, 170), 8, "min", 0, synth.log_lambdas(33, 1, 2 ** 16)),
    ]
    for img, conn, tt, attr, lv in cases:
        t = ctx.build(img, conn, tt, attr)
        want = t.spectrum1d(attr, lv)
        rc, got = noalloc(t, attr, lv)
        assert rc == 0, dcf.lib().dcf_last_error(ctx._h)
        assert np.array_equal(got, want), (img.shape, got, want)
        t.free()
    # a float image: double bins whose summation order varies between calls
    f = synth.vnoise(96, 80).astype(np.float32) * np.float32(0.37)
    t = ctx.build(f, 4, "max", 17)
    lv = np.array([0, 0.16, 0.2, 0.3, 0.5, 1.0, 4.0], np.float32)
    want = t.spectrum1d(17, lv)
    rc, got = noalloc(t, 17, lv)
    assert rc == 0
    assert np.allclose(got, want, rtol=1e-12, atol=0), (got, want)
    # beyond the maximum: refused before anything runs, and the output is untouched
    big = np.zeros(65537, np.float32)
    rc, out = noalloc(t, 17, big)
    assert rc == -1 and np.all(out == -1.0)
    t.free()
    ctx.close()
