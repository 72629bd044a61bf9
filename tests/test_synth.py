"""The seeded generators of the benchmark workloads (disccofan_amd/synth, libdcf_synth.so) on the CPU.

A workload is defined by its generator, so these tests pin the generators: the same call gives the same bytes,
whatever the OpenMP thread count; the seed changes them; shapes, types and ranges are as documented; and small
outputs keep their recorded SHA-256 (taken with this image's gcc and libm; the generators use sin, cos, exp and
sqrt, so another C library may round differently).
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from disccofan_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

GOLDEN_SHA = {
    "vnoise 48x64": "7eec4e9c72461dcd9c18b4414ee24abad8b4e12ee577fdc08483beeff3833cc6",
    "disks u8 50x40": "f0e862ccb3512b7e5f7dd7bf1ea724a8eee88deccd72f6115297b30444bc3afc",
    "disks u16 50x40": "0254281ad98f06ec7b74331ee6734612a5a96d1129eb61bb33cdf25887496538",
    "spheres 12x10x14": "0a73c177fbfdeeb33a6c3b8867833424d95121891676edaf8d56b7dcbdf39448",
    "blobs u16 30x20": "e3d63ab140b25c958ba8157975d25109acf3163fe6c19c32c4fd103b3b98cd1b",
    "batch shapes 16 in [64,300]": "3081774bb34a87f406ef9e45970a79b4db5fd562b6e0f92ce651b2e7d0e90b62",
}


def _small_outputs():
    s = synth.batch_shapes(16, 64, 300)
    return {
        "vnoise 48x64": synth.vnoise(48, 64),
        "disks u8 50x40": synth.disks(50, 40, False),
        "disks u16 50x40": synth.disks(50, 40, True),
        "spheres 12x10x14": synth.spheres(12, 10, 14),
        "blobs u16 30x20": synth.blobs(30, 20),
        "batch shapes 16 in [64,300]": np.array([x[:3] for x in s], dtype=np.int64),
    }


def test_recorded_hashes():
    got = {k: synth.sha256(v) for k, v in _small_outputs().items()}
    assert got == GOLDEN_SHA


def test_types_and_shapes():
    assert synth.vnoise(7, 9).dtype == np.uint16 and synth.vnoise(7, 9).shape == (7, 9)
    assert synth.disks(5, 6).dtype == np.uint8 and synth.disks(5, 6, True).dtype == np.uint16
    assert synth.spheres(3, 4, 5).dtype == np.uint8 and synth.spheres(3, 4, 5).shape == (3, 4, 5)
    q, r = synth.blobs(11, 13, raw=True)
    assert q.dtype == np.uint16 and r.dtype == np.float32 and q.shape == r.shape == (11, 13)
    # the quantisation keeps the order of the float image
    order = np.argsort(r.reshape(-1), kind="stable")
    assert np.all(np.diff(q.reshape(-1)[order].astype(np.int64)) >= 0)


def test_seed_changes_the_output():
    a = synth.vnoise(40, 40)
    assert np.array_equal(a, synth.vnoise(40, 40))
    assert not np.array_equal(a, synth.vnoise(40, 40, seed=synth.SEED["vnoise"] + 1))
    d = synth.disks(40, 40)
    assert not np.array_equal(d, synth.disks(40, 40, seed=synth.SEED["disks"] + 1))


def test_batch_shapes_ranges():
    shapes = synth.batch_shapes(256, 64, 2048)
    W = np.array([s[0] for s in shapes])
    H = np.array([s[1] for s in shapes])
    is16 = np.array([s[2] for s in shapes])
    lam = np.array([s[4] for s in shapes])
    assert W.min() >= 64 and W.max() <= 2048 and H.min() >= 64 and H.max() <= 2048
    assert lam.min() >= 10.0 and lam.max() <= 5000.0
    assert is16.sum() == 128  # half u8, half u16
    assert len(set(s[3] for s in shapes)) == 256  # a seed per image
    assert synth.batch_shapes(256, 64, 2048) == shapes


def test_log_lambdas():
    v = synth.log_lambdas(9, 1.0, 256.0)
    assert v.dtype == np.float32 and v[0] == 0 and v[1] == 1.0 and v[-1] == 256.0
    assert np.all(np.diff(v) > 0)


@pytest.mark.parametrize("threads", ["1", "5"])
def test_thread_count_invariance(threads):
    """the generators run OpenMP loops: their bytes must not depend on the thread count"""
    code = ("import sys; sys.path.insert(0, %r); from disccofan_amd import synth; "
            "print(synth.sha256(synth.vnoise(300, 257)), synth.sha256(synth.disks(130, 170, True)), "
            "synth.sha256(synth.spheres(20, 33, 41)), synth.sha256(synth.blobs(90, 70, raw=True)[1]))" % ROOT)
    env = dict(os.environ, OMP_NUM_THREADS=threads)
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120, check=True)
    here = " ".join([synth.sha256(synth.vnoise(300, 257)), synth.sha256(synth.disks(130, 170, True)),
                     synth.sha256(synth.spheres(20, 33, 41)), synth.sha256(synth.blobs(90, 70, raw=True)[1])])
    assert out.stdout.strip() == here
