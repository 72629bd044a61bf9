"""The C-ABI library loads on a machine without a GPU and exports every entry point of include/dcf.h.
No compute call is made here; status-code behaviour without a device is checked where it is safe."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dcf.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"\b(dcf_[a-z0-9_]+)\s*\(", txt)
    return sorted(set(n for n in names if n != "dcf_output_fn"))


def test_header_declares_the_path():
    names = declared_functions()
    for must in ("dcf_tree_build", "dcf_tree_accumulate", "dcf_filter", "dcf_check", "dcf_spectrum1d",
                 "dcf_spectrum2d", "dcf_tree_parent", "dcf_lvec_read", "dcf_run_operations"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import disccofan_amd as d
    L = d.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(d.EXPORTS) == sorted(declared_functions())


def test_tables_match_reference():
    import disccofan_amd as d
    L = d.lib()
    shorts = [L.dcf_attr_short_name(i).decode() for i in range(18)]
    assert shorts == d.ATTRIBUTES
    assert [L.dcf_attr_group(i) for i in range(18)] == [0, 1, 1, 1, 1, 1, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 4]
    assert [L.dcf_decision_name(i).decode() for i in range(4)] == d.DECISIONS
    assert L.dcf_attr_name(18) is None and L.dcf_decision_name(4) is None


def test_lvec_reader_host_only(tmp_path):
    import numpy as np
    import disccofan_amd as d
    lv = d.lvec_read(os.path.join(ROOT, "tests", "golden", "lvec2.txt"))
    assert lv[0] == 0 and np.allclose(lv[1:], [1, 1.5, 2, 2.5, 3, 5])
    with pytest.raises(d.DcfError) as e:
        d.lvec_read(str(tmp_path / "missing.txt"))
    assert e.value.code == -5
    # no header line: values are read from the start, like the reference's fscanf literal
    p = tmp_path / "nohdr.txt"
    p.write_text("10 20\n30\n")
    assert np.array_equal(d.lvec_read(str(p), 2.0), np.array([0, 5, 10, 15], np.float32))


def test_lvec_reader_matches_oracle(tmp_path):
    import numpy as np
    import disccofan_amd as d
    from oracle.oracle import lvec_read as orc_read
    p = tmp_path / "odd.txt"
    p.write_text("#LambdaVector file. Keep this exact line in all similar lambda vector files\n0.5 1e3\n  7.25\n")
    for s in (1.0, 3.0):
        assert np.array_equal(d.lvec_read(str(p), s), orc_read(str(p), s))
    q = tmp_path / "badhdr.txt"
    q.write_text("#Lambda something else\n1 2 3\n")
    assert np.array_equal(d.lvec_read(str(q)), orc_read(str(q)))


def test_context_without_gpu_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    import disccofan_amd as d
    with pytest.raises(d.DcfError):
        d.Context(0)


def test_wrapper_argument_checks():
    """the Python wrapper refuses, before any library call, what the C side would read out of bounds
    (a pixdim of other than 3 values) or take for a null image (a batch with fewer lambdas or tree
    types than images), and tree types other than max / min"""
    import numpy as np
    import disccofan_amd as d
    assert d._pixdim3((1, 2, 3)).dtype == np.float32 and d._pixdim3([[1, 2, 3]]).tolist() == [1, 2, 3]
    for bad in ((1.0, 1.0), (1, 2, 3, 4), ()):
        with pytest.raises(ValueError):
            d._pixdim3(bad)
    assert d._tree_type_code("max") == 0 and d._tree_type_code("min") == 1
    for bad in ("Max", "minimum", None, 1):
        with pytest.raises(ValueError):
            d._tree_type_code(bad)
    assert d._image_flags(None) == 0 and d._image_flags("min") == d.DCF_IMAGE_MIN_TREE
    with pytest.raises(ValueError):
        d._image_flags("both")
    assert d._batch_args(2, [3, 4.5], None) == ([3.0, 4.5], [None, None])
    assert d._batch_args(2, (1, 2), ["min", None]) == ([1.0, 2.0], ["min", None])
    with pytest.raises(ValueError):
        d._batch_args(3, [1, 2], None)
    with pytest.raises(ValueError):
        d._batch_args(2, [1, 2], ["max"])


def _build_c_caller(tmp_path):
    import subprocess
    exe = str(tmp_path / "c_caller")
    libdir = os.path.join(ROOT, "disccofan_amd", "lib")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "c", "c_caller.c"), "-L", libdir, "-ldcf",
                    "-Wl,-rpath," + libdir, "-o", exe], check=True)
    return exe


def test_c_caller_links_and_reads_tables(tmp_path):
    """include/dcf.h compiles as C99 and a C program links libdcf.so and calls the host-only entry points"""
    import subprocess
    import disccofan_amd as d
    exe = _build_c_caller(tmp_path)
    r = subprocess.run([exe, "tables"], capture_output=True, text=True, check=True)
    rows = [l.split() for l in r.stdout.splitlines() if not l.startswith("decision")]
    assert [row[1] for row in rows] == d.ATTRIBUTES
    lv = os.path.join(ROOT, "tests", "golden", "lvec.txt")
    r = subprocess.run([exe, "lvec", lv], capture_output=True, text=True, check=True)
    vals = [float(x) for x in r.stdout.split()[2:]]
    assert vals == [float(x) for x in d.lvec_read(lv)]


def test_package_asks_for_hardware_queues():
    """importing the package before any HIP runtime has started asks for 16 hardware queues, so that the worker
    streams get queues of their own; a value already in the environment is kept"""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = "import os, disccofan_amd; print(os.environ.get('GPU_MAX_HW_QUEUES'))"
    env = {k: v for k, v in os.environ.items() if k != "GPU_MAX_HW_QUEUES"}
    r = subprocess.run([sys.executable, "-c", code], cwd=root, env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip() == "16"
    env["GPU_MAX_HW_QUEUES"] = "5"
    r = subprocess.run([sys.executable, "-c", code], cwd=root, env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip() == "5"
