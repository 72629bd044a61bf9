"""Convert the reference's own golden files into compact numpy fixtures.

Run once in a container that has the reference checkout (it is the only script that reads it):

    python tests/golden/make_fixtures.py /path/to/reference

Sources (all under the reference's tests/):
  checkimg/input_image.png      Lena 512x512 u8 (PNG rows are top-down; the reference's FreeImage
                                reader stores scanline 0 = bottom row, src/sources/io.c:893-897,
                                so the memory-order image is np.flipud(png))
  checkimg/lena512.fits         the same input already in memory order (cross-check)
  checkimg/GT_2D_check_area.fits    ("check",0), 4-connectivity   (test_sequential.sh:146-148)
  checkimg/GT_2D_check_area_8.fits  ("check",0), 8-connectivity
  checkimg/GT_2D_check_ncomp.fits   ("check",17), 4-connectivity
  checkimg/GT_area_200.png      ("filter",0,200) direct           (test_sequential.sh:125-127)
  checkimg/GT_ncomp_0.2.png     ("filter",17,0.2) direct
  checkimg/GT_tree.txt          ("tree",15) CSV of every levelroot (test_sequential.sh:173-174)
  lvec.txt, lvec2.txt           lambda-vector files

Everything is data: pixel arrays and parsed CSV columns, no reference code.
"""
import os
import shutil
import sys

import numpy as np


def read_fits_f32(path):
    with open(path, "rb") as f:
        raw = f.read()
    hdr = {}
    pos = 0
    while True:
        block = raw[pos:pos + 2880]
        pos += 2880
        done = False
        for i in range(0, 2880, 80):
            card = block[i:i + 80].decode("ascii", "replace")
            key = card[:8].strip()
            if key == "END":
                done = True
                break
            if card[8:10] == "= ":
                hdr[key] = card[10:].split("/")[0].strip()
        if done:
            break
    bitpix = int(hdr["BITPIX"])
    nx, ny = int(hdr["NAXIS1"]), int(hdr["NAXIS2"])
    if bitpix == -32:
        dt = ">f4"
    elif bitpix == 8:
        dt = "u1"
    elif bitpix == 16:
        dt = ">i2"
    else:
        raise ValueError("unsupported BITPIX %d" % bitpix)
    a = np.frombuffer(raw[pos:pos + nx * ny * np.dtype(dt).itemsize], dtype=dt).reshape(ny, nx)
    return a.astype(np.float32 if bitpix == -32 else np.int64)


def read_png(path):
    from PIL import Image
    return np.array(Image.open(path))


def main(ref):
    ck = os.path.join(ref, "tests", "checkimg")
    here = os.path.dirname(os.path.abspath(__file__))
    img = np.flipud(read_png(os.path.join(ck, "input_image.png"))).copy()
    lena = read_fits_f32(os.path.join(ck, "lena512.fits"))
    assert np.array_equal(img.astype(np.int64), lena.astype(np.int64)), "flipud(png) != lena512.fits"
    area4 = read_fits_f32(os.path.join(ck, "GT_2D_check_area.fits"))
    area8 = read_fits_f32(os.path.join(ck, "GT_2D_check_area_8.fits"))
    ncomp4 = read_fits_f32(os.path.join(ck, "GT_2D_check_ncomp.fits"))
    f_area = np.flipud(read_png(os.path.join(ck, "GT_area_200.png"))).copy()
    f_ncomp = np.flipud(read_png(os.path.join(ck, "GT_ncomp_0.2.png"))).copy()

    cols = None
    rows = []
    with open(os.path.join(ck, "GT_tree.txt")) as f:
        cols = f.readline().strip().split(",")
        for line in f:
            line = line.strip()
            if line:
                rows.append([float(v) for v in line.split(",")])
    tab = np.array(rows, dtype=np.float64)
    tree = {("tree_" + c): tab[:, i] for i, c in enumerate(cols)}
    for c in ("index", "x", "y", "z", "parent", "area"):
        tree["tree_" + c] = tree["tree_" + c].astype(np.int64)

    np.savez_compressed(
        os.path.join(here, "lena512.npz"),
        img=img.astype(np.uint8),
        check_area4=area4.astype(np.uint32),
        check_area8=area8.astype(np.uint32),
        check_ncomp4=ncomp4.astype(np.float32),
        filter_area200=f_area.astype(np.uint8),
        filter_ncomp02=f_ncomp.astype(np.uint8),
        **tree,
    )
    for name in ("lvec.txt", "lvec2.txt"):
        shutil.copyfile(os.path.join(ref, "tests", name), os.path.join(here, name))
    print("wrote", os.path.join(here, "lena512.npz"))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
