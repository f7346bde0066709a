"""The .img format (libs/imagelib_fftw3.cpp:57-224) against the reference's own file."""
import os
import struct

import numpy as np

from qstem_amd.imgio import Image, read_img, write_img

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_read_reference_detector_image():
    img = read_img(os.path.join(GOLD, "detector1_2.img"))
    assert img.data.shape == (20, 20) and img.data.dtype == np.float32
    assert abs(img.t - 46.86) < 1e-4 and abs(img.dx - 0.402) < 1e-6 and abs(img.dy - 0.3935) < 1e-6
    assert len(img.params) == 2 + 400 and img.params[0] == 3.0  # [runs, error, image2...] (stemlib.cpp:2737-2745)
    assert img.comment == "detector1"
    assert abs(img.data.mean() - 0.0672667) < 1e-6


def test_roundtrip_is_byte_identical(tmp_path):
    src = os.path.join(GOLD, "detector1_2.img")
    img = read_img(src)
    out = tmp_path / "copy.img"
    write_img(str(out), img)
    assert open(src, "rb").read() == open(out, "rb").read()


def test_complex_header(tmp_path):
    a = (np.arange(12) + 1j * np.arange(12)).reshape(3, 4).astype(np.complex64)
    p = tmp_path / "w.img"
    write_img(str(p), Image(a, t=78.1, dx=0.0625, dy=0.0625, comment="Wavefunction"))
    b = open(p, "rb").read()
    assert struct.unpack("<8i", b[:32]) == (56, 0, 12, 3, 4, 1, 8, 1)
    assert np.array_equal(read_img(str(p)).data, a)
