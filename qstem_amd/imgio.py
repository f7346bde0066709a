"""QSTEM .img files (libs/imagelib_fftw3.cpp:57-224).

Layout: a 56-byte header of 8 int32 (headerSize=56, paramSize, commentSize,
nx, ny, complexFlag, dataSize, version=1) and 3 float64 (t, dx, dy); then
paramSize float64 parameters, the comment (commentSize bytes, no
terminator), then nx*ny row-major values (float32, or complex64 when
complexFlag is set).  Reading skips 56 + commentSize + 8*paramSize bytes;
the reference's ReadImage seeks 56 + commentSize + 2*paramSize and compares
ny with nx (Q7), which this reader does not reproduce.
"""
import struct
from dataclasses import dataclass, field
from typing import List

import numpy as np

HEADER = struct.Struct("<8i3d")


@dataclass
class Image:
    data: np.ndarray
    t: float = 0.0
    dx: float = 1.0
    dy: float = 1.0
    params: List[float] = field(default_factory=list)
    comment: str = ""
    version: int = 1


def write_img(path: str, img: Image) -> None:
    data = np.ascontiguousarray(img.data)
    if np.iscomplexobj(data):
        data = data.astype("<c8")
        cflag, dsize = 1, 8
    else:
        data = data.astype("<f4")
        cflag, dsize = 0, 4
    nx, ny = data.shape
    comment = img.comment.encode()
    params = np.asarray(img.params, dtype="<f8")
    with open(path, "wb") as f:
        f.write(HEADER.pack(56, len(params), len(comment), nx, ny, cflag, dsize, img.version, img.t, img.dx, img.dy))
        f.write(params.tobytes())
        f.write(comment)
        f.write(data.tobytes())


def read_img(path: str) -> Image:
    b = open(path, "rb").read()
    hs, ps, cs, nx, ny, cflag, dsize, ver, t, dx, dy = HEADER.unpack(b[:56])
    params = np.frombuffer(b[56:56 + 8 * ps], "<f8").tolist()
    comment = b[56 + 8 * ps:56 + 8 * ps + cs].decode(errors="replace")
    off = 56 + 8 * ps + cs
    dt = "<c8" if cflag else "<f4"
    n = nx * ny
    data = np.frombuffer(b[off:off + n * np.dtype(dt).itemsize], dt).reshape(nx, ny).copy()
    return Image(data, t, dx, dy, params, comment, ver)
