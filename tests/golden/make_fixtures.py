"""Extract the reference's own test data (tests/data/*.img of the QSTEM tree)
into small fixtures under tests/golden/.

The three .img files are outputs of stem3/debug_data/stem_STO_4x4.qsc with
`tds: yes` (time-seeded), so they are statistical fixtures:
  detector1_2.img   copied verbatim (4.9 KB): a 20x20 ADF (70-133 mrad) image
  SrTiO3.cfg, stem_STO_4x4.qsc  the model and parameters that produced them
  reference_stats.json: exit-wave norm of mulswav_16_2.img and the radial
                    profile / total of diffAvg_0_16.img (the 1.3 MB and 0.6 MB
                    arrays themselves are not kept).
Run in the development container only: python tests/golden/make_fixtures.py
"""
import json
import os
import shutil
import struct

import numpy as np

SRC = "/root/reference/tests/data"
HERE = os.path.dirname(os.path.abspath(__file__))


def read_img(path):
    b = open(path, "rb").read()
    hs, ps, cs, nx, ny, cflag, dsize, ver = struct.unpack("<8i", b[:32])
    t, dx, dy = struct.unpack("<3d", b[32:56])
    params = np.frombuffer(b[56:56 + 8 * ps], "<f8")
    off = 56 + 8 * ps + cs
    dt = "<c8" if cflag else "<f4"
    data = np.frombuffer(b[off:off + nx * ny * dsize], dt).reshape(nx, ny)
    return dict(nx=nx, ny=ny, t=t, dx=dx, dy=dy, params=params, comment=b[56 + 8 * ps:off].decode(), data=data)


def main():
    shutil.copyfile(os.path.join(SRC, "detector1_2.img"), os.path.join(HERE, "detector1_2.img"))
    # the model and parameter file that produced the fixtures (data files of stem3/debug_data)
    dbg = os.path.join(os.path.dirname(os.path.dirname(SRC)), "stem3", "debug_data")
    for name in ("SrTiO3.cfg", "stem_STO_4x4.qsc"):
        shutil.copyfile(os.path.join(dbg, name), os.path.join(HERE, name))
    w = read_img(os.path.join(SRC, "mulswav_16_2.img"))
    d = read_img(os.path.join(SRC, "diffAvg_0_16.img"))
    a = read_img(os.path.join(SRC, "detector1_2.img"))
    wave = w["data"].astype(np.complex128)
    diff = d["data"].astype(np.float64)
    n = diff.shape[0]
    k = (np.arange(n) - n // 2) / (n * d["dx"])
    kr = np.sqrt(k[:, None] ** 2 + k[None, :] ** 2)
    edges = np.linspace(0, kr.max(), 41)
    prof = [float(diff[(kr >= lo) & (kr < hi)].mean()) for lo, hi in zip(edges[:-1], edges[1:])]
    stats = {
        "source": "QSTEM tests/data, outputs of stem3/debug_data/stem_STO_4x4.qsc (tds: yes)",
        "mulswav_16_2": {"nx": w["nx"], "ny": w["ny"], "t": w["t"], "dx": w["dx"], "comment": w["comment"],
                         "norm": float(np.sum(np.abs(wave) ** 2) / wave.size)},
        "diffAvg_0_16": {"nx": d["nx"], "t": d["t"], "dx": d["dx"], "comment": d["comment"],
                         "sum": float(diff.sum()), "radial_edges": edges.tolist(), "radial_mean": prof},
        "detector1_2": {"min": float(a["data"].min()), "max": float(a["data"].max()),
                        "mean": float(a["data"].mean()), "t": a["t"], "dx": a["dx"], "dy": a["dy"],
                        "runs": float(a["params"][0])},
    }
    with open(os.path.join(HERE, "reference_stats.json"), "w") as f:
        json.dump(stats, f, indent=1)
    print(json.dumps({k: v for k, v in stats.items() if k != "diffAvg_0_16"}, indent=1))


if __name__ == "__main__":
    main()
