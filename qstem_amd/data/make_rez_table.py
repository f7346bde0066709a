"""Convert the Rez et al. electron scattering-factor tables shipped with QSTEM
into raw little-endian float64 files that the library and the oracle load.

* ``rez_scatpar_f64.bin`` from python/SF_rez.txt (93 rows x 123 columns, CSV):
  row 0 is the s-grid (1/Angstrom, s = q/2 convention of the reference,
  stem3/scatfactsRez.h:22-25), rows 1..92 are f_e(s) for Z = 1..92.
* ``rez_scatparoffs_f64.bin`` from the scatParOffs table of
  stem3/scatfactsRez.h:122 (same shape): row 0 the same s-grid, row Z the
  derivative of f_e(s) with respect to the ion charge (zero rows for elements
  without ionic data), read by getAtomPotentialOffset3D (stemlib.cpp:1548-1741).

The tables are published physical data; the files are regenerated, not edited.

Usage: python make_rez_table.py [path/to/SF_rez.txt [path/to/scatfactsRez.h]]
"""
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def parse_c_table(path, name, shape=(93, 123)):
    """Numbers of the brace-initialised C array `name` in a header, as a float64 array."""
    text = open(path).read()
    m = re.search(r"\b%s\s*\[[^\]]*\]\s*\[[^\]]*\]\s*=\s*\{" % re.escape(name), text)
    if not m:
        raise ValueError("%s not found in %s" % (name, path))
    depth, i = 1, m.end()
    while depth:
        depth += {"{": 1, "}": -1}.get(text[i], 0)
        i += 1
    body = re.sub(r"//[^\n]*|/\*.*?\*/", "", text[m.end():i - 1], flags=re.S)
    vals = [float(v) for v in re.findall(r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?", body)]
    tab = np.array(vals, dtype=np.float64)
    if tab.size != shape[0] * shape[1]:
        raise ValueError("%s: %d values, expected %d" % (name, tab.size, shape[0] * shape[1]))
    return tab.reshape(shape)


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/python/SF_rez.txt"
    hdr = sys.argv[2] if len(sys.argv) > 2 else "/root/reference/stem3/scatfactsRez.h"
    tab = np.loadtxt(src, delimiter=",", dtype=np.float64)
    assert tab.shape == (93, 123), tab.shape
    out = os.path.join(HERE, "rez_scatpar_f64.bin")
    tab.astype("<f8").tofile(out)
    print(out, tab.shape)
    offs = parse_c_table(hdr, "scatParOffs")
    assert np.array_equal(offs[0], tab[0]), "the two tables share one s-grid"
    out = os.path.join(HERE, "rez_scatparoffs_f64.bin")
    offs.astype("<f8").tofile(out)
    print(out, offs.shape, "elements with ionic data:", [z for z in range(1, 93) if np.any(offs[z])])


if __name__ == "__main__":
    main()
