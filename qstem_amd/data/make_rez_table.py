"""Convert the Rez et al. electron scattering-factor table shipped with QSTEM
(python/SF_rez.txt: 93 rows x 123 columns, CSV) into the raw little-endian
float64 file ``rez_scatpar_f64.bin`` that the library and the oracle load.

Row 0 is the s-grid (1/Angstrom, s = q/2 convention of the reference,
stem3/scatfactsRez.h:22-25), rows 1..92 are f_e(s) for Z = 1..92.  The
table is published physical data; the file is regenerated, not edited.

Usage: python make_rez_table.py [path/to/SF_rez.txt]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/python/SF_rez.txt"
    tab = np.loadtxt(src, delimiter=",", dtype=np.float64)
    assert tab.shape == (93, 123), tab.shape
    out = os.path.join(HERE, "rez_scatpar_f64.bin")
    tab.astype("<f8").tofile(out)
    print(out, tab.shape)


if __name__ == "__main__":
    main()
