"""qstem3's C++ model builder (qstem_amd/csrc/host/model.cpp: .cfg reader, supercell,
Cube cut, Einstein frozen phonons), compiled here with g++ into a small harness
(tests/host_model_harness.cpp), against the Python builder (qstem_amd/specimen.py)
and the Einstein statistics of phononDisplacement (fileio_fftw3.cpp:495-498, 590-594)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "qstem_amd", "csrc", "host")
CFG = os.path.join(ROOT, "tests", "golden", "SrTiO3.cfg")


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("harness") / "host_model_harness")
    subprocess.run([gxx, "-O1", "-std=c++17", "-I", HOST, "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "host_model_harness.cpp"), os.path.join(HOST, "model.cpp"),
                    os.path.join(HOST, "qsc.cpp"), "-o", exe], check=True)
    return exe


def _run(exe, n, tds, T, seed, cube=None):
    args = [exe, CFG, str(n[0]), str(n[1]), str(n[2]), str(int(tds)), str(T), str(seed)]
    if cube is not None:
        args += [str(c) for c in cube]
    out = subprocess.run(args, check=True, capture_output=True, text=True).stdout.splitlines()
    ext = [float(v) for v in out[-1].split()[1:]]
    return np.array([[float(v) for v in line.split()] for line in out[:-1]]), ext


def _lex(a):
    return a[np.lexsort((a[:, 2], a[:, 1], a[:, 0], a[:, 6]))]


def test_cpp_supercell_matches_python(harness):
    from qstem_amd.specimen import build_supercell, read_cfg
    got, ext = _run(harness, (3, 2, 4), False, 300, 1)
    want, wext = build_supercell(read_cfg(CFG), 3, 2, 4)
    assert got.shape == want.shape
    assert np.all(np.diff(got[:, 2]) >= 0)  # z-sorted
    g, w = _lex(got), _lex(want)
    assert np.abs(g[:, :3] - w[:, :3]).max() < 1e-5
    assert np.array_equal(g[:, 6], w[:, 6]) and np.allclose(g[:, 3:6], w[:, 3:6], atol=1e-6)
    assert np.allclose(ext, wext, rtol=1e-12)
    # a Cube equal to the supercell keeps every atom; a smaller one keeps those inside it
    cut, cext = _run(harness, (3, 2, 4), False, 300, 1, cube=(wext[0], wext[1], wext[2]))
    assert len(cut) == len(got) and np.allclose(cext, wext)
    small, _ = _run(harness, (3, 2, 4), False, 300, 1, cube=(5.0, 4.0, 6.0))
    keep = (want[:, 0] < 5.0) & (want[:, 1] < 4.0) & (want[:, 2] < 6.0)
    assert len(small) == int(keep.sum()) > 0


@pytest.mark.parametrize("T", [300.0, 77.0])
def test_cpp_einstein_displacements(harness, T):
    """Each atom's displacement from its lattice site, per element and axis: mean 0 and
    standard deviation sqrt(T/300) sqrt(B/(8 pi^2))/sqrt(3), to the sampling error of
    ~650 samples per element (bound 10 %); two seeds give different configurations,
    one seed the same."""
    from scipy.spatial import cKDTree
    from qstem_amd.specimen import einstein_sigma
    n = (6, 6, 6)
    ref, _ = _run(harness, n, False, T, 1)
    a1, _ = _run(harness, n, True, T, 11)
    a1b, _ = _run(harness, n, True, T, 11)
    a2, _ = _run(harness, n, True, T, 12)
    assert np.array_equal(a1, a1b) and not np.array_equal(a1, a2)
    for Z in np.unique(ref[:, 6]):
        sites = ref[ref[:, 6] == Z, :3]
        moved = a1[a1[:, 6] == Z]
        dist, idx = cKDTree(sites).query(moved[:, :3])
        assert len(np.unique(idx)) == len(sites)  # one atom per site
        u = moved[:, :3] - sites[idx]
        sig = float(einstein_sigma(moved[0, 3], T))
        assert abs(u.mean()) < 4 * sig / np.sqrt(u.size)
        assert abs(u.std() / sig - 1) < 0.1, (Z, u.std(), sig)
