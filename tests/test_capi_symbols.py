"""The C-ABI library loads on a machine without a GPU and exports every
function include/qstem.h declares (no compute calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "qstem.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?(?:int|void|float|char\s*\*|const char\s*\*)\s*\**\s*(qs_\w+)\s*\(", text, re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "qs_create" in names and "qs_run_stem" in names and len(names) >= 10


def test_library_exports_declared_symbols():
    from qstem_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libqstem_hip.so not built: run python -m qstem_amd.build")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding covers every declared entry point
    assert set(declared_functions()) <= set(_lib.SIGNATURES)


def test_record_bins_host_logic():
    """qs_record_bins needs no device: default interval = slices -> bins {0: slices-2, 1: slices-1} (Q4)."""
    from qstem_amd import record_bins
    b, nb = record_bins(20)
    assert nb == 2
    assert list(b[:18]) == [-1] * 18 and b[18] == 0 and b[19] == 1
    b, nb = record_bins(8, outputInterval=2)
    assert nb == 5
    assert list(b) == [-1, 0, -1, 1, -1, 2, 3, 4]


def test_create_without_device_fails_cleanly():
    """No HIP device here: qs_create must return an error code with a message, not crash."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    from qstem_amd import Multislice, QstemError
    from cases import small_config
    with pytest.raises(QstemError):
        Multislice(small_config())


def test_record_bins_repeat_matches_oracle(oracle):
    """qs_record_bins_repeat (host logic, no device) against the oracle's qo_record_bins
    over slab counts, offsets, intervals and repeat counts, including slab offsets past
    the last bin and intervals that do not divide the slice count."""
    from qstem_amd import record_bins
    for slices in (1, 5, 6, 40):
        for cellDiv in (1, 3):
            for interval in (1, 2, 7, slices):
                for reps in (1, 2, 3):
                    for tsc in (0, slices, 2 * slices * cellDiv):
                        got, nb = record_bins(slices, cellDiv, interval, tsc, mulsRepeat1=reps)
                        want, nbo = oracle.record_bins(slices, cellDiv, interval, tsc, mulsRepeat1=reps)
                        assert nb == nbo and list(got) == list(want), (slices, cellDiv, interval, reps, tsc)
