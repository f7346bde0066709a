"""qstem_amd: MI355X-native (gfx950) multislice hot path of QSTEM.

The compute path is libqstem_hip.so (hand-written HIP kernels behind the C ABI
in include/qstem.h); this package is its Python mirror.
"""
from .multislice import (  # noqa: F401
    Detector,
    Multislice,
    QstemError,
    StemConfig,
    device_memory,
    MILLISEC_PICOAMP,
    record_bins,
    run_groups,
    scan_offsets,
)
