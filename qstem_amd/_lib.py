"""ctypes binding of libqstem_hip.so (include/qstem.h).

The library is the product path; there is no CPU fallback.  Loading fails
with an ImportError-style message when the in-tree build is missing.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libqstem_hip.so")

QS_OK = 0
QS_ERR_ARG = 1
QS_ERR_UNSUPPORTED = 2
QS_ERR_HIP = 3
QS_ERR_STATE = 4
QS_MAX_DET = 16
QS_COMPAT_Q1 = 1
QS_COMPAT_Q2 = 2

_F = ctypes.c_float
_I = ctypes.c_int


class QsParams(ctypes.Structure):
    """Mirror of struct qs_params (include/qstem.h)."""
    _fields_ = [
        ("nx", _I), ("ny", _I),
        ("resolutionX", _F), ("resolutionY", _F),
        ("v0", _F),
        ("slices", _I),
        ("sliceThickness", _F),
        ("potNx", _I), ("potNy", _I),
        ("potSizeX", _F), ("potSizeY", _F),
        ("by", _F),
        ("bandlimittrans", _I),
        ("Cs", _F), ("C5", _F), ("df0", _F), ("astigMag", _F), ("astigAngle", _F),
        ("a33", _F), ("a31", _F), ("a44", _F), ("a42", _F), ("a55", _F),
        ("a53", _F), ("a51", _F), ("a66", _F), ("a64", _F), ("a62", _F),
        ("phi33", _F), ("phi31", _F), ("phi44", _F), ("phi42", _F), ("phi55", _F),
        ("phi53", _F), ("phi51", _F), ("phi66", _F), ("phi64", _F), ("phi62", _F),
        ("alpha", _F),
        ("aAIS", _F),
        ("Cc", _F), ("dE_E", _F),
        ("ismoth", _I), ("gaussFlag", _I),
        ("gaussScale", _F),
        ("ndet", _I),
        ("det_rin", _F * QS_MAX_DET), ("det_rout", _F * QS_MAX_DET),
        ("det_shiftx", _F * QS_MAX_DET), ("det_shifty", _F * QS_MAX_DET),
        ("electronScale", _F),
        ("compat", _I),
    ]


class QsAtom(ctypes.Structure):
    _fields_ = [("x", _F), ("y", _F), ("z", _F), ("dw", _F), ("occ", _F), ("q", _F), ("Znum", _I)]


class QsPotParams(ctypes.Structure):
    _fields_ = [("atomRadius", _F), ("potOffsetX", _F), ("potOffsetY", _F), ("czOffset", _F),
                ("centerSlices", _I), ("periodicXY", _I), ("tds", _I), ("compat_q6", _I), ("lut_dr_fix", _I),
                ("compat_q5", _I), ("lut_norm_fix", _I)]


class QsGroup(ctypes.Structure):
    """struct qs_group of include/qstem.h."""
    _fields_ = [("ctx", ctypes.c_void_p), ("ipos", ctypes.POINTER(ctypes.c_int)), ("config", ctypes.POINTER(ctypes.c_int)),
                ("nitem", ctypes.c_int), ("det", ctypes.POINTER(ctypes.c_float)),
                ("int_intensity", ctypes.POINTER(ctypes.c_float))]


# Symbols declared in include/qstem.h, with their ctypes signatures.
_P = ctypes.c_void_p
_FP = ctypes.POINTER(ctypes.c_float)
_IP = ctypes.POINTER(ctypes.c_int)
SIGNATURES = {
    "qs_create": (_I, [ctypes.POINTER(QsParams), ctypes.POINTER(_P)]),
    "qs_destroy": (None, [_P]),
    "qs_last_error": (ctypes.c_char_p, [_P]),
    "qs_set_potential": (_I, [_P, _FP, _I]),
    "qs_set_trans": (_I, [_P, _FP, _I]),
    "qs_get_trans": (_I, [_P, _FP]),
    "qs_make_probe": (_I, [_P, _FP]),
    "qs_make_probe_at": (_I, [_P, ctypes.c_double, ctypes.c_double, _FP]),
    "qs_set_optics": (_I, [_P, ctypes.POINTER(QsParams)]),
    "qs_set_scattering_table": (_I, [_P, ctypes.POINTER(ctypes.c_double)]),
    "qs_set_ion_scattering_table": (_I, [_P, ctypes.POINTER(ctypes.c_double)]),
    "qs_build_potential": (_I, [_P, ctypes.POINTER(QsAtom), _I, ctypes.POINTER(QsPotParams), _FP]),
    "qs_record_bins": (_I, [_I, _I, _I, _I, _IP]),
    "qs_record_bins_repeat": (_I, [_I, _I, _I, _I, _I, _IP]),
    "qs_set_incident_wave": (_I, [_P, _FP]),
    "qs_device_memory": (_I, [ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong), _I]),
    "qs_run_stem": (_I, [_P, _IP, _I, _I, _IP, _I, _I, _FP, _FP, _FP, _FP]),
    "qs_run_stem_cfg": (_I, [_P, _IP, _IP, _I, _I, _IP, _I, _I, _FP, _FP, _FP, _FP]),
    "qs_run_slices": (_I, [_P, _IP, _IP, _I, _I, _I, _IP, _I, _I, _FP, _FP, _FP, _FP]),
    "qs_run_stem_noalloc": (_I, [_P, _IP, _I, _IP, _I, _I, _FP, _FP]),
    "qs_set_configs": (_I, [_P, _I]),
    "qs_select_config": (_I, [_P, _I]),
    "qs_share_trans": (_I, [_P, _P]),
    "qs_run_groups": (_I, [ctypes.POINTER(QsGroup), _I, _IP, _I, _I]),
    "qs_adjoint": (_I, [_P, _IP, _I, _FP, _I, _FP, _FP, _FP]),
    "qs_forward_save": (_I, [_P, _IP, _I, _FP, _I, _FP]),
    "qs_adjoint_saved": (_I, [_P, _FP, _FP]),
    "qs_atom_gradients": (_I, [_P, ctypes.POINTER(QsAtom), _I, ctypes.POINTER(QsPotParams), _FP]),
    "qs_set_batch": (_I, [_P, _I]),
    "qs_get_batch": (_I, [_P]),
    "qs_set_streams": (_I, [_P, _I]),
    "qs_get_streams": (_I, [_P]),
    "qs_last_ckpt_interval": (_I, [_P]),
    "qs_set_graph": (_I, [_P, _I]),
    "qs_set_tuning": (_I, [_P, _I, _I]),
    "qs_get_tuning": (_I, [_P, _IP, _IP, _IP]),
    "qs_autotune": (_I, [_P, _IP, _I, _IP, _IP, _IP, _FP]),
    "qs_last_run_ms": (_F, [_P]),
    "qs_set_profiling": (_I, [_P, _I]),
    "qs_kernel_times": (_I, [_P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                             ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_long)]),
}

_lib = None


def load():
    """Load the in-tree library once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("qstem_amd: %s is missing; build it with `python -m qstem_amd.build`" % LIB_PATH)
    # PyTorch-ROCm ships its own libamdhip64.so.7 / libhsa-runtime64.so.1.  Two
    # HSA runtimes in one process cannot share the device, so when torch is
    # installed it is loaded first and this library binds to the same runtime
    # (the dynamic loader matches the already-loaded SONAMEs).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
