"""ctypes wrapper of the FP64 CPU oracle (oracle/_build/libqs_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the reference that results are compared with.
The product path never uses it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libqs_oracle.so")
MAX_DET = 16
_F, _I, _D = ctypes.c_float, ctypes.c_int, ctypes.c_double


class QoParams(ctypes.Structure):
    _fields_ = [
        ("nx", _I), ("ny", _I), ("resolutionX", _F), ("resolutionY", _F), ("v0", _F),
        ("slices", _I), ("sliceThickness", _F), ("potNx", _I), ("potNy", _I),
        ("potSizeX", _F), ("potSizeY", _F), ("by", _F), ("compat_q1", _I), ("bandlimittrans", _I),
        ("Cs", _F), ("C5", _F), ("df0", _F), ("astigMag", _F), ("astigAngle", _F),
        ("a33", _F), ("a31", _F), ("a44", _F), ("a42", _F), ("a55", _F),
        ("a53", _F), ("a51", _F), ("a66", _F), ("a64", _F), ("a62", _F),
        ("phi33", _F), ("phi31", _F), ("phi44", _F), ("phi42", _F), ("phi55", _F),
        ("phi53", _F), ("phi51", _F), ("phi66", _F), ("phi64", _F), ("phi62", _F),
        ("alpha", _F), ("aAIS", _F), ("Cc", _F), ("dE_E", _F),
        ("ismoth", _I), ("gaussFlag", _I), ("gaussScale", _F), ("compat_q2", _I),
        ("ndet", _I), ("det_rin", _F * MAX_DET), ("det_rout", _F * MAX_DET),
        ("det_shiftx", _F * MAX_DET), ("det_shifty", _F * MAX_DET), ("electronScale", _F),
    ]


class LutGeom(ctypes.Structure):
    _fields_ = [("nx", _I), ("nz", _I), ("nzPerSlice", _I), ("Nr", _I), ("Nz_lut", _I),
                ("dkx", _D), ("dkz", _D), ("smax2", _D)]


class QoAtom(ctypes.Structure):
    _fields_ = [("x", _F), ("y", _F), ("z", _F), ("dw", _F), ("occ", _F), ("q", _F), ("Znum", _I)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


CPU32_PATH = os.path.join(HERE, "_build", "libqs_cpu32.so")
_cpu32 = None


def load_cpu32():
    """The FP32 CPU baseline (qs_cpu32.c): benchmark and test infrastructure."""
    global _cpu32
    if _cpu32 is None:
        if not os.path.exists(CPU32_PATH):
            build()
        _cpu32 = ctypes.CDLL(CPU32_PATH)
        _cpu32.qc_max_threads.restype = _I
    return _cpu32


def cpu32_multislice(cfg, trans, psi0, ipos, bins=None, nbins=0, want_exit=False, nthreads=0):
    """FP32 CPU multislice of the probes ipos through trans (complex [nslice][potNx][potNy]):
    returns (exit waves [n][nx][ny] complex64 or None, det [nbins][ndet][n] float32 or None)."""
    trans = np.ascontiguousarray(trans, dtype=np.complex64)
    psi0 = np.ascontiguousarray(psi0, dtype=np.complex64)
    ipos = np.ascontiguousarray(ipos, dtype=np.int32).reshape(-1, 2)
    n = ipos.shape[0]
    ndet = len(cfg.detectors)
    det = np.zeros((nbins, ndet, n), np.float32) if (bins is not None and ndet) else None
    b = np.ascontiguousarray(bins, dtype=np.int32) if bins is not None else None
    ex = np.empty((n, cfg.nx, cfg.ny), np.complex64) if want_exit else None
    p = params(cfg)
    rc = load_cpu32().qc_multislice(ctypes.byref(p), _p(trans.view(np.float32), _F), trans.shape[0],
                                    _p(psi0.view(np.float32), _F), _p(ipos, _I), n, _p(b, _I), nbins, _p(det, _F),
                                    _p(ex.view(np.float32) if ex is not None else None, _F), int(nthreads))
    if rc != 0:
        raise ValueError("qc_multislice: unsupported size")
    return ex, det


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.qo_wavelength.restype = _D
        _lib.qo_wavelength.argtypes = [_D]
        _lib.qo_sigma.restype = _D
        _lib.qo_sigma.argtypes = [_D]
        _lib.qo_seval.restype = _D
        _lib.qo_record_bins.restype = _I
    return _lib


def _p(a, t=ctypes.c_double):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(t))


def params(cfg):
    """QoParams from a qstem_amd.StemConfig-like object (same field names)."""
    p = QoParams()
    for k in ("nx", "ny", "slices", "potNx", "potNy", "bandlimittrans", "ismoth", "gaussFlag"):
        setattr(p, k, int(getattr(cfg, k)))
    for k in ("resolutionX", "resolutionY", "v0", "sliceThickness", "by", "Cs", "C5", "df0", "astigMag",
              "astigAngle", "alpha", "aAIS", "Cc", "dE_E", "gaussScale", "electronScale"):
        setattr(p, k, float(getattr(cfg, k)))
    p.potSizeX = float(np.float32(cfg.potNx) * np.float32(cfg.resolutionX))
    p.potSizeY = float(np.float32(cfg.potNy) * np.float32(cfg.resolutionY))
    for k, v in getattr(cfg, "aberrations", {}).items():
        setattr(p, k, float(v))
    p.compat_q1 = 1 if (cfg.compat & 1) else 0
    p.compat_q2 = 1 if (cfg.compat & 2) else 0
    p.ndet = len(cfg.detectors)
    for i, d in enumerate(cfg.detectors):
        p.det_rin[i], p.det_rout[i] = d.rInside, d.rOutside
        p.det_shiftx[i], p.det_shifty[i] = d.shiftX, d.shiftY
    return p


def fft2(a, sign):
    a = np.ascontiguousarray(a, dtype=np.complex128).copy()
    load().qo_fft2d(_p(a), a.shape[0], a.shape[1], sign)
    return a


def fft1(a, sign):
    a = np.ascontiguousarray(a, dtype=np.complex128).copy()
    load().qo_fft1d(_p(a), a.shape[0], sign)
    return a


def wavelength(kev):
    return load().qo_wavelength(kev)


def probe(cfg, at=None):
    """probe() at the window centre, or centred at `at` = (x, y) Angstrom (qo_probe_at)."""
    out = np.empty((cfg.nx, cfg.ny), dtype=np.complex128)
    p = params(cfg)
    if at is None:
        load().qo_probe(ctypes.byref(p), _p(out))
    else:
        load().qo_probe_at(ctypes.byref(p), ctypes.c_double(at[0]), ctypes.c_double(at[1]), _p(out))
    return out


def k_tables(cfg):
    kx2 = np.empty(cfg.nx, np.float32)
    ky2 = np.empty(cfg.ny, np.float32)
    km = ctypes.c_float()
    p = params(cfg)
    load().qo_k_tables(ctypes.byref(p), _p(kx2, _F), _p(ky2, _F), ctypes.byref(km))
    return kx2, ky2, km.value


def make_trans(cfg, V):
    V = np.ascontiguousarray(V, dtype=np.float32)
    out = np.empty(V.shape, dtype=np.complex128)
    p = params(cfg)
    load().qo_make_trans(ctypes.byref(p), _p(V, _F), V.shape[0], _p(out))
    return out


def record_bins(slices, cellDiv=1, outputInterval=None, totalSliceCount=0, mulsRepeat1=1):
    outputInterval = slices if outputInterval is None else outputInterval
    out = np.empty(slices * mulsRepeat1, dtype=np.int32)
    nb = load().qo_record_bins(slices, cellDiv, outputInterval, totalSliceCount, mulsRepeat1, _p(out, _I))
    return out, nb


def multislice(cfg, trans, ipos, psi0, bins=None, nbins=0, want_diff=False):
    """Returns (exit waves [n][nx][ny], det [nbins][ndet][n] or None, diff or None)."""
    trans = np.ascontiguousarray(trans, dtype=np.complex128)
    ipos = np.ascontiguousarray(ipos, dtype=np.int32).reshape(-1, 2)
    n = ipos.shape[0]
    psi = np.ascontiguousarray(np.broadcast_to(psi0, (n, cfg.nx, cfg.ny)), dtype=np.complex128).copy()
    det = np.zeros((nbins, len(cfg.detectors), n)) if bins is not None and nbins > 0 else None
    diff = np.zeros((n, cfg.nx, cfg.ny)) if want_diff else None
    b = None if bins is None else np.ascontiguousarray(bins, dtype=np.int32)
    p = params(cfg)
    load().qo_multislice(ctypes.byref(p), _p(trans), trans.shape[0], _p(ipos, _I), n, _p(psi),
                         _p(b, _I), nbins, _p(det), _p(diff))
    return psi, det, diff


def lut_geometry(atomRadius, resX, sliceThickness):
    g = LutGeom()
    load().qo_lut_geometry(_F(atomRadius), _F(resX), _F(sliceThickness), ctypes.byref(g))
    return g


def rez_table():
    path = os.path.join(os.path.dirname(HERE), "qstem_amd", "data", "rez_scatpar_f64.bin")
    return np.fromfile(path, dtype="<f8").reshape(93, 123)


def rez_offs_table():
    """Ion-offset table (scatParOffs): row 0 the s-grid, row Z df_e/dq."""
    path = os.path.join(os.path.dirname(HERE), "qstem_amd", "data", "rez_scatparoffs_f64.bin")
    return np.fromfile(path, dtype="<f8").reshape(93, 123)


def _tapered(tab, g):
    sgrid = np.ascontiguousarray(tab[0].copy())
    rows = np.ascontiguousarray(tab[1:].copy())
    load().qo_taper_scatpar(_p(sgrid), _p(rows), rows.shape[0], _D(g.smax2))
    return sgrid, rows


def lut_norm_fix(g, sliceThickness):
    """Scale of the LUT under the normalisation switch (slice sum -> f_e(0))."""
    f = load().qo_lut_norm_fix
    f.restype = _D
    return float(f(ctypes.byref(g), _F(sliceThickness)))


def atom_luts(g, Zs, Bs, sliceThickness, norm_fix=False, clamp=True):
    """LUT per (Z, B) pair; applies the s-grid taper of the reference first."""
    sgrid, rows = _tapered(rez_table(), g)
    scale = lut_norm_fix(g, sliceThickness) if norm_fix else 1.0
    luts = []
    for Z, B in zip(Zs, Bs):
        frow = np.ascontiguousarray(rows[Z - 1])
        lut = np.zeros((g.Nz_lut, g.Nr), np.float32)
        load().qo_atom_lut2(ctypes.byref(g), _p(sgrid), _p(frow), _p(sgrid), _p(frow), _D(g.dkx), _D(B),
                            _F(sliceThickness), _D(scale), int(clamp), _p(lut, _F))
        luts.append(lut)
    return luts


def ion_luts(g, Zs, Bs, sliceThickness, compat_q5=True, norm_fix=False, clamp=True, resY=None, atomRadius=None):
    """Ion-offset LUT per (Z, B) pair (getAtomPotentialOffset3D): the offset
    table's spline; its q_y sum reads the neutral table's values with that
    spline when compat_q5 (the reference), the offset table's otherwise.  The
    q_y step is dky of resY (= dkx for square pixels)."""
    sg_o, rows_o = _tapered(rez_offs_table(), g)
    sg_n, rows_n = _tapered(rez_table(), g)
    dky = g.dkx
    if resY is not None:
        ny = 2 * 2 * int(np.ceil(np.float32(atomRadius) / np.float32(resY)))
        dky = 0.5 * 2 / float(np.float32(ny * np.float32(resY)))
    scale = lut_norm_fix(g, sliceThickness) if norm_fix else 1.0
    luts = []
    for Z, B in zip(Zs, Bs):
        fo = np.ascontiguousarray(rows_o[Z - 1])
        sgy, fy = (sg_n, np.ascontiguousarray(rows_n[Z - 1])) if compat_q5 else (sg_o, fo)
        lut = np.zeros((g.Nz_lut, g.Nr), np.float32)
        load().qo_atom_lut2(ctypes.byref(g), _p(sg_o), _p(fo), _p(sgy), _p(fy), _D(dky), _D(B), _F(sliceThickness),
                            _D(scale), int(clamp), _p(lut, _F))
        luts.append(lut)
    return luts


def make_slices(cfg, g, luts, kinds, atoms, potOffsetX, potOffsetY, czOffset=0.0, centerSlices=0,
                atomRadius=5.0, periodic=0, offs_luts=None):
    """atoms: structured array or list of (x, y, z, dw, occ, q, Znum), z-sorted.
    offs_luts[k] (None entries allowed): ion-offset LUT of kind k for charged atoms."""
    arr = (QoAtom * len(atoms))()
    for i, a in enumerate(atoms):
        arr[i] = QoAtom(*[float(v) for v in a[:6]], int(a[6]))
    lut_ptrs = (ctypes.POINTER(_F) * len(luts))(*[_p(l, _F) for l in luts])
    offs_ptrs = None
    if offs_luts is not None:
        offs_ptrs = (ctypes.POINTER(_F) * len(offs_luts))(
            *[_p(l, _F) if l is not None else ctypes.POINTER(_F)() for l in offs_luts])
    kinds = np.ascontiguousarray(kinds, dtype=np.int32)
    V = np.zeros((cfg.slices, cfg.potNx, cfg.potNy), np.float32)
    p = params(cfg)
    load().qo_make_slices_q(ctypes.byref(p), ctypes.byref(g), lut_ptrs, offs_ptrs, _p(kinds, _I), arr, len(atoms),
                            _F(potOffsetX), _F(potOffsetY), _F(czOffset), centerSlices, _F(atomRadius), periodic,
                            _p(V, _F))
    return V


def probe_d(cfg, df0, Cs):
    out = np.empty((cfg.nx, cfg.ny), dtype=np.complex128)
    p = params(cfg)
    load().qo_probe_d(ctypes.byref(p), _D(df0), _D(Cs), _p(out))
    return out


def make_trans_d(cfg, V):
    V = np.ascontiguousarray(V, dtype=np.float64)
    out = np.empty(V.shape, dtype=np.complex128)
    p = params(cfg)
    load().qo_make_trans_d(ctypes.byref(p), _p(V), V.shape[0], _p(out))
    return out


def make_slices_d(cfg, g, luts, kinds, xyz, Z, potOffsetX, potOffsetY, czOffset=0.0, centerSlices=0,
                  atomRadius=5.0, dr=None):
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    Z = np.ascontiguousarray(Z, dtype=np.int32)
    kinds = np.ascontiguousarray(kinds, dtype=np.int32)
    lut_ptrs = (ctypes.POINTER(_F) * len(luts))(*[_p(l, _F) for l in luts])
    V = np.zeros((cfg.slices, cfg.potNx, cfg.potNy), np.float64)
    dr = float(np.float32(cfg.resolutionX) / 2) if dr is None else dr
    p = params(cfg)
    load().qo_make_slices_d(ctypes.byref(p), ctypes.byref(g), lut_ptrs, _p(kinds, _I), _p(xyz), _p(Z, _I), len(Z),
                            _D(potOffsetX), _D(potOffsetY), _D(czOffset), centerSlices, _D(atomRadius), _D(dr),
                            _p(V))
    return V


def adjoint(cfg, V, ipos, psi0, target):
    """FP64 adjoint (qo_adjoint): (loss, dL/dV [nslice][potNx][potNy], g_psi0 [nx][ny]) for
    L = sum (|DFT(psi_exit)/N^2|^2 - target)^2 over the probes ipos."""
    V = np.ascontiguousarray(V, dtype=np.float64)
    ipos = np.ascontiguousarray(ipos, dtype=np.int32).reshape(-1, 2)
    psi0 = np.ascontiguousarray(psi0, dtype=np.complex128)
    target = np.ascontiguousarray(target, dtype=np.float64)
    loss = np.zeros(1)
    gV = np.zeros(V.shape)
    gpsi = np.zeros((cfg.nx, cfg.ny), np.complex128)
    p = params(cfg)
    load().qo_adjoint(ctypes.byref(p), _p(V), V.shape[0], _p(ipos, _I), ipos.shape[0], _p(psi0), _p(target),
                      _p(loss), _p(gV), _p(gpsi))
    return float(loss[0]), gV, gpsi


def probe_param_gradients(cfg, gpsi0, h=1e-5):
    """dL/d(df0), dL/d(Cs) from the adjoint's g_psi0: 2 Re sum conj(g) dpsi0/dtheta, with
    dpsi0/dtheta by FP64 central differences of the probe (a smooth function of both)."""
    out = []
    for name in ("df0", "Cs"):
        base = float(getattr(cfg, name))
        step = h * max(1.0, abs(base))
        args = {"df0": float(cfg.df0), "Cs": float(cfg.Cs)}
        args[name] = base + step
        pp = probe_d(cfg, **args)
        args[name] = base - step
        pm = probe_d(cfg, **args)
        out.append(float(2 * np.real(np.vdot(gpsi0, (pp - pm) / (2 * step)))))
    return out


def diffraction_loss(cfg, trans, ipos, psi0, target):
    """L = sum_p sum_k (|FFT(psi_exit)/(nx ny)|^2 - target)^2 with the oracle's FP64 multislice."""
    ex, _, _ = multislice(cfg, trans, ipos, psi0)
    n2 = cfg.nx * cfg.ny
    I = np.abs(np.fft.fft2(ex, axes=(1, 2)) / n2) ** 2
    return float(np.sum((I - target) ** 2)), I
