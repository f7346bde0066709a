"""Python mirror of QSTEM's STEM path over the C ABI.

`Multislice` plays the role of the reference's MULS + WAVEFUNC pair
(libs/data_containers.h:29-308): it owns the transmission functions, the
probe and the wave state of a batch of scan positions on the GPU.  The
method names follow the reference routines they replace:

    set_potential  -> initSTEMSlices   (stem3/stemlib.cpp:2199-2357)
    make_probe     -> probe            (stem3/stemlib.cpp:1952-2191)
    run_stem       -> runMulsSTEM over the scan loop of doSTEM
                      (stem3/stemlib.cpp:2375-2534, stem3/stem3.cpp:2820-2946)

Errors raise QstemError with the library's message instead of the
reference's printf+exit(0).
"""
import ctypes
import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _lib

MILLISEC_PICOAMP = 1e-3 * (1e-12 / 1.6021773e-19)  # stemtypes_fftw3.h:47-49


class QstemError(RuntimeError):
    pass


@dataclass
class Detector:
    """One annular detector line `detector: rin rout name shiftX shiftY` (stem3.cpp:1323-1347)."""
    rInside: float
    rOutside: float
    name: str = "detector"
    shiftX: float = 0.0
    shiftY: float = 0.0


@dataclass
class StemConfig:
    """The MULS fields the STEM path reads, in MULS units (Angstrom, mrad, kV, rad)."""
    nx: int
    ny: int
    resolutionX: float
    resolutionY: float
    v0: float
    slices: int
    sliceThickness: float
    potNx: int
    potNy: int
    by: float = 0.0
    bandlimittrans: int = 1
    Cs: float = 0.0
    C5: float = 0.0
    df0: float = 0.0
    astigMag: float = 0.0
    astigAngle: float = 0.0
    aberrations: dict = field(default_factory=dict)  # a33.., phi33.. in MULS units
    alpha: float = 15.0
    aAIS: float = 0.0
    Cc: float = 0.0
    dE_E: float = 0.0
    ismoth: int = 0
    gaussFlag: int = 0
    gaussScale: float = 0.05
    detectors: List[Detector] = field(default_factory=list)
    electronScale: float = MILLISEC_PICOAMP  # 1 pA * 1 ms, the reference default
    compat: int = 0

    def to_c(self) -> _lib.QsParams:
        p = _lib.QsParams()
        for name in ("nx", "ny", "slices", "potNx", "potNy", "bandlimittrans", "ismoth", "gaussFlag", "compat"):
            setattr(p, name, int(getattr(self, name)))
        for name in ("resolutionX", "resolutionY", "v0", "sliceThickness", "by", "Cs", "C5", "df0", "astigMag",
                     "astigAngle", "alpha", "aAIS", "Cc", "dE_E", "gaussScale", "electronScale"):
            setattr(p, name, float(getattr(self, name)))
        p.potSizeX = float(np.float32(self.potNx) * np.float32(self.resolutionX))
        p.potSizeY = float(np.float32(self.potNy) * np.float32(self.resolutionY))
        for k, v in self.aberrations.items():
            setattr(p, k, float(v))
        if len(self.detectors) > _lib.QS_MAX_DET:
            raise QstemError("at most %d detectors" % _lib.QS_MAX_DET)
        p.ndet = len(self.detectors)
        for i, d in enumerate(self.detectors):
            p.det_rin[i] = d.rInside
            p.det_rout[i] = d.rOutside
            p.det_shiftx[i] = d.shiftX
            p.det_shifty[i] = d.shiftY
        return p


def rez_table() -> np.ndarray:
    """Rez et al. electron scattering factors shipped as data (row 0 = s grid, row Z = f_e(s))."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "rez_scatpar_f64.bin")
    return np.fromfile(path, dtype="<f8").reshape(93, 123)


def rez_ion_table() -> np.ndarray:
    """Ion-offset scattering factors df_e/dq shipped as data (row 0 = s grid, row Z)."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "rez_scatparoffs_f64.bin")
    return np.fromfile(path, dtype="<f8").reshape(93, 123)


def _fptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _iptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _is_device_tensor(x) -> bool:
    """A torch tensor in GPU memory (the library takes its storage pointer)."""
    return type(x).__module__.startswith("torch") and getattr(x, "is_cuda", False)


def _device_f32(t):
    """Contiguous float32 view of a device tensor and its pointer, after the work
    queued on the tensor's current stream has finished (the library's own stream
    does not wait for torch's)."""
    import torch
    t = t.detach().to(torch.float32).contiguous()
    torch.cuda.current_stream(t.device).synchronize()
    return t, ctypes.cast(t.data_ptr(), ctypes.POINTER(ctypes.c_float))


def _atom_array(atoms):
    """(n, 7) x, y, z, dw, occ, q, Z -> ctypes array of struct qs_atom (filled through numpy)."""
    atoms = np.asarray(atoms, dtype=np.float64).reshape(-1, 7)
    arr = (_lib.QsAtom * max(len(atoms), 1))()
    dt = np.dtype({"names": ["x", "y", "z", "dw", "occ", "q", "Znum"],
                   "formats": ["<f4"] * 6 + ["<i4"],
                   "offsets": [getattr(_lib.QsAtom, f).offset for f in ("x", "y", "z", "dw", "occ", "q", "Znum")],
                   "itemsize": ctypes.sizeof(_lib.QsAtom)})
    view = np.frombuffer(arr, dtype=dt, count=len(atoms))
    for j, f in enumerate(("x", "y", "z", "dw", "occ", "q")):
        view[f] = atoms[:, j]
    view["Znum"] = atoms[:, 6].astype(np.int32)
    return arr


def record_bins(slices: int, cellDiv: int = 1, outputInterval: int = 0, totalSliceCount: int = 0,
                mulsRepeat1: int = 1):
    """Bin written by each slice step as its final value (collectIntensity,
    stemlib.cpp:2588-2607); with mulsRepeat1 > 1 ("sequence: r p") the slab's
    slices*mulsRepeat1 steps, recorded under runMulsSTEM's slice index
    totalSliceCount + s*(1 + r) (stemlib.cpp:2429)."""
    lib = _lib.load()
    out = np.empty(slices * mulsRepeat1, dtype=np.int32)
    nb = lib.qs_record_bins_repeat(slices, cellDiv, outputInterval if outputInterval > 0 else slices,
                                   totalSliceCount, mulsRepeat1, _iptr(out))
    if nb < 0:
        raise QstemError("bad bin arguments")
    return out, nb


def scan_offsets(scanXStart, scanXStop, scanXN, scanYStart, scanYStop, scanYN, resX, resY, potNx, potNy, nx, ny):
    """Integer window offsets of all scan positions, in the driver's order i = ix*scanYN + iy
    (stem3.cpp:2825-2871, float arithmetic as the reference)."""
    f = np.float32
    ix = np.repeat(np.arange(scanXN), scanYN).astype(np.float32)
    iy = np.tile(np.arange(scanYN), scanXN).astype(np.float32)
    px = (ix * (f(scanXStop) - f(scanXStart)) / (f(scanXN) * f(resX))).astype(np.int32)
    py = (iy * (f(scanYStop) - f(scanYStart)) / (f(scanYN) * f(resY))).astype(np.int32)
    px = np.minimum(px, potNx - nx)
    py = np.minimum(py, potNy - ny)
    return np.stack([px, py], axis=1).astype(np.int32)


class Multislice:
    """GPU multislice context (one HIP stream)."""

    def __init__(self, cfg: StemConfig):
        self.cfg = cfg
        self._lib = _lib.load()
        self._p = cfg.to_c()
        h = ctypes.c_void_p()
        rc = self._lib.qs_create(ctypes.byref(self._p), ctypes.byref(h))
        if rc != _lib.QS_OK:
            raise QstemError("qs_create failed (%d): %s" % (rc, self._lib.qs_last_error(None).decode()))
        self._h = h
        self._nslice = 0

    def _check(self, rc, what):
        if rc != _lib.QS_OK:
            raise QstemError("%s failed (%d): %s" % (what, rc, self._lib.qs_last_error(self._h).decode()))

    def close(self):
        if getattr(self, "_h", None):
            self._lib.qs_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- potential / transmission ------------------------------------------------
    def set_potential(self, V):
        """Projected potential [nslice][potNx][potNy]: a numpy array, or a torch tensor in
        GPU memory, which the library reads in place (no copy through the host)."""
        if _is_device_tensor(V):
            V, ptr = _device_f32(V)
        else:
            V = np.ascontiguousarray(V, dtype=np.float32)
            ptr = _fptr(V)
        assert V.ndim == 3 and tuple(V.shape[1:]) == (self.cfg.potNx, self.cfg.potNy), V.shape
        self._check(self._lib.qs_set_potential(self._h, ptr, V.shape[0]), "qs_set_potential")
        self._nslice = V.shape[0]

    def set_trans(self, t: np.ndarray):
        t = np.ascontiguousarray(t, dtype=np.complex64)
        assert t.ndim == 3 and t.shape[1:] == (self.cfg.potNx, self.cfg.potNy), t.shape
        self._check(self._lib.qs_set_trans(self._h, _fptr(t.view(np.float32)), t.shape[0]), "qs_set_trans")
        self._nslice = t.shape[0]

    def get_trans(self) -> np.ndarray:
        out = np.empty((self._nslice, self.cfg.potNx, self.cfg.potNy), dtype=np.complex64)
        self._check(self._lib.qs_get_trans(self._h, _fptr(out.view(np.float32))), "qs_get_trans")
        return out

    def set_scattering_table(self, table: Optional[np.ndarray] = None):
        """Rez et al. table (93 x 123 doubles); defaults to qstem_amd/data/rez_scatpar_f64.bin."""
        if table is None:
            table = rez_table()
        table = np.ascontiguousarray(table, dtype=np.float64)
        assert table.shape == (93, 123), table.shape
        self._check(self._lib.qs_set_scattering_table(self._h, table.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
                    "qs_set_scattering_table")
        self._have_table = True

    def set_ion_scattering_table(self, table: Optional[np.ndarray] = None):
        """Ion-offset table (93 x 123 doubles); defaults to qstem_amd/data/rez_scatparoffs_f64.bin."""
        if table is None:
            table = rez_ion_table()
        table = np.ascontiguousarray(table, dtype=np.float64)
        assert table.shape == (93, 123), table.shape
        self._check(self._lib.qs_set_ion_scattering_table(self._h,
                                                          table.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
                    "qs_set_ion_scattering_table")
        self._have_ion_table = True

    @staticmethod
    def _pot_params(atomRadius=5.0, potOffsetX=0.0, potOffsetY=0.0, czOffset=0.0, centerSlices=0, periodicXY=0,
                    tds=0, compat_q6=0, lut_dr_fix=0, compat_q5=1, lut_norm_fix=0):
        return _lib.QsPotParams(atomRadius, potOffsetX, potOffsetY, czOffset, int(centerSlices), int(periodicXY),
                                int(tds), int(compat_q6), int(lut_dr_fix), int(compat_q5), int(lut_norm_fix))

    def build_potential(self, atoms, return_V=False, **pot):
        """make3DSlices + initSTEMSlices on the GPU.  atoms: (n, 7) array of
        x, y, z, dw, occ, q, Z (Cartesian Angstrom); charged atoms add q times their
        element's ion-offset LUT.  pot: the qs_pot_params fields (atomRadius=5.0,
        potOffsetX, potOffsetY, czOffset, centerSlices, periodicXY, tds, compat_q6,
        lut_dr_fix, compat_q5=1, lut_norm_fix)."""
        if not getattr(self, "_have_table", False):
            self.set_scattering_table()
        if not getattr(self, "_have_ion_table", False) and np.any(np.asarray(atoms, dtype=np.float64).reshape(-1, 7)[:, 5]):
            self.set_ion_scattering_table()
        arr = _atom_array(atoms)
        pp = self._pot_params(**pot)
        V = np.empty((self.cfg.slices, self.cfg.potNx, self.cfg.potNy), np.float32) if return_V else None
        self._check(self._lib.qs_build_potential(self._h, arr, len(atoms), ctypes.byref(pp), _fptr(V)),
                    "qs_build_potential")
        self._nslice = self.cfg.slices
        self._build_pot = dict(pot)
        return V

    # -- probe ---------------------------------------------------------------------
    def make_probe(self, at: Optional[Tuple[float, float]] = None) -> np.ndarray:
        """The probe at the window centre, or centred at `at` = (x, y) Angstrom from
        the window's corner (qs_make_probe_at)."""
        out = np.empty((self.cfg.nx, self.cfg.ny), dtype=np.complex64)
        if at is None:
            self._check(self._lib.qs_make_probe(self._h, _fptr(out.view(np.float32))), "qs_make_probe")
        else:
            self._check(self._lib.qs_make_probe_at(self._h, float(at[0]), float(at[1]), _fptr(out.view(np.float32))),
                        "qs_make_probe_at")
        return out

    def set_incident_wave(self, psi: np.ndarray):
        """Incident wave [nx][ny] in place of the probe (plane wave: np.ones)."""
        psi = np.ascontiguousarray(psi, dtype=np.complex64)
        assert psi.shape == (self.cfg.nx, self.cfg.ny), psi.shape
        self._check(self._lib.qs_set_incident_wave(self._h, _fptr(psi.view(np.float32))), "qs_set_incident_wave")

    # -- adjoint ----------------------------------------------------------------------
    def adjoint(self, ipos: np.ndarray, target: np.ndarray, ckpt_every: int = 0, want_grad_V: bool = True,
                want_grad_probe: bool = True) -> dict:
        """Loss and gradients of L = sum (I - target)^2 over all slices (qs_adjoint)."""
        ipos = np.ascontiguousarray(ipos, dtype=np.int32).reshape(-1, 2)
        target = np.ascontiguousarray(target, dtype=np.float32)
        assert target.shape == (ipos.shape[0], self.cfg.nx, self.cfg.ny), target.shape
        loss = np.zeros(1, np.float32)
        gV = np.empty((self._nslice, self.cfg.potNx, self.cfg.potNy), np.float32) if want_grad_V else None
        gp = np.zeros(2, np.float32) if want_grad_probe else None
        rc = self._lib.qs_adjoint(self._h, _iptr(ipos), ipos.shape[0], _fptr(target), int(ckpt_every), _fptr(loss),
                                  _fptr(gV), _fptr(gp))
        self._check(rc, "qs_adjoint")
        out = {"loss": float(loss[0])}
        if gV is not None:
            out["grad_V"] = gV
        if gp is not None:
            out["grad_df0"], out["grad_Cs"] = float(gp[0]), float(gp[1])
        return out

    def forward_save(self, ipos: np.ndarray, target, ckpt_every: int = 0) -> float:
        """Loss of qs_adjoint from its forward pass alone (qs_forward_save); the context
        keeps the checkpoints for a later adjoint_saved.  target: numpy array or torch
        tensor in GPU memory."""
        ipos = np.ascontiguousarray(ipos, dtype=np.int32).reshape(-1, 2)
        if _is_device_tensor(target):
            target, tptr = _device_f32(target)
        else:
            target = np.ascontiguousarray(target, dtype=np.float32)
            tptr = _fptr(target)
        assert tuple(target.shape) == (ipos.shape[0], self.cfg.nx, self.cfg.ny), target.shape
        loss = np.zeros(1, np.float32)
        self._check(self._lib.qs_forward_save(self._h, _iptr(ipos), ipos.shape[0], tptr, int(ckpt_every),
                                              _fptr(loss)), "qs_forward_save")
        return float(loss[0])

    def adjoint_saved(self, want_grad_V: bool = True, want_grad_probe: bool = True, device=None) -> dict:
        """Gradients from the forward pass kept by the last forward_save (qs_adjoint_saved).
        device (a torch CUDA device): grad_V is returned as a tensor there, written by the
        library in place."""
        gp = np.zeros(2, np.float32) if want_grad_probe else None
        if want_grad_V and device is not None:
            import torch
            gV = torch.empty((self._nslice, self.cfg.potNx, self.cfg.potNy), dtype=torch.float32, device=device)
            gptr = ctypes.cast(gV.data_ptr(), ctypes.POINTER(ctypes.c_float))
        else:
            gV = np.empty((self._nslice, self.cfg.potNx, self.cfg.potNy), np.float32) if want_grad_V else None
            gptr = _fptr(gV)
        self._check(self._lib.qs_adjoint_saved(self._h, gptr, _fptr(gp)), "qs_adjoint_saved")
        out = {}
        if gV is not None:
            out["grad_V"] = gV
        if gp is not None:
            out["grad_df0"], out["grad_Cs"] = float(gp[0]), float(gp[1])
        return out

    def atom_gradients(self, atoms, **pot) -> np.ndarray:
        """dL/d(x, y, z) per atom from the last adjoint (same atoms as build_potential;
        the potential parameters default to that call's)."""
        atoms = np.asarray(atoms, dtype=np.float64).reshape(-1, 7)
        arr = _atom_array(atoms)
        kw = dict(getattr(self, "_build_pot", {}))
        kw.update(pot)
        pp = self._pot_params(**kw)
        g = np.zeros((len(atoms), 3), np.float32)
        self._check(self._lib.qs_atom_gradients(self._h, arr, len(atoms), ctypes.byref(pp), _fptr(g)),
                    "qs_atom_gradients")
        return g

    # -- multislice ------------------------------------------------------------------
    def set_batch(self, batch: int):
        self._check(self._lib.qs_set_batch(self._h, int(batch)), "qs_set_batch")

    def set_streams(self, nstreams: int):
        """1 or 2 HIP streams for the probe groups of run_stem (results are unchanged)."""
        self._check(self._lib.qs_set_streams(self._h, int(nstreams)), "qs_set_streams")

    def get_streams(self) -> int:
        return self._lib.qs_get_streams(self._h)

    def set_graph(self, on: bool = True):
        """Replay run_stem's slice loop as a HIP graph (captured once per call shape)."""
        self._check(self._lib.qs_set_graph(self._h, 1 if on else 0), "qs_set_graph")

    def set_tuning(self, rows_per_wave: int = 0, tiles_per_block: int = 0):
        """Launch shape of the two sweeps (0 = library default)."""
        self._check(self._lib.qs_set_tuning(self._h, int(rows_per_wave), int(tiles_per_block)), "qs_set_tuning")

    def get_tuning(self):
        """(rows_per_wave, tiles_per_block, batch) currently set (0 = library default)."""
        r, t, b = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._check(self._lib.qs_get_tuning(self._h, ctypes.byref(r), ctypes.byref(t), ctypes.byref(b)),
                    "qs_get_tuning")
        return r.value, t.value, b.value

    def autotune(self, ipos) -> dict:
        """Time group sizes and sweep launch shapes on this context's potential and probe,
        apply the fastest and return it: {'batch', 'rows_per_wave', 'tiles_per_block', 'rate'}."""
        ipos = np.ascontiguousarray(ipos, dtype=np.int32).reshape(-1, 2)
        b, r, t = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        rate = ctypes.c_float()
        self._check(self._lib.qs_autotune(self._h, _iptr(ipos), ipos.shape[0], ctypes.byref(b), ctypes.byref(r),
                                          ctypes.byref(t), ctypes.byref(rate)), "qs_autotune")
        return {"batch": b.value, "rows_per_wave": r.value, "tiles_per_block": t.value, "rate": rate.value}

    def set_profiling(self, on: bool = True):
        self._check(self._lib.qs_set_profiling(self._h, 1 if on else 0), "qs_set_profiling")

    def kernel_times(self):
        """(ms_rows, ms_cols, n_rows, n_cols): summed launch durations of the two sweeps."""
        a, b = ctypes.c_double(), ctypes.c_double()
        na, nb = ctypes.c_long(), ctypes.c_long()
        self._check(self._lib.qs_kernel_times(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(na),
                                              ctypes.byref(nb)), "qs_kernel_times")
        return a.value, b.value, na.value, nb.value

    @property
    def last_run_ms(self) -> float:
        return float(self._lib.qs_last_run_ms(self._h))

    @property
    def last_ckpt_interval(self) -> int:
        """Checkpoint interval used by the last adjoint / forward_save."""
        return int(self._lib.qs_last_ckpt_interval(self._h))

    def set_configs(self, nconfig: int):
        """Hold nconfig frozen-phonon transmission functions (reload all, configuration 0 first)."""
        self._check(self._lib.qs_set_configs(self._h, int(nconfig)), "qs_set_configs")

    def select_config(self, config: int):
        """Configuration that set_potential / set_trans / build_potential / get_trans address."""
        self._check(self._lib.qs_select_config(self._h, int(config)), "qs_select_config")

    def share_trans(self, src: Optional["Multislice"]):
        """Run on src's transmission functions (same potential grid, any window size)."""
        self._check(self._lib.qs_share_trans(self._h, src._h if src is not None else None), "qs_share_trans")
        self._trans_src = src  # keeps src alive while this context uses its buffers
        if src is not None:
            self._nslice = src._nslice

    def run_stem(self, ipos: np.ndarray, nslice: Optional[int] = None, start: int = 0,
                 bins: Optional[Sequence[int]] = None, nbins: int = 0, det: Optional[np.ndarray] = None,
                 want_det: bool = True, want_diff: bool = False, want_exit: bool = False,
                 want_intensity: bool = False, config: Optional[np.ndarray] = None, first_slice: int = 0) -> dict:
        """Advance len(ipos) probes through transmission slices first_slice ..
        first_slice + nslice - 1 (item i on phonon configuration config[i] when
        given).  Returns a dict with 'det' [nbins][ndet][nprobe], 'diff', 'exit',
        'intensity' as requested."""
        ipos = np.ascontiguousarray(ipos, dtype=np.int32).reshape(-1, 2)
        n = ipos.shape[0]
        if config is not None:
            config = np.ascontiguousarray(config, dtype=np.int32).reshape(-1)
            assert config.shape[0] == n
        nslice = self._nslice - first_slice if nslice is None else nslice
        if bins is None:
            b, nb = record_bins(nslice)
            bins, nbins = b, nb
        bins = np.ascontiguousarray(bins, dtype=np.int32)
        ndet = len(self.cfg.detectors)
        out = {}
        if want_det and ndet > 0:
            if det is None:
                det = np.zeros((nbins, ndet, n), dtype=np.float32)
            det = np.ascontiguousarray(det, dtype=np.float32)
            out["det"] = det
        else:
            det = None
        diff = np.empty((n, self.cfg.nx, self.cfg.ny), dtype=np.float32) if want_diff else None
        exitw = np.empty((n, self.cfg.nx, self.cfg.ny), dtype=np.complex64) if want_exit else None
        inten = np.empty(n, dtype=np.float32) if want_intensity else None
        rc = self._lib.qs_run_slices(self._h, _iptr(ipos), _iptr(config), n, int(start), int(first_slice), _iptr(bins),
                                     int(nslice), int(nbins), _fptr(det), _fptr(diff),
                                     _fptr(exitw.view(np.float32) if exitw is not None else None), _fptr(inten))
        self._check(rc, "qs_run_stem")
        if diff is not None:
            out["diff"] = diff
        if exitw is not None:
            out["exit"] = exitw
        if inten is not None:
            out["intensity"] = inten
        return out


    def run_stem_noalloc(self, ipos: np.ndarray, nslice: Optional[int] = None, bins: Optional[Sequence[int]] = None,
                         nbins: int = 0, want_intensity: bool = False) -> dict:
        """run_stem from the probe through qs_run_stem_noalloc: every device buffer
        of the call lives in the library's static scratch array (no allocation).
        Returns 'det' [nbins][ndet][nprobe] (and 'intensity')."""
        ipos = np.ascontiguousarray(ipos, dtype=np.int32).reshape(-1, 2)
        n = ipos.shape[0]
        nslice = self._nslice if nslice is None else nslice
        if bins is None:
            bins, nbins = record_bins(nslice)
        bins = np.ascontiguousarray(bins, dtype=np.int32)
        ndet = len(self.cfg.detectors)
        det = np.zeros((nbins, ndet, n), dtype=np.float32) if ndet > 0 else None
        inten = np.empty(n, dtype=np.float32) if want_intensity else None
        rc = self._lib.qs_run_stem_noalloc(self._h, _iptr(ipos), n, _iptr(bins), int(nslice), int(nbins), _fptr(det),
                                           _fptr(inten))
        self._check(rc, "qs_run_stem_noalloc")
        out = {}
        if det is not None:
            out["det"] = det
        if inten is not None:
            out["intensity"] = inten
        return out


def run_groups(groups, bins=None, nbins: int = 0, nslice: Optional[int] = None, want_intensity: bool = False):
    """Uneven batch (qs_run_groups): groups = [(Multislice, ipos, config or None), ...]
    whose contexts share one set of transmission functions (share_trans).  Every
    group starts from its probe; slices are interleaved across groups.  Returns a
    list of dicts with 'det' [nbins][ndet][nitem] (and 'intensity')."""
    from ._lib import QsGroup
    ms0 = groups[0][0]
    nslice = ms0._nslice if nslice is None else nslice
    if bins is None:
        bins, nbins = record_bins(nslice)
    bins = np.ascontiguousarray(bins, dtype=np.int32)
    arr = (QsGroup * len(groups))()
    keep, outs = [], []
    for g, (ms, ipos, config) in enumerate(groups):
        ipos = np.ascontiguousarray(ipos, dtype=np.int32).reshape(-1, 2)
        n = ipos.shape[0]
        cfg = None if config is None else np.ascontiguousarray(config, dtype=np.int32).reshape(-1)
        ndet = len(ms.cfg.detectors)
        det = np.zeros((nbins, ndet, n), dtype=np.float32) if ndet > 0 else None
        inten = np.empty(n, dtype=np.float32) if want_intensity else None
        keep += [ipos, cfg, det, inten]
        arr[g].ctx = ms._h
        arr[g].ipos = _iptr(ipos)
        arr[g].config = _iptr(cfg)
        arr[g].nitem = n
        arr[g].det = _fptr(det)
        arr[g].int_intensity = _fptr(inten)
        out = {}
        if det is not None:
            out["det"] = det
        if inten is not None:
            out["intensity"] = inten
        outs.append(out)
    rc = ms0._lib.qs_run_groups(arr, len(groups), _iptr(bins), int(nslice), int(nbins))
    ms0._check(rc, "qs_run_groups")
    return outs


def device_memory(reset_peak: bool = False):
    """(current, peak) bytes of device memory held by the library (all contexts)."""
    lib = _lib.load()
    cur, peak = ctypes.c_longlong(), ctypes.c_longlong()
    lib.qs_device_memory(ctypes.byref(cur), ctypes.byref(peak), 1 if reset_peak else 0)
    return cur.value, peak.value
