"""Python binding of libc3d.so (the C-ABI in include/c3d.h).

`Channel` mirrors the reference's per-step subroutine interface
(compact3d.f: GETUP :879, BCOND :963, RHS1/2/3 :6063/:6784/:7399,
GETUH1/2/3 :8218/:8344/:8478, DPCALC :1037, UPCALC :1052, DIVCHECK :9049,
CFL :9138, CHKMF :9463) on a state that stays resident in HBM.  Host arrays
use the reference's Fortran layouts (U (N1+1,N2+1,N3+1,3), P (N1,N2,N3),
order='F').

There is no CPU fallback: constructing a Channel without the built library or
without a GPU raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libc3d.so")

FP64, MIXED = 0, 1
F_RUH1, F_RUH2, F_RUH3, F_UH1, F_UH2, F_UH3, F_RDP, F_DP, F_HH1, F_HH2, F_HH3, F_TAUW, F_TAUB, F_IWM = range(1, 15)
F_DIV = 15
(S_PRESG, S_PRESG3, S_DMPRESG, S_DMPRESG3, S_FLOW1, S_FLOW3, S_WSM1, S_WSM2, S_WSM3, S_REFINES, S_DIVMAX1,
 S_NTIME, S_CFLMAX, S_WMDT, S_TAOW, S_UINIT, S_USINIT, S_PBAR) = range(18)
ERR_NONFINITE = -3  # C3D_ERR_NONFINITE: the step ran, its DIVMAX or CFLM is NaN or infinite
IWM_NPL = 43  # WMODEL = 2: planes of the integral wall model's state (COMMON /IWMLES/)
NKID = 48  # C3D_NKID of include/c3d.h: entries of the per-kernel timing tables

_dp = C.POINTER(C.c_double)


class C3DError(RuntimeError):
    pass


class C3DDiverged(C3DError):
    """a step's DIVMAX or CFLM is NaN or infinite (C3D_ERR_NONFINITE);
    .divmax and .cflm hold the values"""

    def __init__(self, msg, divmax=float("nan"), cflm=float("nan")):
        super().__init__(msg)
        self.divmax, self.cflm = divmax, cflm


class Config(C.Structure):
    _fields_ = [
        ("n1", C.c_int), ("n2", C.c_int), ("n3", C.c_int),
        ("re", C.c_double), ("alx", C.c_double), ("aly", C.c_double), ("alz", C.c_double),
        ("dt", C.c_double), ("files", C.c_double), ("wmodel", C.c_double), ("cks", C.c_double),
        ("y", _dp), ("precision", C.c_int), ("device", C.c_int), ("stream", C.c_void_p),
        ("cflmax", C.c_double),
    ]


_lib = None


def lib():
    """Load libc3d.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise C3DError(f"{LIB_PATH} not built; run __graft_entry__.build() or make -C compact3d_amd/csrc")
        L = C.CDLL(LIB_PATH)
        L.c3d_create.argtypes = [C.POINTER(Config), C.POINTER(C.c_void_p)]
        L.c3d_last_error.restype = C.c_char_p
        L.c3d_last_error.argtypes = [C.c_void_p]
        L.c3d_footprint_bytes.restype = C.c_longlong
        L.c3d_footprint_bytes.argtypes = [C.POINTER(Config)]
        vp = [C.c_void_p]
        for name in ("c3d_destroy", "c3d_bcond", "c3d_rhs", "c3d_getuh", "c3d_dpcalc", "c3d_upcalc",
                     "c3d_sgs", "c3d_wall_model", "c3d_wallss"):
            getattr(L, name).argtypes = vp
        L.c3d_getup.argtypes = [C.c_void_p, C.c_int]
        L.c3d_step.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        L.c3d_advance.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.c3d_sync.argtypes = [C.c_void_p, _dp, _dp]
        L.c3d_upload_state.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, C.c_double, C.c_double]
        L.c3d_download_state.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp]
        L.c3d_divcheck.argtypes = [C.c_void_p, _dp]
        L.c3d_cfl.argtypes = [C.c_void_p, _dp]
        L.c3d_chkmf.argtypes = [C.c_void_p, _dp, _dp]
        L.c3d_flowrate.argtypes = [C.c_void_p, _dp, _dp]
        L.c3d_stats.argtypes = [C.c_void_p, _dp]
        L.c3d_stats_noalloc.argtypes = [C.c_void_p, _dp]
        L.c3d_cellcentre.argtypes = [C.c_void_p, _dp]
        L.c3d_sgs_profiles.argtypes = [C.c_void_p, _dp, _dp]
        L.c3d_write_flow.argtypes = [C.c_void_p, C.c_char_p, C.c_int]
        L.c3d_write_strain.argtypes = [C.c_void_p, C.c_char_p, C.c_int, C.c_double]
        L.c3d_format_flow.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp]
        L.c3d_format_strain.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, C.c_double]
        L.c3d_get_field.argtypes = [C.c_void_p, C.c_int, _dp]
        L.c3d_set_field.argtypes = [C.c_void_p, C.c_int, _dp]
        L.c3d_get_scalar.argtypes = [C.c_void_p, C.c_int, _dp]
        L.c3d_set_scalar.argtypes = [C.c_void_p, C.c_int, C.c_double]
        L.c3d_set_dt.argtypes = [C.c_void_p, C.c_double]
        L.c3d_apply_op.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp]
        L.c3d_get_mesh.argtypes = [C.c_void_p, _dp, _dp]
        L.c3d_filter.argtypes = [C.c_void_p, C.c_int, _dp, _dp, C.c_int]
        L.c3d_timing.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.c3d_autotune.argtypes = [C.POINTER(Config), C.c_int, C.POINTER(C.c_int), _dp]
        L.c3d_autotune_set.argtypes = [C.POINTER(Config), C.POINTER(C.c_int)]
        L.c3d_autotune_reset.argtypes = []
        L.c3d_autotune_reset.restype = None
        L.c3d_ensemble_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.c_void_p)]
        L.c3d_ensemble_advance.argtypes = [C.c_void_p, C.c_int]
        L.c3d_ensemble_sync.argtypes = [C.c_void_p, _dp, _dp]
        L.c3d_ensemble_destroy.argtypes = [C.c_void_p]
        L.c3d_ensemble_last_error.restype = C.c_char_p
        L.c3d_ensemble_last_error.argtypes = [C.c_void_p]
        L.c3d_timing_select.argtypes = [C.c_void_p, C.c_int]
        L.c3d_kernel_times.argtypes = [C.c_void_p, _dp, C.POINTER(C.c_int)]
        L.c3d_kernel_name.restype = C.c_char_p
        L.c3d_kernel_name.argtypes = [C.c_int]
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    if a.dtype != np.float64 or not (a.flags.f_contiguous or a.ndim == 1):
        raise C3DError("arrays must be float64 in Fortran order")
    return a.ctypes.data_as(_dp)


def format_flow(path, ntime, xc, yc, zc, cell):
    """WRITEUP's FLOW-*.plt text from host arrays (no GPU): xc, yc, zc =
    XC(1:N1-1), YC(1:N2-1), ZC(1:N3-1); cell as Channel.cellcentre()"""
    L = lib()
    xc, yc, zc = (np.ascontiguousarray(a, dtype=np.float64) for a in (xc, yc, zc))
    cell = np.ascontiguousarray(cell, dtype=np.float64)
    if cell.size != 4 * xc.size * yc.size * zc.size:
        raise C3DError("cell must hold 4*(N1-1)*(N2-1)*(N3-1) values")
    if L.c3d_format_flow(os.fsencode(path), int(ntime), xc.size, yc.size, zc.size, _p(xc), _p(yc), _p(zc),
                         _p(cell.reshape(-1))):
        raise C3DError(L.c3d_last_error(None).decode())


def format_strain(path, ntime, xc, zc, tauw, utau):
    """WRITEUP's STRAIN-*.plt text from host arrays (no GPU): tauw indexed
    [K-1, I-1] = TAUW(I,K,1)"""
    L = lib()
    xc, zc = (np.ascontiguousarray(a, dtype=np.float64) for a in (xc, zc))
    tauw = np.ascontiguousarray(tauw, dtype=np.float64)
    if tauw.size != xc.size * zc.size:
        raise C3DError("tauw must hold (N1-1)*(N3-1) values")
    if L.c3d_format_strain(os.fsencode(path), int(ntime), xc.size, zc.size, _p(xc), _p(zc),
                           _p(tauw.reshape(-1)), float(utau)):
        raise C3DError(L.c3d_last_error(None).decode())


def make_config(n1, n2, n3, re, alx, aly, alz, dt, files=0.0, wmodel=0.0, cks=2.5, y=None,
                precision=FP64, device=0, cflmax=0.0):
    cfg = Config()
    cfg.n1, cfg.n2, cfg.n3 = n1, n2, n3
    cfg.re, cfg.alx, cfg.aly, cfg.alz, cfg.dt = re, alx, aly, alz, dt
    cfg.files, cfg.wmodel, cfg.cks = files, wmodel, cks
    cfg._y = None
    if y is not None:
        cfg._y = np.ascontiguousarray(y, dtype=np.float64)
        cfg.y = cfg._y.ctypes.data_as(_dp)
    cfg.precision, cfg.device, cfg.stream = precision, device, None
    cfg.cflmax = cflmax
    return cfg


# c3d_autotune's shape layout (C3D_NSHAPES values); y_cols_fp32 counts the
# FP32 y tile's threads per chunk row, each holding y_pair_fp32 columns
SHAPE_KEYS = ("x_lines", "z_cols_fp64", "z_cols_fp32", "z_seg_fp32", "y_cols_fp64", "y_chunks_fp64",
              "y_cols_fp32", "y_chunks_fp32", "pois_cols", "pois_chunks", "fft_x_lines", "fft_z_lines",
              "y_pair_fp32")
# keys that shape sets written before a key existed may lack, and their value
SHAPE_DEFAULTS = {"y_pair_fp32": 1}


def autotune(steps=4, **kw):
    """c3d_autotune: time candidate tile shapes for this configuration and
    record the fastest for handles created afterwards.  Returns the chosen
    shapes and ms per step before ('ms_rule') and after ('ms_tuned')."""
    L = lib()
    sh = (C.c_int * len(SHAPE_KEYS))()
    ms = np.zeros(2)
    if L.c3d_autotune(C.byref(make_config(**kw)), int(steps), sh, _p(ms)) != 0:
        raise C3DError("c3d_autotune: " + L.c3d_last_error(None).decode())
    out = dict(zip(SHAPE_KEYS, list(sh)))
    out.update(ms_rule=float(ms[0]), ms_tuned=float(ms[1]))
    return out


def autotune_set(shapes, **kw):
    """c3d_autotune_set: record the given shapes (a dict with SHAPE_KEYS, as
    autotune() returns) for this configuration without timing"""
    L = lib()
    sh = (C.c_int * len(SHAPE_KEYS))(*[int(shapes.get(k, SHAPE_DEFAULTS.get(k))) for k in SHAPE_KEYS])
    if L.c3d_autotune_set(C.byref(make_config(**kw)), sh) != 0:
        raise C3DError("c3d_autotune_set: " + L.c3d_last_error(None).decode())


def autotune_reset():
    lib().c3d_autotune_reset()


def footprint_bytes(**kw) -> int:
    return int(lib().c3d_footprint_bytes(C.byref(make_config(**kw))))


class Channel:
    """One channel instance resident on a GPU."""

    def __init__(self, n1, n2, n3, re, alx, aly, alz, dt, files=0.0, wmodel=0.0, cks=2.5, y=None,
                 precision=FP64, device=0, cflmax=0.0):
        self.L = lib()
        self.cfg = make_config(n1, n2, n3, re, alx, aly, alz, dt, files, wmodel, cks, y, precision, device,
                               cflmax)
        h = C.c_void_p()
        rc = self.L.c3d_create(C.byref(self.cfg), C.byref(h))
        if rc != 0:
            raise C3DError("c3d_create: " + self.L.c3d_last_error(None).decode())
        self.h = h
        self.n1, self.n2, self.n3 = n1, n2, n3
        self.files = float(files)
        self.wmodel = float(wmodel)
        self.precision = int(precision)

    def close(self):
        if getattr(self, "h", None):
            self.L.c3d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, rc, what):
        if rc != 0:
            raise C3DError(f"{what}: " + self.L.c3d_last_error(self.h).decode())

    # ---- shapes
    def ushape(self):
        return (self.n1 + 1, self.n2 + 1, self.n3 + 1)

    def pshape(self):
        return (self.n1, self.n2, self.n3)

    # ---- state
    def upload(self, U, P, HH=None, SGSVIS=None, presg=0.0, presg3=0.0):
        self._ck(self.L.c3d_upload_state(self.h, _p(U), _p(P), _p(HH), _p(SGSVIS), presg, presg3), "upload")

    def download(self):
        U = np.zeros(self.ushape() + (3,), order="F")
        P = np.zeros(self.pshape(), order="F")
        HH = np.zeros(self.ushape() + (3,), order="F")
        S = np.zeros(self.ushape(), order="F")
        a, b = C.c_double(), C.c_double()
        self._ck(self.L.c3d_download_state(self.h, _p(U), _p(P), _p(HH), _p(S), C.byref(a), C.byref(b)),
                 "download")
        return dict(U=U, P=P, HH=HH, SGSVIS=S, PRESG=a.value, PRESG3=b.value)

    # ---- reference-named operations
    # ntime: GETUP's NTIME (the integral wall model, WMODEL = 2, counts its
    # steps by it); 0 continues from the last step's
    def getup(self, ntime=0):
        self._ck(self.L.c3d_getup(self.h, ntime), "getup")

    def _ck_checks(self, rc, what, d, c):
        if rc == ERR_NONFINITE:
            raise C3DDiverged(f"{what}: " + self.L.c3d_last_error(self.h).decode(), d.value, c.value)
        self._ck(rc, what)

    def step(self, ntime=0, checks=True):
        """GETUP + DIVCHECK + CFL; (DIVMAX, CFLM).  Raises C3DDiverged when
        either is NaN or infinite."""
        if checks:
            d, c = C.c_double(), C.c_double()
            self._ck_checks(self.L.c3d_step(self.h, ntime, C.byref(d), C.byref(c)), "step", d, c)
            return d.value, c.value
        self._ck(self.L.c3d_step(self.h, ntime, None, None), "step")
        return None

    def advance(self, nsteps, use_graph=True):
        """queue nsteps steps on the handle's stream without waiting"""
        self._ck(self.L.c3d_advance(self.h, int(nsteps), 1 if use_graph else 0), "advance")

    def sync(self):
        """wait for the queued steps; (DIVMAX, CFLM) of the last one.  Raises
        C3DDiverged when either is NaN or infinite."""
        d, c = C.c_double(), C.c_double()
        self._ck_checks(self.L.c3d_sync(self.h, C.byref(d), C.byref(c)), "sync", d, c)
        return d.value, c.value

    def bcond(self):
        self._ck(self.L.c3d_bcond(self.h), "bcond")

    def rhs(self):
        self._ck(self.L.c3d_rhs(self.h), "rhs")

    def getuh(self):
        self._ck(self.L.c3d_getuh(self.h), "getuh")

    def dpcalc(self):
        self._ck(self.L.c3d_dpcalc(self.h), "dpcalc")

    def upcalc(self):
        self._ck(self.L.c3d_upcalc(self.h), "upcalc")

    def sgscs(self):
        self._ck(self.L.c3d_sgs(self.h), "sgs")

    def wall_model(self):
        self._ck(self.L.c3d_wall_model(self.h), "wall_model")

    def wallss(self):
        self._ck(self.L.c3d_wallss(self.h), "wallss")

    def filter(self, filv, width=2):
        """FILTER (width 2) / FILTER4 (width 4) of FILV(N1, N3, n), Fortran order"""
        a = np.asfortranarray(filv, dtype=np.float64)
        out = np.zeros_like(a, order="F")
        self._ck(self.L.c3d_filter(self.h, int(width), _p(a), _p(out), a.shape[2]), "filter")
        return out

    def divcheck(self):
        v = C.c_double()
        self._ck(self.L.c3d_divcheck(self.h, C.byref(v)), "divcheck")
        return v.value

    def cfl(self):
        v = C.c_double()
        self._ck(self.L.c3d_cfl(self.h, C.byref(v)), "cfl")
        return v.value

    def chkmf(self):
        a, b = C.c_double(), C.c_double()
        self._ck(self.L.c3d_chkmf(self.h, C.byref(a), C.byref(b)), "chkmf")
        return a.value, b.value

    def flowrate(self):
        a, b = C.c_double(), C.c_double()
        self._ck(self.L.c3d_flowrate(self.h, C.byref(a), C.byref(b)), "flowrate")
        return a.value, b.value

    def field(self, which):
        if which in (F_RDP, F_DP, F_DIV):
            out = np.zeros(self.pshape(), order="F")
        elif which in (F_TAUW, F_TAUB):
            out = np.zeros((self.n1 + 1, self.n3 + 1, 2), order="F")
        elif which == F_IWM:
            out = np.zeros(IWM_NPL * (self.n3 - 1) * (self.n1 - 1))
            self._ck(self.L.c3d_get_field(self.h, which, _p(out)), "get_field")
            return out.reshape(IWM_NPL, self.n3 - 1, self.n1 - 1)
        else:
            out = np.zeros(self.ushape(), order="F")
        self._ck(self.L.c3d_get_field(self.h, which, _p(out)), "get_field")
        return out

    def set_field(self, which, values):
        """F_IWM ([plane, K-1, I-1] as field(F_IWM) returns it), F_TAUW,
        F_TAUB, F_DIV (the layouts field() returns)"""
        if which == F_IWM:
            a = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        else:
            a = np.asfortranarray(values, dtype=np.float64)
        self._ck(self.L.c3d_set_field(self.h, which, _p(a)), "set_field")

    def scalar(self, which):
        v = C.c_double()
        self._ck(self.L.c3d_get_scalar(self.h, which, C.byref(v)), "get_scalar")
        return v.value

    def set_scalar(self, which, value):
        self._ck(self.L.c3d_set_scalar(self.h, which, float(value)), "set_scalar")

    def set_dt(self, dt):
        self._ck(self.L.c3d_set_dt(self.h, float(dt)), "set_dt")

    def mesh(self):
        """{'Y': Y(0:N2), 'DX1', 'DX3', 'ALY', 'VOL'} as the library built them"""
        y = np.zeros(self.n2 + 1)
        sc = np.zeros(4)
        self._ck(self.L.c3d_get_mesh(self.h, _p(y), _p(sc)), "get_mesh")
        return dict(Y=y, DX1=sc[0], DX3=sc[1], ALY=sc[2], VOL=sc[3])

    def timing(self, on=True, capacity=4096):
        """bracket every following kernel launch with HIP events on the
        handle's stream (per-kernel durations, c3d_kernel_times)"""
        self._ck(self.L.c3d_timing(self.h, 1 if on else 0, int(capacity)), "timing")

    def timing_select(self, name=None):
        """restrict the event recorder to the launches of one kernel (by
        name, as kernel_times() reports it); None: all kernels"""
        kid = -1
        if name is not None:
            names = [self.L.c3d_kernel_name(q).decode() for q in range(NKID)]
            if name not in names:
                raise C3DError("no kernel named %r" % name)
            kid = names.index(name)
        self._ck(self.L.c3d_timing_select(self.h, kid), "timing_select")

    def kernel_times(self):
        """{kernel name: (total ms, launches)} since timing() was enabled"""
        ms = (C.c_double * NKID)()
        cnt = (C.c_int * NKID)()
        self._ck(self.L.c3d_kernel_times(self.h, ms, cnt), "kernel_times")
        return {self.L.c3d_kernel_name(q).decode(): (ms[q], cnt[q]) for q in range(NKID) if cnt[q] > 0}

    STAT_NAMES = ("VM1", "VM2", "VM3", "PM", "WM1", "WM2", "WM3", "VRMS1", "VRMS2", "VRMS3", "VRMS4",
                  "VRMS5", "VRMS6", "PRMS", "PV", "WRMS1", "WRMS2", "WRMS3", "U2V", "VW2")

    def stats(self):
        """ENERGY's plane statistics of the resident field: array (20, N2-1),
        rows in the order of STAT_NAMES (J = 1..N2-1 along each row)"""
        out = np.zeros(20 * (self.n2 - 1))
        self._ck(self.L.c3d_stats(self.h, _p(out)), "stats")
        return out.reshape(20, self.n2 - 1)

    def stats_noalloc(self):
        """stats() through c3d_stats_noalloc (no allocation; N2-1 <= 2048)"""
        out = np.zeros(20 * (self.n2 - 1))
        self._ck(self.L.c3d_stats_noalloc(self.h, _p(out)), "stats_noalloc")
        return out.reshape(20, self.n2 - 1)

    def sgs_profiles(self):
        """(SMAGC(J), BETA(J)) of the last dynamic-model evaluation, J = 1..N2-1
        (BETA only for FILES = 3, else None)"""
        c = np.zeros(self.n2 - 1)
        b = np.zeros(self.n2 - 1) if self.files == 3.0 else None
        self._ck(self.L.c3d_sgs_profiles(self.h, _p(c), _p(b)), "sgs_profiles")
        return c, b

    def cellcentre(self):
        """WRITEUP's cell-centre values of the resident field: array
        (4, N3-1, N2-1, N1-1) of UF, VF, WF, P indexed [q, K-1, J-1, I-1]"""
        n = (self.n1 - 1) * (self.n2 - 1) * (self.n3 - 1)
        out = np.zeros(4 * n)
        self._ck(self.L.c3d_cellcentre(self.h, _p(out)), "cellcentre")
        return out.reshape(4, self.n3 - 1, self.n2 - 1, self.n1 - 1)

    def write_flow(self, path, ntime):
        """WRITEUP's FLOW-*.plt of the resident field to `path`"""
        self._ck(self.L.c3d_write_flow(self.h, os.fsencode(path), int(ntime)), "write_flow")

    def write_strain(self, path, ntime, utau):
        """WRITEUP's STRAIN-*.plt (bottom-wall TAUW) to `path`"""
        self._ck(self.L.c3d_write_strain(self.h, os.fsencode(path), int(ntime), float(utau)), "write_strain")

    def apply_op(self, op, direction, field):
        inp = np.asfortranarray(field, dtype=np.float64)
        out = np.zeros(self.pshape(), order="F")
        self._ck(self.L.c3d_apply_op(self.h, op, direction, _p(inp), _p(out)), "apply_op")
        return out


class Ensemble:
    """Independent channels advanced together: each step of every member is
    one branch of a single captured HIP graph, so one launch per step
    advances all of them and their kernels overlap on the device."""

    def __init__(self, channels):
        self.L = lib()
        self.members = list(channels)
        arr = (C.c_void_p * len(self.members))(*[ch.h for ch in self.members])
        e = C.c_void_p()
        if self.L.c3d_ensemble_create(arr, len(self.members), C.byref(e)) != 0:
            raise C3DError("c3d_ensemble_create: " + self.L.c3d_ensemble_last_error(None).decode())
        self.e = e

    def _ck(self, rc, what):
        if rc != 0:
            raise C3DError(f"{what}: " + self.L.c3d_ensemble_last_error(self.e).decode())

    def advance(self, nsteps):
        """queue nsteps steps of every member without waiting"""
        self._ck(self.L.c3d_ensemble_advance(self.e, int(nsteps)), "ensemble_advance")

    def sync(self):
        """wait; arrays of every member's (DIVMAX, CFLM) after its last step.
        Raises C3DDiverged (with the arrays) when a member's value is NaN
        or infinite."""
        d = np.zeros(len(self.members))
        c = np.zeros(len(self.members))
        rc = self.L.c3d_ensemble_sync(self.e, _p(d), _p(c))
        if rc == ERR_NONFINITE:
            raise C3DDiverged("ensemble_sync: " + self.L.c3d_ensemble_last_error(self.e).decode(), d, c)
        self._ck(rc, "ensemble_sync")
        return d, c

    def close(self):
        if getattr(self, "e", None):
            self.L.c3d_ensemble_destroy(self.e)
            self.e = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
