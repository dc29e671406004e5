"""ctypes bindings for the test oracle (liboracle.so) and the reference build
(_ref/libc3dref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The product library never loads either.

All arrays use the reference's Fortran layouts (compact3d.f:55-60) with
M1=N1, M2=N2, M3=N3:
    U, HH, UH   float64, shape (N1+1, N2+1, N3+1, 3), order='F'
    SGSVIS      float64, shape (N1+1, N2+1, N3+1), order='F'
    P, DP, RDP  float64, shape (N1, N2, N3), order='F'
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libc3dref.so")

# WMODEL = 2: the planes of COMMON /IWMLES/ (compact3d.f:3399-3408) in order
IWM_PLANES = ("utx", "uty", "tauwx", "tauwy", "flt_u", "flt_w", "flt_um", "flt_wm", "flt_p",
              "inte_lu", "inte_luu", "inte_lv", "inte_lvv", "inte_luv",
              "intem_lu", "intem_luu", "intem_lv", "intem_lvv", "intem_luv",
              "unsdy_x", "unsdy_y", "conv_x", "conv_y", "prsgrad_x", "prsgrad_y", "diff_x", "diff_y",
              "lhs_x", "lhs_y", "dudzt_x", "dudzt_y", "dudzb_x", "dudzb_y",
              "flt_us", "tr", "dz", "z0", "ax", "ay", "delv", "deli", "cx", "cy")
IWM_NPL = len(IWM_PLANES)

_dp = C.POINTER(C.c_double)


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.ndim == 1), "need float64 Fortran array"
    return a.ctypes.data_as(_dp)


def build_oracle():
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


def ref_available() -> bool:
    return os.path.exists(REF_SO)


class Oracle:
    """FP64 restatement of the reference hot path (c3d_oracle.c)."""

    _lib = None

    @classmethod
    def lib(cls):
        if cls._lib is None:
            if not os.path.exists(ORACLE_SO):
                build_oracle()
            L = C.CDLL(ORACLE_SO)
            L.c3o_new.restype = C.c_void_p
            L.c3o_new.argtypes = [C.c_int] * 3 + [C.c_double] * 8
            L.c3o_free.argtypes = [C.c_void_p]
            L.c3o_get_vec.argtypes = [C.c_void_p, C.c_int, _dp]
            L.c3o_get_scalar.restype = C.c_double
            L.c3o_get_scalar.argtypes = [C.c_void_p, C.c_int]
            L.c3o_divcheck.restype = C.c_double
            L.c3o_cfl.restype = C.c_double
            for name in ("c3o_set_dt",):
                getattr(L, name).argtypes = [C.c_void_p, C.c_double]
            L.c3o_set_ifwall.argtypes = [C.c_void_p, C.c_int]
            L.c3o_set_threads.argtypes = [C.c_int]
            L.c3o_ctdma.argtypes = [C.c_int, C.c_double, _dp, _dp]
            L.c3o_rhs1.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, C.c_double, _dp]
            L.c3o_rhs2.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp]
            L.c3o_rhs3.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, C.c_double, _dp]
            L.c3o_upcalc.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp]
            L.c3o_getup.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp, C.c_int]
            L.c3o_dft2d.argtypes = [C.c_int, C.c_int, _dp, C.c_int]
            L.c3o_dqrrt.restype = C.c_double
            L.c3o_dqrrt.argtypes = [_dp, C.c_int]
            L.c3o_sgssd_plane.restype = C.c_double
            L.c3o_sgssd_plane.argtypes = [C.c_void_p, _dp, C.c_int, _dp, _dp, _dp]
            L.c3o_wallmodel_iwmles.argtypes = [C.c_void_p, _dp, _dp, C.c_double, C.c_double, C.c_int]
            L.c3o_iwm_state.argtypes = [C.c_void_p, _dp, _dp, C.POINTER(C.c_int)]
            L.c3o_set_cflmax.argtypes = [C.c_void_p, C.c_double]
            L.c3o_iwm_set.argtypes = [C.c_void_p, _dp, C.c_double]
            cls._lib = L
        return cls._lib

    def __init__(self, n1, n2, n3, re, alx, aly, alz, dt, files=0.0, wmodel=0.0, cks=2.5):
        self.L = self.lib()
        self.h = C.c_void_p(self.L.c3o_new(n1, n2, n3, re, alx, aly, alz, dt, files, wmodel, cks))
        self.n1, self.n2, self.n3 = n1, n2, n3
        self.files, self.wmodel = files, wmodel

    def __del__(self):
        if getattr(self, "h", None):
            self.L.c3o_free(self.h)
            self.h = None

    # ---- shapes
    def ushape(self, ncomp=3):
        s = (self.n1 + 1, self.n2 + 1, self.n3 + 1)
        return s + (ncomp,) if ncomp else s

    def pshape(self):
        return (self.n1, self.n2, self.n3)

    def zeros_u(self, ncomp=3):
        return np.zeros(self.ushape(ncomp), order="F")

    def zeros_p(self):
        return np.zeros(self.pshape(), order="F")

    # ---- mesh
    def vec(self, which):
        out = np.zeros(max(self.n1, self.n2, self.n3) + 4)
        n = self.L.c3o_get_vec(self.h, which, _p(out))
        return out[:n].copy()

    def scalar(self, which):
        return self.L.c3o_get_scalar(self.h, which)

    def mesh(self):
        names = ["Y", "DY", "H", "HM", "HC", "HP", "DYM", "DYC", "DYP", "AK1", "AK3",
                 "PMJ", "PCJ", "PPJ", "FJPA", "FJMU", "FJMV"]
        m = {n: self.vec(i) for i, n in enumerate(names)}
        for n, w in (("DX1", 100), ("DX1Q", 101), ("DX3", 102), ("DX3Q", 103), ("ALY", 104), ("VOL", 105)):
            m[n] = self.scalar(w)
        return m

    def set_dt(self, dt):
        self.L.c3o_set_dt(self.h, dt)

    def set_ifwall(self, v):
        self.L.c3o_set_ifwall(self.h, int(v))

    def get_tau(self):
        tw = np.zeros((self.n1 + 1, self.n3 + 1, 2), order="F")
        tb = np.zeros_like(tw)
        self.L.c3o_get_tau(self.h, _p(tw), _p(tb))
        return tw, tb

    def set_tau(self, tw, tb):
        self.L.c3o_set_tau(self.h, _p(np.asfortranarray(tw)), _p(np.asfortranarray(tb)))

    def get_wsm(self):
        w = np.zeros(3)
        self.L.c3o_get_wsm(self.h, _p(w))
        return w

    def set_wsm(self, w):
        self.L.c3o_set_wsm(self.h, _p(np.ascontiguousarray(w, dtype=np.float64)))

    # ---- phases
    def bcond(self, U, SGSVIS):
        self.L.c3o_bcond(self.h, _p(U), _p(SGSVIS))

    def rhs(self, comp, U, P, HH, HH2, RUH, presg, SGSVIS):
        if comp == 1:
            self.L.c3o_rhs1(self.h, _p(U), _p(P), _p(HH), _p(HH2), _p(RUH), presg, _p(SGSVIS))
        elif comp == 2:
            self.L.c3o_rhs2(self.h, _p(U), _p(P), _p(HH), _p(HH2), _p(RUH), _p(SGSVIS))
        else:
            self.L.c3o_rhs3(self.h, _p(U), _p(P), _p(HH), _p(HH2), _p(RUH), presg, _p(SGSVIS))

    def getuh(self, comp, U, UH, RUH):
        f = {1: self.L.c3o_getuh1, 2: self.L.c3o_getuh2, 3: self.L.c3o_getuh3}[comp]
        f(self.h, _p(U), _p(UH), _p(RUH))

    def rhsdp(self, RDP, UH):
        self.L.c3o_rhsdp(self.h, _p(RDP), _p(UH))

    def getdp(self, DP, RDP):
        self.L.c3o_getdp(self.h, _p(DP), _p(RDP))

    def upcalc(self, U, P, UH, DP, presg, presg3):
        a, b = C.c_double(presg), C.c_double(presg3)
        self.L.c3o_upcalc(self.h, _p(U), _p(P), _p(UH), _p(DP), C.byref(a), C.byref(b))
        return a.value, b.value

    def divcheck(self, U, want_div=False):
        div = np.zeros((self.n1 - 1, self.n2 - 1, self.n3 - 1), order="F") if want_div else None
        m = self.L.c3o_divcheck(self.h, _p(U), _p(div))
        return (m, div) if want_div else m

    def cfl(self, U):
        return self.L.c3o_cfl(self.h, _p(U))

    def chkmf(self, U):
        a, b = C.c_double(), C.c_double()
        self.L.c3o_chkmf(self.h, _p(U), C.byref(a), C.byref(b))
        return a.value, b.value

    def wallss(self, U):
        self.L.c3o_wallss(self.h, _p(U))

    def sgscs(self, U, SGSVIS):
        self.L.c3o_sgscs(self.h, _p(U), _p(SGSVIS))

    def plane_stats(self, U, P):
        """ENERGY's plane statistics: array (20, N2M)"""
        out = np.zeros(20 * (self.n2 - 1))
        self.L.c3o_plane_stats(self.h, _p(U), _p(P), _p(out))
        return out.reshape(20, self.n2 - 1)

    # ---- WRITEUP (compact3d.f:757-835), restated in numpy
    @staticmethod
    def writeup_cells(U, P):
        """WRITEUP's cell-centre values (compact3d.f:796-808): array
        (4, N3M, N2M, N1M) of UF, VF, WF, P indexed [q, K-1, J-1, I-1].
        IP, KP wrap periodically (INDICES :446-494); JPA(N2M) = N2M (:473),
        so VF at J = N2M is V(N2M) itself; TERMP = 1 (:794)."""
        n1m, n2m, n3m = P.shape[0] - 1, P.shape[1] - 1, P.shape[2] - 1
        I, J, K = np.arange(1, n1m + 1), np.arange(1, n2m + 1), np.arange(1, n3m + 1)
        IP = np.where(I < n1m, I + 1, 1)
        JP = np.where(J < n2m, J + 1, n2m)
        KP = np.where(K < n3m, K + 1, 1)
        u, v, w = U[..., 0], U[..., 1], U[..., 2]
        uf = 0.5 * (u[np.ix_(I, J, K)] + u[np.ix_(IP, J, K)])
        vf = 0.5 * (v[np.ix_(I, J, K)] + v[np.ix_(I, JP, K)])
        wf = 0.5 * (w[np.ix_(I, J, K)] + w[np.ix_(I, J, KP)])
        pp = P[np.ix_(I - 1, J - 1, K - 1)]
        return np.stack([a.transpose(2, 1, 0) for a in (uf, vf, wf, pp)])

    @staticmethod
    def writeup_coords(n1, n3, alx, alz, Y):
        """XC(1:N1M), YC(1:N2M), ZC(1:N3M) of MESH (:322-333, :363-381):
        XG(I) = (I-1)*ALX/(M1-1) with XG(M1) = ALX, XC(I) = (XG(I+1)+XG(I))/2"""
        def centres(n, al):
            g = lambda i: np.where(i == n, al, (i - 1).astype(np.float64) * al / float(n - 1))
            i = np.arange(1, n)
            return 0.5 * (g(i + 1) + g(i))
        Y = np.asarray(Y, dtype=np.float64)
        return centres(n1, alx), 0.5 * (Y[2:] + Y[1:-1]), centres(n3, alz)

    def sgssd(self, U, SGSVIS):
        """SGSSD; returns (C(J), BETA(J)), J = 0..N2 (1..N2M set)"""
        c = np.zeros(self.n2 + 1)
        b = np.zeros(self.n2 + 1)
        self.L.c3o_sgssd(self.h, _p(U), _p(SGSVIS), _p(c), _p(b))
        return c, b

    def sgssd_plane(self, U, j):
        """SGSSD's plane J: (s1, s2, PA, BETA) with s1, s2 the A..E sums of the
        FILTER and FILTER4 levels"""
        s1, s2, pa = np.zeros(5), np.zeros(5), np.zeros(5)
        beta = self.L.c3o_sgssd_plane(self.h, _p(U), int(j), _p(s1), _p(s2), _p(pa))
        return s1, s2, pa, beta

    @classmethod
    def sd_coefficient(cls, s1, s2):
        pa = np.zeros(5)
        cls.lib().c3o_sd_coefficient(_p(np.ascontiguousarray(s1, dtype=np.float64)),
                                     _p(np.ascontiguousarray(s2, dtype=np.float64)), _p(pa))
        return pa

    @classmethod
    def dqrrt(cls, pa):
        pa = np.ascontiguousarray(pa, dtype=np.float64)
        return cls.lib().c3o_dqrrt(_p(pa), len(pa))

    def sgsdy(self, U, SGSVIS):
        """SGSDY; returns the dynamic coefficient C(J), J = 0..N2 (1..N2M set)"""
        c = np.zeros(self.n2 + 1)
        self.L.c3o_sgsdy(self.h, _p(U), _p(SGSVIS), _p(c))
        return c

    def strain(self, U, j):
        SR = np.zeros((self.n1, self.n3, 6), order="F")
        self.L.c3o_strain(self.h, _p(U), _p(SR), int(j))
        return SR

    def wallmodel_instant(self, U):
        self.L.c3o_wallmodel_instant(self.h, _p(U))

    def wallmodel_werner(self, U):
        self.L.c3o_wallmodel_werner(self.h, _p(U))

    def wallmodel_instant2(self, U):
        self.L.c3o_wallmodel_instant2(self.h, _p(U))

    def set_cflmax(self, v):
        self.L.c3o_set_cflmax(self.h, float(v))

    def wallmodel_iwmles(self, U, P, presg, presg3, ntime):
        self.L.c3o_wallmodel_iwmles(self.h, _p(U), _p(P), float(presg), float(presg3), int(ntime))

    def iwm_state(self):
        """the 43 planes of COMMON /IWMLES/ as [plane, K-1, I-1], wm_dt, and
        the points of the last call whose linear-correction branch needed an
        equilibrium stress of their own"""
        n1m, n3m = self.n1 - 1, self.n3 - 1
        pl = np.zeros(IWM_NPL * n3m * n1m)
        dt = C.c_double(0.0)
        fx = C.c_int(0)
        self.L.c3o_iwm_state(self.h, _p(pl), C.byref(dt), C.byref(fx))
        return pl.reshape(IWM_NPL, n3m, n1m), dt.value, fx.value

    def iwm_set(self, planes, wm_dt):
        self.L.c3o_iwm_set(self.h, _p(np.ascontiguousarray(planes, dtype=np.float64).reshape(-1)), float(wm_dt))

    def filter(self, FILV, four=False):
        n = FILV.shape[2]
        (self.L.c3o_filter4 if four else self.L.c3o_filter)(self.h, _p(FILV), n)

    def getup(self, U, P, HH, SGSVIS, presg, presg3, ntime=1):
        a, b = C.c_double(presg), C.c_double(presg3)
        self.L.c3o_getup(self.h, _p(U), _p(P), _p(HH), _p(SGSVIS), C.byref(a), C.byref(b), ntime)
        return a.value, b.value

    @classmethod
    def ctdma(cls, b, d):
        d = np.ascontiguousarray(d, dtype=np.float64)
        x = np.zeros_like(d)
        cls.lib().c3o_ctdma(len(d), b, _p(d), _p(x))
        return x

    @classmethod
    def dft2d(cls, a, sign):
        """complex 2-D DFT of a (n1, n3) column-major complex array"""
        n1, n3 = a.shape
        flat = np.asarray(a, dtype=np.complex128).reshape(-1, order="F")
        buf = np.empty(2 * flat.size)
        buf[0::2], buf[1::2] = flat.real, flat.imag
        cls.lib().c3o_dft2d(n1, n3, buf.ctypes.data_as(_dp), sign)
        return (buf[0::2] + 1j * buf[1::2]).reshape((n1, n3), order="F")

    @staticmethod
    def set_threads(n):
        Oracle.lib().c3o_set_threads(int(n))


class RefLib:
    """The reference's own hot-path subroutines (compact3d.f, unmodified),
    built by `make -C oracle ref` with M1=M2=M3 fixed by the reference's
    PARAM.H.  Only exists in the build container."""

    _lib = None

    @classmethod
    def lib(cls):
        if cls._lib is None:
            L = C.CDLL(REF_SO)
            # zero the stack below the caller before every reference routine
            # (see ref_stack.c): reads outside local arrays then return 0
            zs = L.refh_zerostack_

            class _Z:
                def __init__(self, lib):
                    self._l = lib

                def __getattr__(self, name):
                    f = getattr(self._l, name)
                    if not name.endswith("_") or name.startswith("refh_"):
                        return f

                    def call(*a):
                        zs()
                        return f(*a)
                    return call
            cls._lib = _Z(L)
        return cls._lib

    @staticmethod
    def param_h():
        # PARAM.H holds the compile-time bounds the library was built with
        path = os.path.join(HERE, "_ref", "PARAM.M")
        if os.path.exists(path):
            return tuple(int(x) for x in open(path).read().split())
        return (33, 33, 33)

    def __init__(self, re, alx, aly, alz, dt, files=0.0, wmodel=0.0):
        self.L = self.lib()
        self.n1, self.n2, self.n3 = self.param_h()
        self.files, self.wmodel = files, wmodel
        d = lambda v: C.byref(C.c_double(v))
        i = lambda v: C.byref(C.c_int(v))
        # MESH writes GRID.PLT into the working directory: run it in a scratch dir
        cwd = os.getcwd()
        with tempfile.TemporaryDirectory() as td:
            os.chdir(td)
            try:
                self.L.refh_setup_(i(self.n1), i(self.n2), i(self.n3), d(re), d(alx), d(aly), d(alz),
                                   d(dt), d(files), d(wmodel))
            finally:
                os.chdir(cwd)

    def set_dt(self, dt):
        self.L.refh_setdt_(C.byref(C.c_double(dt)))

    def set_ifwall(self, v):
        self.L.refh_setifwall_(C.byref(C.c_int(int(v))))

    def mesh(self):
        M1, M2, M3 = self.n1, self.n2, self.n3
        m3mh = (M3 - 1) // 2 + 1
        arrs = dict(Y=np.zeros(M2 + 1), DY=np.zeros(M2 + 1))
        for n in ("H", "HM", "HC", "HP", "DYM", "DYC", "DYP", "PMJ", "PCJ", "PPJ"):
            arrs[n] = np.zeros(M2 + 1)  # index 0 unused (Fortran 1-based)
        arrs["AK1"] = np.zeros(M1 + 1)
        arrs["AK3"] = np.zeros(m3mh + 1)
        sc = np.zeros(6)
        o = lambda a: _p(a[1:]) if True else None
        self.L.refh_getmesh_(_p(arrs["Y"]), _p(arrs["DY"]), o(arrs["H"]), o(arrs["HM"]), o(arrs["HC"]),
                             o(arrs["HP"]), o(arrs["DYM"]), o(arrs["DYC"]), o(arrs["DYP"]), o(arrs["AK1"]),
                             o(arrs["AK3"]), o(arrs["PMJ"]), o(arrs["PCJ"]), o(arrs["PPJ"]), _p(sc))
        for n, v in zip(("DX1", "DX1Q", "DX3", "DX3Q", "ALY", "VOL"), sc):
            arrs[n] = float(v)
        return arrs

    def get_tau(self):
        tw = np.zeros((self.n1 + 1, self.n3 + 1, 2), order="F")
        tb = np.zeros_like(tw)
        self.L.refh_gettau_(_p(tw), _p(tb))
        return tw, tb

    def set_tau(self, tw, tb):
        self.L.refh_settau_(_p(np.asfortranarray(tw)), _p(np.asfortranarray(tb)))

    def get_wsm(self):
        w = np.zeros(4)
        self.L.refh_getwsm_(_p(w))
        return w

    def set_wsm(self, w):
        self.L.refh_setwsm_(_p(np.ascontiguousarray(w, dtype=np.float64)))

    def get_ubc(self):
        u = np.zeros((self.n1 + 1, self.n3 + 1, 3, 2), order="F")
        self.L.refh_getubc_(_p(u))
        return u

    # ---- reference subroutines (argument lists as in compact3d.f)
    @staticmethod
    def _i(v):
        return C.byref(C.c_int(v))

    def bcond(self, U, SGSVIS, ntime=1):
        self.L.bcond_(_p(U), _p(SGSVIS), self._i(ntime))

    def rhs(self, comp, U, P, HH, HH2, RUH, presg, SGSVIS, ntime=1):
        SS = np.zeros_like(HH)
        pg = C.byref(C.c_double(presg))
        if comp == 1:
            self.L.rhs1_(_p(U), _p(P), _p(HH), _p(HH2), _p(SS), _p(SS), _p(RUH), pg, _p(SGSVIS), self._i(ntime))
        elif comp == 2:
            self.L.rhs2_(_p(U), _p(P), _p(HH), _p(HH2), _p(SS), _p(SS), _p(RUH), _p(SGSVIS), self._i(ntime))
        else:
            self.L.rhs3_(_p(U), _p(P), _p(HH), _p(HH2), _p(SS), _p(SS), _p(RUH), pg, _p(SGSVIS), self._i(ntime))

    def getuh(self, comp, U, UH, RUH, SGSVIS, ntime=1):
        f = {1: self.L.getuh1_, 2: self.L.getuh2_, 3: self.L.getuh3_}[comp]
        f(_p(U), _p(UH), _p(RUH), _p(SGSVIS), self._i(ntime))

    def rhsdp(self, RDP, UH, U, ntime=1):
        self.L.rhsdp_(_p(RDP), _p(UH), _p(U), self._i(ntime))

    def upcalc(self, U, P, UH, DP, presg, presg3, ntime=1):
        a, b = C.c_double(presg), C.c_double(presg3)
        self.L.upcalc_(_p(U), _p(P), _p(UH), _p(DP), C.byref(a), C.byref(b), self._i(ntime))
        return a.value, b.value

    def divcheck(self, U, ntime=1):
        t, dm = C.c_double(0.0), C.c_double(0.0)
        self.L.divcheck_(_p(U), C.byref(t), C.byref(dm), self._i(ntime))
        return dm.value

    def cfl(self, U):
        c = C.c_double(0.0)
        self.L.cfl_(_p(U), C.byref(c))
        return c.value

    def wallss(self, U, P):
        """the reference's WALLSS; returns WSM(1:3)"""
        t = C.c_double(0.0)
        self.L.wallss_(_p(U), _p(P), C.byref(t))
        return self.get_wsm()[:3].copy()

    def chkmf(self, U):
        """the reference's CHKMF (run in a scratch directory: it writes
        TIME, FLOW1, FLOW3 to unit 32); returns FLOW1, FLOW3 of COMMON
        /FLOWRATE/"""
        t = C.c_double(0.0)
        cwd = os.getcwd()
        with tempfile.TemporaryDirectory() as td:
            os.chdir(td)
            try:
                self.L.chkmf_(_p(U), C.byref(t))
            finally:
                os.chdir(cwd)
        f = np.zeros(2)
        self.L.refh_getflow_(_p(f))
        return float(f[0]), float(f[1])

    def uhcalc(self, U, UH, P, HH, HH2, presg, presg3, SGSVIS, ntime=1):
        """the reference's UHCALC (compact3d.f:1004-1036): RHS1-3 then GETUH1-3"""
        SS = np.zeros_like(HH)
        SS2 = np.zeros_like(HH)
        a, b = C.c_double(presg), C.c_double(presg3)
        self.L.uhcalc_(_p(U), _p(UH), _p(P), _p(HH), _p(HH2), _p(SS), _p(SS2), C.byref(a), C.byref(b),
                       _p(SGSVIS), self._i(ntime))

    def reset_commons(self):
        """zeroes the COMMON state that routines read from earlier calls:
        TAUW/TAUB (/TAUXYZ/), WSM and WSMO (/WSMEM/), TAOWO and the
        integral wall model's planes and wm_dt, so each case starts from
        the same state whatever ran before it in this process"""
        tw, tb = self.get_tau()
        self.set_tau(np.zeros_like(tw), np.zeros_like(tb))
        self.set_wsm([0.0, 0.0, 0.0])
        self.L.refh_setwsmo_(_p(np.zeros(3)), C.byref(C.c_double(0.0)))
        self.iwm_set(np.zeros((IWM_NPL, self.n3 - 1, self.n1 - 1)), 0.0)
        self.set_ifwall(0)

    def sgscs(self, U, SGSVIS, ntime=1):
        self.L.sgscs_(_p(U), _p(SGSVIS), self._i(ntime))

    def sgsdy(self, U, SGSVIS, ntime=1):
        """the reference's SGSDY; returns SMAGC(1:M2) as C(J) at [J], J = 0..N2"""
        self.L.sgsdy_(_p(U), _p(SGSVIS), self._i(ntime))
        c = np.zeros(self.n2)
        self.L.refh_getsmagc_(_p(c))
        return np.concatenate([[0.0], c])

    def writeup(self, U, P, presg, presg3, ntime, wsmo=(0.0, 0.0, 0.0), taowo=0.0):
        """the reference's WRITEUP, run in a scratch directory; returns the
        text of RESULT/FLOW-*.plt and RESULT/STRAIN-*.plt"""
        self.L.refh_setwsmo_(_p(np.ascontiguousarray(wsmo, dtype=np.float64)), C.byref(C.c_double(taowo)))
        a, b = C.c_double(presg), C.c_double(presg3)
        cwd = os.getcwd()
        with tempfile.TemporaryDirectory() as td:
            os.makedirs(os.path.join(td, "RESULT"))
            os.chdir(td)
            try:
                self.L.writeup_(_p(U), _p(P), C.byref(a), C.byref(b), self._i(ntime))
            finally:
                os.chdir(cwd)
            name = "%06d.plt" % (100000 + ntime)
            flow = open(os.path.join(td, "RESULT", "FLOW-" + name)).read()
            strain = open(os.path.join(td, "RESULT", "STRAIN-" + name)).read()
        return flow, strain

    def sgssd(self, U, SGSVIS, ntime=1):
        """the reference's SGSSD; returns SMAGC(1:M2) as C(J) at [J], J = 0..N2"""
        self.L.sgssd_(_p(U), _p(SGSVIS), self._i(ntime))
        c = np.zeros(self.n2)
        self.L.refh_getsmagc_(_p(c))
        return np.concatenate([[0.0], c])

    def coefficient(self, s1, s2):
        """the reference's COEFFICIENT: PA(1:5) from A1..E1, A2..E2"""
        d = [C.byref(C.c_double(float(v))) for v in list(s1) + list(s2)]
        pa = np.zeros(5)
        self.L.coefficient_(*d, _p(pa))
        return pa

    def dqrrt(self, pa, eps=1e-5):
        """the reference's DQRRT on PA(1:N): BETA"""
        pa = np.ascontiguousarray(pa, dtype=np.float64)
        beta = C.c_double(0.0)
        self.L.dqrrt_(_p(pa), self._i(len(pa)), C.byref(beta), C.byref(C.c_double(eps)))
        return beta.value

    def wallmodel_instant(self, U, SGSVIS):
        tw, tb = self.get_tau()
        self.L.wallmodel_instant_(_p(U), _p(tw), _p(tb), _p(SGSVIS))
        self.set_tau(tw, tb)

    def set_cflmax(self, v):
        self.L.refh_setcflmax_(C.byref(C.c_double(float(v))))

    def iwmles(self, U, P, presg, presg3, ntime):
        """WMODEL = 2: the reference's iwm_* routines under the harness's
        driver (run in a scratch directory: iwm_monitor_smooth appends to
        iwm_track.dat)"""
        a, b = C.c_double(presg), C.c_double(presg3)
        cwd = os.getcwd()
        with tempfile.TemporaryDirectory() as td:
            os.chdir(td)
            try:
                self.L.refh_iwmles_(_p(U), _p(P), C.byref(a), C.byref(b), self._i(ntime))
            finally:
                os.chdir(cwd)

    def iwm_state(self):
        """the 43 planes of COMMON /IWMLES/ as [plane, K-1, I-1] and wm_dt"""
        n1m, n3m = self.n1 - 1, self.n3 - 1
        out = np.zeros((n1m, n3m, IWM_NPL), order="F")
        dt = C.c_double(0.0)
        self.L.refh_getiwm_(_p(out), C.byref(dt))
        return np.ascontiguousarray(out.transpose(2, 1, 0)), dt.value

    def iwm_set(self, planes, wm_dt):
        a = np.asfortranarray(np.asarray(planes, dtype=np.float64).transpose(2, 1, 0))
        self.L.refh_setiwm_(_p(a), C.byref(C.c_double(float(wm_dt))))

    def wallmodel_werner(self, U, SGSVIS, ntime=1):
        tw, tb = self.get_tau()
        self.L.wallmodel_werner_(_p(U), _p(tw), _p(tb), _p(SGSVIS), self._i(ntime))
        self.set_tau(tw, tb)

    def strain(self, U, j):
        SR = np.zeros((self.n1, self.n3, 6), order="F")
        self.L.strain_(_p(U), _p(SR), self._i(j))
        return SR

    def filter(self, FILV, four=False):
        n = FILV.shape[2]
        (self.L.filter4_ if four else self.L.filter_)(_p(FILV), self._i(0), self._i(n))

    def ctdma11(self, b, d):
        n = self.n1 - 1
        A = np.ones(n)
        B = np.full(n, float(b))
        Cc = np.ones(n)
        D = np.ascontiguousarray(d, dtype=np.float64)
        X = np.zeros(n)
        self.L.ctdma11_(_p(A), _p(B), _p(Cc), _p(D), _p(X))
        return X

    def tdmap(self, a, b, c, r):
        n = len(r)
        x = np.zeros(self.n2)
        pad = lambda v: np.concatenate([np.asarray(v, float), np.zeros(self.n2 - n)])
        A, B, Cc, R = pad(a), pad(b), pad(c), pad(r)
        self.L.tdmap_(_p(A), _p(B), _p(Cc), _p(R), self._i(n), _p(x))
        return x[:n]
