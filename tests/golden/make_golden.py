#!/usr/bin/env python3
"""Generate golden vectors from the reference's own subroutines.

Runs the hot-path subroutines of compact3d.f, compiled unmodified into
oracle/_ref/libc3dref.so by `make -C oracle ref` (the reference grid of its
PARAM.H, 32 x 33 x 32), on seeded inputs, and stores a deterministic subsample
of every output plus full-field checksums in tests/golden/ref_<case>.npz.

GETDP needs FFTW, which this image does not have, so the reference cannot run
it; the fixture's UPCALC input DP is the oracle's GETDP of the reference's RDP
(GETDP itself is pinned by closed forms in tests/test_oracle_math.py).

WRITEUP's Tecplot text is stored as checksums and sample lines
(ref_writeup.json).

Every case first zeroes the COMMON state that routines carry between calls
(RefLib.reset_commons), so a fixture does not depend on which cases ran
before it in the same process.

Usage:  python tests/golden/make_golden.py [--out DIR] [case ...]   (build container only)
"""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

from cases import inputs, iwm_inputs  # noqa: E402
from pyoracle import Oracle, RefLib  # noqa: E402

STRIDE = 13

CASES = {
    "dns": dict(re=4200.0, alx=2 * math.pi, aly=2.0, alz=math.pi, dt=1e-3, files=0.0, wmodel=0.0),
    "les": dict(re=5200.0, alx=2 * math.pi, aly=2.0, alz=math.pi, dt=5e-4, files=2.0, wmodel=1.0),
    "dyn": dict(re=5200.0, alx=2 * math.pi, aly=2.0, alz=math.pi, dt=5e-4, files=1.0, wmodel=1.0),
    "sd": dict(re=5200.0, alx=2 * math.pi, aly=2.0, alz=math.pi, dt=5e-4, files=3.0, wmodel=1.0),
    "wer": dict(re=5200.0, alx=2 * math.pi, aly=2.0, alz=math.pi, dt=5e-4, files=2.0, wmodel=4.0),
}


def sub(a, stride=STRIDE):
    return np.ascontiguousarray(a, dtype=np.float64).reshape(-1, order="F")[::stride].copy()


def summ(a):
    a = np.asarray(a, dtype=np.float64)
    return np.array([a.sum(), (a * a).sum(), np.abs(a).max()])


def run_case(name, cfg):
    ref = RefLib(**cfg)
    ref.reset_commons()
    n1, n2, n3 = ref.n1, ref.n2, ref.n3
    orc = Oracle(n1, n2, n3, cks=2.5, **cfg)  # inputs only; outputs come from ref
    U, P, HH2 = inputs(orc, seed=7)
    out = {}
    presg, presg3 = 0.003, -0.001
    S = np.zeros(U.shape[:3], order="F")
    if cfg["files"] == 2.0:
        ref.set_wsm([2.5e-3, 0.0, 4.0e-4])
        ref.sgscs(U, S)
        out["SGSVIS"] = S
        out["STRAIN3"] = ref.strain(U, 3)
    if cfg["files"] == 1.0:
        out["SMAGC"] = ref.sgsdy(U, S)
        out["SGSVIS"] = S
    if cfg["files"] == 3.0:
        out["SMAGC"] = ref.sgssd(U, S)
        out["SGSVIS"] = S
        # the quintic's root selection (DQRRT) on seeded coefficients
        rq = np.random.default_rng(31)
        pa = rq.standard_normal((16, 5)) * np.array([1.0, 3.0, 5.0, 3.0, 1.0])
        out["DQRRT_PA"] = pa
        out["DQRRT_BETA"] = np.array([ref.dqrrt(q) for q in pa])
        s1, s2 = rq.standard_normal(5), rq.standard_normal(5)
        out["COEF_IN"] = np.concatenate([s1, s2])
        out["COEF_PA"] = ref.coefficient(s1, s2)
    if cfg["files"] != 0.0 and cfg["wmodel"] in (1.0, 4.0):
        if cfg["wmodel"] == 1.0:
            ref.wallmodel_instant(U, S)
        else:
            ref.wallmodel_werner(U, S)
        ref.set_ifwall(1)
        tw, tb = ref.get_tau()
        out["TAUW"], out["TAUB"] = tw, tb
        out["WSM"] = ref.get_wsm()
    ref.bcond(U, S)
    out["U_BCOND"] = U.copy(order="F")
    if cfg["files"] != 0.0:
        out["SGSVIS_BCOND"] = S.copy(order="F")
    HH = np.zeros_like(U)
    RUH = [np.zeros(U.shape[:3], order="F") for _ in range(3)]
    for c in (1, 2, 3):
        ref.rhs(c, U, P, HH, HH2, RUH[c - 1], presg if c == 1 else presg3, S)
        out[f"RUH{c}"] = RUH[c - 1]
    out["HH"] = HH
    UH = np.zeros_like(U)
    for c in (1, 2, 3):
        ref.getuh(c, U, UH, RUH[c - 1], S)
    out["UH"] = UH
    RDP = np.zeros((n1, n2, n3), order="F")
    ref.rhsdp(RDP, UH, U)
    out["RDP"] = RDP
    UHb = UH.copy(order="F")
    DP = np.zeros_like(RDP)
    orc.getdp(DP, RDP)
    Pn = P.copy(order="F")
    pg, pg3 = ref.upcalc(U, Pn, UH, DP, presg, presg3)
    out["U_UPCALC"] = U
    out["P_UPCALC"] = Pn
    out["PRESG"] = np.array([pg, pg3])
    out["DIVMAX"] = np.array([ref.divcheck(U)])
    out["CFL"] = np.array([ref.cfl(U)])
    # the diagnostics MAIN calls after the step: CHKMF (with its sums over
    # U(1,J,K,1) and U(I,J,1,3)) and WALLSS (WSM(1:3))
    out["CHKMF"] = np.array(ref.chkmf(U))
    out["WALLSS"] = ref.wallss(U, Pn)
    # DIVCHECK and CFL of the unprojected UH: an O(1) divergence
    out["DIVMAX_UH"] = np.array([ref.divcheck(UHb)])
    out["CFL_UH"] = np.array([ref.cfl(UHb)])
    # line solvers and filters
    rng = np.random.default_rng(99)
    d = rng.uniform(-1, 1, n1 - 1)
    for b in (6.0, 10.0, 22.0):
        out[f"CTDMA11_b{int(b)}"] = ref.ctdma11(b, d)
    out["CTDMA_RHS"] = d
    a, bb, c3, r = (rng.uniform(0.2, 1, n2 - 1), rng.uniform(-4, -3, n2 - 1), rng.uniform(0.2, 1, n2 - 1),
                    rng.uniform(-1, 1, n2 - 1))
    out["TDMAP_IN"] = np.stack([a, bb, c3, r])
    out["TDMAP"] = ref.tdmap(a, bb, c3, r)
    F = np.zeros((n1, n3, 2), order="F")
    F[: n1 - 1, : n3 - 1, :] = rng.uniform(-1, 1, (n1 - 1, n3 - 1, 2))
    out["FILTER_IN"] = F.copy(order="F")
    F1 = F.copy(order="F")
    ref.filter(F1, four=False)
    out["FILTER"] = F1
    F4 = F.copy(order="F")
    ref.filter(F4, four=True)
    out["FILTER4"] = F4
    store = {"config": np.array([n1, n2, n3, cfg["re"], cfg["alx"], cfg["aly"], cfg["alz"], cfg["dt"],
                                 cfg["files"], cfg["wmodel"], 7, presg, presg3, STRIDE])}
    small = {"PRESG", "DIVMAX", "CFL", "WSM", "DIVMAX_UH", "CFL_UH", "CHKMF", "WALLSS", "SMAGC", "CTDMA_RHS",
             "TDMAP_IN", "TDMAP", "FILTER_IN", "FILTER", "FILTER4", "DQRRT_PA", "DQRRT_BETA", "COEF_IN", "COEF_PA"} | {k for k in out if k.startswith("CTDMA11")}
    for k, v in out.items():
        if k in small:
            store[k] = np.asarray(v, dtype=np.float64)
        else:
            store[k + "__sub"] = sub(v)
            store[k + "__sum"] = summ(v)
    path = os.path.join(OUT, f"ref_{name}.npz")
    np.savez_compressed(path, **store)
    print(path, os.path.getsize(path), "bytes")


IWM = dict(re=5200.0, alx=2 * math.pi, aly=2.0, alz=math.pi, dt=5e-4, files=2.0, wmodel=2.0)
IWM_CFLMAX = 0.5


def iwm_drive(lib, U0, P0, first, last):
    """calls NTIME = first..last of the WMODEL = 2 wall model with seeded
    variations of U and P; yields (NTIME, U, P) after each call"""
    for nt in range(first, last + 1):
        U = np.asfortranarray(U0 * (1 + 0.01 * math.sin(nt)))
        P = np.asfortranarray(P0 * (1 + 0.1 * nt))
        yield nt, U, P


def run_iwm():
    """WMODEL = 2 (IWMLES, smooth wall): NTIME = 1..4 from the reference's
    own initialisation (the first compute step, NTIME = 5, reads an
    undefined local of the reference there and is not stored), then
    NTIME = 5..15 from a log-law-consistent state (tests/cases.py
    iwm_inputs), which contains three compute steps."""
    ref = RefLib(**IWM)
    ref.reset_commons()
    n1, n3 = ref.n1, ref.n3
    orc = Oracle(n1, ref.n2, n3, cks=2.5, **IWM)
    ref.set_cflmax(IWM_CFLMAX)
    out = {}
    U0, P0, _ = inputs(orc, seed=7)
    for nt, U, P in iwm_drive(ref, U0, P0, 1, 4):
        ref.iwmles(U, P, 0.003, -0.001, nt)
    pl, wdt = ref.iwm_state()
    tw, tb = ref.get_tau()
    out["INIT_PLANES"], out["INIT_TAUW"], out["INIT_TAUB"] = pl, tw, tb
    out["INIT_WSM"] = np.array([ref.get_wsm()[0], wdt])
    dy = orc.mesh()["DY"]
    U0, P0, pl0, wdt0 = iwm_inputs(n1, n3, dy, IWM["re"], IWM["dt"], seed=1)
    ref.iwm_set(pl0, wdt0)
    wsm = []
    for nt, U, P in iwm_drive(ref, U0, P0, 5, 15):
        ref.iwmles(U, P, 0.003, -0.001, nt)
        wsm.append(ref.get_wsm()[0])
        if nt % 5 == 0:
            pl, wdt = ref.iwm_state()
            out[f"PLANES_{nt}"] = pl
            out[f"WDT_{nt}"] = np.array([wdt])
    tw, tb = ref.get_tau()
    out["TAUW"], out["TAUB"] = tw, tb
    out["WSM"] = np.array(wsm)
    store = {"config": np.array([n1, ref.n2, n3, IWM["re"], IWM["alx"], IWM["aly"], IWM["alz"], IWM["dt"],
                                 IWM["files"], IWM["wmodel"], IWM_CFLMAX, STRIDE])}
    for k, v in out.items():
        if k.startswith("WSM") or k.startswith("INIT_WSM") or k.startswith("WDT"):
            store[k] = np.asarray(v, dtype=np.float64)
        elif "PLANES" in k:
            store[k + "__sub"] = np.stack([sub(p) for p in v])
            store[k + "__sum"] = np.stack([summ(p) for p in v])
        else:
            store[k + "__sub"] = sub(v)
            store[k + "__sum"] = summ(v)
    path = os.path.join(OUT, "ref_iwm.npz")
    np.savez_compressed(path, **store)
    print(path, os.path.getsize(path), "bytes")


GETUP_CASES = {
    "dns": CASES["dns"],
    "les": CASES["les"],
    "wer": CASES["wer"],
    # FILES != 0 with WMODEL = 0: WALL_MODEL calls WALLSS, whose WSM feeds
    # the next step's SGSCS
    "lesw0": dict(re=5200.0, alx=2 * math.pi, aly=2.0, alz=math.pi, dt=5e-4, files=2.0, wmodel=0.0),
}
GETUP_STEPS = 3
GETUP_STRIDE = 53


def ref_getup(ref, orc, U, P, HH, S, presg, presg3, ntime, cfg):
    """one GETUP (compact3d.f:879-960) from the reference's own units: the
    prologue's clears and AB2 shift HH2 <- HH, HH <- 0 (:895-923), the SGS
    model (:925-927), WALL_MODEL's dispatch (:1981-2008), BCOND, UHCALC
    (RHS1-3 then GETUH1-3, :1004-1036), RHSDP, GETDP (the oracle's: FFTW
    is absent here) and UPCALC.  U, P, HH, S are updated in place;
    returns PRESG, PRESG3."""
    n1, n2, n3 = ref.n1, ref.n2, ref.n3
    HH2 = HH.copy(order="F")
    HH[...] = 0.0
    S[...] = 0.0
    UH = np.zeros_like(U)
    if cfg["files"] == 1.0:
        ref.sgsdy(U, S, ntime)
    elif cfg["files"] == 2.0:
        ref.sgscs(U, S, ntime)
    elif cfg["files"] == 3.0:
        ref.sgssd(U, S, ntime)
    if cfg["files"] != 0.0:
        if cfg["wmodel"] == 0.0:
            ref.set_ifwall(0)
            ref.wallss(U, P)
        elif cfg["wmodel"] == 1.0:
            ref.wallmodel_instant(U, S)
            ref.set_ifwall(1)
        elif cfg["wmodel"] == 4.0:
            ref.wallmodel_werner(U, S, ntime)
            ref.set_ifwall(1)
        else:
            raise ValueError(cfg["wmodel"])
    ref.bcond(U, S, ntime)
    ref.uhcalc(U, UH, P, HH, HH2, presg, presg3, S, ntime)
    RDP = np.zeros((n1, n2, n3), order="F")
    ref.rhsdp(RDP, UH, U, ntime)
    DP = np.zeros_like(RDP)
    orc.getdp(DP, RDP)
    return ref.upcalc(U, P, UH, DP, presg, presg3, ntime)


def run_getup():
    """GETUP_STEPS steps of GETUP composed from the reference's units
    (ref_getup) per GETUP_CASES entry, from the seeded inputs with a random
    previous HH; stores U, P, HH and SGSVIS after steps 1 and GETUP_STEPS,
    PRESG/PRESG3 and WSM after each step"""
    store = {"steps": np.array([GETUP_STEPS, GETUP_STRIDE])}
    for name, cfg in GETUP_CASES.items():
        ref = RefLib(**cfg)
        ref.reset_commons()
        orc = Oracle(ref.n1, ref.n2, ref.n3, cks=2.5, **cfg)
        U, P, HH = inputs(orc, seed=11)
        S = np.zeros(U.shape[:3], order="F")
        presg, presg3 = 0.003, -0.001
        if cfg["files"] != 0.0:
            ref.set_wsm([2.5e-3, 0.0, 4.0e-4])
        pg = []
        for nt in range(1, GETUP_STEPS + 1):
            presg, presg3 = ref_getup(ref, orc, U, P, HH, S, presg, presg3, nt, cfg)
            pg.append([presg, presg3] + list(ref.get_wsm()[:3]))
            if nt in (1, GETUP_STEPS):
                for k, v in (("U", U), ("P", P), ("HH", HH), ("SGSVIS", S)):
                    store[f"{name}_{k}{nt}__sub"] = sub(v, GETUP_STRIDE)
                    store[f"{name}_{k}{nt}__sum"] = summ(v)
        store[f"{name}_PRESG"] = np.array(pg)
    path = os.path.join(OUT, "ref_getup.npz")
    np.savez_compressed(path, **store)
    print(path, os.path.getsize(path), "bytes")


def run_writeup():
    """WRITEUP's FLOW and STRAIN files for the writeup_inputs() field:
    sha256 of each text, its line count and a few of its lines"""
    import hashlib
    import json
    from cases import writeup_inputs
    cfg = CASES["dns"]
    ref = RefLib(**cfg)
    ref.reset_commons()
    orc = Oracle(ref.n1, ref.n2, ref.n3, cks=2.5, **cfg)
    U, P, tauw = writeup_inputs(orc)
    tw, tb = ref.get_tau()
    tw[1:ref.n1, 1:ref.n3, 0] = tauw.T
    ref.set_tau(tw, tb)
    out = {"ntime": 4321, "wsmo1": 2.7e-3}
    flow, strain = ref.writeup(U, P, 0.003, -0.001, out["ntime"], wsmo=(out["wsmo1"], 0.0, 0.0))
    for name, text in (("flow", flow), ("strain", strain)):
        lines = text.splitlines(keepends=True)
        out[name] = {"sha256": hashlib.sha256(text.encode()).hexdigest(), "lines": len(lines),
                     "head": lines[:4], "special": lines[2:2 + 16] if name == "flow" else []}
    path = os.path.join(OUT, "ref_writeup.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(path, os.path.getsize(path), "bytes")


OUT = HERE

if __name__ == "__main__":
    # optional case names: dns les ... writeup getup iwm; --out DIR writes the
    # fixtures there instead of tests/golden
    args = sys.argv[1:]
    if "--out" in args:
        i = args.index("--out")
        OUT = args[i + 1]
        del args[i:i + 2]
        os.makedirs(OUT, exist_ok=True)
    only = set(args)
    for name, cfg in CASES.items():
        if not only or name in only:
            run_case(name, cfg)
    if not only or "writeup" in only:
        run_writeup()
    if not only or "getup" in only:
        run_getup()
    if not only or "iwm" in only:
        run_iwm()
