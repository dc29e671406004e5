"""The oracle against golden vectors produced by the reference's own
subroutines (tests/golden/make_golden.py; compact3d.f compiled unmodified).

Every phase the reference can run without FFTW is compared: BCOND, RHS1-3,
GETUH1-3, RHSDP, UPCALC, DIVCHECK and CFL (of the projected field and of the
unprojected UH), CHKMF, WALLSS, GETUP's composition over three steps,
SGSCS/STRAIN, SGSDY (the dynamic
model, its per-plane coefficient included), SGSSD (the scale-dependent
model) with COEFFICIENT and the DQRRT/CHHQR root selection,
WALLMODEL_INSTANT, WALLMODEL_WERNER (both walls), CTDMA11, TDMAP, FILTER,
FILTER4.  The oracle follows the reference's operation
order, so the expected agreement is exact; the bound allows 4 ulps."""
import os

import numpy as np
import pytest

from cases import inputs

HERE = os.path.dirname(os.path.abspath(__file__))
TOL = 1e-15


def _load(name):
    return dict(np.load(os.path.join(HERE, "golden", f"ref_{name}.npz")))


def _sub(a, stride):
    return np.ascontiguousarray(a, dtype=np.float64).reshape(-1, order="F")[::stride]


def _check(fx, key, arr, stride):
    ref = fx[key + "__sub"]
    got = _sub(arr, stride)
    scale = max(np.max(np.abs(ref)), 1e-300)
    err = np.max(np.abs(ref - got)) / scale
    s = np.array([arr.sum(), (arr * arr).sum(), np.abs(arr).max()])
    es = np.max(np.abs(s - fx[key + "__sum"]) / np.maximum(np.abs(fx[key + "__sum"]), 1e-300))
    assert err <= TOL, (key, err)
    assert es <= 1e-12, (key, es)


@pytest.mark.parametrize("name", ["dns", "les", "dyn", "sd", "wer"])
def test_oracle_matches_reference_fixture(name):
    from pyoracle import Oracle
    fx = _load(name)
    c = fx["config"]
    n1, n2, n3 = int(c[0]), int(c[1]), int(c[2])
    re, alx, aly, alz, dt, files, wmodel = (float(x) for x in c[3:10])
    seed, presg, presg3, stride = int(c[10]), float(c[11]), float(c[12]), int(c[13])
    o = Oracle(n1, n2, n3, re, alx, aly, alz, dt, files, wmodel, 2.5)
    U, P, HH2 = inputs(o, seed=seed)
    S = o.zeros_u(0)
    if files == 2.0:
        o.set_wsm([2.5e-3, 0.0, 4.0e-4])
        o.sgscs(U, S)
        _check(fx, "SGSVIS", S, stride)
    if files == 1.0:
        smagc = o.sgsdy(U, S)
        _check(fx, "SGSVIS", S, stride)
        assert np.max(np.abs(smagc - fx["SMAGC"])) <= TOL * np.max(np.abs(fx["SMAGC"]))
    if files == 3.0:
        smagc, _ = o.sgssd(U, S)
        _check(fx, "SGSVIS", S, stride)
        assert np.max(np.abs(smagc - fx["SMAGC"])) <= TOL * np.max(np.abs(fx["SMAGC"]))
        for pa, beta in zip(fx["DQRRT_PA"], fx["DQRRT_BETA"]):
            assert Oracle.dqrrt(pa) == beta
        assert np.array_equal(Oracle.sd_coefficient(fx["COEF_IN"][:5], fx["COEF_IN"][5:]), fx["COEF_PA"])
    if files != 0.0 and wmodel in (1.0, 4.0):
        if wmodel == 1.0:
            o.wallmodel_instant(U)
        else:
            o.wallmodel_werner(U)
        o.set_ifwall(1)
        tw, tb = o.get_tau()
        _check(fx, "TAUW", tw, stride)
        _check(fx, "TAUB", tb, stride)
        assert abs(o.get_wsm()[0] - fx["WSM"][0]) <= TOL * abs(fx["WSM"][0])
    o.bcond(U, S)
    _check(fx, "U_BCOND", U, stride)
    HH = o.zeros_u()
    RUH = [o.zeros_u(0) for _ in range(3)]
    for comp in (1, 2, 3):
        o.rhs(comp, U, P, HH, HH2, RUH[comp - 1], presg if comp == 1 else presg3, S)
        _check(fx, f"RUH{comp}", RUH[comp - 1], stride)
    _check(fx, "HH", HH, stride)
    UH = o.zeros_u()
    for comp in (1, 2, 3):
        o.getuh(comp, U, UH, RUH[comp - 1])
    _check(fx, "UH", UH, stride)
    RDP = o.zeros_p()
    o.rhsdp(RDP, UH)
    _check(fx, "RDP", RDP, stride)
    UHb = UH.copy(order="F")
    DP = o.zeros_p()
    o.getdp(DP, RDP)
    Pn = P.copy(order="F")
    pg, pg3 = o.upcalc(U, Pn, UH, DP, presg, presg3)
    _check(fx, "U_UPCALC", U, stride)
    _check(fx, "P_UPCALC", Pn, stride)
    assert np.max(np.abs(np.array([pg, pg3]) - fx["PRESG"])) <= TOL * np.max(np.abs(fx["PRESG"]))
    assert o.divcheck(U) == pytest.approx(fx["DIVMAX"][0], rel=1e-12, abs=1e-15)
    assert o.cfl(U) == pytest.approx(fx["CFL"][0], rel=1e-15)
    # CHKMF (its U(1,J,K,1) / U(I,J,1,3) sums) and WALLSS of the stepped field
    f1, f3 = o.chkmf(U)
    assert abs(f1 - fx["CHKMF"][0]) <= TOL * abs(fx["CHKMF"][0])
    assert abs(f3 - fx["CHKMF"][1]) <= TOL * max(abs(fx["CHKMF"][1]), abs(fx["CHKMF"][0]))
    o.wallss(U)
    wsm = o.get_wsm()[:3]
    assert np.max(np.abs(wsm - fx["WALLSS"])) <= TOL * np.max(np.abs(fx["WALLSS"])), (wsm, fx["WALLSS"])
    # DIVCHECK and CFL of the unprojected UH, whose divergence is O(1)
    assert fx["DIVMAX_UH"][0] > 1e-3
    assert abs(o.divcheck(UHb) - fx["DIVMAX_UH"][0]) <= TOL * fx["DIVMAX_UH"][0]
    assert abs(o.cfl(UHb) - fx["CFL_UH"][0]) <= TOL * fx["CFL_UH"][0]


@pytest.mark.parametrize("name", ["dns", "les", "wer", "lesw0"])
def test_oracle_getup_matches_reference_composition(name):
    """c3o_getup over three steps against GETUP composed from the
    reference's own units (make_golden.ref_getup): pins the step's
    composition -- the AB2 shift HH2 <- HH of the previous step, the clears,
    the model dispatch and the state carried between steps (WSM, TAUW/TAUB,
    PRESG) -- which the per-phase fixtures do not"""
    import sys
    sys.path.insert(0, HERE + "/golden")
    from make_golden import GETUP_CASES
    from pyoracle import Oracle
    fx = _load("getup")
    nsteps, stride = (int(x) for x in fx["steps"])
    cfg = GETUP_CASES[name]
    o = Oracle(33, 33, 33, cks=2.5, **cfg)
    U, P, HH = inputs(o, seed=11)
    S = o.zeros_u(0)
    presg, presg3 = 0.003, -0.001
    if cfg["files"] != 0.0:
        o.set_wsm([2.5e-3, 0.0, 4.0e-4])
    for nt in range(1, nsteps + 1):
        presg, presg3 = o.getup(U, P, HH, S, presg, presg3, nt)
        ref = fx[f"{name}_PRESG"][nt - 1]
        got = np.array([presg, presg3] + list(o.get_wsm()[:3]))
        assert np.max(np.abs(got - ref)) <= 1e-14 * np.max(np.abs(ref)), (nt, got, ref)
        if nt in (1, nsteps):
            for k, v in (("U", U), ("P", P), ("HH", HH), ("SGSVIS", S)):
                key = f"{name}_{k}{nt}"
                r = fx[key + "__sub"]
                err = np.max(np.abs(_sub(v, stride) - r)) / max(np.max(np.abs(r)), 1e-300)
                assert err <= 4 * TOL, (key, err)
                s = np.array([v.sum(), (v * v).sum(), np.abs(v).max()])
                es = np.max(np.abs(s - fx[key + "__sum"]) / np.maximum(np.abs(fx[key + "__sum"]), 1e-300))
                assert es <= 1e-12, (key, es)


def test_getup_fixture_distinguishes_ab2_slip():
    """the composition fixture fails a plausible slip: a GETUP that keeps
    HH2 from the first step (no AB2 shift) differs from it after three
    steps by far more than the tolerance"""
    import sys
    sys.path.insert(0, HERE + "/golden")
    from make_golden import GETUP_CASES
    from pyoracle import Oracle
    fx = _load("getup")
    nsteps, stride = (int(x) for x in fx["steps"])
    o = Oracle(33, 33, 33, cks=2.5, **GETUP_CASES["dns"])
    U, P, HH = inputs(o, seed=11)
    HH0 = HH.copy(order="F")
    S = o.zeros_u(0)
    presg, presg3 = 0.003, -0.001
    for nt in range(1, nsteps + 1):
        HHk = HH0.copy(order="F")  # the slip: HH2 stays the initial history
        presg, presg3 = o.getup(U, P, HHk, S, presg, presg3, nt)
    r = fx[f"dns_U{nsteps}__sub"]
    err = np.max(np.abs(_sub(U, stride) - r)) / np.max(np.abs(r))
    assert err > 1e3 * TOL, err


def test_line_solvers_and_filters_match_reference():
    from pyoracle import Oracle
    fx = _load("dns")
    d = fx["CTDMA_RHS"]
    for b in (6, 10, 22):
        x = Oracle.ctdma(float(b), d)
        assert np.max(np.abs(x - fx[f"CTDMA11_b{b}"])) <= TOL * np.max(np.abs(x))
    c = fx["config"]
    n1, n2, n3 = int(c[0]), int(c[1]), int(c[2])
    o = Oracle(n1, n2, n3, 4200.0, 2 * np.pi, 2.0, np.pi, 1e-3)
    F = fx["FILTER_IN"].copy(order="F")
    o.filter(F, four=False)
    assert np.max(np.abs(F - fx["FILTER"])) <= TOL
    F = fx["FILTER_IN"].copy(order="F")
    o.filter(F, four=True)
    assert np.max(np.abs(F - fx["FILTER4"])) <= TOL


def test_strain_matches_reference_les():
    from pyoracle import Oracle
    fx = _load("les")
    c = fx["config"]
    n1, n2, n3 = int(c[0]), int(c[1]), int(c[2])
    re, alx, aly, alz, dt, files, wmodel = (float(x) for x in c[3:10])
    o = Oracle(n1, n2, n3, re, alx, aly, alz, dt, files, wmodel, 2.5)
    U, _, _ = inputs(o, seed=int(c[10]))
    _check(fx, "STRAIN3", o.strain(U, 3), int(c[13]))


def test_oracle_iwmles_matches_reference():
    """WMODEL = 2 (IWMLES, smooth wall): the oracle against the reference's
    iwm_init_smooth / iwm_calc_lhs_smooth / iwm_calc_wallstress_smooth over
    NTIME = 1..4 from its own initialisation and NTIME = 5..15 (three compute
    steps) from a log-law-consistent state; every plane of its state, the
    wall stresses and WSM(1)."""
    import sys
    sys.path.insert(0, HERE + "/golden")
    from make_golden import IWM, IWM_CFLMAX, iwm_drive
    from cases import iwm_inputs
    from pyoracle import Oracle
    fx = _load("iwm")
    c = fx["config"]
    stride = int(c[11])
    o = Oracle(int(c[0]), int(c[1]), int(c[2]), cks=2.5, **IWM)
    o.set_cflmax(IWM_CFLMAX)

    def planes(key, pl):
        for q in range(pl.shape[0]):
            got = _sub(pl[q], 1)[::stride]
            ref = fx[key + "__sub"][q]
            assert np.array_equal(got, ref), (key, q)
            s = np.array([pl[q].sum(), (pl[q] * pl[q]).sum(), np.abs(pl[q]).max()])
            assert np.allclose(s, fx[key + "__sum"][q], rtol=1e-12, atol=0), (key, q)

    U0, P0, _ = inputs(o, seed=7)
    for nt, U, P in iwm_drive(o, U0, P0, 1, 4):
        o.wallmodel_iwmles(U, P, 0.003, -0.001, nt)
    pl, wdt, _ = o.iwm_state()
    planes("INIT_PLANES", pl)
    tw, tb = o.get_tau()
    _check(fx, "INIT_TAUW", tw, stride)
    _check(fx, "INIT_TAUB", tb, stride)
    assert o.get_wsm()[0] == fx["INIT_WSM"][0] and wdt == fx["INIT_WSM"][1]
    U0, P0, pl0, wdt0 = iwm_inputs(o.n1, o.n3, o.mesh()["DY"], IWM["re"], IWM["dt"], seed=1)
    o.iwm_set(pl0, wdt0)
    wsm = []
    for nt, U, P in iwm_drive(o, U0, P0, 5, 15):
        o.wallmodel_iwmles(U, P, 0.003, -0.001, nt)
        wsm.append(o.get_wsm()[0])
        pl, wdt, fixed = o.iwm_state()
        assert fixed == 0
        if nt % 5 == 0:
            planes(f"PLANES_{nt}", pl)
            assert wdt == fx[f"WDT_{nt}"][0]
    assert np.array_equal(np.array(wsm), fx["WSM"])
    tw, tb = o.get_tau()
    _check(fx, "TAUW", tw, stride)
    _check(fx, "TAUB", tb, stride)
