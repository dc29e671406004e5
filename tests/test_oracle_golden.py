"""The oracle against golden vectors produced by the reference's own
subroutines (tests/golden/make_golden.py; compact3d.f compiled unmodified).

Every phase the reference can run without FFTW is compared: BCOND, RHS1-3,
GETUH1-3, RHSDP, UPCALC, DIVCHECK, CFL, SGSCS/STRAIN, SGSDY (the dynamic
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
    DP = o.zeros_p()
    o.getdp(DP, RDP)
    Pn = P.copy(order="F")
    pg, pg3 = o.upcalc(U, Pn, UH, DP, presg, presg3)
    _check(fx, "U_UPCALC", U, stride)
    _check(fx, "P_UPCALC", Pn, stride)
    assert np.max(np.abs(np.array([pg, pg3]) - fx["PRESG"])) <= TOL * np.max(np.abs(fx["PRESG"]))
    assert o.divcheck(U) == pytest.approx(fx["DIVMAX"][0], rel=1e-12, abs=1e-15)
    assert o.cfl(U) == pytest.approx(fx["CFL"][0], rel=1e-15)


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
