"""Phase-by-phase parity of one GETUP step (the GPU through the C-ABI
against the oracle), shared by the GPU tests."""
import numpy as np

from cases import inputs, interior_p, interior_u, relerr

EPS = 2.220446049250313e-16


def presg_bound(U, m, n1, n2, n3, comp, dt):
    """what the reference's form of DMPRESG (compact3d.f:1079-1093, 1160-1172)
    can reproduce: a difference of two serial sums of N = N1M*N2M*N3M terms
    u*DY/(DX1*DX3), each carrying a rounding error of about
    eps*S*sqrt(N/3) (S the sum of |terms|, the partial sums growing
    linearly), divided by VOL*DT.  Four times the root-sum-square of the
    two sums' estimates."""
    c = 0 if comp == 1 else 2
    dy = m["DY"][1:n2]
    S = np.sum(np.abs(U[1:n1, 1:n2, 1:n3, c]) * dy[None, :, None]) / (m["DX1"] * m["DX3"])
    N = (n1 - 1) * (n2 - 1) * (n3 - 1)
    return 4.0 * np.sqrt(2.0) * EPS * S * np.sqrt(N / 3.0) / (m["VOL"] * dt)


def div_scale(U, m):
    """the magnitude of the terms of DIVCHECK's sum: max|u| times the
    largest inverse spacing"""
    return np.max(np.abs(U)) * max(m["DX1"], m["DX3"], np.max(1.0 / m["DY"][1:-1]))


def check_phases(cfg, seed=0):
    """one GETUP phase by phase against the oracle: BCOND, RHS1-3 (RUH, HH),
    GETUH1-3, RHSDP, GETDP, UPCALC, DIVCHECK, CFL, CHKMF"""
    from compact3d_amd import c3d
    from pyoracle import Oracle
    n1, n2, n3 = cfg["n1"], cfg["n2"], cfg["n3"]
    orc = Oracle(**cfg)
    ch = c3d.Channel(**cfg)
    U, P, HH2 = inputs(orc, seed=seed)
    presg, presg3 = 0.003, 0.002
    ch.upload(U, P, HH=HH2, presg=presg, presg3=presg3)
    # oracle, following GETUP
    Uo = U.copy(order="F")
    So = orc.zeros_u(0)
    orc.bcond(Uo, So)
    ch.bcond()
    st = ch.download()
    err = relerr(Uo, st["U"])
    print(f"BCOND {err:.3e}")
    assert err == 0.0
    HHo = orc.zeros_u()
    RUH = [orc.zeros_u(0) for _ in range(3)]
    for c in (1, 2, 3):
        orc.rhs(c, Uo, P, HHo, HH2, RUH[c - 1], presg if c == 1 else presg3, So)
    ch.rhs()
    for c in (1, 2, 3):
        j0 = 2 if c == 2 else 1
        e1 = relerr(interior_u(RUH[c - 1], n1, n2, n3, j0=j0),
                    interior_u(ch.field(c3d.F_RUH1 + c - 1), n1, n2, n3, j0=j0))
        e2 = relerr(interior_u(HHo, n1, n2, n3, c - 1, j0=j0),
                    interior_u(ch.field(c3d.F_HH1 + c - 1), n1, n2, n3, j0=j0))
        print(f"RHS{c} {e1:.3e}  HH{c} {e2:.3e}")
        assert e1 <= 1e-12 and e2 <= 1e-12
    UHo = orc.zeros_u()
    for c in (1, 2, 3):
        orc.getuh(c, Uo, UHo, RUH[c - 1])
    ch.getuh()
    for c in (1, 2, 3):
        j0 = 2 if c == 2 else 1
        e = relerr(interior_u(UHo, n1, n2, n3, c - 1, j0=j0),
                   interior_u(ch.field(c3d.F_UH1 + c - 1), n1, n2, n3, j0=j0))
        print(f"GETUH{c} {e:.3e}")
        assert e <= 1e-12
    RDPo = orc.zeros_p()
    DPo = orc.zeros_p()
    orc.rhsdp(RDPo, UHo)
    orc.getdp(DPo, RDPo)
    ch.dpcalc()
    e_rdp = relerr(interior_p(RDPo, n1, n2, n3), interior_p(ch.field(c3d.F_RDP), n1, n2, n3))
    e_dp = relerr(interior_p(DPo, n1, n2, n3), interior_p(ch.field(c3d.F_DP), n1, n2, n3))
    print(f"RHSDP {e_rdp:.3e}  GETDP {e_dp:.3e}")
    assert e_rdp <= 1e-12 and e_dp <= 1e-12
    Po = P.copy(order="F")
    pg, pg3 = orc.upcalc(Uo, Po, UHo, DPo, presg, presg3)
    ch.upcalc()
    st = ch.download()
    eu = [relerr(interior_u(Uo, n1, n2, n3, c, j0=(2 if c == 1 else 1)),
                 interior_u(st["U"], n1, n2, n3, c, j0=(2 if c == 1 else 1))) for c in range(3)]
    ep = relerr(interior_p(Po, n1, n2, n3), interior_p(st["P"], n1, n2, n3))
    print("UPCALC U", ["%.3e" % x for x in eu], f"P {ep:.3e}", "PRESG", pg, st["PRESG"], pg3, st["PRESG3"])
    assert max(eu) <= 1e-12 and ep <= 1e-12
    # PRESG/PRESG3: the GPU forms DMPRESG from the sum of the ADI increment,
    # the reference from a difference of two flow sums (equal in exact
    # arithmetic); they agree to the rounding of that difference
    m = orc.mesh()
    b1 = presg_bound(UHo, m, n1, n2, n3, 1, cfg["dt"])
    b3 = presg_bound(UHo, m, n1, n2, n3, 3, cfg["dt"])
    print(f"PRESG error {abs(pg - st['PRESG']):.3e} (bound {b1:.3e}), "
          f"PRESG3 error {abs(pg3 - st['PRESG3']):.3e} (bound {b3:.3e})")
    assert abs(pg - st["PRESG"]) <= b1 and abs(pg3 - st["PRESG3"]) <= b3
    # the ghost planes follow UBC
    assert relerr(Uo, st["U"]) <= 1e-12
    dmo, dmg = orc.divcheck(Uo), ch.divcheck()
    cfo, cfg_ = orc.cfl(Uo), ch.cfl()
    print("DIVCHECK", dmo, dmg, "CFL", cfo, cfg_)
    assert abs(cfo - cfg_) <= 1e-12 * cfo
    assert dmg <= 1e-10  # post-projection divergence is at roundoff
    f1o, f3o = orc.chkmf(Uo)
    f1g, f3g = ch.chkmf()
    print("CHKMF", f1o, f1g, f3o, f3g)
    # CHKMF's sums run over one plane (N2M*N3M and N1M*N2M terms): the
    # oracle's serial sum carries up to (terms)*eps of rounding, which at
    # 192x129x192 exceeds 1e-12
    t1 = max(1e-12, (n2 - 1) * (n3 - 1) * 2.2e-16)
    t3 = max(1e-12, (n1 - 1) * (n2 - 1) * 2.2e-16)
    assert abs(f1o - f1g) <= t1 * abs(f1o) and abs(f3o - f3g) <= t3 * max(abs(f3o), 1e-300) + 1e-15
    # WALLSS (WSM(1:3)) of the GPU's stepped field (the oracle's differs from
    # it by DT times the DMPRESG rounding above, which is large relative to
    # the near-wall velocity that WALLSS differentiates)
    orc.wallss(st["U"])
    ch.wallss()
    wo = orc.get_wsm()[:3]
    wg = np.array([ch.scalar(c3d.S_WSM1 + q) for q in range(3)])
    ew = np.abs(wo - wg) / np.max(np.abs(wo))
    print("WALLSS", wo, wg, "relative errors", ew)
    # sums over the N1M*N3M wall points, as CHKMF's: (terms)*eps
    assert np.max(ew) <= max(1e-12, (n1 - 1) * (n3 - 1) * 2.2e-16)
    # DIVCHECK and CFL of non-solenoidal fields: the unprojected UH and a
    # uniform[-1,1] field (ghost planes included)
    rng = np.random.default_rng(77 + seed)
    Ur = np.asfortranarray(rng.uniform(-1, 1, Uo.shape))
    for tag, F in (("UH", UHo), ("uniform", Ur)):
        ch.upload(F, Po)
        dmo, dmg = orc.divcheck(F), ch.divcheck()
        cfo, cfg_ = orc.cfl(F), ch.cfl()
        print(f"DIVCHECK {tag}: oracle {dmo:.6e} GPU {dmg:.6e} ({abs(dmo - dmg) / dmo:.2e}); "
              f"CFL {abs(cfo - cfg_) / cfo:.2e}")
        assert dmo > 1e-3
        assert abs(dmo - dmg) <= 1e-12 * dmo and abs(cfo - cfg_) <= 1e-12 * cfo
    ch.close()


def check_fused_div(cfg, precision=0, seed=0):
    """the DIVMAX (and, in mixed mode, the DIV field the next step starts
    from) that c3d_step forms inside its update kernels equal DIVCHECK's
    operator (the oracle's) applied to the GPU's own stepped field, to the
    rounding of the terms (the field is solenoidal, so DIV itself is
    rounding; a wrong index or coefficient would leave errors of the size of
    the terms).  In FP64 mode the DIV field read back is DIVCHECK's."""
    from compact3d_amd import c3d
    from pyoracle import Oracle
    n1, n2, n3 = cfg["n1"], cfg["n2"], cfg["n3"]
    orc = Oracle(**cfg)
    ch = c3d.Channel(**cfg, precision=precision)
    U, P, _ = inputs(orc, seed=seed)
    ch.upload(U, P)
    for nt in (1, 2):
        dmg, cfg_ = ch.step(nt)
    Ug = ch.download()["U"]
    dmo, DIVo = orc.divcheck(Ug, want_div=True)
    DIVg = interior_p(ch.field(c3d.F_DIV), n1, n2, n3)
    scale = div_scale(Ug, orc.mesh())
    e = float(np.max(np.abs(DIVg - DIVo)))
    cfo = orc.cfl(Ug)
    print(f"fused DIV (precision {precision}): max|DIV - oracle| {e:.3e} = {e / scale:.2e} of the terms "
          f"({scale:.3e}); DIVMAX {dmg:.3e} oracle {dmo:.3e}; CFL {abs(cfg_ - cfo) / cfo:.2e}")
    assert e <= 1e-13 * scale
    assert abs(dmg - dmo) <= 1e-13 * scale
    assert abs(cfg_ - cfo) <= 1e-12 * cfo
    ch.close()
