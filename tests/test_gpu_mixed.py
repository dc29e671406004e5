"""Mixed precision (C3D_MIXED) against the FP64 path of the same library,
which the parity tests pin to the oracle at 1e-12.

Mixed mode keeps the state U, V, W, P in FP64 and runs the compact
derivative and interpolation operators, the RHS, the ADI sweeps and their
intermediates in FP32; RDP, the Poisson solve (FFTs and Thomas solves),
the pressure correction, the flow-rate constraint, DIV, DIVMAX and CFL stay
FP64, and an FP64-residual refinement of the projection runs when its
DIVMAX exceeds 1e-10.  The LES terms (SGSVIS, wall model, SGS fluxes) are
formed in FP64 from the FP64 state and added to the FP32 HH.  What is
checked, on the same seeded inputs:
  post-projection DIVMAX of every step                 <= 1e-9
  flow rate after a step equals the one before         <= 1e-12 relative
  velocities after 10 steps, max-norm relative error   <= 1e-5
  pressure after 10 steps                              <= 1e-3
  one RHS phase (RUH), max-norm relative error         <= 1e-5
  LES (Smagorinsky, both wall models), 10 steps: velocities <= 1e-5,
  SGSVIS <= 1e-4
The measured errors are printed (pytest -s) for the record.
"""
import numpy as np
import pytest

from cases import config, inputs, interior_p, interior_u, relerr

pytestmark = pytest.mark.gpu


def _pair(cfg, seed):
    from pyoracle import Oracle
    from compact3d_amd import c3d
    orc = Oracle(**cfg)
    U, P, HH = inputs(orc, seed=seed)
    a = c3d.Channel(**cfg, precision=c3d.FP64)
    b = c3d.Channel(**cfg, precision=c3d.MIXED)
    for ch in (a, b):
        ch.upload(U, P, HH)
    return a, b, U, P


def test_mixed_vs_oracle_10_steps():
    """mixed precision against the FP64 oracle directly (the reference's
    algorithm on the CPU): per-field max and L2 relative errors after 10
    steps, and the mean-velocity profile"""
    from pyoracle import Oracle
    from compact3d_amd import c3d
    n1, n2, n3 = 65, 65, 65
    cfg = config(n1, n2, n3, cks=1.5)
    orc = Oracle(**cfg)
    U, P, _ = inputs(orc, seed=6)
    ch = c3d.Channel(**cfg, precision=c3d.MIXED)
    ch.upload(U, P)
    Uo, Po = U.copy(order="F"), P.copy(order="F")
    HH, S = orc.zeros_u(), orc.zeros_u(0)
    pg = pg3 = 0.0
    for n in range(10):
        pg, pg3 = orc.getup(Uo, Po, HH, S, pg, pg3, n + 1)
        d, _ = ch.step(n + 1)
        assert d <= 1e-9
    st = ch.download()
    out = {}
    for name, c, j0 in (("u", 0, 1), ("v", 1, 2), ("w", 2, 1)):
        a, b = interior_u(Uo, n1, n2, n3, c, j0=j0), interior_u(st["U"], n1, n2, n3, c, j0=j0)
        out[name] = (relerr(a, b), float(np.linalg.norm(a - b) / np.linalg.norm(a)))
    a, b = interior_p(Po, n1, n2, n3), interior_p(st["P"], n1, n2, n3)
    out["P"] = (relerr(a, b), float(np.linalg.norm(a - b) / np.linalg.norm(a)))
    um_o = interior_u(Uo, n1, n2, n3, 0).mean(axis=(0, 2))
    um_g = interior_u(st["U"], n1, n2, n3, 0).mean(axis=(0, 2))
    em = float(np.max(np.abs(um_o - um_g)) / np.max(np.abs(um_o)))
    print("mixed vs oracle after 10 steps, max / L2:", {k: "%.2e / %.2e" % v for k, v in out.items()},
          "mean u %.2e" % em, "PRESG %.3e vs %.3e" % (st["PRESG"], pg))
    assert max(out[k][0] for k in ("u", "v", "w")) <= 1e-5
    assert out["P"][0] <= 1e-3
    assert em <= 1e-5
    ch.close()


def test_mixed_trajectory_10_steps():
    n1, n2, n3 = 65, 65, 65
    cfg = config(n1, n2, n3, cks=1.5)
    a, b, U, P = _pair(cfg, seed=3)
    f0 = b.flowrate()
    for n in range(10):
        da, _ = a.step(n + 1)
        db, _ = b.step(n + 1)
        f1 = b.flowrate()
        print("step", n + 1, "divmax fp64 %.2e mixed %.2e" % (da, db), "flow drift %.2e" % (abs(f1[0] - f0[0]) / abs(f0[0])))
        assert db <= 1e-9
        assert abs(f1[0] - f0[0]) <= 1e-12 * abs(f0[0])
    sa, sb = a.download(), b.download()
    eu = [relerr(interior_u(sa["U"], n1, n2, n3, c), interior_u(sb["U"], n1, n2, n3, c)) for c in range(3)]
    l2 = [float(np.linalg.norm(interior_u(sa["U"] - sb["U"], n1, n2, n3, c)) /
                np.linalg.norm(interior_u(sa["U"], n1, n2, n3, c))) for c in range(3)]
    ep = relerr(interior_p(sa["P"], n1, n2, n3), interior_p(sb["P"], n1, n2, n3))
    from compact3d_amd import c3d
    print("refinement passes:", b.scalar(c3d.S_REFINES))
    print("mixed vs fp64 after 10 steps: U max", ["%.2e" % x for x in eu], "L2", ["%.2e" % x for x in l2],
          "P max %.2e" % ep, "PRESG", sa["PRESG"], sb["PRESG"])
    assert max(eu) <= 1e-5
    assert ep <= 1e-3
    a.close()
    b.close()


def test_mixed_rhs_phase():
    n1, n2, n3 = 49, 41, 33
    cfg = config(n1, n2, n3)
    a, b, U, P = _pair(cfg, seed=7)
    from compact3d_amd import c3d
    for ch in (a, b):
        ch.bcond()
        ch.rhs()
    for f in (c3d.F_RUH1, c3d.F_RUH2, c3d.F_RUH3):
        ra, rb = a.field(f), b.field(f)
        e = relerr(ra, rb)
        print("RUH", f, "%.2e" % e)
        assert e <= 1e-5
    a.close()
    b.close()


def test_mixed_getuh_forms_rdp():
    """Mixed mode forms RHSDP (compact3d.f:8117-8214) in its last ADI sweeps
    (u and v sweep y, z, x; w sweeps y, x, z): after GETUH, RDP equals the
    oracle's RHSDP, which is linear in UH, of the FP32 increments DELTA the
    sweeps stored, plus DIV(U)/DT of the resident state, to 1e-12 of its
    largest value (FP64 arithmetic from the same FP32 increments)"""
    n1, n2, n3 = 49, 41, 33
    cfg = config(n1, n2, n3)
    from pyoracle import Oracle
    from compact3d_amd import c3d
    orc = Oracle(**cfg)
    U, P, HH = inputs(orc, seed=11)
    ch = c3d.Channel(**cfg, precision=c3d.MIXED)
    ch.upload(U, P, HH)
    ch.bcond()
    ch.rhs()
    ch.getuh()
    D = orc.zeros_u()
    for c in range(3):
        D[:, :, :, c] = ch.field(c3d.F_UH1 + c)
    rdp_d = orc.zeros_p()
    orc.rhsdp(rdp_d, D)
    ref = rdp_d + ch.field(c3d.F_DIV) / cfg["dt"]
    got = ch.field(c3d.F_RDP)
    e = relerr(interior_p(ref, n1, n2, n3), interior_p(got, n1, n2, n3))
    print("mixed RDP vs oracle RHSDP(DELTA) + DIV/DT: %.2e" % e)
    assert e <= 1e-12
    ch.close()


def test_mixed_state_roundtrip():
    """the FP64 state survives upload/download in mixed mode bit for bit"""
    n1, n2, n3 = 33, 33, 33
    cfg = config(n1, n2, n3)
    a, b, U, P = _pair(cfg, seed=1)
    sb = b.download()
    assert np.array_equal(interior_u(sb["U"], n1, n2, n3), interior_u(U, n1, n2, n3))
    assert np.array_equal(interior_p(sb["P"], n1, n2, n3), interior_p(P, n1, n2, n3))
    a.close()
    b.close()


@pytest.mark.parametrize("wmodel", [1.0, 0.0, 2.0, 6.0])
def test_mixed_les_trajectory_10_steps(wmodel):
    """wall-modelled (log law; the integral wall model, whose compute steps
    are NTIME 5 and 10) / wall-resolved Smagorinsky LES: mixed against FP64"""
    from compact3d_amd import c3d
    from pyoracle import Oracle
    n1, n2, n3 = 49, 33, 33
    cfg = config(n1, n2, n3, re=20000.0, dt=5e-4, files=2.0, wmodel=wmodel, cks=1.5)
    orc = Oracle(**cfg)
    U, P, _ = inputs(orc, seed=9)
    chs = []
    for prec in (c3d.FP64, c3d.MIXED):
        ch = c3d.Channel(**cfg, precision=prec, cflmax=0.5)
        ch.upload(U, P)
        ch.set_scalar(c3d.S_WSM1, 1e-3)
        ch.set_scalar(c3d.S_WSM3, 1e-4)
        chs.append(ch)
    a, b = chs
    for n in range(10):
        a.step(n + 1)
        db, _ = b.step(n + 1)
        assert db <= 1e-9
    sa, sb = a.download(), b.download()
    eu = [relerr(interior_u(sa["U"], n1, n2, n3, c), interior_u(sb["U"], n1, n2, n3, c)) for c in range(3)]
    es = relerr(interior_u(sa["SGSVIS"], n1, n2, n3), interior_u(sb["SGSVIS"], n1, n2, n3))
    print("mixed LES wmodel", wmodel, "U max", ["%.2e" % x for x in eu], "SGSVIS %.2e" % es)
    assert max(eu) <= 1e-5
    assert es <= 1e-4
    a.close()
    b.close()
