"""x-lines of 512 points (N1-1 = 512): the x-line tiles hold 4 lines of 64
segments in 256 threads.  FP64 phases against the oracle (<= 1e-12), an
FP64 trajectory of 5 steps against the oracle (velocities and P <= 1e-10),
the mixed path against FP64 (velocities <= 1e-5, DIVMAX <= 1e-9 every
step), and LES (Smagorinsky, log-law wall model) in both precisions."""
import pytest

from cases import config, inputs, interior_p, interior_u, relerr

pytestmark = pytest.mark.gpu

GRID = (513, 25, 33)  # 512 x 24 x 32 cells


def test_wide_phases_match_oracle():
    from phases import check_phases
    n1, n2, n3 = GRID
    check_phases(config(n1, n2, n3, cks=1.5), seed=2)


def test_wide_trajectory_fp64_and_mixed():
    from pyoracle import Oracle
    from compact3d_amd import c3d
    n1, n2, n3 = GRID
    cfg = config(n1, n2, n3, cks=1.5)
    orc = Oracle(**cfg)
    U, P, _ = inputs(orc, seed=6)
    HH = orc.zeros_u()
    S = orc.zeros_u(0)
    a = c3d.Channel(**cfg, precision=c3d.FP64)
    b = c3d.Channel(**cfg, precision=c3d.MIXED)
    for ch in (a, b):
        ch.upload(U, P)
    Uo, Po = U.copy(order="F"), P.copy(order="F")
    pg = pg3 = 0.0
    for n in range(5):
        pg, pg3 = orc.getup(Uo, Po, HH, S, pg, pg3, n + 1)
        da, _ = a.step(n + 1)
        db, _ = b.step(n + 1)
        assert da <= 1e-9 and db <= 1e-9
    sa, sb = a.download(), b.download()
    eu = [relerr(interior_u(Uo, n1, n2, n3, c), interior_u(sa["U"], n1, n2, n3, c)) for c in range(3)]
    ep = relerr(interior_p(Po, n1, n2, n3), interior_p(sa["P"], n1, n2, n3))
    em = [relerr(interior_u(sa["U"], n1, n2, n3, c), interior_u(sb["U"], n1, n2, n3, c)) for c in range(3)]
    print("512-point x-lines, 5 steps: FP64 vs oracle U", ["%.2e" % x for x in eu], "P %.2e" % ep,
          "; mixed vs FP64 U", ["%.2e" % x for x in em])
    assert max(eu) <= 1e-10 and ep <= 1e-10
    assert max(em) <= 1e-5
    a.close()
    b.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_wide_les_runs(precision):
    """LES (FILES = 2, WMODEL = 1) on 512-point x-lines: mixed against FP64
    after 5 steps, both finite and divergence-free"""
    from pyoracle import Oracle
    from compact3d_amd import c3d
    n1, n2, n3 = GRID
    cfg = config(n1, n2, n3, cks=1.5, re=20000.0, files=2.0, wmodel=1.0)
    orc = Oracle(**config(n1, n2, n3, cks=1.5))
    U, P, _ = inputs(orc, seed=1)
    chs = [c3d.Channel(**cfg, precision=p) for p in (c3d.FP64, precision)]
    for ch in chs:
        ch.upload(U, P)
    for n in range(5):
        for ch in chs:
            d, cf = ch.step(n + 1)
            assert d <= 1e-9 and cf == cf
    s0, s1 = chs[0].download(), chs[1].download()
    em = [relerr(interior_u(s0["U"], n1, n2, n3, c), interior_u(s1["U"], n1, n2, n3, c)) for c in range(3)]
    print("LES 512-point x-lines, precision", precision, "vs FP64 U", ["%.2e" % x for x in em])
    assert max(em) <= (1e-13 if precision == 0 else 1e-5)
    for ch in chs:
        ch.close()
