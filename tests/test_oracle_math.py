"""Closed-form pins of the oracle for what the reference cannot run here
(GETDP needs FFTW) and independent checks of what it can.  Each check is
chosen so that a plausible slip (sign, index shift, transposed operand,
dropped term, wrong normalisation) fails it."""
import math

import numpy as np
import pytest

from cases import config, inputs


def _o(**kw):
    from pyoracle import Oracle
    return Oracle(**config(**kw))


@pytest.mark.parametrize("b", [6.0, 10.0, 22.0])
@pytest.mark.parametrize("n", [16, 33, 64])
def test_ctdma_equals_dense_cyclic_solve(b, n):
    from pyoracle import Oracle
    rng = np.random.default_rng(n)
    d = rng.uniform(-1, 1, n)
    A = np.diag(np.full(n, b)) + np.diag(np.ones(n - 1), 1) + np.diag(np.ones(n - 1), -1)
    A[0, -1] = A[-1, 0] = 1.0
    x = Oracle.ctdma(b, d)
    assert np.max(np.abs(x - np.linalg.solve(A, d))) <= 1e-15 * np.max(np.abs(x)) * n


def test_dft2d_matches_numpy_fft():
    from pyoracle import Oracle
    rng = np.random.default_rng(1)
    for n1, n3 in ((8, 12), (32, 48), (20, 30)):
        a = rng.standard_normal((n1, n3)) + 1j * rng.standard_normal((n1, n3))
        assert np.allclose(Oracle.dft2d(a, -1), np.fft.fft2(a), rtol=0, atol=1e-12 * np.abs(a).sum())
        assert np.allclose(Oracle.dft2d(a, +1), np.fft.ifft2(a) * n1 * n3, rtol=0, atol=1e-12 * np.abs(a).sum())


def test_modified_wavenumber_is_square_of_compact_d1_symbol():
    """AK1/AK3 (INIWAVE, compact3d.f:9211,9222) must equal |sigma|^2 of the
    compact face<->centre derivative, so that the Poisson operator is exactly
    D1x(D1x) and the projection leaves no divergence."""
    o = _o(n1=49, n2=17, n3=33)
    m = o.mesh()
    n = 48
    th = 2 * np.pi * np.arange(n) / n
    sig = 24 * m["DX1"] * (np.exp(1j * th) - 1) / (22 + 2 * np.cos(th))
    ak1 = m["AK1"][1:n + 1]
    kk = np.where(np.arange(n) <= n // 2, np.arange(n), np.arange(n) - n)
    assert np.allclose(ak1, np.abs(sig[np.abs(kk)]) ** 2, rtol=1e-13, atol=1e-12)


def _apply_lh(o, DP):
    """L_h DP in physical space, built independently with numpy: compact
    D1(face->centre) D1(centre->face) in x and z (exact circulant symbols) plus
    the wall-normal three-point operator of METRICPOISSON."""
    m = o.mesh()
    n1m, n2m, n3m = o.n1 - 1, o.n2 - 1, o.n3 - 1
    f = DP[:n1m, :n2m, :n3m]
    kx = 2 * np.pi * np.arange(n1m) / n1m
    kz = 2 * np.pi * np.arange(n3m) / n3m
    a1 = 288 * m["DX1Q"] * (1 - np.cos(kx)) / (11 + np.cos(kx)) ** 2
    a3 = 288 * m["DX3Q"] * (1 - np.cos(kz)) / (11 + np.cos(kz)) ** 2
    F = np.fft.fft2(f, axes=(0, 2))
    lap = -(a1[:, None, None] + a3[None, None, :]) * F
    out = np.real(np.fft.ifft2(lap, axes=(0, 2)))
    pm, pc, pp = m["PMJ"], m["PCJ"], m["PPJ"]
    for j in range(1, n2m + 1):
        acc = pc[j] * f[:, j - 1, :]
        if j > 1:
            acc = acc + pm[j] * f[:, j - 2, :]
        if j < n2m:
            acc = acc + pp[j] * f[:, j, :]
        out[:, j - 1, :] += acc
    R = np.zeros_like(DP)
    R[:n1m, :n2m, :n3m] = out
    return R


@pytest.mark.parametrize("grid", [(33, 33, 33), (49, 25, 25)])
def test_getdp_recovers_manufactured_solution(grid):
    """GETDP inverts L_h exactly; its gauge is DP's plane mean at J=1 = 0
    (the pinned row of mode (0,0), compact3d.f:9363-9368)."""
    o = _o(n1=grid[0], n2=grid[1], n3=grid[2])
    rng = np.random.default_rng(5)
    DP = o.zeros_p()
    n1m, n2m, n3m = o.n1 - 1, o.n2 - 1, o.n3 - 1
    DP[:n1m, :n2m, :n3m] = rng.uniform(-1, 1, (n1m, n2m, n3m))
    DP[:n1m, 0, :n3m] -= DP[:n1m, 0, :n3m].mean()
    R = _apply_lh(o, DP)
    out = o.zeros_p()
    o.getdp(out, R)
    err = np.max(np.abs(out - DP)) / np.max(np.abs(DP))
    assert err <= 1e-10
    # a wrong sign, a dropped AK term or a missing 1/(N1M*N3M) fails by O(1)


def test_projection_invariants_after_step():
    """After a full step: divergence at roundoff (DIVCHECK) and the flow rate
    through every x-plane equal to the flow rate before it (UPCALC's
    constant-flow-rate constraint, compact3d.f:1079-1128)."""
    o = _o(n1=33, n2=33, n3=33)
    U, P, _ = inputs(o, seed=2)
    HH = o.zeros_u()
    S = o.zeros_u(0)
    m = o.mesh()
    dy = m["DY"][1:o.n2]

    def flow(U):
        return np.einsum("ijk,j->i", U[1:o.n1, 1:o.n2, 1:o.n3, 0], dy) / m["DX3"]

    f0 = flow(U).mean()
    pg = pg3 = 0.0
    for n in range(3):
        pg, pg3 = o.getup(U, P, HH, S, pg, pg3, n + 1)
        assert o.divcheck(U) <= 1e-12 * np.max(np.abs(U))
        assert np.max(np.abs(flow(U).mean() - f0)) <= 1e-13 * abs(f0)


def test_wall_model_solves_log_law():
    """WALLMODEL_INSTANT: tau_w = u_tau^2 with u_tau the damped-Newton root of
    U/u_tau = ln(u_tau*Delta*Re)/kappa + B (kappa=0.41, B=5.2,
    compact3d.f:2370-2371), sampled at the third cell, Delta = DY1+DY2+DY3/2,
    and the stress points along (u, w)."""
    o = _o(n1=17, n2=17, n3=17, re=20000.0, files=2.0, wmodel=1.0)
    U, _, _ = inputs(o, seed=4)
    o.wallmodel_instant(U)
    tw, tb = o.get_tau()
    m = o.mesh()
    dl = m["DY"][1] + m["DY"][2] + m["DY"][3] / 2.0
    u3 = U[1:17, 3, 1:17, 0]
    w3 = U[1:17, 3, 1:17, 2]
    ua = np.sqrt(u3 ** 2 + w3 ** 2)
    tau = np.sqrt(tw[1:17, 1:17, 0] ** 2 + tb[1:17, 1:17, 0] ** 2)
    ut = np.sqrt(tau)
    res = ua / ut - np.log(ut * dl * 20000.0) / 0.41 - 5.2
    # the iteration stops on |du_tau| <= 1e-6 and reports the previous iterate
    assert np.max(np.abs(res)) <= 1e-6 * np.max(ua / ut ** 2 + 1 / (0.41 * ut))
    assert np.allclose(tw[1:17, 1:17, 0] * w3, tb[1:17, 1:17, 0] * u3, rtol=0, atol=1e-15)
    assert o.get_wsm()[0] == pytest.approx(np.mean(tau), rel=1e-13)


def test_strain_of_linear_shear():
    """STRAIN on u = a*y (interior rows): S12 = a/2, all other S_ij = 0."""
    o = _o(n1=17, n2=17, n3=17, cks=1.0)
    m = o.mesh()
    U = o.zeros_u()
    a = 0.7
    yc = 0.5 * (m["Y"][1:17] + m["Y"][2:18])
    U[1:17, 1:17, 1:17, 0] = a * yc[None, :, None]
    U[1:17, 0, 1:17, 0] = -U[1:17, 1, 1:17, 0]
    U[1:17, 17, 1:17, 0] = 2 * a * m["Y"][17] - U[1:17, 16, 1:17, 0]
    for j in (4, 8, 12):
        SR = o.strain(U, j)
        assert np.allclose(SR[:16, :16, 1], a / 2, rtol=1e-12)
        for q in (0, 2, 3, 4, 5):
            assert np.max(np.abs(SR[:16, :16, q])) <= 1e-12


def test_filters_on_fourier_modes():
    """FILTER: Simpson (1,4,1)/6 in x then z -> gain (4+2cos t)/6 per
    direction; FILTER4: (1,4,2,4,1)/12 -> (2+8cos t+2cos 2t)/12."""
    o = _o(n1=33, n2=9, n3=17)
    n1m, n3m = 32, 16
    i = np.arange(n1m)[:, None]
    k = np.arange(n3m)[None, :]
    mx, mz = 3, 2
    f = np.cos(2 * np.pi * mx * i / n1m) * np.cos(2 * np.pi * mz * k / n3m)
    F = np.zeros((33, 17, 1), order="F")
    F[:n1m, :n3m, 0] = f
    G = F.copy(order="F")
    o.filter(G, four=False)
    tx, tz = 2 * np.pi * mx / n1m, 2 * np.pi * mz / n3m
    g = (4 + 2 * np.cos(tx)) / 6 * (4 + 2 * np.cos(tz)) / 6
    assert np.allclose(G[:n1m, :n3m, 0], g * f, atol=1e-14)
    G = F.copy(order="F")
    o.filter(G, four=True)
    g4 = ((2 + 8 * np.cos(tx) + 2 * np.cos(2 * tx)) / 12) * ((2 + 8 * np.cos(tz) + 2 * np.cos(2 * tz)) / 12)
    assert np.allclose(G[:n1m, :n3m, 0], g4 * f, atol=1e-14)


def test_mesh_closed_form():
    o = _o(n1=17, n2=33, n3=17, cks=1.8)
    m = o.mesh()
    j = np.arange(1, 34)
    y = 1 - np.tanh(1.8 * (1 - 2.0 * (j - 1) / 32)) / np.tanh(1.8)
    assert np.allclose(m["Y"][1:34], y, rtol=0, atol=1e-15)
    assert m["ALY"] == 2.0
    dy = np.diff(m["Y"][1:34])
    assert np.allclose(m["DY"][1:33], dy, atol=1e-16)
    assert m["DX1"] == pytest.approx(16 / (2 * math.pi))


def _stats_field(n1, n2, n3, Y, alx, alpha=0.7, beta=0.2, b0=0.3, c=0.9, c2=-0.4, m=3):
    """u = alpha*y + beta (linear in the cell-centre height, wall ghosts on
    the same line), v = b0, w = c sin(2 pi m x/ALX) + c2 (independent of y
    and z), P(J) = J: every plane statistic of ENERGY has a closed form"""
    U = np.zeros((n1 + 1, n2 + 1, n3 + 1, 3), order="F")
    P = np.zeros((n1, n2, n3), order="F")
    yc = 0.5 * (Y[:-1] + Y[1:])  # cell centres, J = 0..N2-1
    for J in range(1, n2):
        U[:, J, :, 0] = alpha * yc[J] + beta
        P[:, J - 1, :] = J
    U[:, 0, :, 0] = alpha * Y[1] + beta        # wall values the face interpolation takes
    U[:, n2, :, 0] = alpha * Y[n2] + beta
    U[:, :, :, 1] = b0
    x = (np.arange(n1 + 1) - 1) * alx / (n1 - 1)
    U[:, :, :, 2] = (c * np.sin(2 * np.pi * m * x / alx) + c2)[:, None, None]
    return U, P, yc


def test_plane_statistics_closed_forms():
    """ENERGY's plane statistics (compact3d.f:10005-10158): means, second
    moments, third moments and the cell-centre vorticity of a field whose
    statistics are known exactly"""
    n1, n2, n3 = 33, 25, 17
    alx = 2 * math.pi
    o = _o(n1=n1, n2=n2, n3=n3, alx=alx, cks=1.5)
    Y = o.vec(0)
    alpha, beta, b0, c, c2, m = 0.7, 0.2, 0.3, 0.9, -0.4, 3
    U, P, yc = _stats_field(n1, n2, n3, Y, alx, alpha, beta, b0, c, c2, m)
    s = o.plane_stats(U, P)
    J = np.arange(1, n2)
    u = alpha * yc[1:n2] + beta
    dx = alx / (n1 - 1)
    ww = 0.5 * c * c + c2 * c2
    g2 = 0.5 * (c * math.sin(2 * math.pi * m * dx / alx) / dx) ** 2  # <(dw/dx)^2>, central difference
    expect = {0: u, 1: np.full(n2 - 1, b0), 2: np.full(n2 - 1, c2), 3: J.astype(float),
              4: np.zeros(n2 - 1), 5: np.zeros(n2 - 1), 6: np.full(n2 - 1, -alpha),
              7: u * u, 8: np.full(n2 - 1, b0 * b0), 9: np.full(n2 - 1, ww), 10: u * b0, 11: u * c2,
              12: np.full(n2 - 1, b0 * c2), 13: (J * J).astype(float), 14: J * b0,
              15: np.zeros(n2 - 1), 16: np.full(n2 - 1, g2), 17: np.full(n2 - 1, alpha * alpha),
              18: u * u * b0, 19: np.full(n2 - 1, b0 * ww)}
    for q, e in expect.items():
        assert np.max(np.abs(s[q] - e)) <= 1e-12 * max(1.0, np.max(np.abs(e))), (q, s[q][:3], e[:3])


def test_dqrrt_selects_the_largest_real_part_in_0_2():
    """DQRRT/CHHQR (compact3d.f:11548-11677): the roots of a quintic with
    known roots; BETA is the largest real part in (0, 2), complex pairs
    included, else 0.  CHHQR deflates at EPS = 1e-5, which leaves errors of
    a few 1e-6 in the roots (the reference's own DQRRT returns the same
    values bit for bit, tests/test_oracle_golden.py), so roots are checked
    to 2e-5."""
    from pyoracle import Oracle
    cases = [([0.3, 1.7, -0.5, 2.5, -4.0], 1.7),
             ([0.9 + 0.4j, 0.9 - 0.4j, -1.0, 3.0, 5.0], 0.9),
             ([-0.2, -1.0, 2.5, 3.0 + 1j, 3.0 - 1j], 0.0),
             ([0.125, 0.25, 0.5, 1.0, 1.5], 1.5)]
    for roots, beta in cases:
        pa = np.real(np.poly(roots))[1:]
        assert abs(Oracle.dqrrt(pa) - beta) <= 2e-5, (roots, Oracle.dqrrt(pa))


def test_two_wall_log_law_is_the_mirror_image():
    """WMODEL = 6 (WALLMODEL_INSTANT at both walls): the top-wall stress of
    a field equals the bottom-wall stress of the field mirrored in y (u, w
    at cell centres J <-> N2-J, v at faces J <-> N2+1-J with its sign
    flipped), and the bottom wall is WALLMODEL_INSTANT's own result"""
    from pyoracle import Oracle
    from cases import config, inputs
    n1, n2, n3 = 33, 33, 33
    cfg = config(n1, n2, n3, re=5200.0, dt=5e-4, files=2.0, wmodel=6.0)
    o = Oracle(**cfg)
    U, _, _ = inputs(o, seed=3)
    M = np.asfortranarray(U[:, ::-1, :, :])      # J -> N2 - J
    M[:, 1:n2 + 1, :, 1] = -U[:, n2:0:-1, :, 1]  # faces J -> N2 + 1 - J
    M[:, 0, :, 1] = 0.0
    o.wallmodel_instant2(U)
    tw, tb = o.get_tau()
    o1 = Oracle(**config(n1, n2, n3, re=5200.0, dt=5e-4, files=2.0, wmodel=1.0))
    o1.wallmodel_instant(M)
    mw, mb = o1.get_tau()
    sl = (slice(1, n1), slice(1, n3))
    for a, b in ((tw[sl + (1,)], mw[sl + (0,)]), (tb[sl + (1,)], mb[sl + (0,)])):
        assert np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(b))
    o2 = Oracle(**config(n1, n2, n3, re=5200.0, dt=5e-4, files=2.0, wmodel=1.0))
    o2.wallmodel_instant(U)
    bw, bb = o2.get_tau()
    assert np.array_equal(tw[:, :, 0], bw[:, :, 0]) and np.array_equal(tb[:, :, 0], bb[:, :, 0])
    assert np.max(np.abs(tw[sl + (1,)])) > 0


def test_iwmles_equilibrium_branch():
    """WMODEL = 2 from its own initialisation: the wall-layer integral starts
    as plug flow (u * Dz), so the first compute step (NTIME = 5) needs linear
    corrections |A| > 1 at every point and takes the equilibrium branch.  The
    reference reads an undefined local there; the oracle computes the
    point's log-law stress, checked here against an independent bisection of
    u/ut = ln(ut Dz Re)/0.4 + 5 (the Newton iteration stops at a step of
    1e-6 and returns the iterate before it: 1e-4 relative in tau)."""
    import math
    from cases import inputs, iwm_equil
    from pyoracle import IWM_PLANES, Oracle
    re = 5200.0
    o = Oracle(24, 17, 20, re, 2 * math.pi, 2.0, math.pi, 5e-4, 2.0, 2.0, 2.5)
    o.set_cflmax(0.5)
    U0, P0, _ = inputs(o, seed=3)
    for nt in range(1, 6):
        o.wallmodel_iwmles(U0, P0, 0.003, -0.001, nt)
    pl, wdt, fixed = o.iwm_state()
    ix = {n: q for q, n in enumerate(IWM_PLANES)}
    assert fixed == 23 * 19
    assert wdt == 5e-4 + 5e-4 + 5e-4 + 5e-4 + 5e-4
    ux, uy, dz = pl[ix["flt_u"]], pl[ix["flt_w"]], pl[ix["dz"]][0, 0]
    dy = o.mesh()["DY"]
    assert dz == dy[1] + dy[2] + dy[3] / 2.0
    vel = np.sqrt(ux * ux + uy * uy)
    ut = iwm_equil(vel, dz, re)
    tx, ty = ut * ut * ux / vel, ut * ut * uy / vel
    assert np.allclose(pl[ix["tauwx"]], tx, rtol=1e-4, atol=0)
    assert np.allclose(pl[ix["tauwy"]], ty, rtol=1e-4, atol=1e-4 * np.abs(tx).max())
    assert np.array_equal(pl[ix["utx"]], np.sqrt(np.abs(pl[ix["tauwx"]])))
    for n in ("ax", "ay", "cx", "cy"):
        assert not pl[ix[n]].any()
    tw, tb = o.get_tau()
    assert np.array_equal(tw[1:24, 1:20, 0], pl[ix["tauwx"]].T)
    assert np.array_equal(tb[1:24, 1:20, 0], pl[ix["tauwy"]].T)
    assert not tw[:, :, 1].any() and not tb[:, :, 1].any()
    assert o.get_wsm()[0] == pytest.approx(np.sqrt(tx * tx + ty * ty).mean(), rel=2e-4)


def test_iwmles_ulp_sensitivity():
    """WMODEL = 2 from its own initialisation: the per-point Newton iteration
    stops at residuals below 1e-6 with a finite-difference Jacobian that is
    nearly singular in the spanwise direction (uty starts at 0), so the
    stopping iterate depends on rounding.  Inputs changed by about 1 ulp move
    the state planes by 1e-11 at the first compute step and by up to ~1e-6
    (the integral of ww) at the second: the scale of any two faithful
    implementations' disagreement."""
    import math
    from cases import config, inputs
    from pyoracle import Oracle

    cfg = config(33, 33, 33, re=5200.0, dt=5e-4, files=2.0, wmodel=2.0)

    def run(pert):
        o = Oracle(**cfg)
        o.set_cflmax(0.5)
        U, P, _ = inputs(o, seed=7)
        if pert:
            rng = np.random.default_rng(1)
            U = np.asfortranarray(U * (1 + 2e-16 * rng.standard_normal(U.shape)))
            P = np.asfortranarray(P * (1 + 2e-16 * rng.standard_normal(P.shape)))
        out = {}
        for nt in range(1, 11):
            o.wallmodel_iwmles(U, P, 0.003, -0.001, nt)
            if nt in (5, 10):
                out[nt] = o.iwm_state()[0].copy()
        return out

    a, b = run(False), run(True)

    def worst(nt):
        return max(np.max(np.abs(a[nt][q] - b[nt][q])) / max(np.max(np.abs(a[nt][q])), 1e-300)
                   for q in range(a[nt].shape[0]))

    assert worst(5) < 1e-9
    assert 1e-8 < worst(10) < 1e-5
    assert math.isfinite(worst(10))
