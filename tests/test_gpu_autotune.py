"""The tile-shape autotuner: it returns shapes within its candidate sets,
the tuned step is not slower than the rule-derived one, and a channel
created with tuned shapes keeps parity with the oracle (the shapes change
only the rounding order of the line solves): velocities after 10 steps
<= 1e-10 relative (FP64), post-projection DIVMAX <= 1e-9 (both precisions)."""
import pytest

from cases import config, inputs, interior_u, relerr

pytestmark = pytest.mark.gpu


def test_autotune_then_parity_fp64():
    from compact3d_amd import c3d
    from pyoracle import Oracle
    n1, n2, n3 = 65, 49, 65
    cfg = config(n1, n2, n3, cks=1.5)
    try:
        tuned = c3d.autotune(steps=3, **cfg, precision=c3d.FP64)
        print("tuned FP64", tuned)
        assert tuned["ms_tuned"] <= tuned["ms_rule"]
        assert tuned["x_lines"] % 8 == 0 and tuned["x_lines"] * (n1 - 1) // 8 <= 256
        assert tuned["z_cols_fp64"] * (n3 - 1) // 8 <= 512
        assert tuned["y_cols_fp64"] * tuned["y_chunks_fp64"] <= 1024
        assert tuned["pois_cols"] % 2 == 0 and tuned["pois_cols"] * tuned["pois_chunks"] <= 1024
        orc = Oracle(**cfg)
        U, P, HH = inputs(orc, seed=4)
        ch = c3d.Channel(**cfg)
        ch.upload(U, P, HH)
        Uo, Po, HHo = U.copy(order="F"), P.copy(order="F"), HH.copy(order="F")
        S = orc.zeros_u(0)
        pg = pg3 = 0.0
        for n in range(10):
            pg, pg3 = orc.getup(Uo, Po, HHo, S, pg, pg3, n + 1)
            d, _ = ch.step(n + 1)
            assert d <= 1e-9
        st = ch.download()
        eu = [relerr(interior_u(Uo, n1, n2, n3, c), interior_u(st["U"], n1, n2, n3, c)) for c in range(3)]
        print("tuned shapes, 10 steps vs oracle: U", ["%.2e" % x for x in eu])
        assert max(eu) <= 1e-10
        ch.close()
    finally:
        c3d.autotune_reset()


def test_autotune_mixed_keeps_divergence():
    from compact3d_amd import c3d
    from pyoracle import Oracle
    n1, n2, n3 = 65, 65, 65
    cfg = config(n1, n2, n3, cks=1.5)
    try:
        tuned = c3d.autotune(steps=3, **cfg, precision=c3d.MIXED)
        print("tuned mixed", tuned)
        assert tuned["ms_tuned"] <= tuned["ms_rule"]
        assert tuned["z_seg_fp32"] in (16, 8, 2)
        orc = Oracle(**cfg)
        U, P, HH = inputs(orc, seed=3)
        a = c3d.Channel(**cfg, precision=c3d.FP64)
        b = c3d.Channel(**cfg, precision=c3d.MIXED)
        for ch in (a, b):
            ch.upload(U, P, HH)
        for n in range(10):
            a.step(n + 1)
            d, _ = b.step(n + 1)
            assert d <= 1e-9
        sa, sb = a.download(), b.download()
        eu = [relerr(interior_u(sa["U"], n1, n2, n3, c), interior_u(sb["U"], n1, n2, n3, c)) for c in range(3)]
        print("tuned mixed vs FP64 after 10 steps: U", ["%.2e" % x for x in eu])
        assert max(eu) <= 1e-5
        a.close()
        b.close()
    finally:
        c3d.autotune_reset()


def test_autotune_set_accepts_candidates_only():
    """c3d_autotune_set records shapes the autotuner could have chosen and
    refuses any other (no kernel ever runs an unchecked tile shape)"""
    from compact3d_amd import c3d
    cfg = config(65, 49, 65, cks=1.5)
    try:
        t = c3d.autotune(steps=2, **cfg, precision=c3d.FP64)
        c3d.autotune_reset()
        c3d.autotune_set(t, **cfg, precision=c3d.FP64)
        ch = c3d.Channel(**cfg)
        ch.close()
        for key, val in (("y_cols_fp64", 7), ("x_lines", 3), ("fft_z_lines", 5)):
            bad = dict(t)
            bad[key] = val
            with pytest.raises(c3d.C3DError):
                c3d.autotune_set(bad, **cfg, precision=c3d.FP64)
    finally:
        c3d.autotune_reset()
