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


@pytest.mark.parametrize("files,wmodel", [(0.0, 0.0), (2.0, 1.0)])
def test_fp32_y_column_pairs_equal_single_columns(files, wmodel):
    """The FP32 y-sweep over neighbouring column pairs (8-byte loads and
    stores, y_pair_fp32 = 2) runs each column's arithmetic exactly as the
    one-column sweep of the same chunks: after 3 mixed steps (DNS, and LES
    with its SGS y terms) and after the stand-alone GETUH phase, every field
    is bit-identical.  Grid 48x97x32: NX = 48 leaves the second half of a
    64-column pair tile empty."""
    import numpy as np
    from compact3d_amd import c3d
    from pyoracle import Oracle
    n1, n2, n3 = 49, 97, 33
    cfg = config(n1, n2, n3, cks=1.5, files=files, wmodel=wmodel)
    try:
        t = c3d.autotune(steps=1, **cfg, precision=c3d.MIXED)
        orc = Oracle(**config(n1, n2, n3, cks=1.5))
        U, P, HH = inputs(orc, seed=5)
        out = {}
        # N2M = 96: C = 16 chunks of 6 points; 32 threads per chunk row
        for pair in (1, 2):
            s = dict(t)
            s.update(y_cols_fp32=32, y_chunks_fp32=16, y_pair_fp32=pair)
            c3d.autotune_set(s, **cfg, precision=c3d.MIXED)
            ch = c3d.Channel(**cfg, precision=c3d.MIXED)
            ch.upload(U, P, HH)
            for n in range(3):
                d, _ = ch.step(n + 1)
                assert d <= 1e-9
            st = ch.download()
            ch.bcond()
            ch.rhs()
            ch.getuh()
            out[pair] = (st["U"], st["P"], [ch.field(f) for f in (c3d.F_UH1, c3d.F_UH2, c3d.F_UH3,
                                                                    c3d.F_HH1, c3d.F_HH2, c3d.F_HH3)])
            ch.close()
        (u1, p1, f1), (u2, p2, f2) = out[1], out[2]
        print("pair vs single: max |dU| %.3e, |dP| %.3e" % (np.max(np.abs(u1 - u2)), np.max(np.abs(p1 - p2))))
        assert np.array_equal(u1, u2) and np.array_equal(p1, p2)
        for a, b in zip(f1, f2):
            assert np.array_equal(a, b)
    finally:
        c3d.autotune_reset()


@pytest.mark.parametrize("files,wmodel", [(0.0, 0.0), (2.0, 1.0)])
def test_mixed_wide_z_update_tiles_equal_narrow(files, wmodel):
    """Mixed mode's z-line update with 32-column tiles of 1024 threads
    (z_cols_fp64 = 32 at NZ = 256: 32 x 32 lines of 8-point segments) gives
    the fields of the 16-column, 512-thread tiles.  Each point's arithmetic
    is the same in both; only the per-block flow-rate partial sums change
    their order, so after 3 steps the fields agree to 1e-12 relative and
    DIVMAX stays <= 1e-9.  NX = 48 leaves half of the last 32-column tile of
    an x row empty."""
    import numpy as np
    from compact3d_amd import c3d
    from pyoracle import Oracle
    n1, n2, n3 = 49, 33, 257
    cfg = config(n1, n2, n3, cks=1.5, files=files, wmodel=wmodel)
    try:
        t = c3d.autotune(steps=1, **cfg, precision=c3d.MIXED)
        orc = Oracle(**config(n1, n2, n3, cks=1.5))
        U, P, HH = inputs(orc, seed=6)
        # 32 x 32 = 1024 threads: a candidate of mixed mode only, where the
        # FP64 z shape serves the update alone
        s = dict(t)
        s.update(z_cols_fp64=16)
        c3d.autotune_set(s, **cfg, precision=c3d.FP64)
        s.update(z_cols_fp64=32)
        with pytest.raises(c3d.C3DError):
            c3d.autotune_set(s, **cfg, precision=c3d.FP64)
        out = {}
        for cols in (16, 32):
            s = dict(t)
            s.update(z_cols_fp64=cols)
            c3d.autotune_set(s, **cfg, precision=c3d.MIXED)
            ch = c3d.Channel(**cfg, precision=c3d.MIXED)
            ch.upload(U, P, HH)
            dm = []
            for n in range(3):
                d, _ = ch.step(n + 1)
                assert d <= 1e-9
                dm.append(d)
            st = ch.download()
            out[cols] = (st["U"], st["P"], dm)
            ch.close()
        (u1, p1, d1), (u2, p2, d2) = out[16], out[32]
        eu, ep = relerr(u1, u2), relerr(p1, p2)
        print("32 vs 16 columns: U %.2e, P %.2e, DIVMAX %s / %s" % (eu, ep, d1, d2))
        assert eu <= 1e-12 and ep <= 1e-12
    finally:
        c3d.autotune_reset()
