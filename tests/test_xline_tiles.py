"""Host-side check of the x-line tile layout the line kernels stage through
(c3d_lines.h, XStage/xstage): for every candidate tile of the autotuner
(c3d_api.cpp cand_lx, the rule xline_step) at the line lengths of the
benchmarked grids (48 to 256) and at 512, the staging's pair layout puts every point of
every valid line of a tile in exactly one LDS slot, at slot p*(L+1) + l, its
global offsets for rows J-1, J, J+1 stay inside the field, and every point a
thread's segment reads (its SEG points and one neighbour each side) is
staged.  The kernels rely on this; a tile outside it would read or write
out of bounds on the GPU."""
import pytest


def xline_step(seg, G):
    return seg if seg * G <= 256 else max(1, seg // 2)


def cand_lx(n):
    seg = 8 if n % 8 == 0 else 2
    G = n // seg
    step = xline_step(seg, G)
    out = []
    L = step
    while L * G <= 256 and L <= 64:
        if 8 * (2 * n * (L + 1) + 4 * G * L) <= 65536:
            out.append((L, seg))
        L += step
    return out


def check_tiles(n, L, SEG, n2m, nz):
    nx = n
    nj = n2m + 2
    fsz = nz * nj * nx
    G = n // SEG
    nt = L * G
    nh = n // 2
    assert nt % nh == 0, "the block must be a multiple of half a line"
    lstep2 = nt // nh
    LS = L + 1
    nlines = n2m * nz
    for blk in range((nlines + L - 1) // L):
        ln0 = blk * L
        staged = {}
        for t in range(nt):
            pq, lq = t % nh, t // nh
            li2 = 2 * pq * LS + lq
            for m in range(SEG // 2):
                ll = lq + m * lstep2
                ln = ln0 + ll
                if not (ll < L and ln < nlines):
                    continue
                jj = 1 + ln % n2m
                go = (ln // n2m * nj + jj) * nx + 2 * pq
                for joff in (-1, 0, 1):
                    off = (go + joff * nx) % (1 << 32)   # 32-bit offsets, summed first
                    assert 0 <= off and off + 1 < fsz
                for e in (0, 1):
                    slot = li2 + LS * e + m * lstep2
                    assert slot not in staged, "a slot staged twice"
                    staged[slot] = (ll, 2 * pq + e)
        for slot, (ll, p) in staged.items():
            assert slot == p * LS + ll
        nvalid = min(L, nlines - ln0)
        assert len(staged) == nvalid * n
        for t in range(nt):
            l, gs = t % L, t // L
            p0 = gs * SEG
            assert p0 + SEG <= n
            if l < nvalid:
                for q in range(-1, SEG + 1):
                    assert ((p0 + q) % n) * LS + l in staged


@pytest.mark.parametrize("n", [48, 64, 96, 128, 192, 256, 512])
def test_xline_tiles_stage_every_point_once(n):
    cands = cand_lx(n)
    assert cands, "no x-line tile for line length %d" % n
    for L, seg in cands:
        for n2m, nz in ((5, 3), (12, 4)):
            check_tiles(n, L, seg, n2m, nz)


def test_lines_of_512_points_take_four_per_tile():
    assert cand_lx(512) == [(4, 8)]
    assert cand_lx(1024) == []
