"""c3d_stats_noalloc (the allocation-free entry point, scratch in a
library-owned device array) against c3d_stats on the same resident field,
through the C-ABI: equal bit for bit on two calls in a row, the second on a
smaller grid (fewer planes than the first left in the shared array), and a
grid beyond the array's plane limit is refused."""
import numpy as np
import pytest

from cases import config

pytestmark = pytest.mark.gpu


def _channel(n1, n2, n3, seed):
    from compact3d_amd import c3d
    from compact3d_amd.initfield import perturbed_poiseuille
    cfg = config(n1, n2, n3, cks=1.5)
    ch = c3d.Channel(**cfg)
    m = ch.mesh()
    U, P = perturbed_poiseuille(n1, n2, n3, m["Y"], m["DX1"], m["DX3"], m["ALY"], seed=seed)
    rng = np.random.default_rng(seed)
    P[: n1 - 1, : n2 - 1, : n3 - 1] = rng.uniform(-1, 1, (n1 - 1, n2 - 1, n3 - 1)) * 1e-2
    ch.upload(U, P)
    return ch


def test_stats_noalloc_equals_stats():
    print("stats_noalloc: 65x97x49 then 33x33x33", flush=True)
    big = _channel(65, 97, 49, seed=1)
    small = _channel(33, 33, 33, seed=2)
    for ch in (big, small):
        ch.step(1)
        a = ch.stats_noalloc()
        b = ch.stats()
        print("planes", ch.n2 - 1, "max |VM1|", float(np.max(np.abs(b[0]))), flush=True)
        assert np.all(np.isfinite(a))
        assert np.array_equal(a, b)
    # the small grid's call after the big one's: only its own planes are read
    again = big.stats_noalloc()
    assert np.array_equal(again, big.stats())
    big.close()
    small.close()


def test_stats_noalloc_refuses_more_planes_than_its_scratch():
    from compact3d_amd import c3d
    print("stats_noalloc: 9x2050x9 refused", flush=True)
    ch = c3d.Channel(**config(9, 2050, 9, cks=1.5))
    with pytest.raises(c3d.C3DError, match="exceeds its scratch"):
        ch.stats_noalloc()
    ch.close()
