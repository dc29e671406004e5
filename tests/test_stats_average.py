"""TAVER's running averages (compact3d.f:10755-11072) in
compact3d_amd.stats: the first call copies, later calls give the arithmetic
mean of all samples (a closed form: the running-mean recursion
XO_n = (XO_{n-1}(n-1) + X_n)/n equals (X_1 + ... + X_n)/n)."""
import numpy as np

from compact3d_amd.stats import TimeAverage


def test_first_call_copies_and_constant_stays():
    t = TimeAverage()
    a = np.arange(40.0).reshape(20, 2)
    t.add(stats=a, wsm=[1.0, 2.0, 3.0], taow=0.5)
    assert np.array_equal(t["stats"], a) and t.navg == 1
    a[0, 0] = 99.0  # the average holds its own copy
    assert t["stats"][0, 0] == 0.0
    for _ in range(5):
        t.add(stats=np.full((20, 2), 3.25), wsm=[1.0, 2.0, 3.0], taow=0.5)
    assert np.allclose(t["wsmo"], [1.0, 2.0, 3.0], rtol=0, atol=0) and t["taowo"] == 0.5


def test_running_mean_is_the_arithmetic_mean():
    t = TimeAverage()
    for n in range(1, 65):
        t.add(stats=np.full((20, 3), float(n)))
        # samples 1..n: mean (n+1)/2, exact in binary for these sizes
        assert np.all(t["stats"] == (n + 1) / 2.0), n
    rng = np.random.default_rng(5)
    xs = rng.standard_normal((50, 20, 7))
    t = TimeAverage()
    for x in xs:
        t.add(stats=x)
    assert np.max(np.abs(t["stats"] - xs.mean(axis=0))) <= 1e-14
    assert np.array_equal(t.profile("VM1"), t["stats"][0]) and np.array_equal(t.profile("VW2"), t["stats"][19])
