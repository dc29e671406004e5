"""The serving loop's scheduling on the CPU, with stand-in channels: admission
by footprint and slots, continuous admission of late arrivals, latency
bookkeeping, and refusal of requests larger than the budget."""
import pytest

from compact3d_amd.serve import Request, Server, poisson_requests, re_from_re_tau, stats


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += max(dt, 1e-6)


class FakeChannel:
    """advance() costs 1 ms of clock per step; records what ran"""

    live = 0
    peak = 0

    def __init__(self, req, clock, log):
        self.req, self.clock, self.log = req, clock, log
        FakeChannel.live += 1
        FakeChannel.peak = max(FakeChannel.peak, FakeChannel.live)
        self.steps = 0

    def mesh(self):
        return None

    def upload(self, U, P):
        pass

    def advance(self, n, use_graph=True):
        self.clock.t += 1e-3 * n
        self.steps += n
        self.log.append((self.req.rid, n))

    def sync(self):
        return 1e-14, 0.1

    def close(self):
        FakeChannel.live -= 1


class NoFields:
    def get(self, req, mesh):
        return None, None


def _server(budget, max_live, chunk=10):
    clock = FakeClock()
    log = []
    srv = Server(budget, max_live=max_live, chunk=chunk, make_channel=lambda r: FakeChannel(r, clock, log),
                 footprint=lambda r: r.pts, fields=NoFields(), clock=clock, sleep=clock.sleep)
    return srv, clock, log


def test_every_request_finishes_with_its_steps():
    FakeChannel.live = FakeChannel.peak = 0
    reqs = [Request(rid=q, grid=(8, 9, 8), re=100.0, steps=25 + 5 * q, t_arrive=0.01 * q) for q in range(6)]
    srv, clock, log = _server(budget=10 ** 9, max_live=4)
    done = srv.run(reqs)
    assert sorted(r.rid for r in done) == list(range(6))
    ran = {}
    for rid, n in log:
        ran[rid] = ran.get(rid, 0) + n
    assert all(ran[r.rid] == r.steps for r in reqs)
    assert FakeChannel.peak <= 4 and FakeChannel.live == 0
    for r in done:
        assert r.t_arrive <= r.t_admit <= r.t_done


def test_admission_respects_the_memory_budget():
    FakeChannel.live = FakeChannel.peak = 0
    big = Request(rid=0, grid=(8, 9, 8), re=100.0, steps=30, t_arrive=0.0)      # 512 "bytes"
    reqs = [big] + [Request(rid=q, grid=(8, 9, 8), re=100.0, steps=10, t_arrive=0.0) for q in range(1, 4)]
    srv, clock, log = _server(budget=2 * big.pts, max_live=64)
    done = srv.run(reqs)
    assert srv.peak_bytes <= 2 * big.pts
    assert FakeChannel.peak == 2
    assert len(done) == 4


def test_late_arrivals_join_while_others_run():
    FakeChannel.live = FakeChannel.peak = 0
    long = Request(rid=0, grid=(8, 9, 8), re=100.0, steps=200, t_arrive=0.0)
    late = Request(rid=1, grid=(8, 9, 8), re=100.0, steps=10, t_arrive=0.05)
    srv, clock, log = _server(budget=10 ** 9, max_live=8)
    done = srv.run([long, late])
    assert [r.rid for r in done] == [1, 0]  # the short late request finishes first
    assert late.t_admit < long.t_done


def test_oversized_request_is_refused():
    srv, clock, log = _server(budget=10, max_live=8)
    with pytest.raises(ValueError):
        srv.run([Request(rid=0, grid=(8, 9, 8), re=100.0, steps=5)])


def test_stats_and_request_stream():
    reqs = poisson_requests(50, 20.0, [(64, 49, 64), (96, 65, 96)], [180, 395, 1000], (20, 40), seed=4)
    assert all(b.t_arrive >= a.t_arrive for a, b in zip(reqs, reqs[1:]))
    assert all(r.precision == 1 for r in reqs)
    reqs64 = poisson_requests(50, 20.0, [(64, 49, 64)], [180, 395, 1000], (20, 40), seed=4, les_precision=0)
    assert all(r.precision == (0 if r.les else 1) for r in reqs64)
    assert abs(re_from_re_tau(180.0) - 4200) / 4200 < 0.05  # the reference's RE for Re_tau 180
    for r in reqs:
        r.t_done = r.t_arrive + 0.5
    s = stats(reqs, 10.0)
    assert s["p50_s"] == pytest.approx(0.5) and s["p99_s"] == pytest.approx(0.5)
    assert s["throughput_mpts_steps_s"] > 0


class DivergingChannel(FakeChannel):
    """a stand-in whose state diverges: sync() reports NaN from its second
    chunk on"""

    def sync(self):
        return (float("nan"), float("nan")) if self.steps > 10 else (1e-14, 0.1)


def test_diverged_request_is_failed_and_evicted():
    FakeChannel.live = FakeChannel.peak = 0
    clock = FakeClock()
    log = []
    bad_rid = 1

    def make(r):
        cls = DivergingChannel if r.rid == bad_rid else FakeChannel
        return cls(r, clock, log)
    reqs = [Request(rid=q, grid=(8, 9, 8), re=100.0, steps=40, t_arrive=0.0) for q in range(3)]
    # the budget holds two instances: the third is admitted when the diverged one is evicted
    srv = Server(2 * reqs[0].pts, max_live=8, chunk=10, make_channel=make, footprint=lambda r: r.pts,
                 fields=NoFields(), clock=clock, sleep=clock.sleep)
    done = srv.run(reqs)
    assert sorted(r.rid for r in done) == [0, 2]
    assert [r.rid for r in srv.failed] == [bad_rid]
    bad = srv.failed[0]
    assert "DIVMAX" in bad.error and bad.result["divmax"] != bad.result["divmax"]
    ran = {}
    for rid, n in log:
        ran[rid] = ran.get(rid, 0) + n
    assert ran[bad_rid] == 20 and ran[0] == 40 and ran[2] == 40
    assert reqs[2].t_admit >= bad.t_done  # admitted into the freed memory
    assert FakeChannel.live == 0


def test_diverged_request_raised_by_the_binding():
    """the binding's C3DDiverged (C3D_ERR_NONFINITE) is handled like a NaN"""
    from compact3d_amd.c3d import C3DDiverged
    FakeChannel.live = FakeChannel.peak = 0
    clock = FakeClock()
    log = []

    class Raising(FakeChannel):
        def sync(self):
            raise C3DDiverged("sync: the state diverged", float("inf"), 1.0)
    srv = Server(10 ** 9, max_live=8, chunk=10, make_channel=lambda r: (Raising if r.rid == 0 else FakeChannel)(
        r, clock, log), footprint=lambda r: r.pts, fields=NoFields(), clock=clock, sleep=clock.sleep)
    done = srv.run([Request(rid=q, grid=(8, 9, 8), re=100.0, steps=30, t_arrive=0.0) for q in range(2)])
    assert [r.rid for r in done] == [1] and [r.rid for r in srv.failed] == [0]
    assert srv.failed[0].result["divmax"] == float("inf")


def test_arrival_admitted_at_the_next_chunk_boundary():
    """a request arriving while instances run is admitted when the next
    instance's chunk completes, before the other live instances' chunks"""
    FakeChannel.live = FakeChannel.peak = 0
    reqs = [Request(rid=q, grid=(8, 9, 8), re=100.0, steps=100, t_arrive=0.0) for q in range(3)]
    late = Request(rid=9, grid=(8, 9, 8), re=100.0, steps=10, t_arrive=0.025)
    srv, clock, log = _server(budget=10 ** 9, max_live=8, chunk=10)
    srv.run(reqs + [late])
    # first queue: 0, 1, 2 (clock 0.03); the late request joins right after
    # the first completion, ahead of the re-queues of instances 1 and 2
    assert log[:5] == [(0, 10), (1, 10), (2, 10), (0, 10), (9, 10)]


def _server_policy(budget, max_live, chunk=10, **kw):
    clock = FakeClock()
    log = []
    srv = Server(budget, max_live=max_live, chunk=chunk, make_channel=lambda r: FakeChannel(r, clock, log),
                 footprint=lambda r: r.pts, fields=NoFields(), clock=clock, sleep=clock.sleep, **kw)
    return srv, clock, log


def test_srpt_serves_the_least_work_first():
    FakeChannel.live = FakeChannel.peak = 0
    steps = [40, 10, 30, 20]
    reqs = [Request(rid=q, grid=(8, 9, 8), re=100.0, steps=s, t_arrive=0.0) for q, s in enumerate(steps)]
    srv, clock, log = _server_policy(10 ** 9, 8, policy="srpt", max_active=1)
    done = srv.run(reqs)
    assert [r.rid for r in done] == [1, 3, 2, 0]
    # one instance at a time, each to the end of its steps
    assert [rid for rid, n in log] == [1, 3, 3, 2, 2, 2, 0, 0, 0, 0]
    assert all(n == 10 for rid, n in log) and FakeChannel.live == 0


def test_srpt_weighs_cells_times_steps():
    """a request of a bigger grid counts its cells: 10 steps of 4x the cells
    rank after 30 steps of the small grid"""
    FakeChannel.live = FakeChannel.peak = 0
    small = Request(rid=0, grid=(8, 9, 8), re=100.0, steps=30, t_arrive=0.0)
    big = Request(rid=1, grid=(16, 9, 16), re=100.0, steps=10, t_arrive=0.0)
    srv, clock, log = _server_policy(10 ** 9, 8, policy="srpt", max_active=1)
    done = srv.run([big, small])
    assert [r.rid for r in done] == [0, 1]


def test_srpt_late_short_request_overtakes_a_long_one():
    FakeChannel.live = FakeChannel.peak = 0
    long = Request(rid=0, grid=(8, 9, 8), re=100.0, steps=200, t_arrive=0.0)
    late = Request(rid=1, grid=(8, 9, 8), re=100.0, steps=10, t_arrive=0.05)
    srv, clock, log = _server_policy(10 ** 9, 8, policy="srpt", max_active=1)
    done = srv.run([long, late])
    assert [r.rid for r in done] == [1, 0]
    # admitted at the next chunk boundary and run before the long one's next chunk
    assert late.t_done - late.t_arrive <= 0.02 + 1e-12
    i = log.index((1, 10))
    assert all(rid == 0 for rid, n in log[:i]) and all(rid == 0 for rid, n in log[i + 1:])
    assert sum(n for rid, n in log if rid == 0) == 200


def test_srpt_mean_latency_not_above_fifo():
    """on a seeded stream of mixed sizes, shortest remaining work first does
    not raise the mean latency over arrival order with every instance
    sharing the (stand-in, serial) device; every request runs all its steps"""
    import random
    rng = random.Random(7)
    base = [Request(rid=q, grid=rng.choice([(8, 9, 8), (16, 9, 16)]), re=100.0, steps=rng.randrange(10, 200, 10),
                    t_arrive=0.01 * q) for q in range(12)]
    lat = {}
    for policy in ("fifo", "srpt"):
        FakeChannel.live = FakeChannel.peak = 0
        reqs = [Request(rid=r.rid, grid=r.grid, re=r.re, steps=r.steps, t_arrive=r.t_arrive) for r in base]
        srv, clock, log = _server_policy(10 ** 9, 16, policy=policy, max_active=1)
        done = srv.run(reqs)
        assert len(done) == 12 and FakeChannel.live == 0
        ran = {}
        for rid, n in log:
            ran[rid] = ran.get(rid, 0) + n
        assert all(ran[r.rid] == r.steps for r in reqs)
        lat[policy] = sum(r.t_done - r.t_arrive for r in done) / len(done)
    assert lat["srpt"] <= lat["fifo"]


def test_unknown_policy_is_refused():
    with pytest.raises(ValueError):
        Server(10, policy="lifo")
    with pytest.raises(ValueError):
        Server(10, policy="srpt", max_active=0)
