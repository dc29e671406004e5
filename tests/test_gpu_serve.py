"""Ensemble and serving on the GPU: the batched driver (c3d_advance, with and
without captured graphs) and the ensemble graph (every member's step a
branch of one graph launch) equal stepping one call at a time, and a
request's result is the same alone and served in a mix of other instances
(bit for bit: an instance's arithmetic does not depend on what else is
resident)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cfg(n1, n2, n3, re=4200.0, les=False):
    return dict(n1=n1, n2=n2, n3=n3, re=re, alx=2 * np.pi, aly=2.0, alz=np.pi, dt=1e-3,
                files=2.0 if les else 0.0, wmodel=1.0 if les else 0.0, cks=1.5)


def _field(ch, cfg, seed):
    from compact3d_amd.initfield import perturbed_poiseuille
    m = ch.mesh()
    return perturbed_poiseuille(cfg["n1"], cfg["n2"], cfg["n3"], m["Y"], m["DX1"], m["DX3"], m["ALY"], seed=seed)


@pytest.mark.parametrize("precision", [0, 1])
def test_advance_and_graphs_equal_stepping(precision):
    from compact3d_amd import c3d
    cfg = _cfg(49, 41, 33)
    chs = [c3d.Channel(**cfg, precision=precision) for _ in range(3)]
    U, P = _field(chs[0], cfg, 5)
    for ch in chs:
        ch.upload(U, P)
    for n in range(7):
        chs[0].step(n + 1)
    chs[1].advance(7, use_graph=False)
    chs[2].advance(3, use_graph=True)
    chs[2].advance(4, use_graph=True)   # odd count: graphs alternate across calls
    d1, _ = chs[1].sync()
    d2, _ = chs[2].sync()
    s = [ch.download() for ch in chs]
    for q in (1, 2):
        assert np.array_equal(s[0]["U"], s[q]["U"]) and np.array_equal(s[0]["P"], s[q]["P"])
        assert s[0]["PRESG"] == s[q]["PRESG"]
    assert d1 <= 1e-9 and d2 <= 1e-9
    for ch in chs:
        ch.close()


@pytest.mark.parametrize("policy", ["fifo", "srpt"])
def test_served_result_equals_alone(policy):
    from compact3d_amd import c3d
    from compact3d_amd.serve import Request, Server
    reqs = [Request(rid=0, grid=(48, 33, 48), re=4200.0, steps=23, seed=1, precision=1, dt=1e-3),
            Request(rid=1, grid=(64, 49, 64), re=10312.5, steps=40, seed=2, precision=0, dt=5e-4, t_arrive=0.001),
            Request(rid=2, grid=(32, 33, 32), re=20000.0, steps=17, seed=3, precision=0, les=True, dt=5e-4,
                    t_arrive=0.002),
            Request(rid=3, grid=(48, 33, 48), re=4200.0, steps=31, seed=1, precision=0, dt=1e-3, t_arrive=0.003)]
    srv = Server(budget_bytes=8 << 30, max_live=8, chunk=6, keep_results=True, policy=policy, max_active=2)
    done = srv.run(reqs)
    assert len(done) == 4
    for r in done:
        ch = c3d.Channel(**r.config(), precision=r.precision)
        U, P = srv.fields.get(r, ch.mesh())
        ch.upload(U, P)
        for n in range(r.steps):
            ch.step(n + 1)
        st = ch.download()
        ch.close()
        print("request", r.rid, "latency %.3f s" % (r.t_done - r.t_arrive), "divmax %.2e" % r.result["divmax"])
        assert np.array_equal(st["U"], r.result["U"]), f"request {r.rid}: U differs alone vs served"
        assert np.array_equal(st["P"], r.result["P"]), f"request {r.rid}: P differs alone vs served"
        assert r.result["divmax"] <= 1e-9


def test_ensemble_equals_alone():
    """members of different grids, precisions and models, advanced by one
    graph launch per step, end bit for bit where each ends stepped alone
    (the integral wall model's member, WMODEL = 2, counts its steps on the
    device: its compute step NTIME = 5 falls inside the ensemble's run)"""
    from compact3d_amd import c3d
    specs = [(_cfg(49, 41, 33), 0, 5), (_cfg(33, 33, 33, les=True, re=20000.0), 1, 6),
             (_cfg(65, 33, 49), 1, 7), (_cfg(33, 33, 33, les=True, re=20000.0), 0, 8),
             (dict(_cfg(33, 33, 33, les=True, re=20000.0), wmodel=2.0, cflmax=0.5), 1, 9)]
    ens_ch, alone = [], []
    for cfg, prec, seed in specs:
        pair = [c3d.Channel(**cfg, precision=prec) for _ in range(2)]
        U, P = _field(pair[0], cfg, seed)
        for ch in pair:
            ch.upload(U, P)
        ens_ch.append(pair[0])
        alone.append(pair[1])
    ens = c3d.Ensemble(ens_ch)
    ens.advance(3)
    ens.advance(4)   # odd counts: the two captured graphs alternate across calls
    dm, cf = ens.sync()
    ens.close()
    for q, ch in enumerate(alone):
        for n in range(7):
            d, c = ch.step(n + 1)
        a, b = ch.download(), ens_ch[q].download()
        print("member", q, "divmax ensemble %.2e alone %.2e" % (dm[q], d))
        assert np.array_equal(a["U"], b["U"]) and np.array_equal(a["P"], b["P"]), f"member {q} differs"
        assert a["PRESG"] == b["PRESG"]
        assert dm[q] == d and cf[q] == c
        assert dm[q] <= 1e-9
    # a member stays usable on its own after the ensemble is released
    ens_ch[0].step(8)
    for ch in ens_ch + alone:
        ch.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_batched_ensemble_equals_alone(precision):
    """members of the same shape run batched (one launch per kernel for all
    of them, each block reading its member's context): three wall-modelled
    LES channels of one grid at different Reynolds numbers (the log law at
    both walls and at the bottom wall), two DNS channels of another grid,
    one resolved-wall LES pair (WMODEL = 0, WALLSS feeds SGSCS) and an
    integral-wall-model member that runs alone; every member ends bit for
    bit where it ends stepped alone"""
    from compact3d_amd import c3d
    specs = [(dict(_cfg(49, 33, 33, les=True, re=20000.0), wmodel=6.0), 1),
             (dict(_cfg(49, 33, 33, les=True, re=40000.0), wmodel=6.0), 2),
             (dict(_cfg(49, 33, 33, les=True, re=90000.0), wmodel=6.0), 3),
             (_cfg(33, 41, 33), 4), (_cfg(33, 41, 33, re=8000.0), 5),
             (dict(_cfg(33, 33, 33, les=True, re=20000.0), wmodel=0.0), 6),
             (dict(_cfg(33, 33, 33, les=True, re=30000.0), wmodel=0.0), 7),
             (dict(_cfg(33, 33, 33, les=True, re=20000.0), wmodel=2.0, cflmax=0.5), 8)]
    ens_ch, alone = [], []
    for cfg, seed in specs:
        pair = [c3d.Channel(**cfg, precision=precision) for _ in range(2)]
        U, P = _field(pair[0], cfg, seed)
        for ch in pair:
            ch.upload(U, P)
            if cfg["files"]:
                ch.set_scalar(c3d.S_WSM1, 1e-3)
                ch.set_scalar(c3d.S_WSM3, 1e-4)
        ens_ch.append(pair[0])
        alone.append(pair[1])
    print("batched ensemble, precision", precision, flush=True)
    ens = c3d.Ensemble(ens_ch)
    ens.advance(5)
    ens.advance(4)
    dm, cf = ens.sync()
    ens.close()
    for q, ch in enumerate(alone):
        for n in range(9):
            d, c = ch.step(n + 1)
        a, b = ch.download(), ens_ch[q].download()
        print("member", q, "divmax ensemble %.2e alone %.2e" % (dm[q], d), flush=True)
        for key in ("U", "P", "HH", "SGSVIS"):
            assert np.array_equal(a[key], b[key]), f"member {q}: {key} differs"
        assert a["PRESG"] == b["PRESG"] and a["PRESG3"] == b["PRESG3"]
        assert dm[q] == d and cf[q] == c
    for ch in ens_ch + alone:
        ch.close()


def test_batched_ensemble_config3_grids_equal_alone():
    """the batched ensemble on the grid set of the 48-member LES benchmark
    (48x33x48, 64x49x64, 96x65x96, 128x65x128, two members each at
    different Re_tau from {1000, 2000, 5200}; Smagorinsky with the log law
    at both walls, mixed precision, dt 5e-4): every member ends bit for bit
    where it ends stepped alone, with its autotuned shapes"""
    from compact3d_amd import c3d
    from compact3d_amd.serve import re_from_re_tau
    grids = [(49, 33, 49), (65, 49, 65), (97, 65, 97), (129, 65, 129)]
    specs = []
    for q, (n1, n2, n3) in enumerate(grids):
        for r, rt in enumerate((1000, 2000, 5200)[q % 2:q % 2 + 2]):
            cfg = dict(_cfg(n1, n2, n3, les=True, re=re_from_re_tau(rt)), wmodel=6.0, dt=5e-4, cks=1.8)
            specs.append((cfg, 10 * q + r))
    try:
        for n1, n2, n3 in grids:
            c3d.autotune(steps=2, **dict(_cfg(n1, n2, n3, les=True), wmodel=6.0, dt=5e-4, cks=1.8),
                         precision=c3d.MIXED)
        ens_ch, alone = [], []
        for cfg, seed in specs:
            pair = [c3d.Channel(**cfg, precision=c3d.MIXED) for _ in range(2)]
            U, P = _field(pair[0], cfg, seed)
            for ch in pair:
                ch.upload(U, P)
            ens_ch.append(pair[0])
            alone.append(pair[1])
        ens = c3d.Ensemble(ens_ch)
        ens.advance(3)
        ens.advance(4)
        dm, cf = ens.sync()
        ens.close()
        for q, ch in enumerate(alone):
            for n in range(7):
                d, c = ch.step(n + 1)
            a, b = ch.download(), ens_ch[q].download()
            print("member", q, specs[q][0]["n1"] - 1, "divmax ensemble %.2e alone %.2e" % (dm[q], d), flush=True)
            for key in ("U", "P", "HH", "SGSVIS"):
                assert np.array_equal(a[key], b[key]), f"member {q}: {key} differs"
            assert dm[q] == d and cf[q] == c and d <= 1e-9
        for ch in ens_ch + alone:
            ch.close()
    finally:
        c3d.autotune_reset()
