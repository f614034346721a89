"""The serving loop (continuous batching of rocking-curve requests).

Each request's result must be the same alone and in a mix with other requests that arrive
while it runs; the bar is bit-identity, since every request runs its own kernels on its own
device state and no schedule decision depends on the other requests.  Requests come from the
seeded uneven-size configuration (N drawn from 23 ... 521), at a coarse dz and a few angles to
keep the test short."""
import numpy as np
import pytest

from trhepd import lib, serve, synth

pytestmark = pytest.mark.gpu


def _structure_with(nb, seed=11):
    for i in range(200):
        st = synth.config(5, seed=seed, index=i)
        if st.bp.nb[0] == nb:
            return st
    raise LookupError(nb)


@pytest.fixture(scope="module")
def reqs(engine):
    out = []
    for nb in (23, 97, 199):
        r = serve.structure_request(engine, _structure_with(nb), dz2=0.1, key="n%d" % nb)
        k = slice(0, 6)
        out.append(serve.Request(r.key, r.pot, r.dz, r.ns, r.method, r.iv, r.G[k], r.F[k], r.W[k], r.nb0))
    return out


def _alone(eng, reqs):
    return [serve.serve(eng, [r])["results"][0][1] for r in reqs]


@pytest.mark.parametrize("precision,period", [(lib.PREC_FP64, 0), (lib.PREC_FP64, -1), (lib.PREC_MIXED, -1)])
def test_same_alone_and_in_the_mix(reqs, precision, period):
    print("precision", precision, "period", period, flush=True)
    with lib.Engine(0, precision=precision) as eng:
        eng.set_renorm_period(period)
        alone = _alone(eng, reqs)
        # the larger requests arrive while the first one runs
        run = serve.serve(eng, reqs, arrivals=[0.0, 0.05, 0.1])
    mix = [r[1] for r in run["results"]]
    print("peak admitted", run["peak_active"], "latencies ms", [round(r[2], 1) for r in run["results"]])
    for a, b in zip(alone, mix):
        assert np.array_equal(a, b)
        assert np.all(np.isfinite(a)) and np.all(a >= 0)


def test_serving_matches_propagate_batch(engine, reqs):
    got = _alone(engine, reqs)
    for r, f2 in zip(reqs, got):
        tid = engine.table_build(r.pot, r.dz, r.ns, r.method)
        try:
            _, ref = engine.propagate_batch(tid, r.iv, r.G, r.F, r.nb0, r.W)
        finally:
            engine.table_release(tid)
        assert np.array_equal(f2, ref)


def test_admission_waits_for_the_hbm_budget(engine, reqs):
    """With a budget that holds the largest request but not two of them, requests are
    admitted one by one as memory frees up, and still give the same results."""
    alone = _alone(engine, reqs)
    r = reqs[-1]
    with lib.Engine(0) as probe:
        # the context keeps its staging scratch after the first table build; a request
        # needs its table and batch on top of that
        probe.table_release(probe.table_build(r.pot, r.dz, r.ns, r.method))
        base = probe.bytes_held
        tid = probe.table_build(r.pot, r.dz, r.ns, r.method)
        bid = probe.batch_create(tid, r.iv, r.G, r.F, r.nb0, r.W)
        need = probe.bytes_held - base
        probe.batch_release(bid)
    print("request bytes", need, "context scratch", base)
    with lib.Engine(0, hbm_budget=base + int(need * 1.5)) as eng:
        run = serve.serve(eng, [r, r, r])
        # a small request fits beside a large one: both are admitted together
        mixed = serve.serve(eng, [r, reqs[1], r])
    assert run["peak_active"] == 1
    for k in range(3):
        assert np.array_equal(run["results"][k][1], alone[-1])
    assert mixed["peak_active"] == 2
    assert np.array_equal(mixed["results"][0][1], alone[-1])
    assert np.array_equal(mixed["results"][1][1], alone[1])
    assert np.array_equal(mixed["results"][2][1], alone[-1])


def test_request_larger_than_the_budget_fails(reqs):
    with lib.Engine(0, hbm_budget=1 << 20) as eng:
        rid = eng.serve_submit(reqs[-1].pot, reqs[-1].dz, reqs[-1].ns, reqs[-1].method, reqs[-1].iv,
                               reqs[-1].G, reqs[-1].F, reqs[-1].nb0, reqs[-1].W)
        eng.serve_step()
        with pytest.raises(lib.TrhepdError) as e:
            eng.serve_result(rid)
    assert e.value.code == lib.E_BUDGET


def test_mix_with_an_n521_request(engine, reqs):
    """An N = 521 request (the largest of the serving configuration) admitted while smaller
    ones run, and a small one arriving while it runs: every result is bit-identical to the
    request served alone and to propagate_batch."""
    st = _structure_with(521)
    r = serve.structure_request(engine, st, dz2=0.1, key="n521")
    k = slice(0, 4)
    big = serve.Request(r.key, r.pot, r.dz, r.ns, r.method, r.iv, r.G[k], r.F[k], r.W[k], r.nb0)
    mix = [reqs[1], big, reqs[0], reqs[2]]
    print("N", [q.nb for q in mix], "slices", [q.ns for q in mix], flush=True)
    with lib.Engine(0) as eng:
        eng.set_renorm_period(-1)
        alone = _alone(eng, mix)
        run = serve.serve(eng, mix, arrivals=[0.0, 0.02, 0.3, 0.35])
        tid = eng.table_build(big.pot, big.dz, big.ns, big.method)
        try:
            _, ref = eng.propagate_batch(tid, big.iv, big.G, big.F, big.nb0, big.W)
        finally:
            eng.table_release(tid)
    print("peak admitted", run["peak_active"], "latencies ms", [round(q[2], 1) for q in run["results"]])
    for a, b in zip(alone, [q[1] for q in run["results"]]):
        assert np.array_equal(a, b)
        assert np.all(np.isfinite(a)) and np.all(a >= 0)
    assert np.array_equal(alone[1], ref)
    assert run["peak_active"] >= 2


def _split(r, parts):
    """the request's angles as `parts` requests of consecutive angles (one structure)"""
    out = []
    for k, idx in enumerate(np.array_split(np.arange(r.nang), parts)):
        out.append(serve.Request("%s/%d" % (r.key, k), r.pot, r.dz, r.ns, r.method, r.iv, r.G[idx], r.F[idx],
                                 r.W[idx], r.nb0))
    return out


@pytest.mark.parametrize("precision,period", [(lib.PREC_FP64, 6), (lib.PREC_FP64, 0), (lib.PREC_MIXED, 6)])
def test_merged_requests_same_bits_as_alone(engine, precision, period):
    """Requests of one structure (an N = 97 scan cut into three) merge into one batch with a
    request of another structure queued between them; every result is bit-identical to the
    request served alone, and the merged ones share admission and completion."""
    r97 = serve.structure_request(engine, _structure_with(97), dz2=0.1, key="m97")
    r23 = serve.structure_request(engine, _structure_with(23), dz2=0.1, key="m23")
    parts = _split(serve.Request(r97.key, r97.pot, r97.dz, r97.ns, r97.method, r97.iv, r97.G[:9], r97.F[:9],
                                 r97.W[:9], r97.nb0), 3)
    mix = [parts[0], r23, parts[1], parts[2]]
    print("precision", precision, "period", period, "angles", [q.nang for q in mix], flush=True)
    with lib.Engine(0, precision=precision) as eng:
        eng.set_renorm_period(period)
        alone = _alone(eng, mix)
        eng.serve_set_merge(1000, 0.0)
        run = serve.serve(eng, mix)
    res = run["results"]
    for a, b in zip(alone, [q[1] for q in res]):
        assert np.array_equal(a, b)
    # the three parts were admitted and finished together (equal service times), the N = 23
    # request on its own
    svc = [q[4] for q in res]
    assert svc[0] == svc[2] == svc[3]
    assert svc[1] != svc[0]


def test_merging_waits_for_its_window_and_respects_the_budget(engine):
    """A later request of the same structure joins the head within the merge window; under a
    budget that holds one request but not two, a merged group is retried smaller and the
    requests run one at a time, with the same results."""
    r97 = serve.structure_request(engine, _structure_with(97), dz2=0.1, key="w97")
    a, b = _split(serve.Request(r97.key, r97.pot, r97.dz, r97.ns, r97.method, r97.iv, r97.G[:8], r97.F[:8],
                                r97.W[:8], r97.nb0), 2)
    with lib.Engine(0) as eng:
        eng.set_renorm_period(6)
        alone = _alone(eng, [a, b])
        eng.serve_set_merge(1000, 150.0)
        run = serve.serve(eng, [a, b], arrivals=[0.0, 0.05])
        with pytest.raises(lib.TrhepdError):
            eng.serve_set_merge(-1, 0.0)
    res = run["results"]
    assert np.array_equal(res[0][1], alone[0]) and np.array_equal(res[1][1], alone[1])
    assert res[0][4] == res[1][4]            # one batch: same service time
    assert res[0][3] >= 150.0                # the head waited out its window
    with lib.Engine(0) as probe:
        probe.set_renorm_period(6)
        probe.table_release(probe.table_build(a.pot, a.dz, a.ns, a.method))
        base = probe.bytes_held
        tid = probe.table_build(a.pot, a.dz, a.ns, a.method)
        bid = probe.batch_create(tid, a.iv, a.G, a.F, a.nb0, a.W)
        one = probe.bytes_held - base
        probe.batch_release(bid)
        bid = probe.batch_create(tid, a.iv, np.concatenate([a.G, b.G]), np.concatenate([a.F, b.F]), a.nb0,
                                 np.concatenate([a.W, b.W]))
        two = probe.bytes_held - base
        probe.batch_release(bid)
        probe.table_release(tid)
    print("bytes: one request %d, the two merged %d" % (one, two))
    assert two > one
    with lib.Engine(0, hbm_budget=base + (one + two) // 2) as eng:
        eng.set_renorm_period(6)
        eng.serve_set_merge(1000, 0.0)
        run = serve.serve(eng, [a, b])
    assert run["peak_active"] == 1
    assert np.array_equal(run["results"][0][1], alone[0]) and np.array_equal(run["results"][1][1], alone[1])
