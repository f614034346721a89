"""The serving client loop (trhepd/serve.py) on the CPU, against a scripted stand-in for the
library's request loop: submission at arrival times, collection of finished requests in
submission order, progress reports, the time limit, and the summary's percentiles."""
import numpy as np
import pytest

from trhepd import serve


class ScriptedLoop:
    """Stands in for lib.Engine's serve_* calls: request k finishes `steps[k]` steps after its
    submission; latency, queueing and service times are derived from the step counts."""

    def __init__(self, steps):
        self.steps = list(steps)
        self.left, self.info, self.next_id, self.calls = {}, {}, 100, 0

    def serve_submit(self, pot, dz, ns, method, iv, G, F, nb0, W):
        rid = self.next_id
        self.next_id += 1
        k = len(self.info)
        self.info[rid] = (k, G.shape)
        self.left[rid] = self.steps[k]
        return rid

    def serve_step(self):
        self.calls += 1
        for rid in self.left:
            if self.left[rid] > 0:
                self.left[rid] -= 1
        active = sum(1 for v in self.left.values() if v > 0)
        done = sum(1 for v in self.left.values() if v == 0)
        return active, 0, done

    def serve_result(self, rid, want_f=False):
        if self.left.get(rid, 1) != 0:
            return None
        del self.left[rid]
        k, shape = self.info[rid]
        f2 = np.full(shape, float(k))
        return None, f2, 10.0 * (k + 1), 1.0 * (k + 1), 9.0 * (k + 1)


def _req(k, nb, nang):
    G = np.zeros((nang, nb), dtype=complex)
    return serve.Request(key="s%d" % k, pot={}, dz=0.1, ns=10, method=5, iv=np.zeros((nb, nb), np.int32), G=G,
                         F=np.zeros((nang, nb, nb), complex), W=np.ones(nang), nb0=1)


def test_results_in_submission_order_and_summary():
    reqs = [_req(k, nb, nang) for k, (nb, nang) in enumerate([(23, 4), (97, 3), (23, 5), (521, 2)])]
    loop = ScriptedLoop([5, 1, 3, 2])
    run = serve.serve(loop, reqs)
    assert [int(r[1][0, 0]) for r in run["results"]] == [0, 1, 2, 3]
    assert run["peak_active"] >= 1
    out = serve.summary(reqs, run)
    assert out["requests"] == 4 and out["points"] == 14
    assert out["p50_ms"] == pytest.approx(np.percentile([10, 20, 30, 40], 50))
    assert out["max_ms"] == 40.0
    assert out["by_beams"]["23"]["requests"] == 2 and out["by_beams"]["23"]["points"] == 9
    assert out["by_beams"]["521"]["latency_ms"]["p50"] == 40.0
    assert out["by_beams"]["97"]["queued_ms"]["p50"] == 2.0
    assert out["by_beams"]["97"]["service_ms"]["p50"] == 18.0


def test_arrivals_progress_and_time_limit():
    reqs = [_req(k, 23, 2) for k in range(3)]
    reports = []
    run = serve.serve(ScriptedLoop([1, 1, 1]), reqs, arrivals=[0.0, 0.02, 0.04],
                      progress=lambda *a: reports.append(a), progress_every_s=0.0)
    assert len(run["results"]) == 3 and all(r is not None for r in run["results"])
    assert run["span_s"] >= 0.04
    assert reports and all(r[1] == 3 for r in reports) and reports[-1][0] <= 3
    # a request that never finishes ends the loop at the time limit
    with pytest.raises(TimeoutError):
        serve.serve(ScriptedLoop([10 ** 9]), [_req(0, 23, 1)], timeout_s=0.05)
