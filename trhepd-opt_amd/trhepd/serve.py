"""Rocking-curve serving on one device: request streams and the client side of the
library's continuous-batching loop (trhepd_serve_submit / _step / _result).

A request is one synthetic structure (seeded, uneven beam count) and its glancing-angle
scan.  The client submits each request when its arrival time comes, keeps stepping the
loop, and collects results as they finish; throughput and latency percentiles come from
the collected latencies (submit to completion, measured by the library)."""
import time
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import driver, lib, synth


@dataclass
class Request:
    """Inputs of one rocking-curve request (one structure, one angle scan)."""
    key: str
    pot: dict
    dz: float
    ns: int
    method: int
    iv: np.ndarray
    G: np.ndarray
    F: np.ndarray
    W: np.ndarray
    nb0: int

    @property
    def nb(self) -> int:
        return self.G.shape[1]

    @property
    def nang(self) -> int:
        return self.G.shape[0]


def structure_request(eng: "lib.Engine", st: "synth.Structure", dz2: float = 0.02, method: int = lib.SP4,
                      bulk_layers: Optional[int] = None, key: str = "") -> Request:
    """The request of a synthetic structure: potential descriptor (set up on the device),
    bulk reflection of `bulk_layers` bulk layers (the deck's ML by default), gamma and
    intensity weights."""
    blocks = synth.bulk_reflection(eng, st, layers=bulk_layers, dz2=dz2, method=method)
    job = driver.SurfaceJob(eng, st.bp, st.surf, dz2, method, build_table=False)
    G, F, W, _ = job.inputs(0, blocks)
    return Request(key, job.pot, job.dz, job.ns, method, job.iv, G, F, W, job.nb0)


def request_pool(eng: "lib.Engine", count: int, seed: int = 1234, cfg: int = 5, dz2: float = 0.02,
                 method: int = lib.SP4) -> List[Request]:
    """`count` seeded structures of configuration cfg (beam counts drawn from its size set)."""
    return [structure_request(eng, synth.config(cfg, seed=seed, index=i), dz2, method, key="cfg%d-%d" % (cfg, i))
            for i in range(count)]


def poisson_arrivals(n: int, rate: float, seed: int = 0) -> np.ndarray:
    """Arrival times (s) of a Poisson stream of mean rate `rate` per second."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, n))


def serve(eng: "lib.Engine", reqs: Sequence[Request], arrivals: Optional[Sequence[float]] = None,
          want_f: bool = False, timeout_s: float = 900.0, progress=None,
          progress_every_s: float = 30.0) -> Dict[str, object]:
    """Runs the request loop until every request has finished.

    Requests are submitted when the wall clock passes their arrival time (all at once when
    `arrivals` is None).  Returns per-request results (f2, latency_ms, optional f) in
    submission order, the maximum number admitted at once, and the span of the run.
    `progress(finished, total, admitted, queued)` is called every `progress_every_s` seconds."""
    arrivals = np.zeros(len(reqs)) if arrivals is None else np.asarray(arrivals, dtype=float)
    order = np.argsort(arrivals, kind="stable")
    ids: Dict[int, int] = {}
    out: List[Optional[tuple]] = [None] * len(reqs)
    t0 = time.perf_counter()
    nxt, left, peak = 0, len(reqs), 0
    t_report, active, queued = progress_every_s, 0, 0
    while left:
        now = time.perf_counter() - t0
        if progress is not None and now >= t_report:
            progress(len(reqs) - left, len(reqs), active, queued)
            t_report = now + progress_every_s
        if now > timeout_s:
            raise TimeoutError("serving loop exceeded %.0f s" % timeout_s)
        while nxt < len(order) and arrivals[order[nxt]] <= now:
            k = int(order[nxt])
            r = reqs[k]
            ids[eng.serve_submit(r.pot, r.dz, r.ns, r.method, r.iv, r.G, r.F, r.nb0, r.W)] = k
            nxt += 1
        active, queued, done = eng.serve_step()
        peak = max(peak, active)
        if done:
            for rid in list(ids):
                res = eng.serve_result(rid, want_f=want_f)
                if res is not None:
                    out[ids.pop(rid)] = res
                    left -= 1
        if not active and not queued and not done and nxt < len(order):
            time.sleep(max(0.0, min(0.05, arrivals[order[nxt]] - (time.perf_counter() - t0))))
    span = time.perf_counter() - t0
    return {"results": out, "span_s": span, "peak_active": peak}


def summary(reqs: Sequence[Request], run: Dict[str, object], first_arrival: float = 0.0) -> Dict[str, object]:
    """Throughput (points/s over the run), latency percentiles of a finished run, and per beam
    count the latency, queueing (submit to admission) and service (admission to completion)
    percentiles."""
    res = run["results"]
    lat = np.array([r[2] for r in res])
    que = np.array([r[3] for r in res])
    srv = np.array([r[4] for r in res])
    pts = sum(r.nang for r in reqs)
    span = run["span_s"] - first_arrival

    def pct(x):
        return {"p50": float(np.percentile(x, 50)), "p99": float(np.percentile(x, 99))}

    buckets = {}
    for nb in sorted({r.nb for r in reqs}):
        k = [i for i, r in enumerate(reqs) if r.nb == nb]
        buckets[str(nb)] = {"requests": len(k), "points": int(sum(reqs[i].nang for i in k)),
                            "latency_ms": pct(lat[k]), "queued_ms": pct(que[k]), "service_ms": pct(srv[k])}
    return {"requests": len(reqs), "points": pts, "span_s": span, "points_per_s": pts / span,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "max_ms": float(lat.max()), "queued_ms": pct(que), "service_ms": pct(srv),
            "peak_admitted": run["peak_active"], "by_beams": buckets}
