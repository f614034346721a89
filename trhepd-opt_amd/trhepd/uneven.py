"""Uneven batches: many seeded structures of different beam counts integrated together.

Every structure gets its own table and one resident batch of all its angles; the complete
run of all batches (every batch on its own stream, slices interleaved) is captured into
one HIP graph and replayed with a single launch.  The stage-product tiles of each batch
come from the library's autotuner (timed per shape at batch creation)."""
import time
from typing import Dict, List, Sequence

import numpy as np

from . import lib, serve, synth


def requests(eng: "lib.Engine", count: int, seed: int = 1234, cfg: int = 4, dz2: float = 0.02,
             method: int = lib.SP4, max_angles: int = 0, log=None) -> List[serve.Request]:
    """`count` seeded structures of the uneven configuration, with their bulk reflections;
    `max_angles` > 0 keeps that many angles of each scan."""
    out = []
    for i in range(count):
        if log and i and i % 16 == 0:
            log("uneven set: %d of %d structures set up" % (i, count))
        r = serve.structure_request(eng, synth.config(cfg, seed=seed, index=i), dz2, method, key="cfg%d-%d" % (cfg, i))
        if max_angles:
            k = slice(0, max_angles)
            r = serve.Request(r.key, r.pot, r.dz, r.ns, r.method, r.iv, r.G[k], r.F[k], r.W[k], r.nb0)
        out.append(r)
    return out


class Resident:
    """The tables and batches of a set of requests on one engine."""

    def __init__(self, eng: "lib.Engine", reqs: Sequence[serve.Request]):
        self.eng = eng
        self.tids, self.bids = [], []
        for r in reqs:
            tid = eng.table_build(r.pot, r.dz, r.ns, r.method)
            self.tids.append(tid)
            self.bids.append(eng.batch_create(tid, r.iv, r.G, r.F, r.nb0, r.W))
        self.gid = None

    def capture(self):
        self.gid = self.eng.graph_create(self.bids)
        return self.gid

    def launch(self):
        self.eng.graph_launch(self.gid)

    def run_eager(self):
        self.eng.batch_run_many(self.bids)

    def results(self) -> List[np.ndarray]:
        return [self.eng.batch_fetch(b, want_f=False)[1] for b in self.bids]

    def release(self):
        if self.gid is not None:
            self.eng.graph_release(self.gid)
            self.gid = None
        for b in self.bids:
            self.eng.batch_release(b)
        for t in self.tids:
            self.eng.table_release(t)
        self.bids, self.tids = [], []


def measure(eng: "lib.Engine", reqs: Sequence[serve.Request], steps: int = 2, warmup: int = 1,
            log=None) -> Dict[str, object]:
    """Points/s of replaying the captured run of all requests."""
    res = Resident(eng, reqs)
    try:
        t0 = time.perf_counter()
        res.capture()
        cap = time.perf_counter() - t0
        if log:
            log("uneven: %d batches captured in %.1f s" % (len(reqs), cap))
        for k in range(warmup):
            res.launch()
            if log:
                log("uneven: warm-up launch %d done" % k)
        t0 = time.perf_counter()
        for k in range(steps):
            res.launch()
            if log:
                log("uneven: timed launch %d done" % k)
        dt = (time.perf_counter() - t0) / steps
        f2 = res.results()
    finally:
        res.release()
    pts = sum(r.nang for r in reqs)
    return {"points": pts, "ms_per_launch": 1e3 * dt, "points_per_s": pts / dt, "capture_s": cap, "f2": f2}
