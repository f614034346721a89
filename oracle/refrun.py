"""Runs the reference's own CPU programs built by oracle/ref.mk (oracle/_ref/).

Test infrastructure only (tests, the fixture scripts and bench.py's CPU baseline): nothing
in the product imports it.  Each run works in a fresh temporary directory holding the deck
files the program reads (bulk.txt; surf.txt and bulkP.b), feeds the program its stdin the
way the reference's scripts do (test/test23/perf-test.sh:14-27: "dz method" for
sim-trhepd-rheed, "dz" for str-lapack), and returns what it wrote.
"""
from __future__ import annotations

import os
import resource
import shutil
import subprocess
import tempfile
import time
from typing import Dict, Optional

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")
SURF = os.path.join(REF_DIR, "surf.exe")
BULK = os.path.join(REF_DIR, "bulk.exe")
LAPACK = os.path.join(REF_DIR, "surf-lapack.exe")


def available(exe: str = SURF) -> bool:
    return os.path.exists(exe)


def _unlimited_stack():
    # the programs keep N x 2N complex work arrays as automatic (stack) arrays
    resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))


def _run(exe: str, files: Dict[str, bytes], stdin: str, threads: int, timeout: float, outputs, extra_env=None):
    if not available(exe):
        raise FileNotFoundError("%s is not built (make -f oracle/ref.mk)" % exe)
    env = dict(os.environ)
    env["OMP_NUM_THREADS"] = str(threads)
    env.setdefault("OMP_STACKSIZE", "512M")
    env.update(extra_env or {})
    d = tempfile.mkdtemp(prefix="trhepd-ref-")
    try:
        for name, data in files.items():
            with open(os.path.join(d, name), "wb") as fh:
                fh.write(data)
        t0 = time.perf_counter()
        p = subprocess.run([exe], input=stdin, cwd=d, env=env, capture_output=True, text=True,
                           timeout=timeout, preexec_fn=_unlimited_stack)
        wall = time.perf_counter() - t0
        if p.returncode != 0:
            raise RuntimeError("%s failed (%d): %s" % (os.path.basename(exe), p.returncode, p.stderr[-2000:]))
        out = {}
        for name in outputs:
            with open(os.path.join(d, name), "rb") as fh:
                out[name] = fh.read()
        return out, p.stdout, wall
    finally:
        shutil.rmtree(d, ignore_errors=True)


def _omp_seconds(stdout: str) -> Optional[float]:
    """The program's own wall time (surfm.f90:95-96 prints the omp_get_wtime delta last)."""
    for tok in reversed(stdout.split()):
        try:
            return float(tok)
        except ValueError:
            continue
    return None


def surf(surf_txt: bytes, bulkp: bytes, dz: float, method: int, threads: int = 8, timeout: float = 3600):
    """sim-trhepd-rheed surf.exe: returns (the .s text, its printed wall time in s, process wall s)."""
    out, so, wall = _run(SURF, {"surf.txt": surf_txt, "bulkP.b": bulkp}, "%r %d\n" % (dz, method), threads,
                         timeout, ["surf-bulkP.s"])
    return out["surf-bulkP.s"].decode(), _omp_seconds(so), wall


def surf_domains(surf_txt: bytes, bulkp: bytes, dz: float, method: int, threads: int = 8, timeout: float = 3600):
    """surf.exe on a deck of several domains (surfm.f90:64-90): returns (the .md text with every
    domain's curve, the .s text of its domain sum)."""
    out, _, _ = _run(SURF, {"surf.txt": surf_txt, "bulkP.b": bulkp}, "%r %d\n" % (dz, method), threads, timeout,
                     ["surf-bulkP.md", "surf-bulkP.s"])
    return out["surf-bulkP.md"].decode(), out["surf-bulkP.s"].decode()


def lapack(surf_txt: bytes, bulkp: bytes, dz: float, threads: int = 8, timeout: float = 3600):
    """str-lapack surf.exe (transfer-matrix method): returns (.s text, printed time, process wall)."""
    out, so, wall = _run(LAPACK, {"surf.txt": surf_txt, "bulkP.b": bulkp}, "%r\n" % dz, threads, timeout,
                         ["surf-bulkP.s"])
    return out["surf-bulkP.s"].decode(), _omp_seconds(so), wall


def bulk(bulk_txt: bytes, threads: int = 8, timeout: float = 3600, extra_env=None):
    """sim-trhepd-rheed bulk.exe: returns (bulkP.b bytes, process wall s)."""
    out, _, wall = _run(BULK, {"bulk.txt": bulk_txt}, "", threads, timeout, ["bulkP.b"], extra_env)
    return out["bulkP.b"], wall
