"""One-rounding-unit sensitivity of the reference's own program on the test521 deck.

For each (int_method, dz) of the N = 521 parity tests this runs the reference's
sim-trhepd-rheed surf.exe (built from its sources by oracle/ref.mk) three times on the deck
in tests/golden/ref/test521: as it is, with the bulk reflection blocks of bulkP.b multiplied
by (1 + 2^-52), and with every surface atom's height multiplied by (1 + 2^-52).  The
largest change of the curve between these runs is how far one rounding of an input moves the
reference's FP64 result: no FP64 implementation can be held closer to the reference's
recorded curve than that, so the parity bars of the N = 521 tests are derived from it
(max(1e-10, 10 x response)).  Also recorded: how far this build of the reference lies from
the authors' recorded curve (perf-<method>-<dz>.txt).

    python tests/golden/make_sensitivity.py [threads]

Writes tests/golden/ref/test521/sensitivity.json.  Needs oracle/_ref (make -f oracle/ref.mk).
"""
import copy
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(REPO, "oracle"), os.path.join(REPO, "trhepd-opt_amd")]
import refrun  # noqa: E402
from trhepd import deckio  # noqa: E402

CASE = os.path.join(HERE, "ref", "test521")
RUNS = [(4, "0.68"), (5, "0.68"), (6, "0.68"), (4, "0.47"), (5, "0.47"), (6, "0.47"),
        (4, "0.15"), (5, "0.15"), (6, "0.15"), (4, "0.1"), (5, "0.1"), (6, "0.1"),
        (4, "0.068"), (5, "0.068"), (6, "0.068"), (4, "0.022"), (5, "0.022"), (6, "0.022"),
        (4, "0.01"), (5, "0.01"), (6, "0.01"),
        # classical RK4 at the steps inside its stability region for this deck (the paper's
        # headline rk4 step is 0.015, test/plots-plus/table.tex:17)
        (7, "0.047"), (7, "0.015")]


def rel(a, b):
    """max per-beam relative difference (floor 1e-12 x the angle's largest intensity), L2"""
    a, b = np.asarray(a), np.asarray(b)
    mx = np.max(np.abs(b), axis=1, keepdims=True)
    r = np.abs(a - b) / np.maximum(np.abs(b), 1e-12 * mx)
    l2 = np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)
    return float(r.max()), float(l2.max())


def main():
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else 8
    surf_txt = open(os.path.join(CASE, "surf.txt"), "rb").read()
    bulkp = open(os.path.join(CASE, "bulkP.b"), "rb").read()
    bp = deckio.BulkP.read(bulkp)
    bp_f = copy.deepcopy(bp)
    for row in bp_f.domains[0].blocks:
        for grp in row:
            for k in range(len(grp)):
                grp[k] = np.asarray(grp[k]) * (1 + 2.0 ** -52)
    bulkp_f = bp_f.to_bytes()
    sd = deckio.SurfDeck.parse(surf_txt.decode())
    sd_z = copy.deepcopy(sd)
    for a in sd_z.atoms:
        a.z = a.z * (1 + 2.0 ** -52)
    surf_z = sd_z.format().encode()
    # the reformatted deck itself (no perturbation) must give the deck's own curve
    surf_fmt = sd.format().encode()
    path = os.path.join(CASE, "sensitivity.json")
    out = {"deck": "test/test521 (1 angle, N = 521)", "threads": threads, "runs": {}}
    if os.path.exists(path):   # keeps what an earlier run measured; measures what is missing
        out["runs"] = json.load(open(path))["runs"]
    for method, dz in RUNS:
        if "%d-%s" % (method, dz) in out["runs"]:
            continue
        t0 = time.perf_counter()
        base = deckio.read_s_text(refrun.surf(surf_txt, bulkp, float(dz), method, threads)[0]).f2
        pf = deckio.read_s_text(refrun.surf(surf_txt, bulkp_f, float(dz), method, threads)[0]).f2
        pz = deckio.read_s_text(refrun.surf(surf_z, bulkp, float(dz), method, threads)[0]).f2
        fmt = deckio.read_s_text(refrun.surf(surf_fmt, bulkp, float(dz), method, threads)[0]).f2
        rec = deckio.read_s(os.path.join(CASE, "perf-%d-%s.txt" % (method, dz))).f2
        e_f, l_f = rel(pf, base)
        e_z, l_z = rel(pz, base)
        e_r, l_r = rel(base, rec)
        e_m, l_m = rel(fmt, base)
        out["runs"]["%d-%s" % (method, dz)] = {
            "ulp_F": [e_f, l_f], "ulp_z": [e_z, l_z], "response": [max(e_f, e_z), max(l_f, l_z)],
            "build_vs_recorded": [e_r, l_r], "reformatted_deck": [e_m, l_m]}
        print("method %d dz %s: one-ulp response %.2e (L2 %.2e), this build vs recorded %.2e (L2 %.2e), %.0f s"
              % (method, dz, max(e_f, e_z), max(l_f, l_z), e_r, l_r, time.perf_counter() - t0), flush=True)
        with open(path, "w") as fh:   # after every case, so that a long sweep can be resumed
            json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
