"""How much the reference's own bulk program moves each of test521's 49 reflection blocks when
one input moves by one ulp: the built reference (oracle/ref.mk) runs test521/bulk.txt as it
is and with the first atom's z raised to the next double.  Per group: max |R' - R| / max |R|.

The groups span 1e-10 (well conditioned) to O(1): the blocks of strongly evanescent groups are
not determined in FP64 by the reference's algorithm (the authors' recorded bulkP.b and a build
of the same sources differ by O(1) there too).  The device's blocks are compared with the
recorded file at a per-group bar set from these responses.

    python tests/golden/make_bulk521_sensitivity.py
"""
import copy
import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(REPO, "oracle"), os.path.join(REPO, "trhepd-opt_amd")]
import refrun  # noqa: E402
from trhepd import deckio  # noqa: E402

D = os.path.join(HERE, "ref", "test521")


def main():
    bd = deckio.BulkDeck.parse(open(os.path.join(D, "bulk.txt")).read())
    base = deckio.BulkP.read(refrun.bulk(bd.format().encode(), threads=8)[0])
    moved = copy.deepcopy(bd)
    a = moved.atoms[0]
    moved.atoms[0] = deckio.Atom(a.ielm, a.ocr, a.x, a.y, math.nextafter(a.z, math.inf))
    pert = deckio.BulkP.read(refrun.bulk(moved.format().encode(), threads=8)[0])
    resp = [float(np.abs(x - y).max() / np.abs(x).max())
            for x, y in zip(base.domains[0].blocks[0][0], pert.domains[0].blocks[0][0])]
    out = {"what": "per-group one-ulp response of the built reference's bulk blocks (atom 1 z + 1 ulp)",
           "response": resp}
    with open(os.path.join(D, "bulk_sensitivity.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    print(" ".join("%.0e" % r for r in resp))


if __name__ == "__main__":
    main()
