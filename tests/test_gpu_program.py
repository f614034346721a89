"""The whole programs on the device, through the host mirror of the reference's drivers:
bulk.txt -> bulkP.b (bulk stage) -> rocking curve (surface slice loop) -> domain sum.

A deck with two identical domains (test23's, duplicated, weights 0.3 / 0.7) must give, after
the domain sum, the single-domain curve: the .md holds both domains' curves bit-identical to
the single-domain run, and the summed .s equals it to the E12.4 of domainsum's output.
"""
import copy
import os

import numpy as np
import pytest

from conftest import REF
from trhepd import deckio, driver, lib

pytestmark = pytest.mark.gpu


def test_two_identical_domains_sum_to_the_single_domain(engine):
    text = open(os.path.join(REF, "test23", "bulk.txt")).read()
    bd1 = deckio.BulkDeck.parse(text)
    bd2 = copy.deepcopy(bd1)
    bd2.ndom, bd2.nb, bd2.rdom = 2, bd1.nb * 2, [0.0, 0.0]
    bd2.ih, bd2.ik = bd1.ih * 2, bd1.ik * 2
    bd2 = deckio.BulkDeck.parse(bd2.format())
    stext = open(os.path.join(REF, "test23", "surf.txt")).read()
    sd1 = deckio.SurfDeck.parse(stext, 1)
    sd2 = deckio.SurfDeck.parse(stext, 2)
    bp1 = driver.bulk_program(engine, bd1)
    bp2 = driver.bulk_program(engine, bd2)
    one = driver.surface_program(engine, bp1, sd1, 0.1, lib.SP4)
    two = driver.surface_program(engine, bp2, sd2, 0.1, lib.SP4, wdom=[0.3, 0.7])
    assert one["md"] is None and two["md"] is not None
    f1 = deckio.read_s_text(one["s"]).f2
    summed = deckio.read_s_text(two["s"])
    assert summed.nb == 23 and list(summed.ih) == list(bd1.ih[:23])
    # E12.4 keeps 4 digits of a mantissa in [0.1, 1): rounding error <= 5e-4 relative
    rel = np.abs(summed.f2 - f1) / np.where(f1 != 0, np.abs(f1), 1.0)
    print("domain sum vs single domain: max rel %.3e" % rel.max())
    assert rel.max() <= 5.01e-4
    # the .md rows of each domain are the single-domain curve
    md = two["md"].splitlines()
    rows = [ln for ln in md if ln and not ln.startswith("#") and "," in ln]
    assert len(rows) == 2 * len(f1)
    assert rows[:len(f1)] == rows[len(f1):]
    assert rows[:len(f1)] == [ln for ln in one["s"].splitlines() if ln and not ln.startswith("#") and "," in ln]
