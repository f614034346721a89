"""The whole programs on the device, through the host mirror of the reference's drivers:
bulk.txt -> bulkP.b (bulk stage) -> rocking curve (surface slice loop) -> domain sum.

A deck with two identical domains (test23's, duplicated, weights 0.3 / 0.7) must give, after
the domain sum, the single-domain curve: the .md holds both domains' curves bit-identical to
the single-domain run, and the summed .s equals it to the E12.4 of domainsum's output.

A deck with two different domains (test23 plus its 60-degree twin domain, weights
0.6 / 0.4) against the reference's own run of it (its surf.exe, built from its sources, on the
bulkP.b its bulk.exe wrote, each domain's blocks from a one-domain run): every domain's curve in the .md within the FP64 bar of 1e-10 per
intensity, the header lines identical, and the summed .s within one unit of its last printed
digit (E12.4 rounding of numbers that agree to 1e-10 can flip that digit).
"""
import copy
import os

import numpy as np
import pytest

import oracle as O
from conftest import REF
from test_post import read_md_rows
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


def test_two_domain_deck_vs_the_reference_run(engine):
    d = os.path.join(REF, "test23dom")
    bp = deckio.BulkP.read(os.path.join(d, "bulkP.b"))
    sd = deckio.SurfDeck.parse(open(os.path.join(d, "surf.txt")).read(), bp.ndom)
    print("two domains: N", bp.nb, "rdom", bp.rdom, "wdom", sd.wdom, flush=True)
    out = driver.surface_program(engine, bp, sd, 0.1, lib.SP4)
    ref_md = open(os.path.join(d, "surf-bulkP.md")).read()
    head = [ln for ln in ref_md.splitlines() if "," not in ln or ln.startswith("#")]
    assert [ln for ln in out["md"].splitlines() if "," not in ln or ln.startswith("#")] == head
    (_, _, _, _, nb), want = read_md_rows(ref_md)
    _, got = read_md_rows(out["md"])
    worst = 0.0
    for dom in range(2):
        for blk_w, blk_g in zip(want[dom], got[dom]):
            for (aw, fw), (ag, fg) in zip(blk_w, blk_g):
                assert aw == ag
                e = O.rel_err(np.array(fg), np.array(fw))
                worst = max(worst, e)
                assert e < 1e-10, "domain %d angle %g: %.3e" % (dom, aw, e)
    s_ref = deckio.read_s_text(open(os.path.join(d, "surf-bulkP.s")).read())
    s_dev = deckio.read_s_text(out["s"])
    assert list(s_dev.ih) == list(s_ref.ih) and np.array_equal(s_dev.angles, s_ref.angles)
    # one unit in the 4th digit of an E12.4 mantissa in [0.1, 1)
    ulp4 = np.where(s_ref.f2 != 0, 10.0 ** (np.floor(np.log10(np.where(s_ref.f2 != 0, s_ref.f2, 1.0))) - 3), 0.0)
    assert np.all(np.abs(s_dev.f2 - s_ref.f2) <= ulp4 * 1.0001)
    same = sum(a == b for a, b in zip(out["s"].splitlines(), open(os.path.join(d, "surf-bulkP.s")).read().splitlines()))
    print("md worst %.3e; .s lines identical %d of %d" % (worst, same, len(out["s"].splitlines())))
