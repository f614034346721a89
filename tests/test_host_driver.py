"""Host-side driver and formats (no device): bulkP.b header reconstruction, .s output
formatting, synthetic decks."""
import os

import numpy as np
import pytest

from conftest import REF
from trhepd import deckio, driver, synth


@pytest.mark.parametrize("case", ["test23", "test47", "test521"])
def test_bulk_header_is_byte_exact(case):
    """bulkio's header/domain records rebuilt from bulk.txt equal the reference's bulkP.b
    (blocks copied over): pins asfparam, scanrg, zoffset, blkibg and blkghk."""
    bd = deckio.BulkDeck.parse(open(os.path.join(REF, case, "bulk.txt")).read())
    h = driver.bulk_header(bd)
    raw = open(os.path.join(REF, case, "bulkP.b"), "rb").read()
    h.domains[0].blocks = deckio.BulkP.read(raw).domains[0].blocks
    assert h.to_bytes() == raw


def test_bulkp_round_trip():
    raw = open(os.path.join(REF, "test47", "bulkP.b"), "rb").read()
    assert deckio.BulkP.read(raw).to_bytes() == raw


@pytest.mark.parametrize("case,name", [("test23", "perf-5-0.1.txt"), ("test521", "perf-6-0.68.txt")])
def test_s_writer_reproduces_reference_text(tmp_path, case, name):
    path = os.path.join(REF, case, name)
    rc = deckio.read_s(path)
    out = tmp_path / "x.s"
    deckio.write_s(str(out), rc.naz, rc.ng, rc.ih, rc.ik, [[(a, f) for a, f in zip(rc.angles, rc.f2)]])
    ref_lines = open(path).read().splitlines()
    got = open(out).read().splitlines()
    assert got[:len(ref_lines) - 1] == ref_lines[:len(ref_lines) - 1]


def test_fortran_e_edge_cases():
    assert deckio.fortran_e(0.0, 23, 15) == "  0.000000000000000E+00"
    assert deckio.fortran_e(1.234e-150, 23, 15).strip() == "0.123400000000000-149"
    assert deckio.fortran_e(-30.0, 12, 4) == " -0.3000E+02"


def test_synthetic_configs_are_seeded():
    a = synth.config(3, seed=5)
    b = synth.config(3, seed=5)
    c = synth.config(3, seed=6)
    assert a.surf.format() == b.surf.format() and a.surf.format() != c.surf.format()
    assert a.bp.nb == [521] and a.bp.ng == 40 and len(a.surf.atoms) == 12
    assert a.bp.ih[0] == 0 and a.bp.ik[0] == 0
    for cfg, n, ang in ((1, 23, 56), (2, 47, 56)):
        s = synth.config(cfg)
        assert s.bp.nb == [n] and s.bp.ng == ang
    # uneven batch draws
    ns = {synth.config(4, index=i).bp.nb[0] for i in range(40)}
    assert ns <= {23, 47, 97, 199, 307} and len(ns) >= 4
