"""Rocking-curve products after the slice loop (trhepd/post.py) against the reference's own
outputs and what the geometry fixes.

- Angular convolution: the reference's recorded curves and their convolutions
  (sim-trhepd-rheed/sample/Ge_001_c4x2/output, script/make_convolution.py; str-orig's
  T4 Al / Si(111) sample, tool/make_convolution.py) are reproduced byte for byte.
- Domain sum (domainsum.f90): the reference ships no multi-domain run, so its own programs,
  built from its sources, ran one (tests/golden/make_domain_fixture.py: test23 plus a second
  domain turned by 60 degrees, the twin of the p3m1 surface, weights 0.6 / 0.4).  Its .md is
  re-summed byte for byte, and re-written byte for byte from its numbers.  Constructions whose answer is known add to that:
  identical domains add with their weights, and a domain rotated by 120 degrees on the
  hexagonal lattice (its beams obtained from the real lattice's geometry, not from the
  reference's beam-matching formula) lands beam for beam on the reference domain.
"""
import math
import os

import numpy as np

from conftest import REF
from trhepd import deckio, post


def test_convolution_beam_reproduces_ge001():
    s = open(os.path.join(REF, "ge001", "surf-bulkP.s")).read()
    _, _, text = post.convolution_beam(s, first_line=5, last_line=74, row_number=2, omega=0.5)
    assert text == open(os.path.join(REF, "ge001", "convolution.txt")).read()


def test_convolution_all_beams_reproduces_t4al():
    s = open(os.path.join(REF, "t4al", "surf-bulkP.s")).read()
    cnv, text = post.convolution(s, omega=0.5)
    assert text == open(os.path.join(REF, "t4al", "convolution.txt")).read()
    assert cnv.shape == (66, 14)


def test_convolution_of_a_constant_is_the_gaussian_mass():
    """Far from the ends a constant curve convolves to itself times the Gaussian's mass
    (1 to within the sampling: the kernel integrates to one)."""
    ang = np.arange(1, 201) * 0.1
    text = "#azimuths,g-angles,beams\n   1 200   1\n#ih,ik\n   0   0\n" + "".join(
        deckio.format_s_row(a, [2.0]) + "\n" for a in ang)
    cnv, _ = post.convolution(text, omega=0.5)
    assert abs(cnv[100, 1] / 2.0 - 1.0) < 1e-6
    _, out, _ = post.convolution_beam(text, first_line=5, last_line=204, row_number=2)
    assert abs(out[100] / 2.0 - 1.0) < 1e-5


def _ref_curve():
    ref = deckio.read_s(os.path.join(REF, "test23", "perf-5-0.1.txt"))
    return ref


def _rot120(h, k):
    """(h, k) of the hexagonal reciprocal lattice (a = b, gamma = 120 deg) rotated by +120 deg,
    from the Cartesian reciprocal basis of the real lattice."""
    a1 = np.array([1.0, 0.0])
    a2 = np.array([math.cos(2 * math.pi / 3), math.sin(2 * math.pi / 3)])
    A = np.array([a1, a2])
    B = 2 * math.pi * np.linalg.inv(A).T        # rows: b1, b2
    c, s = math.cos(2 * math.pi / 3), math.sin(2 * math.pi / 3)
    g = np.array([[c, -s], [s, c]]) @ (h * B[0] + k * B[1])
    hk = np.linalg.solve(B.T, g)
    assert np.allclose(hk, np.round(hk), atol=1e-9)
    return int(round(hk[0])), int(round(hk[1]))


def _md(ih2, ik2, rdom2, w, ref):
    nb = len(ref.ih)
    f = [list(r) for r in ref.f2]
    rows = [[(ref.angles[p], f[p]) for p in range(len(f))]]
    return deckio.format_md(ref.naz, ref.ng, 1, 1, 120.0, [nb, nb], [0.0, rdom2], w,
                            list(ref.ih) + list(ih2), list(ref.ik) + list(ik2), [rows, rows])


def test_domainsum_identical_domains_add_with_weights():
    ref = _ref_curve()
    out = post.domainsum(_md(ref.ih, ref.ik, 0.0, [0.25, 0.5], ref))
    got = deckio.read_s_text(out)
    assert list(got.ih) == list(ref.ih) and list(got.ik) == list(ref.ik)
    want = 0.75 * np.array(ref.f2)
    assert np.allclose(got.f2, want, rtol=1e-3, atol=1e-300)
    # the domain sum writes E12.4 values
    assert all(float(deckio.fortran_e(v, 12, 4)) == g for v, g in zip(want[3], got.f2[3]))


def test_domainsum_rotated_domain_maps_onto_reference_beams():
    ref = _ref_curve()
    rot = [_rot120(h, k) for h, k in zip(ref.ih, ref.ik)]
    ih2 = [r[0] for r in rot]
    ik2 = [r[1] for r in rot]
    # domain 2 is domain 1 turned by +120 deg: its beam R g carries domain 1's intensity of g
    out = post.domainsum(_md(ih2, ik2, 120.0, [1.0, 1.0], ref))
    got = deckio.read_s_text(out)
    assert np.allclose(got.f2, 2.0 * np.array(ref.f2), rtol=1e-3, atol=1e-300)
    # turned the other way, the beams do not coincide with the same partners
    bad = post.domainsum(_md(ih2, ik2, -120.0, [1.0, 1.0], ref))
    assert not np.allclose(deckio.read_s_text(bad).f2, 2.0 * np.array(ref.f2), rtol=1e-3)


def test_domainsum_refuses_single_domain():
    import pytest
    with pytest.raises(ValueError):
        post.domainsum(open(os.path.join(REF, "ge001", "surf-bulkP.s")).read())


DOM = os.path.join(REF, "test23dom")


def read_md_rows(md_text):
    """The numbers of a .md text: header values, and per domain the list of (angle, f2) rows
    of each azimuth / glancing-angle block."""
    lines = md_text.splitlines()
    naz, ng, ndom, nh, nk = (int(v) for v in lines[1].split()[:5])
    nb = [int(lines[3 + d].split()[0]) for d in range(ndom)]
    k = 4 + ndom
    assert lines[3 + ndom] == "#ih,ik"
    for d in range(ndom):
        vals = []
        while len(vals) < 2 * nb[d]:
            vals += lines[k].split()
            k += 1
    nrep1, nrep2 = (naz, ng) if ng > naz else (ng, naz)
    out = []
    for d in range(ndom):
        blocks = []
        for _ in range(nrep1):
            rows = []
            for _ in range(nrep2):
                v = [deckio._parse_fnum(t) for t in lines[k].split(",") if t.strip()]
                rows.append((v[0], v[1:1 + nb[d]]))
                k += 1
            assert lines[k] == ""
            k += 1
            blocks.append(rows)
        out.append(blocks)
    return (naz, ng, nh, nk, nb), out


def test_domainsum_reproduces_the_reference_two_domain_run():
    md = open(os.path.join(DOM, "surf-bulkP.md")).read()
    assert post.domainsum(md) == open(os.path.join(DOM, "surf-bulkP.s")).read()


def test_format_md_reproduces_the_reference_md():
    md = open(os.path.join(DOM, "surf-bulkP.md")).read()
    (naz, ng, nh, nk, nb), rows = read_md_rows(md)
    bp = deckio.BulkP.read(os.path.join(DOM, "bulkP.b"))
    sd = deckio.SurfDeck.parse(open(os.path.join(DOM, "surf.txt")).read(), bp.ndom)
    assert (bp.ndom, bp.nb, sd.wdom) == (2, [23, 23], [0.6, 0.4])
    deg = 180.0 / math.pi
    out = deckio.format_md(naz, ng, bp.nh, bp.nk, bp.gam * deg, bp.nb, [r * deg for r in bp.rdom], sd.wdom,
                           bp.ih, bp.ik, rows)
    assert out == md
    # the two domains' curves differ: the second is not a copy of the first
    f1 = np.array([r[1] for r in rows[0][0]])
    f2 = np.array([r[1] for r in rows[1][0]])
    assert np.max(np.abs(f1 - f2)) > 1e-3 * np.max(f1)


def test_domainsum_weights_override_on_the_reference_run():
    """Weights 1 / 0 give domain 1's curve, at the E12.4 of the domain sum's output."""
    md = open(os.path.join(DOM, "surf-bulkP.md")).read()
    _, rows = read_md_rows(md)
    got = deckio.read_s_text(post.domainsum(md, wdom=[1.0, 0.0]))
    want = np.array([r[1] for r in rows[0][0]])
    assert np.array_equal(got.f2, np.vectorize(lambda v: float(deckio.fortran_e(v, 12, 4)))(want))
