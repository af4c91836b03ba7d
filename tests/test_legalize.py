"""Abacus legalisation against the reference's AbacusLegalizer
(Legalization/legalizer.cpp:58-680) on the reference's own outputs, and the
legality checker.  Host code only: runs without a GPU."""
import numpy as np
import pytest

from eplace_amd.design import synthetic
from eplace_amd.legalize import check_legal, legalize_abacus
from golden_utils import Golden


def design_with_start(name):
    g = Golden(name)
    d = g.design()
    key = "input_positions" if "input_positions" in g.a else "final_positions"
    return g, d, g[key][: d.num_nodes]


@pytest.mark.parametrize("name", ["smoke_default", "legal_synth_c2", "legal_rand20k"])
def test_abacus_matches_reference(name):
    g, d, start = design_with_start(name)
    out, st = legalize_abacus(d, start, mode=0)
    ref = g["legal_positions"][: d.num_nodes]
    diff = np.abs(out - ref).max()
    assert diff == 0.0, f"max |diff| {diff} over {st['cells']} cells"


@pytest.mark.parametrize("name", ["smoke_default", "legal_synth_c2", "legal_rand20k"])
def test_reference_result_is_legal(name):
    g, d, start = design_with_start(name)
    chk = check_legal(d, g["legal_positions"][: d.num_nodes])
    assert chk["legal"], chk


@pytest.mark.parametrize("name", ["legal_synth_c2", "legal_rand20k"])
def test_corrected_mode_is_legal(name):
    """Mode 1 (row 0 searched, site rounding from the core origin, float x)
    is legal; its displacement stays within 10% of the reference's
    (measured: +5.4% on legal_synth_c2, rounding pushes more clusters into
    their neighbours)."""
    g, d, start = design_with_start(name)
    out0, st0 = legalize_abacus(d, start, mode=0)
    out1, st1 = legalize_abacus(d, start, mode=1)
    chk = check_legal(d, out1)
    assert chk["legal"], chk
    assert st1["total_displacement"] <= st0["total_displacement"] * 1.10


def test_checker_finds_planted_overlaps():
    g, d, start = design_with_start("legal_synth_c2")
    legal = g["legal_positions"][: d.num_nodes].copy()
    assert check_legal(d, legal)["legal"]
    bad = legal.copy()
    # move cell 1 onto cell 0 (same row), shift cell 2 off the site grid, lift cell 3 off its row
    bad[1] = bad[0] + np.array([0.5, 0.0], np.float32)
    bad[2, 0] += 0.3
    bad[3, 1] += 0.4
    chk = check_legal(d, bad)
    assert not chk["legal"]
    assert chk["overlaps"] >= 1 and chk["overlap_area"] > 0
    assert chk["off_site"] >= 1
    assert chk["off_row"] >= 1


def test_checker_sees_macro_blockage():
    d = synthetic(600, 650, seed=5, num_macros=2, util=0.6, name="mb")
    mac = np.where(d.node_is_macro())[0]
    assert len(mac) == 2
    out, _ = legalize_abacus(d, mode=1)
    assert check_legal(d, out)["legal"]
    # a std cell dropped inside a macro is reported
    xy = out.copy()
    cell = np.where(~d.node_is_macro().astype(bool))[0][0]
    xy[cell] = xy[mac[0]]
    chk = check_legal(d, xy)
    assert chk["in_blockage"] >= 1
