"""The mixed-size flow after mGP (main/ePlace_main.cpp:174-219): mLG on the
placer's own rand() stream, FILLERONLY and cGP, against the reference's run
of the same design (tests/golden/flow_mixed.npz)."""
import numpy as np
import pytest

from golden_utils import Golden
from eplace_amd import EPlacer2D, EplaceNesterovOpt
from eplace_amd.flow import place
from eplace_amd.legalize import SAMacroLegalizer, check_legal
from placement_checks import cells_legal, macro_overlap_extent

CASE = "flow_mixed"


def say(msg):
    print(msg, flush=True)


@pytest.mark.gpu
def test_stages_after_mgp_match_reference():
    g = Golden(CASE)
    d = g.design()
    n = d.num_nodes
    say(f"[mixed] {CASE}: {n} nodes, mGP then mLG / FILLERONLY / cGP")
    gp = EPlacer2D(d, 1.0)
    try:
        # fillers come from the same rand() stream as the reference's
        np.testing.assert_array_equal(gp.read("positions")[n:], g["fillers_init"])
        gp.initialization()
        r = EplaceNesterovOpt(gp).opt()
        k = int(g.s["iterations"])
        assert abs(r["iterations"] - k) <= max(5, 0.05 * k)
        assert abs(r["final_hpwl"] / g.s["final_hpwl"] - 1) < 5e-3
        # from the reference's mGP result the annealing is the reference's,
        # bit for bit: same positions, same rand() state (fillers drawn, mGP
        # draws nothing)
        gp.write("positions", g["final_positions"])
        gp.mGPIterationCount = int(g.s["mixed_mgp_iteration_count"])
        m = SAMacroLegalizer(gp)
        m.setTargetDensity(1.0)
        st = m.legalization()
        assert st["final_overlap"] == 0 and st["macros"] == int(d.node_is_macro().sum())
        np.testing.assert_array_equal(gp.read("positions")[:n], g["mixed_mlg_positions"])
        assert gp.calcHPWL() == pytest.approx(g.s["mixed_mlg_hpwl"], rel=1e-12)
        # FILLERONLY re-draws the fillers from the same stream
        gp.switch2FillerOnly()
        np.testing.assert_array_equal(gp.read("positions")[n:], g["mixed_fo_fillers"])
        fo = EplaceNesterovOpt(gp).opt()
        assert fo["iterations"] == int(g.s["mixed_fo_iterations"]) == 20
        pos = gp.read("positions")
        np.testing.assert_array_equal(pos[:n], g["mixed_mlg_positions"])  # only fillers move
        ref = g["mixed_fo_positions"]
        say(f"  FILLERONLY: filler offset from the reference {np.abs(pos[n:] - ref[n:]).mean():.4g}, filler travel "
            f"{np.abs(ref[n:] - g['mixed_fo_fillers']).mean():.4g}, overflow {fo['final_overflow']:.4f} vs "
            f"{g.s['mixed_fo_overflow']:.4f}")
        # 20 steps from the same fillers with this run's penalty factor
        # (mGP's lambda differs from the reference's by float rounding)
        assert np.abs(pos[n:] - ref[n:]).mean() < 1e-3 * np.abs(ref[n:] - g["mixed_fo_fillers"]).mean()
        assert abs(fo["final_overflow"] - g.s["mixed_fo_overflow"]) < 1e-3
        gp.switch2cGP()
        cg = EplaceNesterovOpt(gp).opt()
        kc = int(g.s["mixed_cgp_iterations"])
        say(f"  cGP gpu {cg['iterations']} iterations, HPWL {cg['final_hpwl']:.6g}; reference {kc}, "
            f"{g.s['mixed_cgp_hpwl']:.6g}")
        assert abs(cg["iterations"] - kc) <= 2
        assert cg["final_overflow"] < 0.07
        assert abs(gp.calcHPWL() / g.s["mixed_cgp_hpwl"] - 1) < 1e-3
        macros = d.node_is_macro().astype(bool)
        np.testing.assert_array_equal(gp.read("positions")[:n][macros], g["mixed_mlg_positions"][macros])
    finally:
        gp.close()


@pytest.mark.gpu
@pytest.mark.parametrize("qp", [False, True])
def test_full_mixed_flow_is_legal(qp):
    """mGP, mLG, FILLERONLY, cGP and Abacus on the same design (with QP
    first, or from the .pl as the reference's run): the cells are legal; from
    the same start the HPWL is within 1% of the reference's flow, and QP
    does not make it worse than that.  mLG stops when the union of the
    integer-truncated macro rectangles covers their total area
    (legalizer.cpp:1243-1260), so macro pairs may keep an overlap below one
    unit, as in the reference; from the reference's start they keep none."""
    g = Golden(CASE)
    d = g.design()
    say(f"[mixed] full flow on {CASE}, qp={qp}")
    out = place(d, 1.0, qp=qp, legal=True)
    assert out["mlg"]["final_overlap"] == 0
    assert out["fillers"]["iterations"] == 20 and out["cgp"]["final_overflow"] < 0.07
    chk = out["legal"]["check"]
    assert cells_legal(chk), chk
    ext = macro_overlap_extent(d, out["positions"])
    say(f"  macro overlap extent {ext:.3f}")
    assert ext < 1.0
    if not qp:
        assert chk["legal"], chk
    gap = out["cgp"]["final_hpwl"] / g.s["mixed_cgp_hpwl"] - 1
    say(f"  cGP HPWL {out['cgp']['final_hpwl']:.6g} vs reference {g.s['mixed_cgp_hpwl']:.6g} ({100 * gap:+.2f}%)")
    if qp:
        assert gap < 1e-2
    else:
        assert abs(gap) < 1e-2
    back = check_legal(d)  # the design holds the legal positions
    assert cells_legal(back) and (qp or back["legal"]), back
