"""The workload configurations at their full sizes on the GPU, checked
through properties that do not need a reference run of the same size:

* 200K std cells, 512x512 bins, FP32 and split-FP16 transforms: one
  gradient evaluation against the FP64 oracle on the same positions
  (fixed-point density maps bit-exact, fields and gradients within 1e-5
  relative);
* 500K std cells + 60 movable macros, 1024x1024 bins, split-FP16
  transforms: mGP -> macro legalisation -> FILLERONLY -> cGP -> Abacus
  reaches the stage targets and legal standard cells;
* 2M std cells, 2048x2048 bins, split-FP16 transforms: the learned
  controller and plain backtracking both reach overflow 0.10, the learned
  run in fewer gradient evaluations and within 1% of the plain HPWL.
  (With bins 1.34 rows high the overflow of plain backtracking levels off
  between 0.08 and 0.10 by 1000-3000 iterations, so 0.07 is not asked.)

mGP on the macro-heavy 500K design (macros cover about 40% of the die)
stops at its iteration cap, as the reference's loop does when the target
is not reached; the test asks for the overflow to fall below 0.5.  (Two
hundred macros of 3-12% of the die width would cover more than the whole
die, so the 2M case is run with standard cells only.)  The annealing macro
legaliser measures overlap on integer-truncated rectangles
(legalizer.cpp), so macros with fractional sizes may still overlap by less
than one unit; the test asks for that and no more.
"""
import sys
import time

import numpy as np
import pytest

from eplace_amd import EPlacer2D, synthetic
from eplace_amd.flow import place
from oracle.oracle import Oracle
from placement_checks import macro_overlap_extent

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    n = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (n if n > 0 else 1.0))


@pytest.mark.parametrize("mode", [0, 2], ids=["fp32", "split-fp16"])
def test_200k_cells_512_bins_against_fp64_oracle(mode):
    say(f"[config] 200K cells, 512x512 bins, poisson_mode {mode} vs the FP64 oracle")
    d = synthetic(200000, 210000, seed=21, num_io=1000, util=0.7, name="cells200k")
    gp = EPlacer2D(d, 1.0, bins=512, poisson_mode=mode)
    try:
        orc = Oracle(d, 1.0, mode="fp64", bins=512, exp="fast", wl="WA")
        orc.set("positions", orc.get("positions"))
        np.testing.assert_array_equal(gp.read("positions").reshape(-1), orc.get("positions").astype(np.float32))
        gp.totalGradientUpdate()
        shift = gp.info()["fixed_shift"]
        orc.run("scatter_fixed", shift)
        np.testing.assert_array_equal(gp.read("node_q").reshape(-1), orc.get("node_q"))
        np.testing.assert_array_equal(gp.read("filler_q").reshape(-1), orc.get("filler_q"))
        orc.run("total_gradient_update", 1, shift)
        for nm in ("ex", "ey"):
            e = rel(gp.read(nm), orc.map(nm))
            say(f"  {nm} rel err {e:.2e}")
            assert e < 1e-5
        for nm in ("grad_density", "grad_wirelength"):
            e = rel(gp.read(nm), orc.get(nm).reshape(-1, 2))
            say(f"  {nm} rel err {e:.2e}")
            assert e < 1e-5
    finally:
        gp.close()


def run_mixed(cells, nets, macros, bins, util, seed, legal, **kw):
    t = time.perf_counter()
    d = synthetic(cells, nets, seed=seed, num_io=1000, util=util, num_macros=macros, name=f"mixed{cells}")
    say(f"  design: {d.num_nodes} nodes, {d.num_pins} pins ({time.perf_counter() - t:.1f}s)")
    t = time.perf_counter()
    out = place(d, 1.0, qp=False, legal=legal, bins=bins, poisson_mode=2, verbose=True, **kw)
    ext = macro_overlap_extent(d, out["positions"])
    say(f"  flow {time.perf_counter() - t:.1f}s: mGP {out['mgp']['iterations']} it (overflow "
        f"{out['mgp']['final_overflow']:.4f}), mLG overlap {out['mlg']['final_overlap']} "
        f"({out['mlg']['seconds']:.1f}s), cGP {out['cgp']['iterations']} it (overflow "
        f"{out['cgp']['final_overflow']:.4f}), HPWL {out['hpwl']:.6g}, macro overlap extent {ext:.3f}")
    assert out["mgp"]["final_overflow"] < 0.5
    assert out["mlg"]["macros"] == macros and out["mlg"]["final_overlap"] == 0
    assert ext < 1.0
    assert out["fillers"]["iterations"] == 20
    assert out["cgp"]["final_overflow"] < 0.07
    if legal:
        chk = out["legal"]["check"]
        say(f"  Abacus: {chk}")
        for k in ("overlaps", "off_row", "off_site", "out_of_core", "in_blockage"):
            assert chk[k] == 0, chk
    return out


def test_500k_cells_60_macros_mixed_flow():
    say("[config] 500K cells + 60 macros, 1024x1024 bins, mGP -> mLG -> cGP -> Abacus")
    run_mixed(500000, 520000, 60, 1024, 0.75, 3, True, method="bktrk")


def test_2m_cells_learned_controller_against_plain():
    say("[config] 2M cells, 2048x2048 bins, learned controller vs plain backtracking to overflow 0.10")
    from eplace_amd import EplaceNesterovOpt
    from eplace_amd.controller import Controller
    t = time.perf_counter()
    d = synthetic(2000000, 2100000, seed=5, num_io=1000, util=0.8, name="cells2m")
    gp = EPlacer2D(d, 1.0, bins=2048, poisson_mode=2)
    say(f"  design and engine: {d.num_pins} pins, {gp.info()['num_fillers']} fillers "
        f"({time.perf_counter() - t:.1f}s)")
    try:
        init = gp.read("positions").copy()
        res = {}
        for method, kw in (("bktrk", {}), ("learned", {"controller": Controller.load()})):
            gp.write("positions", init)
            gp.initialization()
            t = time.perf_counter()
            r = EplaceNesterovOpt(gp, method=method, target_overflow=0.10, log=False, **kw).opt()
            res[method] = r
            say(f"  {method}: {r['iterations']} iterations, {r['evaluations']} evaluations in "
                f"{time.perf_counter() - t:.2f}s, overflow {r['final_overflow']:.4f}, HPWL {r['final_hpwl']:.6g}")
            assert r["final_overflow"] < 0.10 and r["iterations"] < 1000
        assert res["learned"]["evaluations"] < res["bktrk"]["evaluations"]
        gap = res["learned"]["final_hpwl"] / res["bktrk"]["final_hpwl"] - 1
        say(f"  learned HPWL gap {100 * gap:+.3f}%")
        # the controller was trained on 50K-300K-cell designs; at 2M cells its
        # HPWL lands about half a percent above plain (1M cells: -0.25%)
        assert abs(gap) < 1e-2
    finally:
        gp.close()
