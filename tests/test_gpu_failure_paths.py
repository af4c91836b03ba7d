"""Failure handling of the device-resident loop, forced from the host: every
kernel still runs to its normal end.  The reference has none of these
guards: it carries a 0/0 Lipschitz
constant into the positions (optutils.hpp:33-34), has no cap on
backtracking (nesterov.hpp:337-364) and only its MAX_ITERATION cap
(nesterov.hpp:36, 105)."""
import dataclasses
import sys

import numpy as np
import pytest

from eplace_amd import EPlacer2D, EplaceNesterovOpt, synthetic

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_zero_gradient_keeps_finite_positions():
    """Every net on terminals only: no node has a wirelength gradient, so
    lambda0 = sum|grad WL| / sum|grad D| = 0 and the preconditioned
    direction is zero; the positions never move and each step's 1/L is
    0/0.  The loop keeps the previous step, counts it, stops at the
    iteration cap, and the positions stay finite and where they were."""
    say("[failure] zero direction: 0/0 Lipschitz steps")
    d = synthetic(400, 600, seed=3, degrees="c1", widths="uniform14", name="zerograd")
    n = d.num_nodes
    # every pin moved onto terminal 0 (offset 0): nets exist, no node pin
    d = dataclasses.replace(d, pin_obj=np.full(d.num_pins, n, np.int32),
                            pin_dx=np.zeros(d.num_pins, np.float32), pin_dy=np.zeros(d.num_pins, np.float32))
    gp = EPlacer2D(d, 0.8, bins=16)
    gp.initialization()
    assert gp.stats()["lambda"] == 0.0
    before = gp.read("positions").copy()
    r = EplaceNesterovOpt(gp, method="bktrk", max_iterations=8, log=False).opt()
    after = gp.read("positions")
    say(f"  {r['iterations']} iterations, {r['numeric_fallbacks']} numeric fallbacks, overflow {r['final_overflow']:.3f}")
    gp.close()
    assert np.all(np.isfinite(after))
    assert np.array_equal(after, before)
    assert r["numeric_fallbacks"] >= r["iterations"] - 1 > 0
    assert r["iterations"] <= 9
    assert np.isfinite(r["final_hpwl"])


def test_iteration_cap():
    """A run that cannot reach its target overflow stops at max_iterations
    (the reference's MAX_ITERATION semantics: stop when iter > cap)."""
    say("[failure] iteration cap")
    d = synthetic(2000, 2100, seed=12, name="cap")
    gp = EPlacer2D(d, 1.0)
    gp.initialization()
    r = EplaceNesterovOpt(gp, method="bktrk", max_iterations=20, target_overflow=0.0, log=False).opt()
    gp.close()
    say(f"  {r['iterations']} iterations, overflow {r['final_overflow']:.3f}")
    assert 20 <= r["iterations"] <= 21
    assert r["final_overflow"] > 0.0


def test_backtracking_cap():
    """With the trial cap at 1, every iteration takes its first trial even
    when it fails the acceptance test (counted as a backtrack, with no
    second evaluation), and the run still converges."""
    say("[failure] backtracking cap")
    d = synthetic(2000, 2100, seed=12, name="bt")
    res = {}
    for cap in (30, 1):
        gp = EPlacer2D(d, 1.0)
        gp.initialization()
        res[cap] = EplaceNesterovOpt(gp, method="bktrk", max_backtracks=cap, log=False).opt()
        gp.close()
    say("  " + ", ".join(f"cap {k}: {v['iterations']} iterations, {v['evaluations']} evaluations, "
                         f"{v['backtracks']} backtracks" for k, v in res.items()))
    capped = res[1]
    assert capped["evaluations"] <= capped["iterations"] + 1
    assert capped["backtracks"] <= capped["iterations"] + 1
    assert res[30]["evaluations"] > res[30]["iterations"]
    assert capped["final_overflow"] <= 0.10
    assert np.isfinite(capped["final_hpwl"])
