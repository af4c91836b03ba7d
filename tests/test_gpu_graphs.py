"""Evaluations replayed from captured HIP graphs against the same kernels
launched one by one (the optimiser launches evaluations eagerly when their
phases are timed): identical iterates, bit for bit, over several runs, each
run replaying the cached graphs many times with host reads and eager work
(moves, acceptance tests, combine-only re-evaluations) in between, for
plain, backtracking and learned runs and for trials taken as accepted at
the backtracking limit (max_backtracks=1)."""
import numpy as np
import pytest

from eplace_amd import EPlacer2D, EplaceNesterovOpt, synthetic
from eplace_amd.controller import Controller

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("method,max_backtracks", [("bktrk", 30), ("vanilla", 30), ("learned", 30), ("bktrk", 1)])
def test_graph_replays_match_eager_launches(method, max_backtracks):
    say(f"[graphs] method={method} max_backtracks={max_backtracks}")
    d = synthetic(4000, 4200, seed=11, num_macros=4, util=0.7, name="graphs")
    gp = EPlacer2D(d, 1.0, bins=64)
    ctrl = Controller.load() if method == "learned" else None
    try:
        init = gp.read("positions").copy()
        runs = []
        for profile in (True, False, False, True, False):
            gp.write("positions", init)
            gp.initialization()
            r = EplaceNesterovOpt(gp, method=method, max_iterations=40, max_backtracks=max_backtracks, log=False,
                                  controller=ctrl).opt(profile=profile)
            runs.append((profile, r["iterations"], r["evaluations"], r["final_hpwl"], gp.read("positions").copy(),
                         gp.read("grad_total").copy(), r["backtracks"]))
        ref = runs[0]
        for profile, it, ev, hp, pos, gt, bt in runs[1:]:
            say(f"  profile={profile}: {it} iterations, {ev} evaluations, {bt} backtracks, HPWL {hp:.9g}")
            assert (it, ev, bt) == (ref[1], ref[2], ref[6])
            assert hp == ref[3]
            np.testing.assert_array_equal(pos, ref[4])
            np.testing.assert_array_equal(gt, ref[5])
        if method == "bktrk":
            assert ref[6] > 0  # rejected trials were taken
    finally:
        gp.close()
        if ctrl is not None:
            ctrl.close()
