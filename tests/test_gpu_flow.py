"""The global-placement -> legalisation flow on the GPU engine: the GP
result legalised by Abacus (reference semantics) passes the legality
checker."""
import sys

import numpy as np
import pytest

from eplace_amd import EPlacer2D, EplaceNesterovOpt, synthetic
from eplace_amd.legalize import check_legal, legalize_abacus

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("cells", [20000, 200000])
def test_gp_then_abacus_is_legal(cells):
    say(f"[flow] {cells} cells: GP -> Abacus -> check")
    d = synthetic(cells, int(cells * 1.05), seed=31, num_io=200, util=0.7, name="flow")
    gp = EPlacer2D(d, 1.0, poisson_mode=2)
    gp.initialization()
    r = EplaceNesterovOpt(gp, method="bktrk").opt()
    assert r["final_overflow"] < 0.1
    xy = gp.read("positions")[: d.num_nodes]
    out, st = legalize_abacus(d, xy, mode=0)
    chk = check_legal(d, out)
    say(f"  GP HPWL {r['final_hpwl']:.6e}, legalised {st['cells']} cells, mean displacement "
        f"{st['total_displacement'] / st['cells']:.3f}, max {st['max_displacement']:.2f}; {chk}")
    assert chk["legal"], chk
    # the legalised placement moved cells by about a site on average
    assert st["total_displacement"] / st["cells"] < 5.0
