"""The annealing macro legaliser (ep_legalize_macros, host code) against the
reference's own SAMacroLegalizer output (tests/golden/mlg_*.npz, written by
tests/golden/gen_golden.py from the reference's legalizer.cpp): identical
positions, cost terms and rand() stream position, from the same centres and
rand() state."""
import ctypes

import numpy as np
import pytest

from golden_utils import Golden
from eplace_amd import _lib as L
from eplace_amd.design import synthetic
from eplace_amd.legalize import legalize_macros

CASES = ["mlg_mixed", "mlg_m52", "mlg_mm62", "mlg_mm63"]


def glibc_rand_after(draws):
    """Value of the next rand() of an unseeded glibc stream after `draws`
    draws (the C library the reference links)."""
    libc = ctypes.CDLL(None)
    libc.srand(1)
    for _ in range(draws):
        libc.rand()
    return libc.rand()


def macro_rects(d, xy):
    m = d.node_is_macro().astype(bool)
    w, h = d.node_w[m], d.node_h[m]
    l = (xy[m, 0] - np.float32(0.5) * w).astype(np.float32)
    b = (xy[m, 1] - np.float32(0.5) * h).astype(np.float32)
    return l, b, (l + w).astype(np.float32), (b + h).astype(np.float32)


@pytest.mark.parametrize("name", CASES)
def test_bit_exact_with_reference(name):
    g = Golden(name)
    d = g.design()
    xy = g["input_positions"]
    out, st = legalize_macros(d, xy, float(g["target_density"]), rand_skip=int(g["rand_skip"]))
    np.testing.assert_array_equal(out, g["mlg_positions"])
    assert st["final_hpwl"] == g.s["mlg_hpwl"]
    assert st["final_overlap"] == g.s["mlg_overlap"] == 0
    assert st["overlap_free"] == g.s["mlg_overlap_free"] == 1
    assert st["cell_area_covered"] == g.s["mlg_cell_covered"]
    assert st["bins_x"] == g.s["mlg_bins"] and st["macros"] == g.s["mlg_macros"]
    # the annealing drew exactly as many rand() values as the reference
    assert glibc_rand_after(int(g["rand_skip"]) + st["rand_draws"]) == g.s["mlg_rand_next"]


def test_multi_round_case_runs_several_rounds():
    g = Golden("mlg_mm63")
    d = g.design()
    _, st = legalize_macros(d, g["input_positions"], 1.0, rand_skip=int(g["rand_skip"]))
    assert st["outer_iterations"] >= 5 and st["perturbations"] > 100_000


@pytest.mark.parametrize("name", CASES)
def test_macros_do_not_overlap(name):
    """The reference's stop test: no overlap between the integer-truncated
    macro rectangles; cells are not moved."""
    g = Golden(name)
    d = g.design()
    xy = g["input_positions"]
    out, _ = legalize_macros(d, xy, float(g["target_density"]), rand_skip=int(g["rand_skip"]))
    l, b, r, t = (a.astype(np.int64) for a in macro_rects(d, out))
    for i in range(len(l)):
        ox = np.maximum(0, np.minimum(r[i], r) - np.maximum(l[i], l))
        oy = np.maximum(0, np.minimum(t[i], t) - np.maximum(b[i], b))
        ov = ox * oy
        ov[i] = 0
        assert ov.max() == 0
    cells = ~d.node_is_macro().astype(bool)
    np.testing.assert_array_equal(out[cells], xy[cells])
    llx, lly, urx, ury = d.core_region()
    fl, fb, fr, ft = macro_rects(d, out)
    assert fl.min() >= llx and fb.min() >= lly and fr.max() <= urx and ft.max() <= ury


def test_no_macros_is_a_no_op():
    d = synthetic(500, 520, seed=3, num_io=10)
    xy = np.stack([d.node_x, d.node_y], 1)
    out, st = legalize_macros(d, xy)
    np.testing.assert_array_equal(out, xy)
    assert st["macros"] == 0 and st["rand_draws"] == 0 and st["overlap_free"] == 1


def test_rand_stream_moves_the_result():
    """Another rand() state gives another (still legal) annealing path."""
    g = Golden("mlg_m52")
    d = g.design()
    a, sa = legalize_macros(d, g["input_positions"], 0.9, rand_skip=777)
    b, sb = legalize_macros(d, g["input_positions"], 0.9, rand_skip=778)
    assert not np.array_equal(a, b)
    assert sa["final_overlap"] == sb["final_overlap"] == 0


def test_bad_arguments():
    lib = L.load()
    d = synthetic(100, 110, seed=1)
    from eplace_amd.placer import design_struct
    sd, keep = design_struct(d)
    out = np.zeros((d.num_nodes, 2), np.float32)
    p = L.EPMlgParams(target_density=0.0, max_outer=0, max_inner=0, rand_seed=1, rand_skip=0)
    assert lib.ep_legalize_macros(ctypes.byref(sd), None, ctypes.byref(p), out.ctypes.data, None) == L.EP_ERR_ARG
    assert b"target_density" in lib.ep_mlg_last_error()
    p.target_density = 1.0
    assert lib.ep_legalize_macros(None, None, ctypes.byref(p), out.ctypes.data, None) == L.EP_ERR_ARG
    del keep
