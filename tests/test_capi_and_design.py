"""CPU checks of the boundary and the host logic: the C-ABI library loads and
exports every symbol include/eplace_amd.h declares; Bookshelf I/O round
trips; the generator is deterministic; argument validation rejects bad input
without touching a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

from eplace_amd import _lib as L
from eplace_amd.design import read_bookshelf, synthetic, write_bookshelf


def header_symbols(root):
    txt = open(os.path.join(root, "include", "eplace_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ep_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol(root):
    lib = L.load()
    names = header_symbols(root)
    assert len(names) >= 18
    for n in names:
        assert hasattr(lib, n), n
        assert n in L.SYMBOLS, f"{n} not bound in eplace_amd/_lib.py"


def test_bookshelf_roundtrip(tmp_path):
    d = synthetic(200, 260, seed=8, num_io=12, degrees="c1", widths="uniform14")
    aux = write_bookshelf(d, str(tmp_path))
    e = read_bookshelf(aux)
    assert e.num_nodes == d.num_nodes and e.num_terminals == d.num_terminals
    np.testing.assert_array_equal(e.net_start, d.net_start)
    np.testing.assert_array_equal(e.pin_obj, d.pin_obj)
    np.testing.assert_array_equal(e.node_w, d.node_w)
    np.testing.assert_allclose(e.pin_dx, d.pin_dx, rtol=1e-6)
    np.testing.assert_allclose(e.node_x, d.node_x, atol=2e-5)
    np.testing.assert_array_equal(e.row_bottom, d.row_bottom)
    assert e.core_region() == d.core_region()


def test_reads_reference_smoke_format():
    smoke = "/root/reference/examples/smoke/smoke.aux"
    if not os.path.exists(smoke):
        pytest.skip("reference tree not present")
    d = read_bookshelf(smoke)
    assert (d.num_nodes, d.num_terminals, d.num_nets, d.num_pins) == (100, 4, 184, 368)
    assert d.term_ni.all()


def test_generator_deterministic_and_recipe():
    a = synthetic(20000, 21000, seed=3)
    b = synthetic(20000, 21000, seed=3)
    np.testing.assert_array_equal(a.pin_obj, b.pin_obj)
    np.testing.assert_array_equal(a.node_x, b.node_x)
    deg = np.diff(a.net_start)
    assert abs((deg == 2).mean() - 0.55) < 0.02 and abs((deg == 3).mean() - 0.20) < 0.02
    assert deg.max() <= 200
    assert 3.4 < a.num_pins / a.num_nets < 4.2  # ~3.8 pins per net
    assert abs(a.node_w.sum() / (a.meta["die"] ** 2) - 0.7) < 0.01
    m = synthetic(3000, 3200, seed=4, num_macros=5, util=0.75)
    assert m.node_is_macro().sum() == 5


def test_create_rejects_bad_input_before_any_gpu_work():
    from eplace_amd import EPlacer2D
    d = synthetic(50, 60, seed=1, degrees="c1")
    d.pin_obj = d.pin_obj.copy()
    d.pin_obj[0] = d.num_nodes + d.num_terminals + 5
    with pytest.raises(L.EPError, match="pin_obj out of range"):
        EPlacer2D(d, 1.0)
    d = synthetic(50, 60, seed=1, degrees="c1")
    with pytest.raises(L.EPError, match="bins out of range"):
        EPlacer2D(d, 1.0, bins=4096)
    with pytest.raises(L.EPError, match="target_density"):
        EPlacer2D(d, 0.0)


def test_create_rejects_odd_bins_and_unknown_poisson_mode():
    from eplace_amd import EPlacer2D
    d = synthetic(50, 60, seed=1, degrees="c1")
    with pytest.raises(L.EPError, match="even"):
        EPlacer2D(d, 1.0, bins=(15, 16))
    with pytest.raises(L.EPError, match="poisson_mode"):
        EPlacer2D(d, 1.0, poisson_mode=4)
