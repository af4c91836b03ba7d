"""Geometric checks of placements shared by the flow tests."""
import numpy as np


def macro_overlap_extent(d, xy):
    """Largest min(x overlap, y overlap) over macro pairs (0: none)."""
    m = d.node_is_macro().astype(bool)
    w, h = d.node_w[m].astype(np.float64), d.node_h[m].astype(np.float64)
    l, b = xy[m, 0] - w / 2, xy[m, 1] - h / 2
    worst = 0.0
    for i in range(len(l)):
        ox = np.minimum(l[i] + w[i], l + w) - np.maximum(l[i], l)
        oy = np.minimum(b[i] + h[i], b + h) - np.maximum(b[i], b)
        ox[i] = 0.0
        e = np.where((ox > 0) & (oy > 0), np.minimum(ox, oy), 0.0)
        worst = max(worst, float(e.max()))
    return worst


def cells_legal(chk):
    """The standard cells of a legality check are legal (macro pairs aside)."""
    return all(chk[k] == 0 for k in ("overlaps", "off_row", "off_site", "out_of_core", "in_blockage"))
