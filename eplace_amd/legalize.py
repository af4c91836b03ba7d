"""Legalisation through the library's host entry points: standard cells
(AbacusLegalizer, Legalization/legalizer.cpp:58-680), macros
(SAMacroLegalizer, legalizer.cpp:701-1530) and a legality checker."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .design import Design
from .placer import design_struct


def _centres(d: Design, xy):
    if xy is None:
        return np.ascontiguousarray(np.stack([d.node_x, d.node_y], 1), np.float32)
    xy = np.ascontiguousarray(xy, np.float32).reshape(-1, 2)
    if len(xy) < d.num_nodes:
        raise ValueError(f"need {d.num_nodes} centres, got {len(xy)}")
    return np.ascontiguousarray(xy[: d.num_nodes])


def legalize_abacus(d: Design, xy=None, mode: int = 0, site_step: float | None = None):
    """Legal node centres [N, 2] and stats; mode 0 = reference semantics, 1 = corrected."""
    lib = L.load()
    sd, keep = design_struct(d)
    inp = _centres(d, xy)
    out = np.zeros_like(inp)
    p = L.EPLegalParams(site_step=float(site_step or d.site_width), mode=int(mode))
    st = L.EPLegalStats()
    rc = lib.ep_legalize_abacus(C.byref(sd), inp.ctypes.data, C.byref(p), out.ctypes.data, C.byref(st))
    if rc != L.EP_OK:
        raise L.EPError(rc, lib.ep_host_last_error().decode())
    del keep
    return out, {k: getattr(st, k) for k, _ in L.EPLegalStats._fields_ if k != "pad"}


def check_legal(d: Design, xy=None, site_step: float | None = None) -> dict:
    lib = L.load()
    sd, keep = design_struct(d)
    inp = _centres(d, xy)
    r = L.EPLegalCheck()
    rc = lib.ep_check_legal(C.byref(sd), inp.ctypes.data, float(site_step or d.site_width), C.byref(r))
    if rc != L.EP_OK:
        raise L.EPError(rc, lib.ep_host_last_error().decode())
    del keep
    out = {k: getattr(r, k) for k, _ in L.EPLegalCheck._fields_}
    out["legal"] = all(out[k] == 0 for k in ("overlaps", "off_row", "off_site", "out_of_core", "in_blockage",
                                               "macro_overlaps"))
    return out


def detailed_place(d: Design, xy=None, passes: int = 2, site_step: float | None = None,
                   ism_time_limit: float = 0.0):
    """DetailedPlacer::detailedPlacement (DetailedPlacement/detailed.cpp:74-85)
    on legal node centres: `passes` rounds of local reordering, independent-set
    matching and global swap.  Returns the centres [N, 2] and stats (HPWL
    before, after and after each move of each round)."""
    lib = L.load()
    sd, keep = design_struct(d)
    inp = _centres(d, xy)
    out = np.zeros_like(inp)
    p = L.EPDPParams(site_step=float(site_step or d.site_width), ism_time_limit=float(ism_time_limit),
                     passes=int(passes))
    st = L.EPDPStats()
    rc = lib.ep_detailed_place(C.byref(sd), inp.ctypes.data, C.byref(p), out.ctypes.data, C.byref(st))
    if rc != L.EP_OK:
        raise L.EPError(rc, lib.ep_host_last_error().decode())
    del keep
    n = min(int(passes), L.EP_DP_MAX_PASSES)
    stats = {k: getattr(st, k) for k in ("hpwl_before", "hpwl_after", "seconds", "seconds_lr", "seconds_ism",
                                         "seconds_gs", "lr_windows", "ism_matchings", "gs_moves")}
    for k in ("hpwl_after_lr", "hpwl_after_ism", "hpwl_after_gs"):
        stats[k] = list(getattr(st, k))[:n]
    return out, stats


def legalize_macros(d: Design, xy=None, target_density: float = 1.0, rand_seed: int = 1, rand_skip: int = 0,
                    max_outer: int = 0, max_inner: int = 0):
    """SAMacroLegalizer(db).legalization() on node centres xy (default: the
    design's) with a rand() stream of seed rand_seed after rand_skip draws.
    Returns the centres after legalisation [N, 2] and the stats."""
    lib = L.load()
    sd, keep = design_struct(d)
    inp = _centres(d, xy)
    out = np.zeros_like(inp)
    p = L.EPMlgParams(target_density=float(target_density), max_outer=int(max_outer), max_inner=int(max_inner),
                      rand_seed=int(rand_seed), rand_skip=int(rand_skip))
    st = L.EPMlgStats()
    rc = lib.ep_legalize_macros(C.byref(sd), inp.ctypes.data, C.byref(p), out.ctypes.data, C.byref(st))
    if rc != L.EP_OK:
        raise L.EPError(rc, lib.ep_mlg_last_error().decode())
    del keep
    return out, {k: getattr(st, k) for k, _ in L.EPMlgStats._fields_}


class SAMacroLegalizer:
    """SAMacroLegalizer (Legalization/legalizer.h:304-420) on a placer, as the
    reference runs it between mGP and FILLERONLY (ePlace_main.cpp:174-186):
    node positions are read from and written back to the placer and the
    annealing draws from the placer's rand() stream."""

    def __init__(self, placer):
        self.placer = placer
        self.targetDensity = 1.0
        self.stats = None

    def setTargetDensity(self, td: float) -> None:
        self.targetDensity = float(td)

    def legalization(self, max_outer: int = 0, max_inner: int = 0) -> dict:
        pl = self.placer
        sd, keep = design_struct(pl.design)
        p = L.EPMlgParams(target_density=self.targetDensity, max_outer=int(max_outer), max_inner=int(max_inner),
                          rand_seed=1, rand_skip=0)
        st = L.EPMlgStats()
        pl._check(pl.lib.ep_macro_legalize(pl.h, C.byref(sd), C.byref(p), C.byref(st)))
        del keep
        self.stats = {k: getattr(st, k) for k, _ in L.EPMlgStats._fields_}
        return self.stats
