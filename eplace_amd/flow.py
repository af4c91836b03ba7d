"""The reference driver's placement flow (main/ePlace_main.cpp:121-259) over
the library:

    QP            QPPlacer::quadraticPlacement          (unless qp=False, -noQP)
    mGP           EplaceNesterovOpt on every object     (unless mgp=False, -nomGP)
    mLG           SAMacroLegalizer                      (designs with macros, unless mlg=False, -nomLG)
    FILLERONLY    switch2FillerOnly + EplaceNesterovOpt (after mLG, unless cgp=False, -nocGP)
    cGP           switch2cGP + EplaceNesterovOpt        (after FILLERONLY)
    legalisation  AbacusLegalizer                       (legal=True, -internalLegal)
    detailed      DetailedPlacer::detailedPlacement     (dp=True after legal=True, -internalDP,
                                                         ePlace_main.cpp:271-283)

The design's node centres are updated with the result, as the reference
writes its database.
"""
from __future__ import annotations

import time

import numpy as np

from .design import Design
from .legalize import SAMacroLegalizer, check_legal, detailed_place, legalize_abacus
from .placer import EPlacer2D, EplaceNesterovOpt
from .qplace import quadratic_placement


def place(d: Design, target_density: float = 1.0, target_overflow: float = 0.10, *, qp: bool = True,
          qp_iterations: int = 20, mgp: bool = True, mlg: bool = True, cgp: bool = True, legal: bool = False,
          dp: bool = False, method: str = "vanilla", controller=None, bins=0, wirelength: str = "WA", poisson_mode: int = 0,
          device: int = 0, verbose: bool = False) -> dict:
    """Runs the flow on d and returns per-stage results; d.node_x/node_y hold
    the final node centres (the legal ones when legal=True, after detailed
    placement when dp=True as well).  method
    "learned" takes its step / lambda controller from `controller`."""
    say = print if verbose else (lambda *a, **k: None)
    out: dict = {}
    t0 = time.perf_counter()
    if qp:
        out["qp"] = quadratic_placement(d, max_iterations=qp_iterations, device=device)
        say(f"QP: {out['qp']['iterations']} iterations, HPWL {out['qp']['final_hpwl']:.6g}")
    gp = EPlacer2D(d, target_density, bins=bins, wirelength=wirelength, poisson_mode=poisson_mode, device=device)
    try:
        gp.initialization()
        if mgp:
            out["mgp"] = EplaceNesterovOpt(gp, method=method, controller=controller, target_overflow=target_overflow,
                                           log=False).opt()
            say(f"mGP: {out['mgp']['iterations']} iterations, HPWL {out['mgp']['final_hpwl']:.6g}")
        macros = int(d.node_is_macro().sum())
        if macros > 0 and mlg:
            m = SAMacroLegalizer(gp)
            m.setTargetDensity(target_density)
            out["mlg"] = m.legalization()
            say(f"mLG: {macros} macros, overlap {out['mlg']['final_overlap']}, HPWL {out['mlg']['final_hpwl']:.6g}")
            if cgp:
                gp.switch2FillerOnly()
                out["fillers"] = EplaceNesterovOpt(gp, method=method, controller=controller, log=False).opt()
                gp.switch2cGP()
                out["cgp"] = EplaceNesterovOpt(gp, method=method, controller=controller, log=False).opt()
                say(f"cGP: {out['cgp']['iterations']} iterations, HPWL {out['cgp']['final_hpwl']:.6g}")
        xy = gp.read("positions")[: d.num_nodes].copy()
        out["hpwl"] = gp.calcHPWL()
    finally:
        gp.close()
    if legal:
        lxy, lst = legalize_abacus(d, xy, mode=0)
        out["legal"] = dict(lst, check=check_legal(d, lxy))
        xy = lxy
        if dp:
            dxy, dst = detailed_place(d, xy)
            out["dp"] = dict(dst, check=check_legal(d, dxy))
            say(f"DP: HPWL {dst['hpwl_before']:.6g} -> {dst['hpwl_after']:.6g}")
            xy = dxy
    d.node_x[:] = xy[:, 0]
    d.node_y[:] = xy[:, 1]
    out["positions"] = xy
    out["seconds"] = time.perf_counter() - t0
    return out
