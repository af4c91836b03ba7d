"""Host-side mirror of the reference's placement API over the C ABI.

``EPlacer2D`` keeps the method names and call order of EPlacer_2D
(EPlace/eplace.h:95-200); ``EplaceNesterovOpt`` mirrors
FirstOrderOptimizer::opt / EplaceNesterovOpt (Optimization/opt.hpp:29-36,
nesterov.hpp:59-415) but runs the loop device-resident through
ep_nesterov_run.  Positions and gradients are (n, 2) float32 arrays in the
reference's stage order.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .design import Design

__all__ = ["EPlacer2D", "EplaceNesterovOpt", "MGP", "FILLERONLY", "CGP"]

MGP, FILLERONLY, CGP = L.EP_STAGE_MGP, L.EP_STAGE_FILLERONLY, L.EP_STAGE_CGP


def _ptr(a):
    return a.ctypes.data if a is not None and a.size else None


def design_struct(d: Design):
    """The ep_design view of a Design; keep the returned arrays alive with it."""
    a = lambda v, dt: np.ascontiguousarray(v, dtype=dt)
    arrays = dict(
        node_w=a(d.node_w, np.float32), node_h=a(d.node_h, np.float32),
        node_x=a(d.node_x, np.float32), node_y=a(d.node_y, np.float32),
        term_w=a(d.term_w, np.float32), term_h=a(d.term_h, np.float32),
        term_x=a(d.term_x, np.float32), term_y=a(d.term_y, np.float32),
        term_ni=a(d.term_ni, np.uint8), net_start=a(d.net_start, np.int64),
        pin_obj=a(d.pin_obj, np.int32), pin_dx=a(d.pin_dx, np.float32), pin_dy=a(d.pin_dy, np.float32),
        row_bottom=a(d.row_bottom, np.float64), row_height=a(d.row_height, np.float64),
        row_x0=a(d.row_x0, np.float64), row_x1=a(d.row_x1, np.float64),
    )
    sd = L.EPDesign(num_nodes=d.num_nodes, num_terminals=d.num_terminals, num_nets=d.num_nets,
                    num_rows=len(d.row_bottom), num_pins=d.num_pins, common_row_height=d.common_row_height,
                    **{k: _ptr(v) for k, v in arrays.items()})
    return sd, arrays


# timed phases of a gradient evaluation, in the order the library reports them
# (scatter and density finalise; on the second stream the wiring kernels and
# the node sums of their pin gradients; then the Poisson solve, the field
# gather and the combine, which includes any wait for the second stream)
PHASES = ("scatter", "finalize", "wirelength", "node_sums", "poisson", "field_gather", "combine")
# phase_seconds of ep_nesterov_run also times the transform GEMM launches alone
RUN_PHASES = PHASES + ("poisson_gemms",)

class EPlacer2D:
    """GPU global-placement evaluator (EPlacer_2D)."""

    def __init__(self, design: Design, target_density: float = 1.0, *, bins=0, wirelength="WA",
                 exp="reference", filler_seed: int = 1, filler_xy=None, filler_count: int = -1,
                 poisson_mode: int = 0, device: int = 0):
        self.lib = L.load()
        self.design = design
        a = lambda v, dt: np.ascontiguousarray(v, dtype=dt)
        sd, self._arrays = design_struct(design)
        bx, by = (bins, bins) if np.isscalar(bins) else bins
        self._filler_xy = None if filler_xy is None else a(filler_xy, np.float32).reshape(-1)
        p = L.EPParams(target_density=float(target_density), bins_x=int(bx), bins_y=int(by),
                       wirelength_model=0 if wirelength == "WA" else 1, exp_mode=0 if exp == "reference" else 1,
                       filler_seed=int(filler_seed), filler_xy=_ptr(self._filler_xy),
                       filler_count=int(filler_count), poisson_mode=int(poisson_mode), device=int(device))
        h = C.c_void_p()
        rc = self.lib.ep_create(C.byref(sd), C.byref(p), C.byref(h))
        self.h = h.value
        if rc != L.EP_OK:
            msg = self.lib.ep_last_error(self.h).decode() if self.h else "ep_create failed"
            if self.h:
                self.lib.ep_destroy(self.h)
            self.h = None
            raise L.EPError(rc, msg)
        self.placementStage = MGP
        self.mGPIterationCount = 0

    # -- lifetime --------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            self.lib.ep_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc != L.EP_OK:
            raise L.EPError(rc, self.lib.ep_last_error(self.h).decode())

    # -- info --------------------------------------------------------------
    def info(self) -> dict:
        i = L.EPInfo()
        self._check(self.lib.ep_get_info(self.h, C.byref(i)))
        return {k: getattr(i, k) for k, _ in L.EPInfo._fields_}

    def stats(self) -> dict:
        s = L.EPStats()
        self._check(self.lib.ep_get_stats(self.h, C.byref(s)))
        out = {k: getattr(s, k) for k, _ in L.EPStats._fields_}
        out["lambda"] = out.pop("lambda_")
        return out

    @property
    def globalDensityOverflow(self) -> float:
        return self.stats()["overflow"]

    @property
    def lambda_(self) -> float:
        return self.stats()["lambda"]

    @property
    def lastHPWL(self) -> float:
        return self.stats()["last_hpwl"]

    # -- EPlacer_2D surface ------------------------------------------------
    def initialization(self):
        """First evaluation and lambda0 (eplace.cpp:39-50; fillers and bins
        were built by the constructor)."""
        self.totalGradientUpdate()
        self.penaltyFactorInitilization()

    def totalGradientUpdate(self):
        self._check(self.lib.ep_total_gradient_update(self.h))

    def stage_size(self) -> int:
        return int(self.lib.ep_stage_size(self.h))

    def getPosition(self) -> np.ndarray:
        n = self.stage_size()
        out = np.zeros((n, 2), np.float32)
        self._check(self.lib.ep_get_position(self.h, _ptr(out), n))
        return out

    def setPosition(self, xy) -> None:
        xy = np.ascontiguousarray(xy, np.float32).reshape(-1, 2)
        self._check(self.lib.ep_set_position(self.h, _ptr(xy), xy.shape[0]))

    def getGradient(self) -> np.ndarray:
        n = self.stage_size()
        out = np.zeros((n, 2), np.float32)
        self._check(self.lib.ep_get_gradient(self.h, _ptr(out), n))
        return out

    def penaltyFactorInitilization(self):
        self._check(self.lib.ep_penalty_factor_init(self.h))

    def updatePenaltyFactor(self):
        self._check(self.lib.ep_update_penalty_factor(self.h))

    def calcHPWL(self) -> float:
        """HPWL of the current positions (PlaceDB::calcHPWL, placedb.cpp:424-435)."""
        import ctypes
        v = ctypes.c_double()
        self._check(self.lib.ep_hpwl(self.h, ctypes.byref(v)))
        return v.value

    def switch2FillerOnly(self):
        self._check(self.lib.ep_switch_to_filler_only(self.h))
        self.placementStage = FILLERONLY

    def switch2cGP(self):
        self._check(self.lib.ep_switch_to_cgp(self.h, int(self.mGPIterationCount)))
        self.placementStage = CGP

    def showInfo(self) -> str:
        s = self.stats()
        return f"Overflow: {s['overflow']}\npenalty factor: {s['lambda']:.6f}\nHPWL: {s['last_hpwl']}\n"

    # -- raw arrays (parity tests) -----------------------------------------
    def read(self, which: str) -> np.ndarray:
        i = self.info()
        nb = i["bins_x"] * i["bins_y"]
        O, N, P = i["num_objects"], i["num_nodes"], i["num_pins"]
        table = {
            "positions": (L.EP_ARR_POSITIONS, np.float32, 2 * O, (O, 2)),
            "node_q": (L.EP_ARR_NODE_Q, np.int64, nb, (i["bins_x"], i["bins_y"])),
            "filler_q": (L.EP_ARR_FILLER_Q, np.int64, nb, (i["bins_x"], i["bins_y"])),
            "rho": (L.EP_ARR_RHO, np.float32, nb, (i["bins_x"], i["bins_y"])),
            "phi": (L.EP_ARR_PHI, np.float32, nb, (i["bins_x"], i["bins_y"])),
            "ex": (L.EP_ARR_EX, np.float32, nb, (i["bins_x"], i["bins_y"])),
            "ey": (L.EP_ARR_EY, np.float32, nb, (i["bins_x"], i["bins_y"])),
            "grad_density": (L.EP_ARR_GRAD_DENSITY, np.float32, 2 * O, (O, 2)),
            "grad_wirelength": (L.EP_ARR_GRAD_WL, np.float32, 2 * N, (N, 2)),
            "grad_total": (L.EP_ARR_GRAD_TOTAL, np.float32, 2 * O, (O, 2)),
            "term_map": (L.EP_ARR_TERM_MAP, np.float32, nb, (i["bins_x"], i["bins_y"])),
            "base_map": (L.EP_ARR_BASE_MAP, np.float32, nb, (i["bins_x"], i["bins_y"])),
            "pin_grad": (L.EP_ARR_PIN_GRAD, np.float32, 2 * P, (P, 2)),
            "main": (L.EP_ARR_MAIN, np.float32, 2 * O, (O, 2)),
        }
        code, dt, n, shape = table[which]
        out = np.zeros(n, dt)
        self._check(self.lib.ep_read(self.h, code, _ptr(out), n))
        return out.reshape(shape)

    def write(self, which: str, values) -> None:
        table = {"positions": (L.EP_ARR_POSITIONS, np.float32), "ex": (L.EP_ARR_EX, np.float32),
                 "ey": (L.EP_ARR_EY, np.float32), "rho": (L.EP_ARR_RHO, np.float32)}
        code, dt = table[which]
        v = np.ascontiguousarray(values, dt).reshape(-1)
        self._check(self.lib.ep_write(self.h, code, _ptr(v), v.size))

    def poisson(self, rho, with_phi=True):
        rho = np.ascontiguousarray(rho, np.float32)
        ex, ey = np.zeros_like(rho), np.zeros_like(rho)
        phi = np.zeros_like(rho) if with_phi else None
        self._check(self.lib.ep_poisson(self.h, _ptr(rho), _ptr(phi) if with_phi else None, _ptr(ex), _ptr(ey)))
        return phi, ex, ey

    def last_timings(self) -> dict:
        t = np.zeros(len(PHASES), np.float64)
        self._check(self.lib.ep_last_timings(self.h, _ptr(t)))
        return dict(zip(PHASES, t.tolist()))


class EplaceNesterovOpt:
    """FirstOrderOptimizer<VECTOR_3D>::opt() for one stage, device-resident.

    method: "vanilla" (default), "bktrk" (-bktrk), or "learned" (backtracking
    whose first trial step and lambda multiplier come from `controller`).
    """

    METHODS = {"vanilla": 0, "bktrk": 1, "learned": 2, "bb": 3}

    def __init__(self, placer: EPlacer2D, method: str = "vanilla", target_overflow: float = 0.10,
                 max_iterations: int = 1000, max_backtracks: int = 30, sync_every: int = 16, controller=None,
                 log: bool = True):
        if method not in self.METHODS:
            raise ValueError(f"method must be one of {sorted(self.METHODS)}")
        if method == "learned" and controller is None:
            raise ValueError("the learned method needs a controller")
        self.placer = placer
        self.method = method
        self.target_overflow = target_overflow
        self.max_iterations = max_iterations
        self.max_backtracks = max_backtracks
        self.sync_every = sync_every
        self.controller = controller
        self.keep_log = log
        self.log = None
        self.result = None

    def opt(self, profile: bool = False) -> dict:
        cap = self.max_iterations + 2 if self.keep_log else 0
        log = np.zeros((max(cap, 1), L.EP_LOG_COLS), np.float64)
        phases = np.zeros(len(RUN_PHASES) + 1, np.float64)
        p = L.EPOptParams(method=self.METHODS[self.method], target_overflow=self.target_overflow,
                          max_iterations=self.max_iterations, max_backtracks=self.max_backtracks,
                          sync_every=self.sync_every, log_capacity=cap, log=_ptr(log) if cap else None,
                          phase_seconds=_ptr(phases) if profile else None,
                          controller=self.controller.h if self.controller is not None else None)
        r = L.EPRunStats()
        self.placer._check(self.placer.lib.ep_nesterov_run(self.placer.h, C.byref(p), C.byref(r)))
        res = {k: getattr(r, k) for k, _ in L.EPRunStats._fields_ if k != "pad"}
        if profile:
            n = max(phases[-1], 1.0)
            res["phase_seconds"] = dict(zip(RUN_PHASES, (phases[:-1] / n).tolist()))
            res["phase_evaluations"] = int(phases[-1])
        self.log = log[: res["iterations"]] if cap else None
        if self.placer.placementStage == MGP:
            self.placer.mGPIterationCount = res["iterations"]
        self.result = res
        return res
