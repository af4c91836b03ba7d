"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE: imported only by tests/, the package smoke check and the
bench's CPU-baseline leg, as the checker or the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build() -> None:
    root = os.path.dirname(_HERE)
    subprocess.run(["make", "-s", "-f", "oracle/Makefile"], cwd=root, check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.c_void_p, C.c_float, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                 C.c_void_p, C.c_int, C.c_uint32]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_info_i.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_info_d.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_set_scalar.argtypes = [C.c_void_p, C.c_char_p, C.c_double]
        L.orc_get.restype = C.c_int64
        L.orc_get.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p]
        L.orc_set.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p]
        L.orc_run.argtypes = [C.c_void_p, C.c_char_p, C.c_int, C.c_int]
        L.orc_hpwl.restype = C.c_double
        L.orc_hpwl.argtypes = [C.c_void_p]
        L.orc_stop_condition.argtypes = [C.c_void_p, C.c_float]
        L.orc_poisson_f64.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float] + [C.c_void_p] * 4 + [C.c_int]
        L.orc_glibc_rand.argtypes = [C.c_uint32, C.c_int, C.c_void_p]
        _lib = L
    return _lib


class _Design(C.Structure):
    _fields_ = [
        ("num_nodes", C.c_int32), ("num_terminals", C.c_int32), ("num_nets", C.c_int32),
        ("num_rows", C.c_int32), ("num_pins", C.c_int64),
        ("node_w", C.c_void_p), ("node_h", C.c_void_p), ("node_x", C.c_void_p), ("node_y", C.c_void_p),
        ("term_w", C.c_void_p), ("term_h", C.c_void_p), ("term_x", C.c_void_p), ("term_y", C.c_void_p),
        ("term_ni", C.c_void_p), ("net_start", C.c_void_p), ("pin_obj", C.c_void_p),
        ("pin_dx", C.c_void_p), ("pin_dy", C.c_void_p),
        ("row_bottom", C.c_void_p), ("row_height", C.c_void_p), ("row_x0", C.c_void_p), ("row_x1", C.c_void_p),
        ("common_row_height", C.c_double),
    ]


def _ptr(a):
    return a.ctypes.data if a is not None and len(a) else None


class Oracle:
    """One oracle placer.  mode "ref" (reference float arithmetic) or "fp64"."""

    INFO_I = ["N", "F", "O", "NT", "E", "nx", "ny", "stage", "iter", "mgp_iterations", "bktrk_count",
              "fixed_shift", "P"]
    INFO_D = ["bsx", "bsy", "bin_area", "core_llx", "core_lly", "core_urx", "core_ury", "total_row_area",
              "std_area", "macro_area", "overflow", "invgx", "invgy", "lambda", "last_hpwl", "hpwl_bound",
              "wl_value", "td", "last_step", "ns_a", "filler_w", "filler_h"]

    def __init__(self, design, target_density, *, mode="ref", bins=0, exp="fast", wl="WA",
                 filler_xy=None, filler_count=-1, seed=1):
        self._keep = []
        arr = lambda a, dt: np.ascontiguousarray(a, dtype=dt)
        d = design
        fields = dict(
            node_w=arr(d.node_w, np.float32), node_h=arr(d.node_h, np.float32),
            node_x=arr(d.node_x, np.float32), node_y=arr(d.node_y, np.float32),
            term_w=arr(d.term_w, np.float32), term_h=arr(d.term_h, np.float32),
            term_x=arr(d.term_x, np.float32), term_y=arr(d.term_y, np.float32),
            term_ni=arr(d.term_ni, np.uint8), net_start=arr(d.net_start, np.int64),
            pin_obj=arr(d.pin_obj, np.int32), pin_dx=arr(d.pin_dx, np.float32), pin_dy=arr(d.pin_dy, np.float32),
            row_bottom=arr(d.row_bottom, np.float64), row_height=arr(d.row_height, np.float64),
            row_x0=arr(d.row_x0, np.float64), row_x1=arr(d.row_x1, np.float64),
        )
        self._keep.append(fields)
        sd = _Design(num_nodes=d.num_nodes, num_terminals=d.num_terminals, num_nets=d.num_nets,
                     num_rows=len(d.row_bottom), num_pins=d.num_pins,
                     common_row_height=d.common_row_height,
                     **{k: _ptr(v) for k, v in fields.items()})
        self._sd = sd
        bx, by = (bins, bins) if np.isscalar(bins) else bins
        fxy = None if filler_xy is None else arr(filler_xy, np.float32).reshape(-1)
        self._keep.append(fxy)
        self.h = lib().orc_create(C.byref(sd), float(target_density), int(bx), int(by),
                                  0 if mode == "ref" else 1, 0 if exp == "fast" else 1,
                                  0 if wl == "WA" else 1, _ptr(fxy), int(filler_count), int(seed))
        self.mode = mode

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def info(self):
        a = np.zeros(len(self.INFO_I), np.int64)
        b = np.zeros(len(self.INFO_D), np.float64)
        lib().orc_info_i(self.h, a.ctypes.data)
        lib().orc_info_d(self.h, b.ctypes.data)
        out = dict(zip(self.INFO_I, a.tolist()))
        out.update(dict(zip(self.INFO_D, b.tolist())))
        return out

    def get(self, name):
        n = lib().orc_get(self.h, name.encode(), None)
        if n < 0:
            raise KeyError(name)
        dt = np.int64 if name.endswith("_q") else np.float64
        out = np.zeros(n, dt)
        lib().orc_get(self.h, name.encode(), out.ctypes.data)
        return out

    def set(self, name, values):
        dt = np.int64 if name.endswith("_q") else np.float64
        v = np.ascontiguousarray(values, dtype=dt).reshape(-1)
        want = {"stage_positions_clamped": "stage_positions"}.get(name, name)
        n = lib().orc_get(self.h, want.encode(), None)
        if n >= 0 and v.size != n:
            raise ValueError(f"{name}: got {v.size} values, the oracle holds {n}")
        if not lib().orc_set(self.h, name.encode(), v.ctypes.data):
            raise KeyError(name)

    def set_scalar(self, name, v):
        lib().orc_set_scalar(self.h, name.encode(), float(v))

    def run(self, what, a0=0, a1=-1):
        if not lib().orc_run(self.h, what.encode(), int(a0), int(a1)):
            raise KeyError(what)

    def map(self, name):
        i = self.info()
        return self.get(name).reshape(i["nx"], i["ny"])

    def hpwl(self):
        return lib().orc_hpwl(self.h)

    def stop_condition(self, target_overflow):
        return bool(lib().orc_stop_condition(self.h, float(target_overflow)))


def poisson_f64(rho, bsx, bsy, method="fast"):
    rho = np.ascontiguousarray(rho, np.float64)
    nx, ny = rho.shape
    phi, ex, ey = (np.zeros_like(rho) for _ in range(3))
    lib().orc_poisson_f64(nx, ny, bsx, bsy, rho.ctypes.data, phi.ctypes.data, ex.ctypes.data, ey.ctypes.data,
                          1 if method == "fast" else 0)
    return phi, ex, ey


def glibc_rand(n, seed=1):
    out = np.zeros(n, np.int32)
    lib().orc_glibc_rand(seed, n, out.ctypes.data)
    return out
