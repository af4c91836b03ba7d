"""Quadratic initial placement (QPPlacer, QPlace/qplace.cpp:20-295) through
the library's device entry points.

``quadratic_placement(design)`` is ``QPPlacer(db).quadraticPlacement()``: it
moves the design's movable nodes (centres in ``design.node_x/node_y``) to the
quadratic-placement solution, as the reference moves its database's modules.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .design import Design
from .placer import design_struct

FLOAT_EPS = float(np.finfo(np.float32).eps)  # Eigen's default BiCGSTAB tolerance for float


def _err(lib, rc):
    if rc != L.EP_OK:
        raise L.EPError(rc, lib.ep_qp_last_error().decode())


def quadratic_placement(d: Design, max_iterations: int = 20, solver_iterations: int = 100,
                        target_error: float = 1e-6, start_at_center: bool = True, device: int = 0) -> dict:
    """Runs QP on the device and writes the result into d.node_x / d.node_y.

    max_iterations is -IPiteCount (qplace.cpp:35-42).  Returns the stats and
    the per-iteration log (HPWL, X error, Y error, BiCGSTAB iterations)."""
    lib = L.load()
    sd, keep = design_struct(d)
    xy = np.ascontiguousarray(np.stack([d.node_x, d.node_y], 1), np.float32)
    log = np.zeros((max(max_iterations, 0) + 1, 4), np.float64)
    p = L.EPQpParams(max_iterations=int(max_iterations), solver_iterations=int(solver_iterations),
                     target_error=float(target_error), solver_tolerance=FLOAT_EPS,
                     start_at_center=1 if start_at_center else 0, device=int(device))
    st = L.EPQpStats()
    _err(lib, lib.ep_qplace(C.byref(sd), C.byref(p), xy.ctypes.data, log.ctypes.data, C.byref(st)))
    del keep
    d.node_x[:] = xy[:, 0]
    d.node_y[:] = xy[:, 1]
    out = {k: getattr(st, k) for k, _ in L.EPQpStats._fields_ if k != "pad"}
    out["log"] = log[: st.iterations].copy()
    return out


def qp_matrix(d: Design, xy=None, device: int = 0) -> dict:
    """QPPlacer::createSparseMatrix at the given centres (default: the
    design's): off-diagonal CSR over 2N rows (X then Y), diagonal, b."""
    lib = L.load()
    sd, keep = design_struct(d)
    if xy is None:
        xy = np.stack([d.node_x, d.node_y], 1)
    xy = np.ascontiguousarray(xy, np.float32).reshape(-1, 2)[: d.num_nodes]
    out = L.EPQpMatrixOut()
    _err(lib, lib.ep_qp_matrix(C.byref(sd), xy.ctypes.data, int(device), C.byref(out)))
    n, nnz = d.num_nodes, int(out.nnz)
    arrs = dict(rowptr=np.zeros(2 * n + 1, np.int32), col=np.zeros(max(nnz, 1), np.int32),
                val=np.zeros(max(nnz, 1), np.float32), diag=np.zeros(2 * n, np.float32),
                b=np.zeros(2 * n, np.float32))
    out.capacity = max(nnz, 1)
    for k, a in arrs.items():
        setattr(out, k, a.ctypes.data)
    _err(lib, lib.ep_qp_matrix(C.byref(sd), xy.ctypes.data, int(device), C.byref(out)))
    del keep
    arrs["col"], arrs["val"] = arrs["col"][:nnz], arrs["val"][:nnz]
    arrs["hpwl"] = float(out.hpwl)
    return arrs


def qp_solve(rowptr, col, val, b, x0, diag=None, ndim: int = 1, max_iterations: int = 100,
             tolerance: float = FLOAT_EPS, device: int = 0):
    """BiCGSTAB (identity preconditioner) on ndim stacked systems of one CSR.
    Returns (x, error[ndim], iterations[ndim])."""
    lib = L.load()
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    x = np.ascontiguousarray(x0, np.float32).copy()
    n = len(b) // ndim
    dg = None if diag is None else np.ascontiguousarray(diag, np.float32)
    sys_ = L.EPQpSystem(n=n, ndim=ndim, nnz=len(col), rowptr=rowptr.ctypes.data,
                        col=col.ctypes.data if len(col) else None, val=val.ctypes.data if len(val) else None,
                        diag=None if dg is None else dg.ctypes.data, b=b.ctypes.data)
    err = np.zeros(ndim, np.float32)
    its = np.zeros(ndim, np.int32)
    _err(lib, lib.ep_qp_solve(C.byref(sys_), x.ctypes.data, int(max_iterations), float(tolerance), int(device),
                              err.ctypes.data, its.ctypes.data))
    return x, err, its


def qp_solve_noalloc(n: int, ndim: int, nnz: int, rowptr: int, col: int, val: int, b: int, x: int,
                     diag: int = 0, max_iterations: int = 100, tolerance: float = FLOAT_EPS, device: int = 0,
                     stream: int = 0):
    """ep_qp_solve_noalloc: BiCGSTAB on a system already in device memory.
    rowptr, col, val, b, diag (0: none) and x are device addresses on
    `device`; x holds the guess and receives the solution.  The library
    allocates nothing: its work space is a fixed static device array of
    EP_QP_NOALLOC_MAX_ROWS rows.  Returns (error[ndim], iterations[ndim])."""
    lib = L.load()
    sys_ = L.EPQpSystem(n=int(n), ndim=int(ndim), nnz=int(nnz), rowptr=rowptr or None, col=col or None,
                        val=val or None, diag=diag or None, b=b or None)
    err = np.zeros(ndim, np.float32)
    its = np.zeros(ndim, np.int32)
    _err(lib, lib.ep_qp_solve_noalloc(C.byref(sys_), x or None, int(max_iterations), float(tolerance), int(device),
                                      stream or None, err.ctypes.data, its.ctypes.data))
    return err, its
