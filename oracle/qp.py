"""CPU restatement of the reference's quadratic initial placement (numpy).

TEST INFRASTRUCTURE: imported only by tests/ and the bench's CPU-baseline
leg, as the checker of the library's quadratic placer (ep_qplace,
ep_qp_matrix, ep_qp_solve).  Never part of the product path.

Follows
  * QPPlacer::quadraticPlacement          QPlace/qplace.cpp:20-103
  * QPPlacer::createSparseMatrix          QPlace/qplace.cpp:123-280
    (X pairs contain a bound pin; Y pairs pass the reference's condition
    pin2 == ymin || pin1 == ymax || pin2 == ymax, qplace.cpp:228)
  * QPPlacer::updateModuleLocation        QPlace/qplace.cpp:283-295
  * Net::calcBoundPin                     PlaceCommon/objects.cpp:75-140
  * PlaceDB::moveNodesCenterToCenter      PlaceDB/placedb.cpp:591-598
  * PlaceDB::setModuleCenter_2D (clamp)   PlaceDB/placedb.cpp:253-290
  * Eigen's bicgstab() with the identity preconditioner
    (Library/eigen-git-mirror/Eigen/src/IterativeLinearSolvers/BiCGSTAB.h:28-110):
    the van der Vorst BiCGSTAB iteration with Eigen's restart when
    |r0 . r| < eps^2 |r0|^2, the first restart resetting the count.
Differences, all below the reference's own float rounding: the diagonal and
b are summed in double and rounded once (the reference adds into float);
dot products and row sums are accumulated in double and rounded to float
(Eigen accumulates in float).

Pinned by tests/test_qp_oracle.py against the reference's own output
(tests/golden/qp_*.npz: its assembled matrices, b and its positions after
QP) and by properties of the systems (symmetry, row sums equal to the
fixed-pin weights, the exact solution of small systems).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

F32 = np.float32
EPS = 1.0e-15  # global.h:39
MIN_DISTANCE = 25.0  # qplace.h:20
FLOAT_EPS = F32(np.finfo(np.float32).eps)


def core_region(d):
    """(llx, lly, urx, ury) as float32 (placedb.cpp:52-77)."""
    return d.core_region()


def clamp(x, w, lo, hi):
    """setModuleCenter_2D's clamp of movable nodes (placedb.cpp:257-281):
    double comparisons, float assignments, the second test on the first's
    result."""
    x = np.asarray(x, F32).copy()
    w64 = np.asarray(w, F32).astype(np.float64)
    m = x.astype(np.float64) - 0.5 * w64 < float(lo)
    x[m] = (float(lo) + 0.5 * w64[m] + EPS).astype(F32)
    m = x.astype(np.float64) + 0.5 * w64 > float(hi)
    x[m] = (float(hi) - 0.5 * w64[m] - EPS).astype(F32)
    return x


def core_center(d):
    llx, lly, urx, ury = core_region(d)
    return F32(0.5 * float(F32(urx + llx))), F32(0.5 * float(F32(ury + lly)))


def pin_positions(d, cx, cy):
    """Pin::calculateAbsolutePos (objects.cpp:454-459): centre + offset in
    float; terminals keep their own centres."""
    n = d.num_nodes
    ox = np.concatenate([np.asarray(cx, F32), np.asarray(d.term_x, F32)])
    oy = np.concatenate([np.asarray(cy, F32), np.asarray(d.term_y, F32)])
    return (ox[d.pin_obj] + d.pin_dx).astype(F32), (oy[d.pin_obj] + d.pin_dy).astype(F32)


def _first_extreme(vals, starts, counts, nets_of_pin, mode):
    """Pin id of the first strict minimum / maximum of each net (the
    sequential scan of calcBoundPin keeps the first of equal values)."""
    E = len(starts)
    out = np.full(E, -1, np.int64)
    has = counts > 0
    v = vals.astype(np.float64)
    if mode == "min":
        ext = np.full(E, np.inf)
        np.minimum.at(ext, nets_of_pin, np.where(np.isnan(v), np.inf, v))
    else:
        ext = np.full(E, -np.inf)
        np.maximum.at(ext, nets_of_pin, np.where(np.isnan(v), -np.inf, v))
    hit = v == ext[nets_of_pin]
    idx = np.where(hit, np.arange(len(v)), len(v))
    first = np.full(E, len(v), np.int64)
    np.minimum.at(first, nets_of_pin, idx)
    ok = has & (first < len(v))
    out[ok] = first[ok]
    return out, ext


def bound_pins(d, px, py):
    """Per net (xmin, xmax, ymin, ymax) pin ids and the HPWL
    (calcBoundPin / calcNetBoundPins, placedb.cpp:462-470)."""
    starts = d.net_start[:-1]
    counts = np.diff(d.net_start)
    nets_of_pin = np.repeat(np.arange(len(starts)), counts)
    b0, mnx = _first_extreme(px, starts, counts, nets_of_pin, "min")
    b1, mxx = _first_extreme(px, starts, counts, nets_of_pin, "max")
    b2, mny = _first_extreme(py, starts, counts, nets_of_pin, "min")
    b3, mxy = _first_extreme(py, starts, counts, nets_of_pin, "max")
    has = counts > 0
    hpwl = float(np.sum((mxx[has] - mnx[has]) + (mxy[has] - mny[has])))
    return np.stack([b0, b1, b2, b3], 1), hpwl


def _pairs(d, B, dim):
    """Unordered pin pairs (a, c) of the reference's X or Y condition, one
    entry per pair, pins of the same module skipped."""
    P = d.num_pins
    counts = np.diff(d.net_start)
    net = np.repeat(np.arange(len(counts)), counts)
    q = np.arange(P)
    lo, hi = (B[net, 0], B[net, 1]) if dim == 0 else (B[net, 2], B[net, 3])
    if dim == 0:
        # pairs containing xmin, then pairs containing xmax but not xmin
        m1 = (lo >= 0) & (q != lo)
        m2 = (hi >= 0) & (hi != lo) & (q != hi) & (q != lo)
        a = np.concatenate([lo[m1], hi[m2]])
        c = np.concatenate([q[m1], q[m2]])
    else:
        # pairs containing ymax; pairs whose later pin is ymin, without ymax
        m1 = (hi >= 0) & (q != hi)
        m2 = (lo >= 0) & (lo != hi) & (q < lo) & (q != hi)
        a = np.concatenate([hi[m1], lo[m2]])
        c = np.concatenate([q[m1], q[m2]])
    keep = d.pin_obj[a] != d.pin_obj[c]
    a, c = a[keep], c[keep]
    return a, c, net[a]


def _c1(d):
    p = np.diff(d.net_start).astype(F32).astype(np.float64)
    with np.errstate(divide="ignore"):
        return (1.0 / (p - 1.0)).astype(F32)


def weights(c1, dist):
    """weightX / weightY (qplace.cpp:174-186)."""
    dist = np.asarray(dist, F32)
    ge = (dist.astype(np.float64) - 25.0 > EPS) | (np.abs(dist.astype(np.float64) - 25.0) < EPS)
    with np.errstate(divide="ignore", invalid="ignore"):
        near = (c1.astype(np.float64) / MIN_DISTANCE).astype(F32)
        far = (c1 / np.where(ge, dist, F32(1))).astype(F32)
    return np.where(ge, far, near).astype(F32)


def assemble(d, cx, cy):
    """createSparseMatrix at node centres (cx, cy): per direction the
    off-diagonal entries (row, col, value), the diagonal and b, float32."""
    n = d.num_nodes
    px, py = pin_positions(d, cx, cy)
    B, hpwl = bound_pins(d, px, py)
    c1 = _c1(d)
    out = {"hpwl": hpwl, "bound": B}
    for dim, pos, off in ((0, px, d.pin_dx), (1, py, d.pin_dy)):
        a, c, net = _pairs(d, B, dim)
        ma, mc = d.pin_obj[a].astype(np.int64), d.pin_obj[c].astype(np.int64)
        w = weights(c1[net], np.abs((pos[a] - pos[c]).astype(F32)))
        w64 = w.astype(np.float64)
        diag = np.zeros(n)
        b = np.zeros(n)
        mova, movc = ma < n, mc < n
        both = mova & movc
        np.add.at(diag, ma[mova], w64[mova])
        np.add.at(diag, mc[movc], w64[movc])
        oa, oc = off[a], off[c]
        np.add.at(b, ma[both], -1.0 * w64[both] * (oa[both] - oc[both]).astype(F32).astype(np.float64))
        np.add.at(b, mc[both], -1.0 * w64[both] * (oc[both] - oa[both]).astype(F32).astype(np.float64))
        fa = ~mova & movc  # a on a terminal
        np.add.at(b, mc[fa], (w[fa] * (pos[a][fa] - oc[fa]).astype(F32)).astype(F32).astype(np.float64))
        fc = mova & ~movc
        np.add.at(b, ma[fc], (w[fc] * (pos[c][fc] - oa[fc]).astype(F32)).astype(F32).astype(np.float64))
        rows = np.concatenate([ma[both], mc[both]])
        cols = np.concatenate([mc[both], ma[both]])
        vals = np.concatenate([-w[both], -w[both]]).astype(F32)
        out["xy"[dim]] = {"row": rows, "col": cols, "val": vals, "diag": diag.astype(F32), "b": b.astype(F32)}
    return out


def csr(n, row, col, val, diag=None):
    """Full matrix (off-diagonal entries plus the diagonal) as float64 CSR,
    duplicates summed."""
    A = sp.coo_matrix((np.asarray(val, np.float64), (row, col)), shape=(n, n)).tocsr()
    if diag is not None:
        A = A + sp.diags(np.asarray(diag, np.float64))
    A.sum_duplicates()
    return A.tocsr()


def bicgstab(A, b, x0, max_iter=100, tol=FLOAT_EPS):
    """Eigen's bicgstab() (BiCGSTAB.h:28-110), identity preconditioner, float
    vectors, double-accumulated products.  Returns (x, error, iterations,
    restarts)."""
    def mv(v):
        return (A @ v.astype(np.float64)).astype(F32)

    def dot(u, v):
        return F32(np.dot(u.astype(np.float64), v.astype(np.float64)))

    b = np.asarray(b, F32)
    x = np.asarray(x0, F32).copy()
    tol = F32(tol)
    r = (b - mv(x)).astype(F32)
    r0 = r.copy()
    r0_sq = dot(r0, r0)
    rhs_sq = dot(b, b)
    if rhs_sq == 0:
        return np.zeros_like(x), float(tol), 0, 0
    rho = alpha = w = F32(1)
    v = np.zeros_like(x)
    p = np.zeros_like(x)
    tol2 = F32(tol * tol * rhs_sq)
    eps2 = F32(FLOAT_EPS * FLOAT_EPS)
    i = restarts = 0
    r_sq = dot(r, r)
    while r_sq > tol2 and i < max_iter:
        rho_old = rho
        rho = dot(r0, r)
        if abs(rho) < F32(eps2 * r0_sq):
            r = (b - mv(x)).astype(F32)
            r0 = r.copy()
            rho = r0_sq = dot(r, r)
            if restarts == 0:
                i = 0
            restarts += 1
        beta = F32(F32(rho / rho_old) * F32(alpha / w))
        p = (r + beta * (p - w * v)).astype(F32)
        v = mv(p)
        alpha = F32(rho / dot(r0, v))
        s = (r - alpha * v).astype(F32)
        t = mv(s)
        tmp = dot(t, t)
        w = F32(dot(t, s) / tmp) if tmp > 0 else F32(0)
        x = (x + (alpha * p + w * s).astype(F32)).astype(F32)
        r = (s - w * t).astype(F32)
        r_sq = dot(r, r)
        i += 1
    return x, float(np.sqrt(F32(r_sq / rhs_sq))), i, restarts


def quadratic_placement(d, max_iterations=20, solver_iterations=100, target_error=1e-6, start_at_center=True):
    """quadraticPlacement: node centres after QP and the per-iteration log
    (HPWL, X error, Y error, BiCGSTAB iterations)."""
    n = d.num_nodes
    llx, lly, urx, ury = core_region(d)
    if start_at_center:
        ccx, ccy = core_center(d)
        cx = clamp(np.full(n, ccx, F32), d.node_w, llx, urx)
        cy = clamp(np.full(n, ccy, F32), d.node_h, lly, ury)
    else:
        cx, cy = np.asarray(d.node_x, F32).copy(), np.asarray(d.node_y, F32).copy()
    log = []
    it = 0
    while True:
        sysm = assemble(d, cx, cy)
        res = []
        for key, guess in (("x", cx), ("y", cy)):
            s = sysm[key]
            A = csr(n, s["row"], s["col"], s["val"], s["diag"])
            res.append(bicgstab(A, s["b"], guess, solver_iterations))
        cx = clamp(res[0][0], d.node_w, llx, urx)
        cy = clamp(res[1][0], d.node_h, lly, ury)
        px, py = pin_positions(d, cx, cy)
        _, hp = bound_pins(d, px, py)
        log.append((hp, res[0][1], res[1][1], res[0][2] + res[1][2]))
        if abs(res[0][1]) < target_error and abs(res[1][1]) < target_error and it > 4:
            break
        if it >= max_iterations:
            break
        it += 1
    return cx, cy, np.array(log)
