"""The device quadratic placer (ep_qplace / ep_qp_matrix / ep_qp_solve)
against the reference's QPPlacer output (tests/golden/qp_*.npz) and the
oracle (oracle/qp.py)."""
import ctypes as C

import numpy as np
import pytest

from golden_utils import Golden
from oracle import qp as Q
from eplace_amd import _lib as L
from eplace_amd import qplace as QP
from eplace_amd.design import synthetic

CASES = ["qp_smoke", "qp_synth2k", "qp_mixed3k"]


def relmax(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def lib_dense(m, n, dim):
    """One direction of ep_qp_matrix's output as a float64 CSR."""
    rp = m["rowptr"]
    r0, r1 = rp[dim * n], rp[dim * n + n]
    rows = np.repeat(np.arange(n), np.diff(rp[dim * n: dim * n + n + 1]))
    return Q.csr(n, rows, m["col"][r0:r1], m["val"][r0:r1], m["diag"][dim * n: dim * n + n])


def core_tol(d, frac):
    llx, lly, urx, ury = d.core_region()
    return frac * float(max(urx - llx, ury - lly))


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_gpu_matrix_matches_reference(name):
    g = Golden(name)
    d = g.design()
    n = d.num_nodes
    print(f"ep_qp_matrix on {name} ({n} nodes)", flush=True)
    pos = g["qp_matrix_positions"]
    m = QP.qp_matrix(d, pos)
    assert abs(m["hpwl"] - g.s["qp_matrix_hpwl"]) <= 1e-9 * g.s["qp_matrix_hpwl"]
    orc = Q.assemble(d, pos[:, 0], pos[:, 1])
    for dim, key in ((0, "x"), (1, "y")):
        ref = Q.csr(n, g[f"qp_a{key}_row"], g[f"qp_a{key}_col"], g[f"qp_a{key}_val"]).toarray()
        mine = lib_dense(m, n, dim).toarray()
        assert relmax(mine, ref) < 2e-6, key
        assert relmax(m["b"][dim * n: dim * n + n], g["qp_b" + key]) < 2e-6, key
        # same terms as the oracle, both summed in double: equal to a float ulp
        assert relmax(m["diag"][dim * n: dim * n + n], orc[key]["diag"]) < 1e-7
        assert relmax(m["b"][dim * n: dim * n + n], orc[key]["b"]) < 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_gpu_solve_reference_system(name):
    """BiCGSTAB on the reference's own X and Y systems (one call, ndim 2)."""
    g = Golden(name)
    d = g.design()
    n = d.num_nodes
    print(f"ep_qp_solve on {name}", flush=True)
    mats = [Q.csr(n, g[f"qp_a{k}_row"], g[f"qp_a{k}_col"], g[f"qp_a{k}_val"]) for k in "xy"]
    rowptr = np.concatenate([mats[0].indptr, mats[1].indptr[1:] + mats[0].nnz]).astype(np.int32)
    col = np.concatenate([m.indices for m in mats]).astype(np.int32)
    val = np.concatenate([m.data for m in mats]).astype(np.float32)
    b = np.concatenate([g["qp_bx"], g["qp_by"]])
    x0 = np.concatenate([g["qp_guess_x"], g["qp_guess_y"]])
    x, err, its = QP.qp_solve(rowptr, col, val, b, x0, ndim=2, max_iterations=100)
    for dim in range(2):
        xd = x[dim * n: dim * n + n]
        bd = b[dim * n: dim * n + n]
        true = np.linalg.norm(mats[dim] @ xd.astype(np.float64) - bd) / np.linalg.norm(bd)
        assert err[dim] < 1e-5 and true < 1e-4
        xo, eo, io, _ = Q.bicgstab(mats[dim], bd, x0[dim * n: dim * n + n], 100)
        assert abs(int(its[dim]) - io) <= max(3, io // 5)
    # X systems are anchored by IO pins: a unique solution, reached by both
    xo = Q.bicgstab(mats[0], g["qp_bx"], g["qp_guess_x"], 100)[0]
    assert np.abs(x[:n] - xo).max() < core_tol(d, 1e-5)


@pytest.mark.gpu
def test_gpu_solve_zero_rhs_and_dims():
    """b = 0 gives x = 0 with the tolerance as the error (Eigen's early
    return); a diagonal passed apart is added to the CSR entries."""
    rowptr = np.array([0, 1, 2], np.int32)
    col = np.array([1, 0], np.int32)
    val = np.array([-1.0, -1.0], np.float32)
    diag = np.array([2.0, 2.0], np.float32)
    x, err, its = QP.qp_solve(rowptr, col, val, np.zeros(2, np.float32), np.ones(2, np.float32), diag=diag)
    assert np.all(x == 0) and its[0] == 0 and err[0] == pytest.approx(QP.FLOAT_EPS)
    x, err, its = QP.qp_solve(rowptr, col, val, np.array([1.0, 0.0], np.float32), np.zeros(2, np.float32),
                              diag=diag)
    np.testing.assert_allclose(x, [2 / 3, 1 / 3], rtol=1e-6)


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_gpu_qplace_first_iteration(name):
    g = Golden(name)
    d = g.design()
    print(f"ep_qplace (1 iteration) on {name}", flush=True)
    st = QP.quadratic_placement(d, max_iterations=0)
    ref = g["qp_positions_it0"]
    tol = core_tol(d, 1e-4)
    assert np.abs(d.node_x - ref[:, 0]).max() < tol
    assert np.abs(d.node_y - ref[:, 1]).max() < tol
    assert st["iterations"] == 1
    assert abs(st["log"][0, 0] / g["qp_log"][0, 2] - 1) < 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_gpu_qplace_full(name):
    """The reference's default QP (-IPiteCount 20 from the core centre):
    same iteration count, per-iteration and final HPWL within 0.5%."""
    g = Golden(name)
    d = g.design()
    print(f"ep_qplace on {name}", flush=True)
    st = QP.quadratic_placement(d)
    ref_log = g["qp_log"]
    assert st["iterations"] == len(ref_log)
    assert np.max(np.abs(st["log"][:, 0] / ref_log[:, 2] - 1)) < 5e-3
    assert abs(st["final_hpwl"] / g.s["qp_hpwl"] - 1) < 5e-3
    llx, lly, urx, ury = d.core_region()
    assert np.all(d.node_x - d.node_w / 2 >= llx - 1e-3) and np.all(d.node_x + d.node_w / 2 <= urx + 1e-3)


@pytest.mark.gpu
def test_gpu_qplace_deterministic_and_matches_oracle_50k():
    """50K cells, 3 outer iterations: two device runs are identical, and the
    HPWL trajectory follows the oracle's (unconverged 100-iteration solves
    from identical systems, float rounding apart)."""
    d = synthetic(50_000, 52_500, seed=41, num_io=500, util=0.7, name="qp50k")
    e = synthetic(50_000, 52_500, seed=41, num_io=500, util=0.7, name="qp50k")
    print("ep_qplace on qp50k (3 iterations, twice)", flush=True)
    s1 = QP.quadratic_placement(d, max_iterations=2)
    s2 = QP.quadratic_placement(e, max_iterations=2)
    assert np.array_equal(d.node_x, e.node_x) and np.array_equal(d.node_y, e.node_y)
    assert np.array_equal(s1["log"], s2["log"])
    f = synthetic(50_000, 52_500, seed=41, num_io=500, util=0.7, name="qp50k")
    cx, cy, log = Q.quadratic_placement(f, max_iterations=2)
    assert len(log) == len(s1["log"]) == 3
    np.testing.assert_allclose(s1["log"][:, 0], log[:, 0], rtol=1e-2)


@pytest.mark.gpu
def test_gpu_qplace_large_net_and_clamp():
    """A design with a 300-pin net (its bound pins found by a block each)
    and movable macros: the matrix equals the oracle's, the first
    iteration's HPWL follows it and every node ends inside the core."""
    d = synthetic(8000, 8400, seed=43, num_io=64, num_macros=6, util=0.6, name="qp_big")
    # add one 300-pin net over the first 300 cells
    extra = np.arange(300, dtype=np.int32)
    d.pin_obj = np.concatenate([d.pin_obj, extra]).astype(np.int32)
    d.pin_dx = np.concatenate([d.pin_dx, np.zeros(300, np.float32)])
    d.pin_dy = np.concatenate([d.pin_dy, np.zeros(300, np.float32)])
    d.net_start = np.concatenate([d.net_start, [d.net_start[-1] + 300]]).astype(np.int64)
    e = synthetic(8000, 8400, seed=43, num_io=64, num_macros=6, util=0.6, name="qp_big")
    for k in ("pin_obj", "pin_dx", "pin_dy", "net_start"):
        setattr(e, k, getattr(d, k).copy())
    print("ep_qplace on qp_big (300-pin net)", flush=True)
    pos = np.stack([d.node_x, d.node_y], 1)
    m = QP.qp_matrix(d, pos)
    orc = Q.assemble(d, pos[:, 0], pos[:, 1])
    n = d.num_nodes
    for dim, key in ((0, "x"), (1, "y")):
        mine = lib_dense(m, n, dim)
        ref = Q.csr(n, orc[key]["row"], orc[key]["col"], orc[key]["val"], orc[key]["diag"])
        assert abs(mine - ref).max() <= 2e-6 * abs(ref).max()
    st = QP.quadratic_placement(d, max_iterations=0)
    cx, cy, log = Q.quadratic_placement(e, max_iterations=0)
    assert abs(st["log"][0, 0] / log[0, 0] - 1) < 1e-3
    llx, lly, urx, ury = d.core_region()
    assert np.all(d.node_y - d.node_h / 2 >= lly - 1e-3) and np.all(d.node_y + d.node_h / 2 <= ury + 1e-3)


def test_qp_rejects_bad_arguments():
    """Argument checks run before any device call."""
    lib = L.load()
    rowptr = np.array([0, 1, 5], np.int32)  # last entry != nnz
    col = np.array([1, 0], np.int32)
    val = np.ones(2, np.float32)
    b = np.ones(2, np.float32)
    x = np.zeros(2, np.float32)
    s = L.EPQpSystem(n=2, ndim=1, nnz=2, rowptr=rowptr.ctypes.data, col=col.ctypes.data, val=val.ctypes.data,
                     diag=None, b=b.ctypes.data)
    assert lib.ep_qp_solve(C.byref(s), x.ctypes.data, 100, 1e-7, 0, None, None) == L.EP_ERR_ARG
    assert b"nnz" in lib.ep_qp_last_error()
    rowptr[2] = 2
    col[0] = 7  # column out of range
    assert lib.ep_qp_solve(C.byref(s), x.ctypes.data, 100, 1e-7, 0, None, None) == L.EP_ERR_ARG
    assert b"column" in lib.ep_qp_last_error()
    p = L.EPQpParams(max_iterations=20, solver_iterations=100, target_error=1e-6, solver_tolerance=1e-7,
                     start_at_center=1, device=0)
    assert lib.ep_qplace(None, C.byref(p), x.ctypes.data, None, None) == L.EP_ERR_ARG
