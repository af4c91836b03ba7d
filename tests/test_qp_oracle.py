"""The quadratic-placement oracle (oracle/qp.py) pinned against the
reference's own QPPlacer output (tests/golden/qp_*.npz, written by
tests/golden/gen_golden.py from the reference's qplace.cpp built with its
vendored Eigen) and against an independent loop restatement of
createSparseMatrix's pair conditions on the smallest designs."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from golden_utils import Golden
from oracle import qp as Q

CASES = ["qp_smoke", "qp_synth2k", "qp_mixed3k"]


def ref_matrix(g, d, key):
    n = d.num_nodes
    tag = "qp_a" + key
    return Q.csr(n, g[tag + "_row"], g[tag + "_col"], g[tag + "_val"])


def oracle_matrix(m, n, key):
    s = m[key]
    return Q.csr(n, s["row"], s["col"], s["val"], s["diag"])


def loop_restatement(d, cx, cy, y_symmetric=False):
    """qplace.cpp:145-272 written as the reference's own loops: pairs j < k
    in net order, the X condition on all four bound-pin tests, the Y
    condition exactly as written (pin2 == ymin || pin1 == ymax ||
    pin2 == ymax), or with y_symmetric the X-style Y condition.  Dense
    float64 A and b per direction."""
    n = d.num_nodes
    px, py = Q.pin_positions(d, cx, cy)
    B, _ = Q.bound_pins(d, px, py)
    A = {0: np.zeros((n, n)), 1: np.zeros((n, n))}
    b = {0: np.zeros(n), 1: np.zeros(n)}
    for e in range(d.num_nets):
        s, t = int(d.net_start[e]), int(d.net_start[e + 1])
        p = t - s
        if p < 2:
            continue
        c1 = np.float32(1.0 / (float(np.float32(p)) - 1.0))
        xmin, xmax, ymin, ymax = B[e]
        for j in range(s, t):
            for k in range(j + 1, t):
                m1, m2 = int(d.pin_obj[j]), int(d.pin_obj[k])
                if m1 == m2:
                    continue
                for dim, pos, off, ok in (
                        (0, px, d.pin_dx, j == xmin or j == xmax or k == xmin or k == xmax),
                        (1, py, d.pin_dy, (k == ymin or j == ymax or k == ymax) if not y_symmetric
                         else (j == ymin or j == ymax or k == ymin or k == ymax))):
                    if not ok:
                        continue
                    dist = np.float32(abs(np.float32(pos[j] - pos[k])))
                    w = float(c1 / dist) if dist >= 25.0 else float(np.float32(float(c1) / 25.0))
                    f1, f2 = m1 >= n, m2 >= n
                    if not f1 and not f2:
                        A[dim][m1, m1] += w
                        A[dim][m2, m2] += w
                        A[dim][m1, m2] -= w
                        A[dim][m2, m1] -= w
                        b[dim][m1] += -w * float(np.float32(off[j] - off[k]))
                        b[dim][m2] += -w * float(np.float32(off[k] - off[j]))
                    elif f1 and not f2:
                        A[dim][m2, m2] += w
                        b[dim][m2] += float(np.float32(np.float32(w) * np.float32(pos[j] - off[k])))
                    elif not f1 and f2:
                        A[dim][m1, m1] += w
                        b[dim][m1] += float(np.float32(np.float32(w) * np.float32(pos[k] - off[j])))
    return A, b


def relmax(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


@pytest.mark.parametrize("name", CASES)
def test_assembly_matches_reference(name):
    g = Golden(name)
    d = g.design()
    pos = g["qp_matrix_positions"]
    m = Q.assemble(d, pos[:, 0], pos[:, 1])
    assert abs(m["hpwl"] - g.s["qp_matrix_hpwl"]) <= 1e-9 * g.s["qp_matrix_hpwl"]
    for key in "xy":
        ref = ref_matrix(g, d, key).toarray()
        mine = oracle_matrix(m, d.num_nodes, key).toarray()
        # float sums in another order: a few float32 ulps of the largest entry
        assert relmax(mine, ref) < 2e-6, key
        assert relmax(m[key]["b"], g["qp_b" + key]) < 2e-6, key


@pytest.mark.parametrize("name", ["qp_smoke", "qp_synth2k"])
def test_assembly_matches_loop_restatement(name):
    g = Golden(name)
    d = g.design()
    pos = g["qp_matrix_positions"]
    m = Q.assemble(d, pos[:, 0], pos[:, 1])
    A, b = loop_restatement(d, pos[:, 0], pos[:, 1])
    for dim, key in ((0, "x"), (1, "y")):
        mine = oracle_matrix(m, d.num_nodes, key).toarray()
        assert relmax(mine, A[dim]) < 1e-6, key
        assert relmax(m[key]["b"], b[dim]) < 1e-6, key


def test_y_condition_quirk_is_kept():
    """The reference's Y condition drops pairs whose earlier pin is ymin
    (neither being ymax); the X-style condition would not match its
    matrix, the condition as written does."""
    g = Golden("qp_synth2k")
    d = g.design()
    pos = g["qp_matrix_positions"]
    ref = ref_matrix(g, d, "y").toarray()
    A, _ = loop_restatement(d, pos[:, 0], pos[:, 1])
    S, _ = loop_restatement(d, pos[:, 0], pos[:, 1], y_symmetric=True)
    assert relmax(A[1], ref) < 1e-6
    assert relmax(S[1], ref) > 1e-2


@pytest.mark.parametrize("name", CASES)
def test_matrix_properties(name):
    """Symmetric, and each row sums to the weights of its pairs with fixed
    pins (movable-movable pairs cancel), which are >= 0."""
    g = Golden(name)
    d = g.design()
    n = d.num_nodes
    for key in "xy":
        A = ref_matrix(g, d, key)
        assert abs(A - A.T).max() == 0
        rs = np.asarray(A.sum(axis=1)).ravel()
        assert rs.min() > -1e-4 * np.abs(A.diagonal()).max()


@pytest.mark.parametrize("name", CASES)
def test_bicgstab_solves_reference_system(name):
    """The restated BiCGSTAB on the reference's X system from its guess:
    its reported error is the true relative residual, and on these small
    systems it reaches the direct solution."""
    g = Golden(name)
    d = g.design()
    for key in "xy":
        A = ref_matrix(g, d, key)
        b = g["qp_b" + key]
        x, err, its, _ = Q.bicgstab(A, b, g["qp_guess_" + key], 100)
        true = np.linalg.norm(A @ x.astype(np.float64) - b) / np.linalg.norm(b)
        assert err < 1e-5
        assert true < 1e-4
        xs = spla.spsolve(A.tocsc(), b.astype(np.float64))
        # disconnected cells make A singular: compare where the direct solve is finite
        ok = np.isfinite(xs)
        if ok.all():
            assert relmax(x, xs) < 1e-3


def test_bicgstab_zero_rhs():
    """Eigen returns x = 0 and keeps the tolerance as the error when b = 0."""
    import scipy.sparse as sp
    A = sp.csr_matrix(np.array([[2.0, -1.0], [-1.0, 2.0]]))
    x, err, its, _ = Q.bicgstab(A, np.zeros(2, np.float32), np.ones(2, np.float32))
    assert np.all(x == 0) and its == 0 and err == pytest.approx(float(Q.FLOAT_EPS))


@pytest.mark.parametrize("name", CASES)
def test_first_iteration_matches_reference(name):
    g = Golden(name)
    d = g.design()
    cx, cy, log = Q.quadratic_placement(d, max_iterations=0)
    ref = g["qp_positions_it0"]
    # 1e-4 of the core: float rounding of a solve whose Y system is singular
    # where cells have no Y pair to a fixed pin
    llx, lly, urx, ury = d.core_region()
    tol = 1e-4 * float(max(urx - llx, ury - lly))
    assert np.abs(cx - ref[:, 0]).max() < tol
    assert np.abs(cy - ref[:, 1]).max() < tol
    assert abs(log[0, 0] - g["qp_log"][0, 2]) <= 1e-5 * g["qp_log"][0, 2]


@pytest.mark.parametrize("name", CASES)
def test_full_qp_matches_reference(name):
    """Later iterations re-pick bound pins from positions that agree to
    float rounding; near-ties then pick other pins and cells that the Y
    pairs leave without a fixed anchor keep solver-dependent positions, so
    the whole run is compared by its iteration count and HPWL."""
    g = Golden(name)
    d = g.design()
    cx, cy, log = Q.quadratic_placement(d)
    ref_log = g["qp_log"]
    assert len(log) == len(ref_log)
    assert abs(log[-1, 0] / g.s["qp_hpwl"] - 1) < 5e-3
    assert np.max(np.abs(log[:, 0] / ref_log[:, 2] - 1)) < 5e-3
    llx, lly, urx, ury = d.core_region()
    assert np.all(cx - d.node_w / 2 >= llx - 1e-3) and np.all(cx + d.node_w / 2 <= urx + 1e-3)
    assert np.all(cy - d.node_h / 2 >= lly - 1e-3) and np.all(cy + d.node_h / 2 <= ury + 1e-3)
