"""ep_qp_solve_noalloc, the allocation-free variant of ep_qp_solve, through
the C ABI: on the same systems its solutions, errors and iteration counts
equal ep_qp_solve's bit for bit, over two calls in a row (the second on a
smaller system, in the work space the first one used), and it rejects a
system larger than its work space or with a column index out of range
before solving.  Device buffers and the stream come from the HIP runtime the
library itself links (libamdhip64.so.7, through ctypes)."""
import ctypes as C
import sys

import numpy as np
import pytest

import eplace_amd._lib as L
import eplace_amd.qplace as QP
from eplace_amd import synthetic

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def system(cells, seed):
    d = synthetic(cells, int(cells * 1.05), seed=seed, name=f"noalloc{cells}")
    rng = np.random.default_rng(seed)
    xy = np.stack([d.node_x, d.node_y], 1).astype(np.float32)
    xy += rng.normal(0, 2.0, xy.shape).astype(np.float32)
    m = QP.qp_matrix(d, xy)
    x0 = np.concatenate([xy[:, 0], xy[:, 1]]).astype(np.float32)
    return d.num_nodes, m, x0


class Hip:
    """The few HIP runtime calls the test needs, from the runtime that
    libeplace_amd.so is linked against (loaded first, so dlopen returns it)."""

    def __init__(self):
        L.load()
        self.h = C.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [C.c_void_p, C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipStreamCreate.argtypes = [C.c_void_p]
        self.h.hipStreamDestroy.argtypes = [C.c_void_p]
        self.h.hipDeviceSynchronize.argtypes = []
        self.owned = []

    def check(self, rc, what):
        assert rc == 0, f"{what} failed with HIP error {rc}"

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        self.check(self.h.hipMalloc(C.byref(p), max(a.nbytes, 4)), "hipMalloc")
        self.owned.append(p.value)
        self.check(self.h.hipMemcpy(p.value, a.ctypes.data, a.nbytes, 1), "hipMemcpy H2D")
        return p.value

    def download(self, p, like):
        out = np.empty_like(like)
        self.check(self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, 2), "hipMemcpy D2H")
        return out

    def stream(self):
        s = C.c_void_p()
        self.check(self.h.hipStreamCreate(C.byref(s)), "hipStreamCreate")
        return s.value

    def close(self, streams=()):
        self.check(self.h.hipDeviceSynchronize(), "hipDeviceSynchronize")
        for s in streams:
            self.h.hipStreamDestroy(s)
        for p in self.owned:
            self.h.hipFree(p)
        self.owned = []


def device_solve(hip, n, m, x0, stream=0, bad_col=False):
    col = np.array(m["col"], np.int32)
    if bad_col:
        col[len(col) // 2] = n
    d = {k: hip.upload(np.asarray(m[k])) for k in ("rowptr", "val", "diag", "b")}
    d["col"] = hip.upload(col)
    x = hip.upload(x0.astype(np.float32))
    err, its = QP.qp_solve_noalloc(n, 2, len(col), d["rowptr"], d["col"], d["val"], d["b"], x, diag=d["diag"],
                                   max_iterations=100, stream=stream)
    return hip.download(x, x0.astype(np.float32)), err, its


def test_noalloc_solve_equals_allocating_solve_twice():
    hip = Hip()
    st = hip.stream()
    for cells, seed in ((20000, 11), (3000, 12)):
        say(f"[noalloc] {cells} cells seed {seed}")
        n, m, x0 = system(cells, seed)
        xr, er, ir = QP.qp_solve(m["rowptr"], m["col"], m["val"], m["b"], x0, diag=m["diag"], ndim=2,
                                 max_iterations=100)
        xa, ea, ia = device_solve(hip, n, m, x0, stream=st)
        say(f"  allocating: iterations {ir.tolist()} error {er.tolist()}; no-alloc: {ia.tolist()} {ea.tolist()}")
        np.testing.assert_array_equal(ia, ir)
        np.testing.assert_array_equal(ea, er)
        np.testing.assert_array_equal(xa, xr)
        assert ir.min() > 0 and np.all(np.isfinite(xa))
    hip.close([st])


def test_noalloc_rejects_oversize_and_bad_columns():
    hip = Hip()
    lib = L.load()
    say("[noalloc] rejections")
    # too many rows: rejected on the host before any device work
    dummy = hip.upload(np.zeros(16, np.int32))
    n = L.EP_QP_NOALLOC_MAX_ROWS // 2 + 1
    with pytest.raises(L.EPError, match="exceeds the allocation-free work space"):
        QP.qp_solve_noalloc(n, 2, 0, dummy, 0, 0, dummy, dummy)
    # a column index out of range: found by the checking kernel
    n, m, x0 = system(2000, 13)
    with pytest.raises(L.EPError, match="column index out of range"):
        device_solve(hip, n, m, x0, bad_col=True)
    # the work space still serves a good call afterwards (null stream)
    xr, er, ir = QP.qp_solve(m["rowptr"], m["col"], m["val"], m["b"], x0, diag=m["diag"], ndim=2,
                             max_iterations=100)
    xa, ea, ia = device_solve(hip, n, m, x0)
    np.testing.assert_array_equal(xa, xr)
    assert lib.ep_qp_last_error().decode() == ""
    hip.close()
