"""Pins of the FP64 oracle by mathematics, independent of the reference's
outputs: closed-form Poisson modes, finite differences of the smooth
wirelength models, the gamma -> 0 limit, charge conservation, brute force.
"""
import numpy as np
import pytest

from eplace_amd import synthetic
from oracle.oracle import Oracle, glibc_rand, poisson_f64


@pytest.mark.parametrize("nx,ny,bx,by", [(16, 16, 1.0, 1.0), (32, 16, 1.5, 2.0), (64, 64, 0.7, 0.7)])
@pytest.mark.parametrize("u0,v0", [(1, 0), (0, 3), (2, 5), (7, 1)])
@pytest.mark.parametrize("method", ["fast", "direct"])
def test_poisson_single_mode(nx, ny, bx, by, u0, v0, method):
    """rho = cos(pi u0 (x+1/2)/Nx) cos(pi v0 (y+1/2)/Ny) is an eigenmode:
    phi = rho / (wx^2 + wy^2), Ex = wx phi_hat sin cos, Ey = wy phi_hat cos sin."""
    x = np.arange(nx)[:, None] + 0.5
    y = np.arange(ny)[None, :] + 0.5
    cx, sx = np.cos(np.pi * u0 * x / nx), np.sin(np.pi * u0 * x / nx)
    cy, sy = np.cos(np.pi * v0 * y / ny), np.sin(np.pi * v0 * y / ny)
    rho = cx * cy + 0.7  # plus a constant: only the discarded DC term moves
    wx = np.pi * u0 / nx
    wy = np.pi * v0 / ny * by / bx
    den = wx * wx + wy * wy
    phi, ex, ey = poisson_f64(rho, bx, by, method)
    np.testing.assert_allclose(phi, cx * cy / den, atol=1e-9 / den)
    np.testing.assert_allclose(ex, wx / den * sx * cy, atol=1e-9 / den)
    np.testing.assert_allclose(ey, wy / den * cx * sy, atol=1e-9 / den)


def test_poisson_linearity_and_fast_equals_direct():
    rng = np.random.default_rng(0)
    a, b = rng.random((32, 64)), rng.random((32, 64))
    pa, pb, pab = (poisson_f64(r, 1.2, 0.9) for r in (a, b, 2 * a - 3 * b))
    for k in range(3):
        np.testing.assert_allclose(pab[k], 2 * pa[k] - 3 * pb[k], atol=1e-10)
    d = poisson_f64(a, 1.2, 0.9, "direct")
    for k in range(3):
        np.testing.assert_allclose(pa[k], d[k], atol=1e-10)


def small_design(seed=0, macros=0):
    return synthetic(300, 400, seed=seed, degrees="c1", widths="uniform14", num_macros=macros, util=0.6)


def wl_value(o, xy):
    o.set("positions", xy)
    o.run("wirelength")
    return o.info()["wl_value"]


@pytest.mark.parametrize("wl,exp", [("WA", "exp"), ("WA", "fast"), ("LSE", "exp"), ("LSE", "fast")])
def test_wirelength_gradient_is_derivative(wl, exp):
    """Central differences of the FP64 WA/LSE value reproduce the per-node
    gradient (both exponent forms; the reference's (1+a/1024)^1024 is smooth
    for a > -300)."""
    d = small_design(1)
    o = Oracle(d, 1.0, mode="fp64", wl=wl, exp=exp)
    o.set_scalar("invgx", 0.35)
    o.set_scalar("invgy", 0.55)
    xy = o.get("positions")
    o.set("positions", xy)
    o.run("wirelength")
    g = o.get("grad_wirelength").reshape(-1, 2)
    rng = np.random.default_rng(2)
    for node in rng.choice(d.num_nodes, 12, replace=False):
        for ax in range(2):
            h = 1e-3
            p = xy.copy()
            p[2 * node + ax] += h
            fp = wl_value(o, p)
            p[2 * node + ax] -= 2 * h
            fm = wl_value(o, p)
            num = (fp - fm) / (2 * h)
            assert abs(num - g[node, ax]) <= 2e-3 * max(1.0, abs(g[node, ax])), (node, ax, num, g[node, ax])


def test_wa_tends_to_hpwl():
    """As 1/gamma grows the WA value tends to HPWL from below."""
    d = small_design(3)
    o = Oracle(d, 1.0, mode="fp64", exp="exp")
    prev = -1.0
    for ig in (0.5, 2.0, 8.0, 32.0, 128.0):
        o.set_scalar("invgx", ig)
        o.set_scalar("invgy", ig)
        o.run("wirelength")
        wl, hp = o.info()["wl_value"], o.info()["hpwl_bound"]
        assert wl <= hp * (1 + 1e-12)
        assert wl >= prev
        prev = wl
    assert abs(wl - hp) < 0.02 * hp
    assert abs(o.hpwl() - hp) < 1e-9 * hp


@pytest.mark.parametrize("macros", [0, 3])
def test_scatter_conserves_charge(macros):
    """Every object inside the core deposits its area (macros x td) exactly
    once, local smoothing included; fixed point and float agree."""
    d = small_design(4, macros)
    td = 0.9
    o = Oracle(d, td, mode="fp64")
    i = o.info()
    wh = o.get("wh").reshape(-1, 2)
    xy = o.get("positions").reshape(-1, 2)
    # keep inflated footprints inside the core so nothing spills
    lo = np.array([i["core_llx"], i["core_lly"]]) + np.maximum(wh, [i["bsx"], i["bsy"]]) / 2 + 1e-3
    hi = np.array([i["core_urx"], i["core_ury"]]) - np.maximum(wh, [i["bsx"], i["bsy"]]) / 2 - 1e-3
    o.set("positions", np.clip(xy, lo, hi).reshape(-1))
    o.run("scatter")
    macro = np.r_[d.node_is_macro(), np.zeros(i["F"], np.uint8)].astype(bool)
    filler = np.r_[np.zeros(d.num_nodes, bool), np.ones(i["F"], bool)]
    area = (wh[:, 0].astype(np.float32) * wh[:, 1].astype(np.float32)).astype(np.float64)
    want_node = (area * np.where(macro, td, 1.0))[~filler].sum()
    # bin and rectangle edges are float (as in the reference): 1e-6
    assert abs(o.get("node_map").sum() - want_node) < 1e-6 * want_node
    assert abs(o.get("filler_map").sum() - area[filler].sum()) < 1e-6 * area[filler].sum()
    shift = 40
    o.run("scatter_fixed", shift)
    q = o.get("node_q").sum() * 2.0 ** -shift
    assert abs(q - want_node) < 1e-5 * want_node


def test_scatter_brute_force_tiny():
    """Each bin's charge equals a brute-force overlap sum over objects."""
    d = synthetic(30, 20, seed=5, degrees="c1", widths="uniform14", util=0.5)
    o = Oracle(d, 1.0, mode="fp64", bins=8)
    o.run("scatter")
    i = o.info()
    xy = o.get("positions").reshape(-1, 2)
    wh = o.get("wh").reshape(-1, 2)
    bx, by = i["bsx"], i["bsy"]
    node = np.zeros((8, 8))
    ex = np.array([i["core_llx"] + k * bx for k in range(9)])
    ey = np.array([i["core_lly"] + k * by for k in range(9)])
    for o_ in range(i["N"]):
        w, h = wh[o_]
        sx = w / bx if w < bx else 1.0
        sy = h / by if h < by else 1.0
        ew, eh = max(w, bx), max(h, by)
        l, r = xy[o_, 0] - ew / 2, xy[o_, 0] + ew / 2
        b, t = xy[o_, 1] - eh / 2, xy[o_, 1] + eh / 2
        for a in range(8):
            for c in range(8):
                ox = max(0.0, min(r, ex[a + 1]) - max(l, ex[a]))
                oy = max(0.0, min(t, ey[c + 1]) - max(b, ey[c]))
                node[a, c] += sx * sy * ox * oy
    np.testing.assert_allclose(o.map("node_map"), node, rtol=1e-5, atol=1e-6)


def test_glibc_rand_matches_libc():
    import ctypes
    libc = ctypes.CDLL("libc.so.6")
    for seed in (1, 7, 12345):
        libc.srand(seed)
        want = [libc.rand() for _ in range(2000)]
        np.testing.assert_array_equal(glibc_rand(2000, seed), want)


def test_gamma_schedule_closed_form():
    d = small_design(6)
    o = Oracle(d, 1.0, mode="fp64")
    bsx = o.info()["bsx"]
    for tau, mult in ((1.5, 0.1), (0.05, 10.0), (0.1, 10 ** 1.0), (0.55, 10 ** -((0.55 - 0.1) * 20 / 9 - 1))):
        o.set_scalar("overflow", tau)
        o.run("gamma")
        np.testing.assert_allclose(o.info()["invgx"], 0.125 / bsx * mult, rtol=1e-12)


def test_penalty_update_closed_form():
    d = small_design(7)
    o = Oracle(d, 1.0, mode="fp64")
    hp = o.hpwl()
    for last, want in ((hp + 10.0, 1.05), (hp - 350000.0 / 2, 1.05 ** 0.5), (hp - 1e7, 0.95), (hp, 1.05)):
        o.set_scalar("lambda", 2.0)
        o.set_scalar("last_hpwl", last)
        o.run("update_penalty")
        np.testing.assert_allclose(o.info()["lambda"], 2.0 * want, rtol=1e-12)
