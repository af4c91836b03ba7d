"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Bars: the fixed-point density maps are bit-exact; fields, potential and
gradients within 1e-5 relative (norm-wise) of the FP64 oracle evaluated on
the same inputs.
"""
import sys

import numpy as np
import pytest

from eplace_amd import EPlacer2D, EplaceNesterovOpt, synthetic
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    n = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (n if n > 0 else 1.0))


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


CASES = {
    "c1_16": dict(kw=dict(num_cells=100, num_nets=184, seed=1, degrees="c1", widths="uniform14"), td=0.8, bins=16),
    "c2_small_64": dict(kw=dict(num_cells=3000, num_nets=3200, seed=2), td=1.0, bins=64),
    "macro_64": dict(kw=dict(num_cells=3000, num_nets=3300, seed=4, num_macros=6, util=0.75), td=1.0, bins=64),
    "rect_128x64": dict(kw=dict(num_cells=5000, num_nets=5200, seed=5), td=0.9, bins=(128, 64)),
}


def make(case, exp="reference", wl="WA", mode="fp64"):
    c = CASES[case]
    d = synthetic(**c["kw"], name=case)
    gp = EPlacer2D(d, c["td"], bins=c["bins"], exp=exp, wirelength=wl)
    orc = Oracle(d, c["td"], mode=mode, bins=c["bins"], exp="fast" if exp == "reference" else "exp", wl=wl)
    # The device state is centres; lower-left corners are derived as
    # c - w/2.  The oracle keeps the corner a filler was placed at
    # (Module::coor), which can differ by one ulp; re-derive it so both
    # sides scatter identical rectangles.
    orc.set("positions", orc.get("positions"))
    return d, gp, orc


@pytest.mark.parametrize("case", list(CASES))
def test_density_fixed_point_bit_exact(case):
    say(f"[density] case={case}")
    d, gp, orc = make(case)
    np.testing.assert_array_equal(gp.read("positions").reshape(-1), orc.get("positions").astype(np.float32))
    np.testing.assert_array_equal(gp.read("term_map").reshape(-1), orc.get("term_map").astype(np.float32))
    np.testing.assert_array_equal(gp.read("base_map").reshape(-1), orc.get("base_map").astype(np.float32))
    gp.totalGradientUpdate()
    shift = gp.info()["fixed_shift"]
    orc.run("scatter_fixed", shift)
    np.testing.assert_array_equal(gp.read("node_q").reshape(-1), orc.get("node_q"))
    np.testing.assert_array_equal(gp.read("filler_q").reshape(-1), orc.get("filler_q"))
    orc.run("overflow")
    st = gp.stats()
    # per-bin terms are float on the GPU, double in the oracle; summed in FP64
    assert abs(st["overflow"] - orc.info()["overflow"]) <= 1e-6 * max(1.0, abs(orc.info()["overflow"]))
    orc.run("rho")
    assert rel(gp.read("rho"), orc.map("rho")) < 1e-6


@pytest.mark.parametrize("case", ["c2_small_64", "macro_64", "rect_128x64"])
def test_scatter_and_gather_after_large_moves(case):
    """A second evaluation after every object is mirrored across the core
    (all of them change tile): the density maps are bit-exact and the field
    gather matches the oracle."""
    say(f"[resort] case={case}")
    d, gp, orc = make(case)
    gp.totalGradientUpdate()  # sorts
    pos = gp.read("positions").copy()
    llx, lly, urx, ury = d.core_region()
    pos[:, 0] = np.float32(llx + urx) - pos[:, 0]
    pos[:, 1] = np.float32(lly + ury) - pos[:, 1]
    gp.write("positions", pos)
    moved = gp.read("positions").copy()
    orc.set("positions", moved.reshape(-1).astype(np.float64))
    gp.totalGradientUpdate()  # reuses the order
    shift = gp.info()["fixed_shift"]
    orc.run("scatter_fixed", shift)
    np.testing.assert_array_equal(gp.read("node_q").reshape(-1), orc.get("node_q"))
    np.testing.assert_array_equal(gp.read("filler_q").reshape(-1), orc.get("filler_q"))
    orc.set("ex", gp.read("ex").reshape(-1).astype(np.float64))
    orc.set("ey", gp.read("ey").reshape(-1).astype(np.float64))
    orc.run("gather")
    e = rel(gp.read("grad_density"), orc.get("grad_density").reshape(-1, 2))
    say(f"  grad_density rel err {e:.2e}")
    assert e < 1e-5


@pytest.mark.parametrize("method", ["bktrk", "vanilla"])
def test_density_exact_after_a_run(method):
    """After a Nesterov run (whose backtracking trials count the scatter's
    keys inside their moves) a plain evaluation at the final positions still
    gives the integer oracle's maps bit for bit: the sort's key totals were
    left at zero."""
    say(f"[density] after a {method} run")
    d, gp, orc = make("c2_small_64")
    gp.initialization()
    EplaceNesterovOpt(gp, method=method, max_iterations=6, log=False).opt()
    gp.totalGradientUpdate()
    orc.set("positions", gp.read("positions").reshape(-1).astype(np.float64))
    orc.run("scatter_fixed", gp.info()["fixed_shift"])
    np.testing.assert_array_equal(gp.read("node_q").reshape(-1), orc.get("node_q"))
    np.testing.assert_array_equal(gp.read("filler_q").reshape(-1), orc.get("filler_q"))
    gp.close()


POISSON_GRIDS = [(16, 16), (10, 6), (64, 64), (128, 64), (96, 160), (256, 256), (512, 512), (1024, 1024),
                 (2048, 2048)]
# poisson_mode: 0 FP32 MFMA, 1 split-BF16 (2 planes, 3 products), 2 split-FP16
# (power-of-two operand scales, 3 products), 3 split-BF16 (3 planes, 6 products)
POISSON_MODE_NAMES = {0: "fp32", 1: "split-bf16", 2: "split-fp16", 3: "split-bf16x3"}


def density_sample(n, kind, seed=7):
    """A charge map of shape n: uniform noise (flat spectrum), a clustered map
    as in the first iterations (most charge in a few bins), or a nearly even
    map as near convergence."""
    rng = np.random.default_rng(seed)
    if kind == "uniform":
        return rng.random(n).astype(np.float32) * 2.0
    nx, ny = n
    x, y = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    if kind == "clustered":
        r2 = ((x - 0.45 * nx) / (0.05 * nx + 1)) ** 2 + ((y - 0.55 * ny) / (0.04 * ny + 1)) ** 2
        return (400.0 * np.exp(-r2) + 0.02 * rng.random(n)).astype(np.float32)
    if kind == "even":
        return np.clip(1.0 + 0.05 * rng.standard_normal(n), 0, None).astype(np.float32)
    raise ValueError(kind)


_POISSON_DESIGN = None


def poisson_errors(n, mode, kind="uniform"):
    global _POISSON_DESIGN
    if _POISSON_DESIGN is None:
        _POISSON_DESIGN = synthetic(200, 200, seed=9, name="p")
    gp = EPlacer2D(_POISSON_DESIGN, 1.0, bins=n, poisson_mode=mode)
    try:
        i = gp.info()
        rho = density_sample(n, kind)
        phi, ex, ey = gp.poisson(rho)
    finally:
        gp.close()
    from oracle.oracle import poisson_f64
    rphi, rex, rey = poisson_f64(rho.astype(np.float64), i["bin_step_x"], i["bin_step_y"], "fast")
    out = {nm: rel(a, b) for a, b, nm in ((ex, rex, "ex"), (ey, rey, "ey"), (phi, rphi, "phi"))}
    say(f"  {POISSON_MODE_NAMES[mode]} grid {n} {kind}: " + ", ".join(f"{k} {v:.2e}" for k, v in out.items()))
    return out


@pytest.mark.parametrize("n", POISSON_GRIDS)
@pytest.mark.parametrize("mode", [0, 2, 3])
def test_poisson_vs_fp64(n, mode):
    """FP32 MFMA, split-FP16 and 3-plane split-BF16 transforms: field and
    potential within 1e-5 (relative, norm-wise) of FP64 on every grid."""
    say(f"[poisson {POISSON_MODE_NAMES[mode]}] grid={n}")
    for nm, e in poisson_errors(n, mode).items():
        assert e < 1e-5, (nm, e)


@pytest.mark.parametrize("n", [(64, 64), (512, 512), (1024, 1024), (2048, 2048)])
@pytest.mark.parametrize("kind", ["clustered", "even"])
def test_poisson_split_fp16_density_shapes(n, kind):
    """The split-FP16 operand scales follow the data: a map with most charge
    in a few bins (peak 400) and a nearly even one both stay within 1e-5."""
    say(f"[poisson split-fp16] grid={n} {kind}")
    for nm, e in poisson_errors(n, 2, kind).items():
        assert e < 1e-5, (nm, e)


def test_poisson_split_fp16_extreme_scales():
    """Operand scales are powers of two taken from the data: a charge peak of
    1e6 does not overflow fp16, a map scaled by 1e-6 keeps its precision, and
    a zero map gives zero fields."""
    say("[poisson split-fp16] extreme scales")
    d = synthetic(200, 200, seed=9, name="p")
    gp = EPlacer2D(d, 1.0, bins=128, poisson_mode=2)
    try:
        i = gp.info()
        from oracle.oracle import poisson_f64
        base = density_sample((128, 128), "clustered")
        for scale in (1e6 / 400.0, 1e-6):
            rho = (base * np.float32(scale)).astype(np.float32)
            phi, ex, ey = gp.poisson(rho)
            rphi, rex, rey = poisson_f64(rho.astype(np.float64), i["bin_step_x"], i["bin_step_y"], "fast")
            for a, b, nm in ((ex, rex, "ex"), (ey, rey, "ey"), (phi, rphi, "phi")):
                e = rel(a, b)
                say(f"  scale {scale:.3g} {nm} {e:.2e}")
                assert np.all(np.isfinite(a)) and e < 1e-5, (scale, nm, e)
        phi, ex, ey = gp.poisson(np.zeros((128, 128), np.float32))
        assert not np.any(ex) and not np.any(ey) and not np.any(phi)
    finally:
        gp.close()


@pytest.mark.parametrize("n", POISSON_GRIDS)
def test_poisson_split_bf16_vs_fp64(n):
    """Two-plane split-BF16 (16 significant bits per operand): within 3e-5 of
    FP64.  Kept as a measured mode; the placement runs use split-FP16."""
    say(f"[poisson split-bf16] grid={n}")
    for nm, e in poisson_errors(n, 1).items():
        assert e < 3e-5, (nm, e)


def test_split_placement_matches_fp32():
    """Full runs with split-FP16 and split-BF16 transforms land within 0.5%
    HPWL of FP32."""
    say("[split run] c2_small_64")
    c = CASES["c2_small_64"]
    d = synthetic(**c["kw"], name="c2")
    res = []
    for mode in (0, 2, 1):
        gp = EPlacer2D(d, c["td"], bins=c["bins"], poisson_mode=mode)
        gp.initialization()
        r = EplaceNesterovOpt(gp).opt()
        say(f"  mode {mode}: {r['iterations']} iterations, hpwl {r['final_hpwl']:.3f}, tau {r['final_overflow']:.4f}")
        assert r["final_overflow"] < 0.1
        res.append(r)
    for r in res[1:]:
        assert abs(r["final_hpwl"] - res[0]["final_hpwl"]) <= 0.005 * res[0]["final_hpwl"]


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("exp", ["reference", "exp"])
def test_wirelength_gradient(case, exp):
    say(f"[wirelength] case={case} exp={exp}")
    d, gp, orc = make(case, exp=exp)
    gp.totalGradientUpdate()
    st = gp.stats()
    orc.set_scalar("invgx", st["inv_gamma_x"])
    orc.set_scalar("invgy", st["inv_gamma_y"])
    orc.run("wirelength")
    gw = gp.read("grad_wirelength")
    e = rel(gw, orc.get("grad_wirelength").reshape(-1, 2))
    say(f"  grad_wl rel err {e:.2e}")
    assert e < 1e-5
    assert abs(st["hpwl"] - orc.info()["hpwl_bound"]) <= 1e-9 * orc.info()["hpwl_bound"]
    assert abs(st["wirelength"] - orc.info()["wl_value"]) <= 1e-5 * orc.info()["wl_value"]


@pytest.mark.parametrize("case", ["c2_small_64", "macro_64"])
def test_lse_gradient(case):
    say(f"[lse] case={case}")
    d, gp, orc = make(case, wl="LSE")
    gp.totalGradientUpdate()
    st = gp.stats()
    orc.set_scalar("invgx", st["inv_gamma_x"])
    orc.set_scalar("invgy", st["inv_gamma_y"])
    orc.run("wirelength")
    e = rel(gp.read("grad_wirelength"), orc.get("grad_wirelength").reshape(-1, 2))
    say(f"  grad_lse rel err {e:.2e}")
    assert e < 1e-5


@pytest.mark.parametrize("case", list(CASES))
def test_gather_and_combine(case):
    say(f"[gather] case={case}")
    d, gp, orc = make(case)
    gp.totalGradientUpdate()
    gp.penaltyFactorInitilization()
    gp.totalGradientUpdate()
    st = gp.stats()
    # same field and wirelength gradient on both sides, then gather + combine
    orc.set("ex", gp.read("ex").reshape(-1).astype(np.float64))
    orc.set("ey", gp.read("ey").reshape(-1).astype(np.float64))
    orc.run("gather")
    e = rel(gp.read("grad_density"), orc.get("grad_density").reshape(-1, 2))
    say(f"  grad_density rel err {e:.2e}")
    assert e < 1e-5
    orc.set("grad_wirelength", gp.read("grad_wirelength").reshape(-1).astype(np.float64))
    orc.set("grad_density", gp.read("grad_density").reshape(-1).astype(np.float64))
    orc.set_scalar("lambda", st["lambda"])
    orc.run("combine")
    e = rel(gp.read("grad_total"), orc.get("grad_total").reshape(-1, 2))
    say(f"  grad_total rel err {e:.2e}")
    assert e < 1e-6


@pytest.mark.parametrize("case", list(CASES))
def test_total_gradient_end_to_end(case):
    say(f"[total] case={case}")
    d, gp, orc = make(case)
    gp.totalGradientUpdate()
    shift = gp.info()["fixed_shift"]
    orc.run("total_gradient_update", 1, shift)
    for nm in ("ex", "ey"):
        e = rel(gp.read(nm), orc.map(nm))
        say(f"  {nm} rel err {e:.2e}")
        assert e < 1e-5
    e = rel(gp.read("grad_density"), orc.get("grad_density").reshape(-1, 2))
    say(f"  grad_density rel err {e:.2e}")
    assert e < 1e-5
    e = rel(gp.read("grad_wirelength"), orc.get("grad_wirelength").reshape(-1, 2))
    say(f"  grad_wl rel err {e:.2e}")
    assert e < 1e-5


def test_lambda0_and_penalty_update():
    say("[lambda]")
    d, gp, orc = make("c2_small_64")
    gp.initialization()
    orc.run("total_gradient_update", 1, gp.info()["fixed_shift"])
    orc.run("penalty_init")
    st = gp.stats()
    assert abs(st["lambda"] - orc.info()["lambda"]) < 1e-5 * orc.info()["lambda"]
    assert abs(st["last_hpwl"] - orc.info()["last_hpwl"]) < 1e-9 * orc.info()["last_hpwl"]
    # move everything a little and update lambda on both sides
    xy = gp.getPosition() + np.float32(0.37)
    gp.setPosition(xy)
    orc.set("stage_positions_clamped", xy.reshape(-1).astype(np.float64))
    np.testing.assert_array_equal(gp.read("positions").reshape(-1), orc.get("positions").astype(np.float32))
    gp.updatePenaltyFactor()
    orc.run("update_penalty")
    st = gp.stats()
    assert abs(st["lambda"] - orc.info()["lambda"]) < 1e-5 * orc.info()["lambda"]


# Final-HPWL bar per design: 0.5% (the target) where the design is large
# enough for the run-to-run spread of the algorithm itself to be well below
# it; on the 100-cell design the reference-float and FP64 oracle runs of the
# same algorithm already differ by 1.2%, so 2% there.
@pytest.mark.parametrize("case,tol", [("c2_small_64", 0.005), ("c1_16", 0.02)])
@pytest.mark.parametrize("method", ["vanilla", "bktrk", "bb"])
def test_nesterov_run_matches_oracle(method, case, tol):
    say(f"[nesterov] method={method} case={case}")
    d, gp, orc = make(case)
    gp.initialization()
    shift = gp.info()["fixed_shift"]
    orc.run("total_gradient_update", 1, shift)
    orc.run("penalty_init")
    orc.run("opt_init")
    k = 0
    while not orc.stop_condition(0.1):
        orc.run("opt_step", {"vanilla": 0, "bktrk": 1, "bb": 2}[method], shift)
        k += 1
    orc.run("wrap_up")
    r = EplaceNesterovOpt(gp, method=method).opt()
    say(f"  gpu iters {r['iterations']} hpwl {r['final_hpwl']:.3f}; oracle iters {k} hpwl {orc.hpwl():.3f}")
    assert abs(r["iterations"] - k) <= max(5, 0.05 * k)
    assert abs(r["final_hpwl"] - orc.hpwl()) <= tol * orc.hpwl()


def oracle_rows(orc, shift, method, K):
    rows = []
    for _ in range(K):
        orc.run("opt_step", {"vanilla": 0, "bktrk": 1, "bb": 2}[method], shift)
        i = orc.info()
        rows.append(dict(tau=i["overflow"], lam=i["lambda"], hpwl=i["last_hpwl"], step=i["last_step"],
                         bk=i["bktrk_count"]))
    return rows


# Per-row bars of the GPU against the FP64 oracle over the first K
# iterations (same fixed-point density on both sides; the field differs by
# the FP32 transform's ~1e-6, which the iteration amplifies).  Measured on
# MI355X: at most 2.3e-7 for overflow, lambda and HPWL and 9.1e-7 for the
# step on both designs and methods.
TRAJ_TOL = dict(tau=1e-6, lam=1e-6, hpwl=1e-6, step=5e-6)


@pytest.mark.parametrize("case", ["c1_16", "c2_small_64"])
@pytest.mark.parametrize("method", ["vanilla", "bktrk", "bb"])
def test_trajectory_first_iterations(case, method):
    """Row-by-row agreement with the oracle for the first 12 iterations
    (before the runs separate: the dynamics are chaotic): overflow, lambda,
    HPWL and the Lipschitz step to about six digits, and under backtracking
    the same rejected trials at the same iterations (nesterov.hpp:310-369,
    375-415) and the Barzilai-Borwein step (231-302).  Vanilla and BB rows
    log lambda and HPWL one iteration late (the penalty update at y_{k+1}
    runs inside the next evaluation), so those two are compared with the
    oracle's previous row."""
    say(f"[trajectory] case={case} method={method}")
    d, gp, orc = make(case)
    gp.initialization()
    shift = gp.info()["fixed_shift"]
    orc.run("total_gradient_update", 1, shift)
    orc.run("penalty_init")
    orc.run("opt_init")
    K = 12
    otraj = oracle_rows(orc, shift, method, K)
    opt = EplaceNesterovOpt(gp, method=method, max_iterations=K - 1)
    opt.opt()
    log = opt.log
    assert len(log) >= K
    worst = {k: 0.0 for k in TRAJ_TOL}
    for k in range(K):
        gl, o = log[k], dict(otraj[k])
        if method in ("vanilla", "bb"):
            if k == 0:
                continue
            o["lam"], o["hpwl"] = otraj[k - 1]["lam"], otraj[k - 1]["hpwl"]
        g = dict(tau=gl[1], lam=gl[2], hpwl=gl[3], step=gl[4])
        say(f"  k={k:2d} tau {g['tau']:.7f}/{o['tau']:.7f} step {g['step']:.6g}/{o['step']:.6g} "
            f"lam {g['lam']:.6g}/{o['lam']:.6g} hpwl {g['hpwl']:.7g}/{o['hpwl']:.7g} bk {int(gl[7])}/{o['bk']}")
        for nm in TRAJ_TOL:
            worst[nm] = max(worst[nm], abs(g[nm] - o[nm]) / abs(o[nm]))
        if method == "bktrk":
            assert int(gl[7]) == o["bk"], (k, int(gl[7]), o["bk"])
    say("  worst relative differences: " + ", ".join(f"{k} {v:.1e}" for k, v in worst.items()))
    for nm, tol in TRAJ_TOL.items():
        assert worst[nm] <= tol, (nm, worst[nm])


def test_hpwl_keeps_wirelength_value():
    """calcHPWL runs the wiring kernels without gradients; the WA value of the
    last evaluation stays readable afterwards."""
    say("[hpwl keeps wl value]")
    d, gp, orc = make("c2_small_64")
    gp.totalGradientUpdate()
    wl = gp.stats()["wirelength"]
    assert wl > 0
    h = gp.calcHPWL()
    assert h > 0 and gp.stats()["wirelength"] == wl
