"""The reference's main flow through the C++ adapter
eplace_amd/host/eplace_amd_adapter.h (main/ePlace_main.cpp:121-219).

oracle/adapter_main.cpp drives QP, mGP and the mixed-size stages through the
adapter from a PlaceDB that the reference's own Bookshelf parser filled;
oracle/Makefile.ref links it with the reference's unmodified Parser /
PlaceCommon / PlaceDB sources and with eplace_amd/libeplace_amd.so.  The CPU
test builds it (only where the reference tree is); the GPU tests run the
built binary on designs written from the committed fixtures and compare with
the reference's own runs of the same designs (final-HPWL bars as in
tests/test_gpu_golden.py).
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from golden_utils import Golden
from eplace_amd.design import write_bookshelf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REFERENCE = "/root/reference"
ADAPTER = os.path.join(ROOT, "oracle", "_ref", "adapter_main")
LIB = os.path.join(ROOT, "eplace_amd", "libeplace_amd.so")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_adapter_builds_against_reference_headers():
    if not os.path.isdir(REFERENCE):
        pytest.skip("the reference tree is not on this machine")
    if not os.path.exists(LIB):
        pytest.skip("libeplace_amd.so is not built")
    r = subprocess.run(["make", "-s", "-j8", "-f", "oracle/Makefile.ref", f"REF={REFERENCE}", "adapter"], cwd=ROOT,
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    assert os.access(ADAPTER, os.X_OK)
    dyn = subprocess.run(["readelf", "-d", ADAPTER], capture_output=True, text=True, check=True).stdout
    assert "libeplace_amd.so" in dyn and "$ORIGIN/../../eplace_amd" in dyn
    # argument errors come back as an exit code before any device call
    r = subprocess.run([ADAPTER], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "-aux is required" in r.stderr


def test_detailed_placement_through_adapter(tmp_path):
    """detailedPlacementAMD (ePlace_main.cpp:271-283) on a PlaceDB that the
    reference's parser filled and its AbacusLegalizer legalised: the
    library's result equals the reference DetailedPlacer's, run in the same
    process from the same legal positions, bit for bit.  No device work
    (-noGP), so this runs on the CPU."""
    if not os.access(ADAPTER, os.X_OK):
        pytest.skip("oracle/_ref/adapter_main is not built (it needs the reference tree at build time)")
    g = Golden("dp_synth_c2")
    d = g.design()
    d.node_x[:] = g["input_positions"][:, 0]
    d.node_y[:] = g["input_positions"][:, 1]
    aux = write_bookshelf(d, str(tmp_path / "in"))
    out = tmp_path / "out"
    r = subprocess.run([ADAPTER, "-aux", aux, "-noQP", "-noGP", "-internalDP", "-out", str(out)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-2000:])
    s = dict((k, float(v)) for k, v in (ln.split() for ln in (out / "scalars.txt").read_text().splitlines()))
    assert s["done"] == 1
    assert s["dp_max_diff_vs_ref"] == 0.0
    assert s["dp_hpwl"] == s["dp_ref_hpwl"]
    assert s["dp_lib_hpwl_before"] == pytest.approx(s["legal_hpwl"], rel=1e-12)
    assert s["dp_hpwl"] < s["legal_hpwl"]


def run_adapter(tmp_path, name, args, timeout=300):
    if not os.access(ADAPTER, os.X_OK):
        pytest.skip("oracle/_ref/adapter_main is not built (it needs the reference tree at build time)")
    g = Golden(name)
    d = g.design()
    aux = write_bookshelf(d, str(tmp_path / "in"))
    out = tmp_path / "out"
    cmd = [ADAPTER, "-aux", aux, "-targetDensity", repr(g.target_density), "-out", str(out)] + args
    say(f"[adapter] {name}: {' '.join(os.path.basename(a) if a == aux else a for a in cmd[1:])}")
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-2000:])
    s = {}
    for line in (out / "scalars.txt").read_text().splitlines():
        k, v = line.split()
        s[k] = float(v)
    assert s.get("done") == 1
    xy = np.fromfile(out / "node_centers.f32", np.float32).reshape(-1, 2)
    assert xy.shape == (d.num_nodes, 2)
    # the library's HPWL of the parsed positions is the PlaceDB's: pins,
    # modules and offsets reached the library with the right numbering
    assert s["ep_hpwl_initial"] == pytest.approx(s["hpwl_before_gp"], rel=1e-9)
    # the write-back put the library's positions into the PlaceDB
    assert s["mgp_hpwl"] == pytest.approx(s["mgp_final_hpwl"], rel=1e-6)
    return g, s, xy


# (fixture, adapter flags, final-HPWL bar): QP runs on the GPU where the
# reference's run had QP
CASES = [("smoke_default", [], 0.015), ("smoke_bktrk", ["-bktrk"], 0.015), ("smoke_bb", ["-bb"], 0.015),
         ("synth_c2", ["-noQP"], 0.005), ("synth_c2_bb", ["-noQP", "-bb"], 0.005)]


@pytest.mark.gpu
@pytest.mark.parametrize("name,args,tol", CASES, ids=[c[0] for c in CASES])
def test_main_flow_through_adapter(tmp_path, name, args, tol):
    g, s, _ = run_adapter(tmp_path, name, args)
    iters = int(g.s["iterations"])
    say(f"  adapter {int(s['mgp_iterations'])} iterations, HPWL {s['mgp_hpwl']:.3f}; "
        f"reference {iters}, {g.s['final_hpwl']:.3f}")
    if "-noQP" not in args:
        # QP on the GPU from the core centre, as the reference's run
        assert s["qp_iterations"] >= 1
    assert abs(s["mgp_iterations"] - iters) <= max(3, 0.05 * iters)
    assert abs(s["mgp_hpwl"] - g.s["final_hpwl"]) <= tol * g.s["final_hpwl"]
    assert s["mgp_final_overflow"] <= 0.10


@pytest.mark.gpu
def test_mixed_flow_through_adapter(tmp_path):
    """mGP, mLG, FILLERONLY and cGP (ePlace_main.cpp:141-219) through the
    adapter, against the reference's run of the same design; bars as in
    tests/test_gpu_mixed_flow.py."""
    g, s, _ = run_adapter(tmp_path, "flow_mixed", ["-noQP", "-mixed"])
    k = int(g.s["iterations"])
    say(f"  mGP {int(s['mgp_iterations'])} iterations (reference {k}); mLG overlap {s['mlg_final_overlap']:.0f}; "
        f"FILLERONLY {int(s['fo_iterations'])}; cGP {int(s['cgp_iterations'])} iterations, HPWL {s['cgp_hpwl']:.6g} "
        f"vs reference {g.s['mixed_cgp_hpwl']:.6g}")
    assert abs(s["mgp_iterations"] - k) <= max(5, 0.05 * k)
    assert abs(s["mgp_hpwl"] / g.s["final_hpwl"] - 1) < 5e-3
    assert s["mlg_final_overlap"] == 0
    assert s["fo_iterations"] == 20
    assert s["cgp_final_overflow"] < 0.07
    assert s["cgp_hpwl"] == pytest.approx(s["cgp_final_hpwl"], rel=1e-6)
    assert abs(s["cgp_hpwl"] / g.s["mixed_cgp_hpwl"] - 1) < 1e-2
    # the macros stay where mLG put them through FILLERONLY and cGP
    assert s["macros"] == int(g.design().node_is_macro().sum()) > 0
    assert s["macros_moved_after_mlg"] == 0


@pytest.mark.gpu
def test_learned_controller_through_adapter(tmp_path):
    """Step::Learned with the shipped controller weights on the 2000-cell
    design: it stops at the overflow target and the HPWL stays within 1% of
    the reference's plain run."""
    p = np.load(os.path.join(ROOT, "eplace_amd", "controller_params.npy")).astype(np.float32)
    f = tmp_path / "ctrl.f32"
    p.tofile(f)
    g, s, _ = run_adapter(tmp_path, "synth_c2", ["-noQP", "-learned", str(f)])
    say(f"  learned {int(s['mgp_iterations'])} iterations, HPWL {s['mgp_hpwl']:.3f}; reference plain "
        f"{int(g.s['iterations'])}, {g.s['final_hpwl']:.3f}")
    assert s["mgp_final_overflow"] <= 0.10
    assert abs(s["mgp_hpwl"] / g.s["final_hpwl"] - 1) < 1e-2
