"""The C++ host layer (eplace_amd/host): its Bookshelf reader against the
Python reader, and the ep_place command against the Python API on the GPU."""
import os
import subprocess
import sys

import numpy as np
import pytest

from eplace_amd.design import load_pl, read_bookshelf, synthetic, write_bookshelf, write_pl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "eplace_amd", "bin", "ep_place")

FIELDS = [("node_w", np.float32), ("node_h", np.float32), ("node_x", np.float32), ("node_y", np.float32),
          ("term_w", np.float32), ("term_h", np.float32), ("term_x", np.float32), ("term_y", np.float32),
          ("term_ni", np.uint8), ("net_start", np.int64), ("pin_obj", np.int32), ("pin_dx", np.float32),
          ("pin_dy", np.float32), ("row_bottom", np.float64), ("row_height", np.float64),
          ("row_x0", np.float64), ("row_x1", np.float64)]


def exe():
    if not os.path.exists(EXE):
        pytest.skip("ep_place is not built (make -C eplace_amd/host)")
    return EXE


def read_dump(path):
    out = {}
    with open(path, "rb") as f:
        for name, dt in FIELDS:
            n = int(np.frombuffer(f.read(8), np.int64)[0])
            out[name] = np.frombuffer(f.read(n * np.dtype(dt).itemsize), dt)
        assert f.read() == b""
    return out


def cpp_parse(aux, tmp_path, loadpl=None):
    dump = str(tmp_path / "design.bin")
    cmd = [exe(), "-aux", aux, "-dumpDesign", dump] + (["-loadpl", loadpl] if loadpl else [])
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return read_dump(dump)


def assert_same(cpp, d):
    for name, dt in FIELDS:
        np.testing.assert_array_equal(cpp[name], np.asarray(getattr(d, name), dt), err_msg=name)


@pytest.mark.parametrize("kw", [
    dict(num_cells=300, num_nets=380, seed=3, num_io=20, degrees="c1", widths="uniform14"),
    dict(num_cells=800, num_nets=900, seed=4, num_macros=3, util=0.7),
])
def test_cpp_reader_matches_python_reader(tmp_path, kw):
    d = synthetic(**kw, name="cli")
    aux = write_bookshelf(d, str(tmp_path / "in"))
    assert_same(cpp_parse(aux, tmp_path), read_bookshelf(aux))


def test_cpp_loadpl_clamps_like_python(tmp_path):
    d = synthetic(num_cells=200, num_nets=240, seed=6, num_io=10, name="clamp")
    aux = write_bookshelf(d, str(tmp_path / "in"))
    # Push some cells outside the core on every side; movable nodes are
    # clamped into the core, terminals are not (placedb.cpp:200-240).
    llx, lly, urx, ury = d.core_region()
    e = read_bookshelf(aux)
    e.node_x[:10] = llx - 50.0
    e.node_x[10:20] = urx + 50.0
    e.node_y[20:30] = lly - 7.0
    e.node_y[30:40] = ury + 7.0
    e.term_x[:3] = urx + 100.0
    pl = str(tmp_path / "moved.pl")
    write_pl(e, pl)
    ref = read_bookshelf(aux)
    load_pl(ref, pl)
    got = cpp_parse(aux, tmp_path, loadpl=pl)
    assert_same(got, ref)
    assert got["node_x"][:10].min() >= np.float32(llx)
    assert got["term_x"][:3].min() > urx


def test_cli_usage_and_bad_input(tmp_path):
    r = subprocess.run([exe()], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
    r = subprocess.run([exe(), "-aux", str(tmp_path / "missing.aux")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "cannot open" in r.stderr


def test_cli_rejects_unknown_poisson_mode(tmp_path):
    """An unknown -poissonMode is refused by ep_create's argument checks."""
    d = synthetic(num_cells=100, num_nets=120, seed=1, name="pm")
    aux = write_bookshelf(d, str(tmp_path / "in"))
    r = subprocess.run([exe(), "-aux", aux, "-noQP", "-poissonMode", "7", "-outputPath", str(tmp_path / "out")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "poisson_mode must be" in r.stderr


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="checks the no-device error path")
def test_cli_without_device_fails_cleanly(tmp_path):
    d = synthetic(num_cells=100, num_nets=120, seed=1, name="nodev")
    aux = write_bookshelf(d, str(tmp_path / "in"))
    r = subprocess.run([exe(), "-aux", aux, "-noQP", "-outputPath", str(tmp_path / "out")], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 1
    assert "ep_create" in r.stderr
    # QP runs first by default (ePlace_main.cpp:121-126)
    r = subprocess.run([exe(), "-aux", aux, "-outputPath", str(tmp_path / "out")], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 1
    assert "ep_qplace" in r.stderr


@pytest.mark.gpu
def test_cli_matches_python_api(tmp_path):
    from eplace_amd import EPlacer2D, EplaceNesterovOpt, quadratic_placement

    d = synthetic(num_cells=3000, num_nets=3200, seed=2, name="cli3k")
    aux = write_bookshelf(d, str(tmp_path / "in"))
    print("ep_place on cli3k (3000 cells, QP, bins 64)", flush=True)
    r = subprocess.run([exe(), "-aux", aux, "-bins", "64", "-internalLegal", "-outputPath", str(tmp_path / "out")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "; legal (" in r.stdout, r.stdout
    hpwl_cli = float([ln for ln in r.stdout.splitlines() if ln.startswith("Final HPWL")][0].split()[-1])

    e = read_bookshelf(aux)
    qp = quadratic_placement(e)
    assert "INFO:  initial HPWL" in r.stdout
    qlines = [ln for ln in r.stdout.splitlines() if "at iteration number" in ln]
    assert len(qlines) == qp["iterations"]
    assert abs(float(qlines[-1].split()[-1]) / qp["final_hpwl"] - 1) < 1e-6
    gp = EPlacer2D(e, 1.0, bins=64)
    gp.initialization()
    res = EplaceNesterovOpt(gp).opt()
    assert res["stopped"] == 1
    assert abs(hpwl_cli - res["final_hpwl"]) <= 1e-6 * res["final_hpwl"]

    out = read_bookshelf(aux)
    load_pl(out, str(tmp_path / "out" / "cli3k" / "cli3k-eGP.pl"))
    xy = gp.read("positions")[: e.num_nodes]
    np.testing.assert_allclose(out.node_x, xy[:, 0], atol=1e-3)
    np.testing.assert_allclose(out.node_y, xy[:, 1], atol=1e-3)


def test_cli_learned_parameters_checked_before_device_work(tmp_path):
    """-learned reads the controller's saved parameters (a float32 .npy file)
    before any device work: the shipped file is accepted, a file that is not
    .npy or holds the wrong number of floats is refused with a message."""
    d = synthetic(num_cells=100, num_nets=120, seed=1, name="lrn")
    aux = write_bookshelf(d, str(tmp_path / "in"))
    shipped = os.path.join(ROOT, "eplace_amd", "controller_params.npy")
    r = subprocess.run([exe(), "-aux", aux, "-learned", shipped, "-dumpDesign", str(tmp_path / "d.bin")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    bad = tmp_path / "bad.npy"
    bad.write_bytes(b"not numpy")
    r = subprocess.run([exe(), "-aux", aux, "-learned", str(bad)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "not a .npy file" in r.stderr
    np.save(tmp_path / "short.npy", np.zeros(10, np.float32))
    r = subprocess.run([exe(), "-aux", aux, "-learned", str(tmp_path / "short.npy")], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1 and "found 10" in r.stderr
    np.save(tmp_path / "f64.npy", np.zeros(5150, np.float64))
    r = subprocess.run([exe(), "-aux", aux, "-learned", str(tmp_path / "f64.npy")], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1 and "float32" in r.stderr


@pytest.mark.gpu
def test_cli_learned_controller_matches_python_api(tmp_path):
    """ep_place -learned -poissonMode 2 runs the same device loop as the
    Python API with the same controller and split-FP16 transforms: the same
    iterations and final HPWL."""
    from eplace_amd import EPlacer2D, EplaceNesterovOpt
    from eplace_amd.controller import Controller

    d = synthetic(num_cells=20000, num_nets=21000, seed=7, name="lrn20k")
    aux = write_bookshelf(d, str(tmp_path / "in"))
    shipped = os.path.join(ROOT, "eplace_amd", "controller_params.npy")
    print("ep_place -learned on 20000 cells", flush=True)
    r = subprocess.run([exe(), "-aux", aux, "-noQP", "-learned", shipped, "-poissonMode", "2", "-outputPath",
                        str(tmp_path / "out")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    line = [ln for ln in r.stdout.splitlines() if ln.startswith("mGP finished")][0]
    iters, overflow = int(line.split()[2]), float(line.split()[-1])
    hpwl_cli = float([ln for ln in r.stdout.splitlines() if ln.startswith("Final HPWL")][0].split()[-1])
    e = read_bookshelf(aux)
    gp = EPlacer2D(e, 1.0, poisson_mode=2)
    gp.initialization()
    ctrl = Controller.load()
    res = EplaceNesterovOpt(gp, method="learned", controller=ctrl, log=False).opt()
    gp.close()
    print(f"  cli {iters} iterations, overflow {overflow:.4f}, HPWL {hpwl_cli:.6e}; python {res['iterations']}, "
          f"{res['final_hpwl']:.6e}", flush=True)
    assert overflow <= 0.10
    assert iters == res["iterations"]
    assert abs(hpwl_cli - res["final_hpwl"]) <= 1e-6 * res["final_hpwl"]
