"""Generate golden fixtures from the reference placer (run in the container
that has the reference tree; the fixtures are committed, the reference is not).

    python tests/golden/gen_golden.py [gp] [bb] [poisson] [legal] [qp] [mlg] [flow] [dp] [grid]   (default: all)

Builds oracle/_ref/ref_harness from the reference's own sources
(oracle/Makefile.ref), runs it on the reference's smoke design and on seeded
synthetic designs written as Bookshelf, and stores its dumps as .npz files
(inputs and outputs only) next to this script.
"""
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
sys.path.insert(0, ROOT)

from eplace_amd.design import synthetic, write_bookshelf  # noqa: E402

# Keys kept per gp run (everything else the harness writes is dropped).
KEEP = ["nodes_whxy", "nodes_macro", "terms_whxyll", "terms_ni", "pin_obj", "pin_offset", "net_start", "rows",
        "fillers_init", "fillers_wh", "init_ll", "init_positions", "init_map_node", "init_map_filler",
        "init_map_terminal", "init_map_base", "init_map_ex", "init_map_ey", "init_map_phi", "init_grad_density",
        "init_grad_wirelength", "init_grad_total", "it0_main", "it1_main", "it2_main", "it0_db_positions",
        "it1_db_positions", "it2_db_positions", "trajectory", "final_positions", "legal_positions",
        "input_positions", "qp_ax_row", "qp_ax_col", "qp_ax_val", "qp_ay_row", "qp_ay_col", "qp_ay_val", "qp_bx",
        "qp_by", "qp_guess_x", "qp_guess_y", "qp_matrix_positions", "qp_positions", "mlg_positions", "mixed_mlg_positions", "mixed_fo_fillers",
        "mixed_fo_positions", "mixed_cgp_positions", "lr0", "ism0", "gs0", "lr1", "ism1", "gs1"]


def run(args, out):
    if os.path.exists(out):
        shutil.rmtree(out)
    with open(os.path.join(tempfile.gettempdir(), "ref_harness.log"), "w") as log:
        subprocess.run([HARNESS] + args + ["-out", out], check=True, stdout=log, stderr=log,
                       cwd=tempfile.gettempdir())


def pack(name, out, extra=None, keep=None):
    arrays = {}
    for k in keep or KEEP:
        f = os.path.join(out, k + ".npy")
        if os.path.exists(f):
            arrays[k] = np.load(f)
    scal = {}
    for line in open(os.path.join(out, "scalars.txt")):
        k, v = line.split()
        scal[k] = float(v)
    arrays["scalar_names"] = np.array(sorted(scal))
    arrays["scalar_values"] = np.array([scal[k] for k in sorted(scal)])
    for k, v in (extra or {}).items():
        arrays[k] = np.asarray(v)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
    print(f"{name}: {len(arrays)} arrays, iterations {scal.get('iterations')}, final HPWL {scal.get('final_hpwl')}")


def main(parts):
    subprocess.run(["make", "-s", "-j8", "-f", "oracle/Makefile.ref", f"REF={REF}"], cwd=ROOT, check=True)
    tmp = tempfile.mkdtemp(prefix="golden_")
    if "gp" in parts:
        gp_fixtures(tmp)
    if "bb" in parts:
        bb_fixtures(tmp)
    if "poisson" in parts:
        poisson_fixtures(tmp)
    if "legal" in parts:
        legal_fixtures(tmp)
    if "dp" in parts:
        dp_fixtures(tmp)
    if "grid" in parts:
        grid_fixtures(tmp)
    if "qp" in parts:
        qp_fixtures(tmp)
    if "mlg" in parts:
        mlg_fixtures(tmp)
    if "flow" in parts:
        flow_fixtures(tmp)
    shutil.rmtree(tmp)


def gp_fixtures(tmp):
    smoke = os.path.join(REF, "examples", "smoke", "smoke.aux")
    # 1. the reference's own smoke flow (QP -> mGP -> Abacus), default and -bktrk
    run(["gp", "-aux", smoke, "-targetDensity", "0.8", "-targetOverflow", "0.1", "-legal"], tmp + "/s")
    pack("smoke_default", tmp + "/s")
    run(["gp", "-aux", smoke, "-targetDensity", "0.8", "-targetOverflow", "0.1", "-bktrk"], tmp + "/sb")
    pack("smoke_bktrk", tmp + "/sb")
    run(["gp", "-aux", smoke, "-targetDensity", "0.8", "-targetOverflow", "0.1", "-LSE", "-iters", "3"],
        tmp + "/sl")
    pack("smoke_lse", tmp + "/sl")
    # 2. seeded synthetic designs, no QP (the .pl is the start for both sides)
    designs = {
        "synth_c1": dict(num_cells=100, num_nets=184, seed=11, degrees="c1", widths="uniform14", td=0.8),
        "synth_c2": dict(num_cells=2000, num_nets=2100, seed=12, td=1.0),
        "synth_mixed": dict(num_cells=2000, num_nets=2200, seed=13, num_macros=4, util=0.75, td=1.0),
    }
    for name, kw in designs.items():
        td = kw.pop("td")
        d = synthetic(**kw, name=name)
        aux = write_bookshelf(d, os.path.join(tmp, name))
        run(["gp", "-aux", aux, "-targetDensity", str(td), "-targetOverflow", "0.1", "-noQP",
             "-loadpl", aux[:-4] + ".pl", "-dumpEvery", "1000"], tmp + "/out_" + name)
        pack(name, tmp + "/out_" + name, {"target_density": td})


def grid_fixtures(tmp):
    # 8. the reference at the grids the workload configurations name, through
    # the harness's copy of binInitialization for a given grid (-bins):
    # config 1 as specified (100 cells, 16x16 bins, target density 0.8, QP ->
    # mGP -> Abacus, then detailed placement from the GP result), and a
    # scaled-down config 3 (20000 cells + 60 macros with sides of 3-12% of the
    # die, utilisation 0.75, 256x256 bins: mGP -> mLG -> FILLERONLY -> cGP)
    # seed 2: every cell has a pin (the reference's global swap needs that)
    d = synthetic(num_cells=100, num_nets=184, seed=2, degrees="c1", widths="uniform14", name="c1_16")
    aux = write_bookshelf(d, os.path.join(tmp, "c1_16"))
    run(["gp", "-aux", aux, "-targetDensity", "0.8", "-targetOverflow", "0.1", "-bins", "16", "-legal",
         "-dumpEvery", "1000"], tmp + "/out_c1_16")
    gp_xy = np.load(tmp + "/out_c1_16/final_positions.npy")[:100]
    f = os.path.join(tmp, "c1_16.f32")
    np.ascontiguousarray(gp_xy, np.float32).tofile(f)
    run(["dp", "-aux", aux, "-noQP", "-centers", f], tmp + "/dp_c1_16")
    extra = {"target_density": 0.8, "bins": 16, "dp_positions": np.load(tmp + "/dp_c1_16/gs1.npy")}
    for line in open(tmp + "/dp_c1_16/scalars.txt"):
        k, v = line.split()
        if k.startswith("dp_") or k.endswith("_hpwl"):
            extra["dpflow_" + k] = float(v)
    pack("c1_16_flow", tmp + "/out_c1_16", extra)
    d = synthetic(num_cells=20000, num_nets=21000, seed=41, num_macros=60, util=0.75, name="mixed20k")
    aux = write_bookshelf(d, os.path.join(tmp, "mixed20k"))
    run(["gp", "-aux", aux, "-targetDensity", "1.0", "-targetOverflow", "0.1", "-noQP", "-loadpl",
         aux[:-4] + ".pl", "-bins", "256", "-mixed", "-dumpEvery", "100000"], tmp + "/out_mixed20k")
    pack("mixed20k_256", tmp + "/out_mixed20k", {"target_density": 1.0, "bins": 256},
         keep=["nodes_whxy", "nodes_macro", "terms_whxyll", "terms_ni", "pin_obj", "pin_offset", "net_start", "rows",
               "trajectory"])


def bb_fixtures(tmp):
    # the Barzilai-Borwein variant (-bb, nesterov.hpp:231-302): the smoke flow
    # after QP, and the 2000-cell synthetic design from its .pl
    smoke = os.path.join(REF, "examples", "smoke", "smoke.aux")
    run(["gp", "-aux", smoke, "-targetDensity", "0.8", "-targetOverflow", "0.1", "-bb"], tmp + "/sbb")
    pack("smoke_bb", tmp + "/sbb")
    d = synthetic(num_cells=2000, num_nets=2100, seed=12, name="synth_c2")
    aux = write_bookshelf(d, os.path.join(tmp, "synth_c2_bb"))
    run(["gp", "-aux", aux, "-targetDensity", "1.0", "-targetOverflow", "0.1", "-noQP", "-loadpl",
         aux[:-4] + ".pl", "-bb"], tmp + "/out_synth_c2_bb")
    pack("synth_c2_bb", tmp + "/out_synth_c2_bb", {"target_density": 1.0})


def poisson_fixtures(tmp):
    # 3. the Poisson solve alone on random densities (square and non-square bins)
    rng = np.random.default_rng(5)
    fft = {}
    for (nx, ny, bx, by) in [(16, 16, 1.25, 1.25), (32, 64, 1.5, 2.0), (128, 128, 0.8, 0.8)]:
        rho = (rng.random((nx, ny)) * 1.5).astype(np.float32)
        f = os.path.join(tmp, "rho.f32")
        rho.tofile(f)
        run(["fft", "-nx", str(nx), "-ny", str(ny), "-bx", str(bx), "-by", str(by), "-in", f], tmp + "/fft")
        key = f"{nx}x{ny}"
        fft[key + "_rho"] = rho
        fft[key + "_bins"] = np.array([bx, by], np.float32)
        for k in ("ex", "ey", "phi"):
            fft[key + "_" + k] = np.load(os.path.join(tmp, "fft", k + ".npy"))
    np.savez_compressed(os.path.join(HERE, "poisson.npz"), **fft)
    print("poisson: ", sorted(fft))


def legal_fixtures(tmp):
    # 4. AbacusLegalizer on given centres: the 2000-cell design's GP result,
    # and a 20000-cell design from uniformly random centres
    legal_cases = {
        "legal_synth_c2": (synthetic(num_cells=2000, num_nets=2100, seed=12, name="legal_synth_c2"),
                           np.load(os.path.join(HERE, "synth_c2.npz"))["final_positions"][:2000]),
    }
    d = synthetic(num_cells=20000, num_nets=21000, seed=21, util=0.7, name="legal_rand20k")
    llx, lly, urx, ury = d.core_region()
    rng = np.random.default_rng(22)
    cx = rng.uniform(llx + d.node_w / 2, urx - d.node_w / 2).astype(np.float32)
    cy = rng.uniform(lly + d.node_h / 2, ury - d.node_h / 2).astype(np.float32)
    legal_cases["legal_rand20k"] = (d, np.stack([cx, cy], 1))
    for name, (d, xy) in legal_cases.items():
        aux = write_bookshelf(d, os.path.join(tmp, name))
        f = os.path.join(tmp, name + ".f32")
        np.ascontiguousarray(xy, np.float32).tofile(f)
        run(["legal", "-aux", aux, "-noQP", "-centers", f], tmp + "/out_" + name)
        pack(name, tmp + "/out_" + name)


def connect_all(d):
    """The design with a two-pin net (centre pins) from every node without
    pins to its index neighbour: GlobalSwapDP reads the optimal region of
    every cell, which for a cell without nets indexes an empty vector
    (placedb.cpp:1480-1533), so the reference's detailed placer needs every
    cell connected."""
    import dataclasses
    n = d.num_nodes
    cnt = np.bincount(d.pin_obj[d.pin_obj < n], minlength=n)
    lone = np.flatnonzero(cnt == 0)
    extra = np.stack([lone, np.where(lone + 1 < n, lone + 1, lone - 1)], 1).reshape(-1).astype(np.int32)
    k = len(lone)
    return dataclasses.replace(
        d, net_start=np.concatenate([d.net_start, d.net_start[-1] + 2 * np.arange(1, k + 1)]).astype(np.int64),
        pin_obj=np.concatenate([d.pin_obj, extra]).astype(np.int32),
        pin_dx=np.concatenate([d.pin_dx, np.zeros(2 * k, np.float32)]),
        pin_dy=np.concatenate([d.pin_dy, np.zeros(2 * k, np.float32)]))


def dp_fixtures(tmp):
    # 7. DetailedPlacer after AbacusLegalizer (ePlace_main.cpp:249-283 with
    # -internalLegal -internalDP): the smoke design's GP result, the
    # 2000-cell design's GP result (every cell connected), and a 20000-cell
    # design from legalised random centres
    d20 = connect_all(synthetic(num_cells=20000, num_nets=21000, seed=31, util=0.7, name="dp_rand20k"))
    llx, lly, urx, ury = d20.core_region()
    rng = np.random.default_rng(32)
    xy20 = np.stack([rng.uniform(llx + d20.node_w / 2, urx - d20.node_w / 2),
                     rng.uniform(lly + d20.node_h / 2, ury - d20.node_h / 2)], 1).astype(np.float32)
    cases = {
        "dp_smoke": (Golden_design("smoke_default"), np.load(os.path.join(HERE, "smoke_default.npz"))["final_positions"][:100]),
        "dp_synth_c2": (connect_all(synthetic(num_cells=2000, num_nets=2100, seed=12, name="dp_synth_c2")),
                        np.load(os.path.join(HERE, "synth_c2.npz"))["final_positions"][:2000]),
        "dp_rand20k": (d20, xy20),
    }
    for name, (d, xy) in cases.items():
        aux = d if isinstance(d, str) else write_bookshelf(d, os.path.join(tmp, name))
        f = os.path.join(tmp, name + ".f32")
        np.ascontiguousarray(xy, np.float32).tofile(f)
        run(["dp", "-aux", aux, "-noQP", "-centers", f], tmp + "/out_" + name)
        pack(name, tmp + "/out_" + name, {"input_positions": xy})


def Golden_design(name):
    """The reference's own smoke design (its .aux path)."""
    return os.path.join(REF, "examples", "smoke", "smoke.aux")


def qp_log(path):
    """The reference's per-iteration QP line (qplace.cpp:89): iteration,
    max(X error, Y error), HPWL."""
    rows = []
    for line in open(path):
        if "at iteration number" in line:
            f = line.replace(",", " ").split()
            rows.append((float(f[4]), float(f[7]), float(f[9])))
    return np.array(rows, np.float64)


def qp_fixtures(tmp):
    # 5. QPPlacer: createSparseMatrix at the .pl positions, then
    # quadraticPlacement from the core centre (the reference's default flow)
    smoke = os.path.join(REF, "examples", "smoke", "smoke.aux")
    cases = {
        "qp_smoke": (smoke, None),
        "qp_synth2k": (None, synthetic(2000, 2100, seed=31, num_io=100, init="random", name="qp_synth2k")),
        "qp_mixed3k": (None, synthetic(3000, 3200, seed=32, num_io=200, num_macros=4, util=0.7, init="random",
                                       name="qp_mixed3k")),
    }
    for name, (aux, d) in cases.items():
        args = ["qp", "-noQP"]
        if d is not None:
            aux = write_bookshelf(d, os.path.join(tmp, name))
            args += ["-loadpl", aux[:-4] + ".pl"]
        run(args + ["-aux", aux, "-IPiteCount", "0"], tmp + "/out0_" + name)
        it0 = np.load(os.path.join(tmp, "out0_" + name, "qp_positions.npy"))
        run(args + ["-aux", aux], tmp + "/out_" + name)
        log = qp_log(os.path.join(tempfile.gettempdir(), "ref_harness.log"))
        pack(name, tmp + "/out_" + name, {"qp_log": log, "qp_positions_it0": it0})


def mlg_fixtures(tmp):
    # 6. SAMacroLegalizer on given node centres after K rand() draws: the
    # reference's own mGP result of synth_mixed, and macro-heavy designs
    # started at the core centre (one, two and seven annealing rounds)
    def macro_design(seed, cells, macros, util, side=(0.03, 0.12), cap=0.45):
        d = synthetic(cells, int(cells * 1.05), seed=seed, num_io=cells // 50, num_macros=macros, util=util,
                      macro_side=side, macro_area_cap=cap, init="center", name=f"mlg{seed}")
        return d, np.stack([d.node_x, d.node_y], 1)
    mixed = synthetic(num_cells=2000, num_nets=2200, seed=13, num_macros=4, util=0.75, name="mlg_mixed")
    cases = {
        "mlg_mixed": (mixed, np.load(os.path.join(HERE, "synth_mixed.npz"))["final_positions"][:mixed.num_nodes],
                      0, 1.0),
        "mlg_m52": macro_design(52, 4000, 20, 0.7) + (777, 0.9),
        "mlg_mm62": macro_design(62, 3000, 25, 0.6, (0.08, 0.2), 0.6) + (99, 1.0),
        "mlg_mm63": macro_design(63, 3000, 60, 0.55, (0.04, 0.12), 0.6) + (99, 1.0),
    }
    for name, (d, xy, skip, td) in cases.items():
        aux = write_bookshelf(d, os.path.join(tmp, name))
        f = os.path.join(tmp, name + ".f32")
        np.ascontiguousarray(xy, np.float32).tofile(f)
        run(["mlg", "-aux", aux, "-noQP", "-centers", f, "-randSkip", str(skip), "-targetDensity", str(td)],
            tmp + "/out_" + name)
        pack(name, tmp + "/out_" + name, {"rand_skip": skip, "target_density": td})


def flow_fixtures(tmp):
    # 7. the mixed-size flow after mGP (ePlace_main.cpp:174-219): mLG,
    # FILLERONLY and cGP, from the .pl (no QP)
    d = synthetic(3000, 3200, seed=71, num_macros=10, util=0.7, name="flow_mixed")
    aux = write_bookshelf(d, os.path.join(tmp, "flow_mixed"))
    run(["gp", "-aux", aux, "-targetDensity", "1.0", "-targetOverflow", "0.1", "-noQP", "-loadpl", aux[:-4] + ".pl",
         "-mixed"], tmp + "/out_flow_mixed")
    pack("flow_mixed", tmp + "/out_flow_mixed", {"target_density": 1.0})


if __name__ == "__main__":
    main(set(sys.argv[1:]) or {"gp", "poisson", "legal", "qp", "mlg", "flow", "dp", "grid"})
