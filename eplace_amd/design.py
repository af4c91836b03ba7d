"""Placement design container, Bookshelf I/O and seeded synthetic generator.

The container is the structure-of-arrays form of the reference's PlaceDB
(PlaceDB/placedb.h:52-185): movable nodes in ``dbNodes`` order, fixed
terminals in ``dbTerminals`` order, pins sorted by net (the order the
Bookshelf parser adds them, Parser/parser.cpp:473-680) and site rows in file
order (placedb.cpp:52-77 derives the core region from them).

Positions are module centres; pin offsets are relative to the centre
(Pin::calculateAbsolutePos, PlaceCommon/objects.cpp:454-459).
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

__all__ = ["Design", "read_bookshelf", "write_bookshelf", "synthetic", "write_pl"]


@dataclass
class Design:
    name: str
    node_names: list
    node_w: np.ndarray  # float32 [N]
    node_h: np.ndarray
    node_x: np.ndarray  # centres
    node_y: np.ndarray
    term_names: list
    term_w: np.ndarray  # float32 [T]
    term_h: np.ndarray
    term_x: np.ndarray  # centres
    term_y: np.ndarray
    term_ni: np.ndarray  # uint8 [T]
    net_start: np.ndarray  # int64 [E+1]
    pin_obj: np.ndarray  # int32 [P]; < N node, else terminal (index - N)
    pin_dx: np.ndarray  # float32 [P]
    pin_dy: np.ndarray
    row_bottom: np.ndarray  # float64 [R]
    row_height: np.ndarray
    row_x0: np.ndarray
    row_x1: np.ndarray
    site_width: float = 1.0
    meta: dict = field(default_factory=dict)

    @property
    def num_nodes(self) -> int:
        return len(self.node_w)

    @property
    def num_terminals(self) -> int:
        return len(self.term_w)

    @property
    def num_nets(self) -> int:
        return len(self.net_start) - 1

    @property
    def num_pins(self) -> int:
        return len(self.pin_obj)

    @property
    def common_row_height(self) -> float:
        # PlaceDB::commonRowHeight is the last row's height (parser.cpp:100-273).
        return float(self.row_height[-1])

    def core_region(self):
        """(llx, lly, urx, ury) as float32, placedb.cpp:52-77."""
        f = np.float32
        bottom = f(self.row_bottom[0])
        top = f(f(self.row_bottom[-1]) + f(self.row_height[-1]))
        left = f(np.min(self.row_x0.astype(np.float32)))
        right = f(np.max(self.row_x1.astype(np.float32)))
        return left, bottom, right, top

    def node_is_macro(self) -> np.ndarray:
        """PlaceDB::addNode (placedb.cpp:102-111): h > row height."""
        h = self.node_h.astype(np.float32)
        return ((h - np.float32(self.common_row_height)).astype(np.float64) > 1e-15).astype(np.uint8)

    def validate(self) -> None:
        n, t, p = self.num_nodes, self.num_terminals, self.num_pins
        assert self.net_start[0] == 0 and self.net_start[-1] == p
        assert np.all(np.diff(self.net_start) >= 0)
        assert np.all((self.pin_obj >= 0) & (self.pin_obj < n + t))
        for a in (self.node_w, self.node_h, self.node_x, self.node_y):
            assert a.dtype == np.float32 and a.shape == (n,)
        for a in (self.pin_dx, self.pin_dy):
            assert a.dtype == np.float32 and a.shape == (p,)


# --------------------------------------------------------------------------
# Bookshelf reader (the subset Parser/parser.cpp reads)
# --------------------------------------------------------------------------

def _aux_files(aux_path):
    with open(aux_path) as f:
        toks = f.read().replace(":", " ").split()
    out = {}
    for t in toks:
        ext = os.path.splitext(t)[1]
        if ext in (".nodes", ".nets", ".pl", ".scl", ".wts"):
            out[ext] = os.path.join(os.path.dirname(aux_path), t)
    return out


def read_bookshelf(aux_path: str, pl_path: str | None = None) -> Design:
    """Parse .nodes/.nets/.pl/.scl like Parser/parser.cpp:44-820."""
    files = _aux_files(aux_path)
    # .scl (parser.cpp:114-273)
    rows = []
    site_width = 1.0
    with open(files[".scl"]) as f:
        next(f)
        for line in f:
            tk = line.split()
            if not tk:
                continue
            if tk[0] == "CoreRow":
                rows.append({"bottom": 0.0, "height": 0.0, "x0": 0.0, "x1": 0.0})
            elif tk[0] == "Coordinate":
                rows[-1]["bottom"] = float(tk[2])
            elif tk[0] == "Height":
                rows[-1]["height"] = float(tk[2])
            elif tk[0] == "Sitewidth":
                site_width = float(tk[2])
            elif len(tk) >= 6 and tk[3] in ("Numsites", "NumSites"):
                origin, nsites = float(tk[2]), float(tk[5])
                rows[-1]["x0"] = origin
                rows[-1]["x1"] = nsites * site_width + origin
    # .nodes (parser.cpp:287-457)
    node_names, node_wh, term_names, term_wh, term_ni = [], [], [], [], []
    with open(files[".nodes"]) as f:
        next(f)
        header = 0
        for line in f:
            if line.startswith("#"):
                continue
            if line.startswith("NumNodes") or line.startswith("NumTerminals"):
                header += 1
                if header == 2:
                    break
        for line in f:
            tk = line.split()
            if not tk:
                continue
            typ = tk[3] if len(tk) > 3 else ""
            w, h = float(tk[1]), float(tk[2])
            if typ in ("terminal", "terminal_NI"):
                term_names.append(tk[0])
                term_wh.append((w, h))
                term_ni.append(1 if typ == "terminal_NI" else 0)
            else:
                node_names.append(tk[0])
                node_wh.append((w, h))
    nidx = {nm: i for i, nm in enumerate(node_names)}
    nn = len(node_names)
    tidx = {nm: nn + i for i, nm in enumerate(term_names)}
    # .nets (parser.cpp:473-680)
    net_start, pin_obj, pin_dx, pin_dy = [0], [], [], []
    with open(files[".nets"]) as f:
        lines = iter(f.readlines()[1:])
        for line in lines:
            if line.startswith("#"):
                continue
            if line.startswith("NumNets") or line.startswith("NumPins"):
                continue
            tk = line.replace(":", " : ").split()
            if not tk:
                continue
            if tk[0] != "NetDegree":
                raise ValueError(f"unsupported .nets line: {line!r}")
            deg = int(tk[2])
            for _ in range(deg):
                pl = next(lines).split()
                name = pl[0]
                dx = float(pl[3]) if len(pl) > 3 else 0.0
                dy = float(pl[4]) if len(pl) > 4 else 0.0
                pin_obj.append(nidx[name] if name in nidx else tidx[name])
                pin_dx.append(dx)
                pin_dy.append(dy)
            net_start.append(len(pin_obj))
    node_w = np.array([w for w, _ in node_wh], np.float32)
    node_h = np.array([h for _, h in node_wh], np.float32)
    term_w = np.array([w for w, _ in term_wh], np.float32).reshape(-1)
    term_h = np.array([h for _, h in term_wh], np.float32).reshape(-1)
    d = Design(
        name=os.path.splitext(os.path.basename(aux_path))[0],
        node_names=node_names, node_w=node_w, node_h=node_h,
        node_x=np.zeros(nn, np.float32), node_y=np.zeros(nn, np.float32),
        term_names=term_names, term_w=term_w, term_h=term_h,
        term_x=np.zeros(len(term_names), np.float32), term_y=np.zeros(len(term_names), np.float32),
        term_ni=np.array(term_ni, np.uint8),
        net_start=np.array(net_start, np.int64), pin_obj=np.array(pin_obj, np.int32),
        pin_dx=np.array(pin_dx, np.float32), pin_dy=np.array(pin_dy, np.float32),
        row_bottom=np.array([r["bottom"] for r in rows]), row_height=np.array([r["height"] for r in rows]),
        row_x0=np.array([r["x0"] for r in rows]), row_x1=np.array([r["x1"] for r in rows]),
        site_width=site_width,
    )
    load_pl(d, pl_path or files[".pl"])
    d.validate()
    return d


def load_pl(d: Design, pl_path: str) -> None:
    """ReadPLFile (parser.cpp:697-820) + setModuleLocation_2D (placedb.cpp:200-240).

    The .pl holds lower-left corners; movable nodes are clamped into the core
    region, terminals are not.  Centres = ll + 0.5*w in float32.
    """
    f32 = np.float32
    llx, lly, urx, ury = d.core_region()
    nidx = {nm: i for i, nm in enumerate(d.node_names)}
    tidx = {nm: i for i, nm in enumerate(d.term_names)}
    with open(pl_path) as f:
        next(f)
        for line in f:
            if not line.strip() or line.startswith("#"):
                continue
            for ch in "(),:=\r":
                line = line.replace(ch, " ")
            tk = line.split()
            if len(tk) < 3:
                continue
            name, x, y = tk[0], f32(float(tk[1])), f32(float(tk[2]))
            if name in nidx:
                i = nidx[name]
                w, h = d.node_w[i], d.node_h[i]
                if x < llx:
                    x = llx
                if f32(x + w) > urx:
                    x = f32(float(f32(urx - w)) - 1e-15)
                if y < lly:
                    y = lly
                if f32(y + h) > ury:
                    y = f32(float(f32(ury - h)) - 1e-15)
                d.node_x[i] = f32(x + f32(f32(0.5) * w))
                d.node_y[i] = f32(y + f32(f32(0.5) * h))
            elif name in tidx:
                i = tidx[name]
                d.term_x[i] = f32(x + f32(f32(0.5) * d.term_w[i]))
                d.term_y[i] = f32(y + f32(f32(0.5) * d.term_h[i]))
            else:
                raise KeyError(f"module {name} not in design")


# --------------------------------------------------------------------------
# Bookshelf writer (format of placedb.cpp:1022-1323)
# --------------------------------------------------------------------------

def _fmt(v) -> str:
    return np.format_float_positional(np.float32(v), trim="-")


def write_pl(d: Design, path: str) -> None:
    with open(path, "w") as f:
        f.write("UCLA pl 1.0\n\n")
        for i, nm in enumerate(d.node_names):
            llx = np.float32(d.node_x[i] - np.float32(0.5) * d.node_w[i])
            lly = np.float32(d.node_y[i] - np.float32(0.5) * d.node_h[i])
            f.write(f"{nm} {_fmt(llx)} {_fmt(lly)} : N\n")
        for i, nm in enumerate(d.term_names):
            llx = np.float32(d.term_x[i] - np.float32(0.5) * d.term_w[i])
            lly = np.float32(d.term_y[i] - np.float32(0.5) * d.term_h[i])
            tag = " /FIXED_NI" if d.term_ni[i] else " /FIXED"
            f.write(f"{nm} {_fmt(llx)} {_fmt(lly)} : N{tag}\n")


def write_bookshelf(d: Design, out_dir: str) -> str:
    """Write .aux/.nodes/.nets/.wts/.pl/.scl; returns the .aux path."""
    os.makedirs(out_dir, exist_ok=True)
    b = os.path.join(out_dir, d.name)
    with open(b + ".aux", "w") as f:
        f.write(f"RowBasedPlacement : {d.name}.nodes {d.name}.nets {d.name}.wts {d.name}.pl {d.name}.scl\n")
    with open(b + ".nodes", "w") as f:
        f.write("UCLA nodes 1.0\n\n")
        f.write(f"NumNodes : {d.num_nodes + d.num_terminals}\n")
        f.write(f"NumTerminals : {d.num_terminals}\n")
        for i, nm in enumerate(d.node_names):
            f.write(f"{nm} {_fmt(d.node_w[i])} {_fmt(d.node_h[i])}\n")
        for i, nm in enumerate(d.term_names):
            typ = "terminal_NI" if d.term_ni[i] else "terminal"
            f.write(f"{nm} {_fmt(d.term_w[i])} {_fmt(d.term_h[i])} {typ}\n")
    names = list(d.node_names) + list(d.term_names)
    with open(b + ".nets", "w") as f:
        f.write("UCLA nets 1.0\n\n")
        f.write(f"NumNets : {d.num_nets}\nNumPins : {d.num_pins}\n")
        ns = d.net_start
        for e in range(d.num_nets):
            f.write(f"NetDegree : {ns[e + 1] - ns[e]} n{e}\n")
            for p in range(ns[e], ns[e + 1]):
                f.write(f"{names[d.pin_obj[p]]} B : {_fmt(d.pin_dx[p])} {_fmt(d.pin_dy[p])}\n")
    with open(b + ".wts", "w") as f:
        f.write("UCLA wts 1.0\n")
    write_pl(d, b + ".pl")
    with open(b + ".scl", "w") as f:
        f.write("UCLA scl 1.0\n\n")
        f.write(f"NumRows : {len(d.row_bottom)}\n\n")
        for r in range(len(d.row_bottom)):
            nsites = int(round((d.row_x1[r] - d.row_x0[r]) / d.site_width))
            f.write("CoreRow Horizontal\n")
            f.write(f" Coordinate : {d.row_bottom[r]:g}\n")
            f.write(f" Height : {d.row_height[r]:g}\n")
            f.write(f" Sitewidth : {d.site_width:g}\n")
            f.write(f" Sitespacing : {d.site_width:g}\n")
            f.write(" Siteorient : N\n Sitesymmetry : Y\n")
            f.write(f" SubrowOrigin : {d.row_x0[r]:g} Numsites : {nsites}\n")
            f.write("End\n")
    return b + ".aux"


# --------------------------------------------------------------------------
# Seeded synthetic generator
# --------------------------------------------------------------------------

def _tail_degrees(rng, n, lo=11, hi=200, mean=14.0):
    """Power-law degrees on [lo, hi] with exponent chosen for the given mean."""
    d = np.arange(lo, hi + 1, dtype=np.float64)
    a_lo, a_hi = 1.0, 6.0
    for _ in range(60):
        a = 0.5 * (a_lo + a_hi)
        p = d ** (-a)
        m = (d * p).sum() / p.sum()
        if m > mean:
            a_lo = a
        else:
            a_hi = a
    p = d ** (-a)
    return rng.choice(d.astype(np.int64), size=n, p=p / p.sum())


def synthetic(num_cells: int, num_nets: int, *, seed: int = 0, num_io: int = 4,
              util: float = 0.7, widths: str = "lognormal", degrees: str = "c2",
              locality: float = 0.9, window: float = 0.05, num_macros: int = 0,
              macro_side=(0.03, 0.12), macro_pins=(20, 80), macro_area_cap: float = 0.45,
              init: str = "center", name: str | None = None) -> Design:
    """Seeded synthetic design following the workload recipes.

    widths   "uniform14": U{1..4} sites; "lognormal": median 3, clipped 1..40.
    degrees  "c1": 2:60%, 3:25%, 4-5:15%;  "c2": 2:55%, 3:20%, 4-10:20%,
             11-200 power-law 5% (tail mean 14, giving ~3.8 pins/net).
    locality fraction of pins drawn from a window of ``window`` x cells
             around a per-net random start (cell index locality).
    IO pins  zero-size fixed terminal_NI pins on the die boundary; with 4 IOs
             they sit at the edge midpoints.
    Macros   movable, square-ish with side U(macro_side) x die width, total
             area capped at ``macro_area_cap`` x die area by uniform rescale.
    init     "center": cells around the die centre (sigma 5% of the die);
             "random": uniform over the core.
    """
    rng = np.random.default_rng(seed)
    f32 = np.float32
    if widths == "uniform14":
        w = rng.integers(1, 5, size=num_cells).astype(np.float32)
    else:
        w = np.clip(np.round(np.exp(rng.normal(np.log(3.0), 0.6, size=num_cells))), 1, 40).astype(np.float32)
    h = np.ones(num_cells, np.float32)
    cell_area = float(w.sum())
    # die: square, integer side, sized for the utilisation (macros included)
    side0 = np.sqrt(cell_area / util)
    mw = np.zeros(0, np.float32)
    if num_macros:
        s = rng.uniform(macro_side[0], macro_side[1], size=num_macros)
        asp = rng.uniform(0.8, 1.25, size=num_macros)
        # iterate the die size because the macro sides scale with it
        side = side0
        for _ in range(20):
            mws = s * side * np.sqrt(asp)
            mhs = s * side / np.sqrt(asp)
            scale = min(1.0, np.sqrt(macro_area_cap * side * side / (mws * mhs).sum()))
            mws, mhs = mws * scale, mhs * scale
            side = np.sqrt((cell_area + (mws * mhs).sum()) / util)
        mw = np.maximum(np.round(mws), 2).astype(np.float32)
        mh = np.maximum(np.round(mhs), 2).astype(np.float32)
        side = np.sqrt((cell_area + float((mw * mh).sum())) / util)
    W = int(np.ceil(side0 if not num_macros else side))
    H = W
    node_w = np.concatenate([w, mw]) if num_macros else w
    node_h = np.concatenate([h, mh]) if num_macros else h
    n = len(node_w)
    # IO terminals on the boundary
    if num_io == 4:
        tx = np.array([W / 2, 0, W / 2, W], np.float32)
        ty = np.array([0, H / 2, H, H / 2], np.float32)
    else:
        per = rng.uniform(0, 4 * W, size=num_io)
        side_id = (per // W).astype(int)
        t = per % W
        tx = np.where(side_id == 0, t, np.where(side_id == 1, W, np.where(side_id == 2, W - t, 0))).astype(np.float32)
        ty = np.where(side_id == 0, 0, np.where(side_id == 1, t, np.where(side_id == 2, H, H - t))).astype(np.float32)
    # degrees
    if degrees == "c1":
        r = rng.random(num_nets)
        deg = np.where(r < 0.60, 2, np.where(r < 0.85, 3, rng.integers(4, 6, size=num_nets)))
    else:
        r = rng.random(num_nets)
        deg = np.where(r < 0.55, 2, np.where(r < 0.75, 3, rng.integers(4, 11, size=num_nets)))
        tail = r >= 0.95
        deg[tail] = _tail_degrees(rng, int(tail.sum()))
    deg = deg.astype(np.int64)
    net_start = np.zeros(num_nets + 1, np.int64)
    np.cumsum(deg, out=net_start[1:])
    P = int(net_start[-1])
    net_of_pin = np.repeat(np.arange(num_nets), deg)
    # pin -> object with index locality among std cells
    win = max(int(window * num_cells), 8)
    wstart = rng.integers(0, max(num_cells - win, 1), size=num_nets)
    local = rng.random(P) < locality
    pin_obj = np.where(local, wstart[net_of_pin] + rng.integers(0, win, size=P),
                       rng.integers(0, num_cells, size=P)).astype(np.int64)
    # macro pins: each macro takes pins from random nets
    if num_macros:
        for m in range(num_macros):
            k = int(rng.integers(macro_pins[0], macro_pins[1] + 1))
            pins = rng.integers(0, P, size=k)
            pin_obj[pins] = num_cells + m
    # IO pins: the first pin of num_io nets spread over the netlist
    if num_io:
        nets = rng.choice(num_nets, size=min(num_io, num_nets), replace=False)
        pin_obj[net_start[nets]] = n + np.arange(len(nets))
    pin_obj = pin_obj.astype(np.int32)
    is_node = pin_obj < n
    pw = np.where(is_node, node_w[np.minimum(pin_obj, n - 1)], 0).astype(np.float32)
    ph = np.where(is_node, node_h[np.minimum(pin_obj, n - 1)], 0).astype(np.float32)
    pin_dx = (rng.uniform(-0.5, 0.5, size=P) * pw).astype(np.float32)
    pin_dy = (rng.uniform(-0.5, 0.5, size=P) * ph).astype(np.float32)
    # initial placement
    if init == "random":
        x = rng.uniform(0, W, size=n)
        y = rng.uniform(0, H, size=n)
    else:
        x = W / 2 + rng.normal(0, 0.05 * W, size=n)
        y = H / 2 + rng.normal(0, 0.05 * H, size=n)
    x = np.clip(x, node_w / 2, W - node_w / 2).astype(np.float32)
    y = np.clip(y, node_h / 2, H - node_h / 2).astype(np.float32)
    nt = len(tx)
    d = Design(
        name=name or f"synth_{num_cells}_{seed}",
        node_names=[f"o{i}" for i in range(n)],
        node_w=node_w.astype(np.float32), node_h=node_h.astype(np.float32), node_x=x, node_y=y,
        term_names=[f"p{i}" for i in range(nt)],
        term_w=np.zeros(nt, np.float32), term_h=np.zeros(nt, np.float32),
        term_x=tx.astype(np.float32), term_y=ty.astype(np.float32), term_ni=np.ones(nt, np.uint8),
        net_start=net_start, pin_obj=pin_obj, pin_dx=pin_dx, pin_dy=pin_dy,
        row_bottom=np.arange(H, dtype=np.float64), row_height=np.ones(H),
        row_x0=np.zeros(H), row_x1=np.full(H, float(W)), site_width=1.0,
        meta={"seed": seed, "util": util, "widths": widths, "degrees": degrees,
              "num_macros": num_macros, "macro_area_cap": macro_area_cap, "die": W},
    )
    d.validate()
    return d
