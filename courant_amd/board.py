"""Board model -> simulation front-end over include/courant_board.h.

The board path of the reference's web front-end, with its names in snake case:

  Board, import_kicad, stackup / apply_stackup   web/src/import/kicad.ts:51-220, load.ts:14-29
  port_nets                                      web/src/import/load.ts:31-38
  copper_at, copper_row, copper_edges           web/src/import/kicad.ts:277-387
  lambda_over_20, mesh_axis, with_pml, estimate  web/src/mesh/mesher.ts:47-156
  plan_board, build_board_sim                    web/src/import/boardsim.ts:22-59, 110-331
  gcpw_board                                     web/src/scenarios/gcpwboard.ts:20-48

All geometry, meshing and building runs in the C++ of libcourant.so; this module
converts arguments and results.  Options and results are dicts keyed like the
reference's objects (fmax, df, hmin, feed, monitor, air_above, air_side, npml, lossy,
thick_copper, y_window, ports=[{"net", "edge", "offset"}], excite).
"""
import ctypes as C
import math

import numpy as np

from . import lib as L

_P, _D, _I, _DP = C.c_void_p, C.c_double, C.c_int, C.POINTER(C.c_double)
_IP, _CP = C.POINTER(C.c_int), C.c_char_p
_U8P = C.POINTER(C.c_ubyte)
NAN = float("nan")
COPPER, DIELECTRIC = 0, 1
XLO, XHI = 0, 1


class AxisFeatures(C.Structure):
    """cem_axis_features (courant_board.h)."""
    _fields_ = [("lo", _D), ("hi", _D), ("nsheet", _I), ("sheet_x", _DP), ("sheet_metal_hi", _IP),
                ("nplanes", _I), ("planes", _DP), ("nspans", _I), ("spans", _DP), ("ncaps", _I), ("caps", _DP),
                ("nrefine", _I), ("refine", _DP), ("nextra", _I), ("extra", _DP)]


class BoardOptions(C.Structure):
    """cem_board_options (courant_board.h)."""
    _fields_ = [("fmax", _D), ("df", _D), ("hmin", _D), ("feed", _D), ("monitor", _D), ("air_above", _D),
                ("air_side", _D), ("npml", _I), ("lossy", _I), ("thick_copper", _I), ("y_window", _D),
                ("excite", _I)]


class BoardPort(C.Structure):
    """cem_board_port (courant_board.h)."""
    _fields_ = [("net_name", _CP), ("net", _I), ("edge", _I), ("offset", _D)]


SIGNATURES = {
    "cem_board_create": (_P, []),
    "cem_board_import_kicad": (_P, [_CP]),
    "cem_board_destroy": (None, [_P]),
    "cem_board_clone": (_P, [_P]),
    "cem_board_add_layer": (_I, [_P, _CP, _I, _D, _D, _D, _D, _I]),
    "cem_board_clear_layers": (_I, [_P]),
    "cem_board_add_track": (_I, [_P, _D, _D, _D, _D, _D, _CP, _I]),
    "cem_board_add_via": (_I, [_P, _D, _D, _D, _D, _CP, _CP, _I]),
    "cem_board_add_poly": (_I, [_P, _I, _DP, _CP, _I]),
    "cem_board_set_net": (_I, [_P, _I, _CP]),
    "cem_board_set_outline": (_I, [_P, _D, _D, _D, _D]),
    "cem_board_count": (_I, [_P, _I]),
    "cem_board_layer": (_I, [_P, _I, _CP, _I, _IP, _DP, _IP]),
    "cem_board_track": (_I, [_P, _I, _DP, _CP, _I, _IP]),
    "cem_board_via": (_I, [_P, _I, _DP, _CP, _CP, _I, _IP]),
    "cem_board_poly": (_I, [_P, _I, _DP, _I, _CP, _I, _IP]),
    "cem_board_net": (_I, [_P, _I, _IP, _CP, _I]),
    "cem_board_outline": (_I, [_P, _DP]),
    "cem_board_copper_at": (_I, [_P, _CP, _D, _D]),
    "cem_board_copper_row": (_I, [_P, _CP, _D, _I, _DP, _U8P]),
    "cem_board_copper_edges": (_I, [_P, _CP, _I, _DP, _IP, _I]),
    "cem_lambda_over_20": (_D, [_D, _D]),
    "cem_mesh_axis": (_I, [C.POINTER(AxisFeatures), _D, _D, _D, _DP, _I, _DP, _CP, _I]),
    "cem_mesh_with_pml": (_I, [_DP, _I, _I, _D, _I, _I, _DP]),
    "cem_mesh_estimate": (_I, [_DP, _I, _DP, _I, _DP, _I, _D, _I, _I, _DP]),
    "cem_board_plan": (_I, [_P, C.POINTER(BoardOptions), C.POINTER(BoardPort), _I, _IP, _DP, _DP, _DP]),
    "cem_board_build": (_P, [_P, C.POINTER(BoardOptions), C.POINTER(BoardPort), _I, _IP, _DP]),
}

_default = None


def _engine(cem=None):
    """The product library with the board functions bound (the default build if cem is None)."""
    global _default
    if cem is None:
        if _default is None:
            _default = L.Courant()
        cem = _default
    if not getattr(cem, "_board_bound", False):
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(cem.lib, name)
            fn.restype = res
            fn.argtypes = args
        cem._board_bound = True
    return cem


def _arr(v):
    a = np.ascontiguousarray(np.asarray(v, dtype=np.float64))
    return a, a.ctypes.data_as(_DP)


def _err(cem):
    return L.CourantError(cem.last_error())


def _opt(v):
    return NAN if v is None else float(v)


class Board:
    """A board model (kicad.ts:67-75) held by the library."""

    def __init__(self, cem, handle):
        if not handle:
            raise _err(cem)
        self.cem = cem
        self.lib = cem.lib
        self.h = handle

    def __del__(self):
        try:
            if self.h:
                self.lib.cem_board_destroy(self.h)
                self.h = None
        except Exception:
            pass

    @classmethod
    def empty(cls, cem=None):
        cem = _engine(cem)
        return cls(cem, cem.lib.cem_board_create())

    def copy(self):
        return Board(self.cem, self.lib.cem_board_clone(self.h))

    # ---- building a model ----
    def add_layer(self, name, type, thickness, eps_r=None, loss_tangent=None, roughness=None, mask=False):
        t = {"copper": COPPER, "dielectric": DIELECTRIC}.get(type, type)
        rc = self.lib.cem_board_add_layer(self.h, name.encode(), t, thickness, _opt(eps_r), _opt(loss_tangent),
                                          _opt(roughness), 1 if mask else 0)
        if rc < 0:
            raise _err(self.cem)
        return self

    def add_track(self, a, b, width, layer, net):
        self.lib.cem_board_add_track(self.h, a[0], a[1], b[0], b[1], width, layer.encode(), int(net))
        return self

    def add_via(self, at, drill, size, layers, net):
        self.lib.cem_board_add_via(self.h, at[0], at[1], drill, size, layers[0].encode(), layers[1].encode(), int(net))
        return self

    def add_poly(self, pts, layer, net):
        a, p = _arr(np.asarray(pts, dtype=np.float64).reshape(-1))
        self.lib.cem_board_add_poly(self.h, len(a) // 2, p, layer.encode(), int(net))
        return self

    def set_net(self, n, name):
        self.lib.cem_board_set_net(self.h, int(n), name.encode())
        return self

    def set_outline(self, lo, hi):
        self.lib.cem_board_set_outline(self.h, lo[0], lo[1], hi[0], hi[1])
        return self

    # ---- reading it back ----
    def _count(self, what):
        return self.lib.cem_board_count(self.h, what)

    @property
    def layers(self):
        """Bottom to top; eps_r / loss_tangent / roughness None where not given."""
        out = []
        for i in range(self._count(0)):
            name = C.create_string_buffer(256)
            t, m = C.c_int(), C.c_int()
            v = (C.c_double * 6)()
            self.lib.cem_board_layer(self.h, i, name, 256, C.byref(t), v, C.byref(m))
            opt = [None if math.isnan(x) else x for x in v[1:4]]
            out.append({"name": name.value.decode(), "type": "copper" if t.value == COPPER else "dielectric",
                        "thickness": v[0], "eps_r": opt[0], "loss_tangent": opt[1], "roughness": opt[2],
                        "mask": bool(m.value), "z0": v[4], "z1": v[5]})
        return out

    @property
    def copper_layers(self):
        return [l["name"] for l in self.layers if l["type"] == "copper"]

    @property
    def thickness(self):
        ls = self.layers
        return ls[-1]["z1"] if ls else 0.0

    @property
    def tracks(self):
        out = []
        for i in range(self._count(1)):
            v = (C.c_double * 5)()
            layer = C.create_string_buffer(256)
            net = C.c_int()
            self.lib.cem_board_track(self.h, i, v, layer, 256, C.byref(net))
            out.append({"a": (v[0], v[1]), "b": (v[2], v[3]), "width": v[4], "layer": layer.value.decode(),
                        "net": net.value})
        return out

    @property
    def vias(self):
        out = []
        for i in range(self._count(2)):
            v = (C.c_double * 4)()
            la, lb = C.create_string_buffer(256), C.create_string_buffer(256)
            net = C.c_int()
            self.lib.cem_board_via(self.h, i, v, la, lb, 256, C.byref(net))
            out.append({"at": (v[0], v[1]), "drill": v[2], "size": v[3],
                        "layers": (la.value.decode(), lb.value.decode()), "net": net.value})
        return out

    @property
    def polys(self):
        out = []
        for i in range(self._count(3)):
            layer = C.create_string_buffer(256)
            net = C.c_int()
            n = self.lib.cem_board_poly(self.h, i, None, 0, layer, 256, C.byref(net))
            xy = np.zeros(2 * n)
            self.lib.cem_board_poly(self.h, i, xy.ctypes.data_as(_DP), n, None, 0, None)
            out.append({"pts": xy.reshape(n, 2), "layer": layer.value.decode(), "net": net.value})
        return out

    @property
    def nets(self):
        out = {}
        for i in range(self._count(4)):
            num = C.c_int()
            name = C.create_string_buffer(256)
            self.lib.cem_board_net(self.h, i, C.byref(num), name, 256)
            out[num.value] = name.value.decode()
        return out

    @property
    def outline(self):
        v = (C.c_double * 4)()
        self.lib.cem_board_outline(self.h, v)
        return (v[0], v[1]), (v[2], v[3])

    # ---- copper geometry (kicad.ts:297-387) ----
    def copper_at(self, layer, p):
        return self.lib.cem_board_copper_at(self.h, layer.encode(), p[0], p[1]) == 1

    def copper_row(self, layer, y, xs):
        """copper_at(layer, (x, y)) for every x, from the row rasteriser."""
        a, p = _arr(xs)
        out = np.zeros(len(a), np.uint8)
        self.lib.cem_board_copper_row(self.h, layer.encode(), y, len(a), p, out.ctypes.data_as(_U8P))
        return out.astype(bool)

    def copper_edges(self, layer, axis):
        """[(x, "lo" | "hi")]: axis-aligned copper boundaries and the side the metal is on."""
        n = self.lib.cem_board_copper_edges(self.h, layer.encode(), axis, None, None, 0)
        x = np.zeros(n)
        m = (C.c_int * max(n, 1))()
        self.lib.cem_board_copper_edges(self.h, layer.encode(), axis, x.ctypes.data_as(_DP), m, n)
        return [(float(x[i]), "hi" if m[i] else "lo") for i in range(n)]

    # ---- stackup (load.ts:14-29) ----
    def stackup(self):
        """Layer rows top to bottom."""
        keys = ("name", "type", "thickness", "eps_r", "loss_tangent", "roughness", "mask")
        return [{k: l[k] for k in keys} for l in reversed(self.layers)]

    def apply_stackup(self, rows):
        """A copy of the board with its layers replaced by `rows` (top to bottom, same copper layers)."""
        cu = [r["name"] for r in rows if r["type"] == "copper"]
        have = list(reversed(self.copper_layers))
        if cu != have:
            raise ValueError(f"stackup copper layers {', '.join(cu)} do not match the board's {', '.join(have)}")
        b = self.copy()
        b.lib.cem_board_clear_layers(b.h)
        for r in reversed(rows):
            b.add_layer(r["name"], r["type"], r["thickness"], r.get("eps_r"), r.get("loss_tangent"),
                        r.get("roughness"), r.get("mask", False))
        return b


def port_nets(board, off_lo=0.0, off_hi=0.0):
    """Names of the nets with an x-directed track crossing both port planes (offsets measured
    into the board from its edges) (load.ts:31-38)."""
    (x0, _), (x1, _) = board.outline
    a, b = x0 + off_lo, x1 - off_hi
    tracks = board.tracks

    def across(net, xc):
        return any(t["net"] == net and abs(t["a"][1] - t["b"][1]) < 1e-9 and
                   min(t["a"][0], t["b"][0]) <= xc + 1e-6 and max(t["a"][0], t["b"][0]) >= xc - 1e-6 for t in tracks)

    return [name or str(n) for n, name in board.nets.items() if n and across(n, a) and across(n, b)]


def import_kicad(text, cem=None):
    """A .kicad_pcb text as a Board (kicad.ts:86-219)."""
    cem = _engine(cem)
    return Board(cem, cem.lib.cem_board_import_kicad(text.encode()))


def gcpw_board(length, er, cem=None):
    """The app's built-in GCPW line as a board model (web/src/scenarios/gcpwboard.ts:20-48):
    500 um strip, 200 um gaps on a 254 um core, 1 oz copper, a solid bottom plane and a row of
    0.3 mm stitching vias at 0.8 mm pitch on each side, barrels 300 um beyond the gaps."""
    h, w, s, tcu = 254e-6, 500e-6, 200e-6, 35e-6
    drill, pad, pitch, fence = 0.3e-3, 0.6e-3, 0.8e-3, 300e-6
    a = w / 2
    b = a + s
    yv = b + fence + drill / 2
    W = yv + 1.5e-3
    edge = W + 0.1e-3
    bd = Board.empty(cem)
    bd.add_layer("B.Cu", "copper", tcu)
    bd.add_layer("core", "dielectric", h, eps_r=er, loss_tangent=0.0037)
    bd.add_layer("F.Cu", "copper", tcu)
    bd.add_track((0, 0), (length, 0), w, "F.Cu", 1)
    i = 0
    while (i + 0.5) * pitch < length - 1e-9:
        for y in (-yv, yv):
            bd.add_via(((i + 0.5) * pitch, y), drill, pad, ("F.Cu", "B.Cu"), 2)
        i += 1

    def rect(x0, y0, x1, y1):
        return [(x0, y0), (x1, y0), (x1, y1), (x0, y1)]

    bd.add_poly(rect(0, -W, length, W), "B.Cu", 2)
    bd.add_poly(rect(0, -W, length, -b), "F.Cu", 2)
    bd.add_poly(rect(0, b, length, W), "F.Cu", 2)
    bd.set_outline((0, -edge), (length, edge))
    bd.set_net(1, GCPW_NET)
    bd.set_net(2, "GND")
    return bd


GCPW_NET = "RF"


# ---------------------------------------------------------------------------
# mesher (mesher.ts:47-156)
# ---------------------------------------------------------------------------
def lambda_over_20(fmax, er, cem=None):
    return _engine(cem).lib.cem_lambda_over_20(fmax, er)


def mesh_axis(lo, hi, hmin, hmax, sheet_edges=(), planes=(), spans=(), max_cell=(), refine=None, extra=(),
              ratio=1.3, cem=None):
    """Node lines of one axis: sheet_edges [(x, "lo" | "hi")], planes [x], spans [(a, b, n)],
    max_cell [(a, b, h)], refine [x] (None: every plane and sheet edge), extra [x].
    Returns (nodes, report) with report cells, min_cell, max_cell, max_ratio, warnings."""
    cem = _engine(cem)
    keep = []

    def arr(v, cols=None):
        a = np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(-1))
        keep.append(a)
        return a.ctypes.data_as(_DP)

    f = AxisFeatures()
    f.lo, f.hi = lo, hi
    f.nsheet = len(sheet_edges)
    f.sheet_x = arr([e[0] for e in sheet_edges])
    mh = np.ascontiguousarray([1 if e[1] == "hi" else 0 for e in sheet_edges] or [0], dtype=np.int32)
    keep.append(mh)
    f.sheet_metal_hi = mh.ctypes.data_as(_IP)
    f.nplanes, f.planes = len(planes), arr(list(planes))
    f.nspans, f.spans = len(spans), arr([list(s) for s in spans])
    f.ncaps, f.caps = len(max_cell), arr([list(m) for m in max_cell])
    f.nrefine = -1 if refine is None else len(refine)
    f.refine = arr(list(refine or []))
    f.nextra, f.extra = len(extra), arr(list(extra))
    rep = (C.c_double * 5)()
    warn = C.create_string_buffer(8192)
    n = cem.lib.cem_mesh_axis(C.byref(f), hmin, hmax, ratio, None, 0, rep, warn, 8192)
    if n < 0:
        raise _err(cem)
    nodes = np.zeros(n)
    cem.lib.cem_mesh_axis(C.byref(f), hmin, hmax, ratio, nodes.ctypes.data_as(_DP), n, None, None, 0)
    warnings = [w for w in warn.value.decode().split("\n") if w]
    return nodes, {"cells": int(rep[0]), "min_cell": rep[1], "max_cell": rep[2], "max_ratio": rep[3],
                   "warnings": warnings}


def with_pml(nodes, n, h, lo=True, hi=True, cem=None):
    cem = _engine(cem)
    a, p = _arr(nodes)
    out = np.zeros(len(a) + 2 * n)
    m = cem.lib.cem_mesh_with_pml(p, len(a), n, h, 1 if lo else 0, 1 if hi else 0, out.ctypes.data_as(_DP))
    return out[:m]


def estimate(x, y, z, t_sim, backend="gpu", npml=8, cem=None):
    """Pre-run estimate: cells, points, dt, steps, bytes, cell_steps (mesher.ts:139-156)."""
    cem = _engine(cem)
    xa, xp = _arr(x)
    ya, yp = _arr(y)
    za, zp = _arr(z)
    out = (C.c_double * 6)()
    if cem.lib.cem_mesh_estimate(xp, len(xa), yp, len(ya), zp, len(za), t_sim, 1 if backend == "gpu" else 0,
                                 npml, out) < 0:
        raise ValueError("estimate needs at least 2 nodes per axis")
    return {"cells": int(out[0]), "points": int(out[1]), "dt": out[2], "steps": int(out[3]), "bytes": out[4],
            "cell_steps": out[5]}


# ---------------------------------------------------------------------------
# board -> simulation (boardsim.ts:110-331)
# ---------------------------------------------------------------------------
def _options(fmax, df, hmin, feed=None, monitor=None, air_above=None, air_side=None, npml=8, lossy=False,
             thick_copper=False, y_window=0.0, excite=0):
    o = BoardOptions()
    o.fmax, o.df, o.hmin = fmax, df, hmin
    o.feed, o.monitor, o.air_above, o.air_side = _opt(feed), _opt(monitor), _opt(air_above), _opt(air_side)
    o.npml = -1 if npml is None else int(npml)
    o.lossy, o.thick_copper = int(bool(lossy)), int(bool(thick_copper))
    o.y_window = float(y_window or 0.0)
    o.excite = int(excite or 0)
    return o


def _ports(ports):
    arr = (BoardPort * max(len(ports), 1))()
    names = []
    for k, p in enumerate(ports):
        net = p["net"]
        if isinstance(net, str):
            names.append(net.encode())
            arr[k].net_name = names[-1]
            arr[k].net = 0
        else:
            arr[k].net_name = None
            arr[k].net = int(net)
        arr[k].edge = {"xlo": XLO, "xhi": XHI}[p["edge"]]
        arr[k].offset = float(p.get("offset", 0.0) or 0.0)
    return arr, names


def plan_board(board, ports, **opts):
    """Mesh of a board run without building it: {"cells": (nx, ny, nz), "mesh": {"x", "y", "z"}}."""
    lib = board.lib
    o = _options(**opts)
    pa, keep = _ports(ports)
    cells = (C.c_int * 3)()
    if lib.cem_board_plan(board.h, C.byref(o), pa, len(ports), cells, None, None, None) < 0:
        raise _err(board.cem)
    x, y, z = (np.zeros(cells[a] + 1) for a in range(3))
    lib.cem_board_plan(board.h, C.byref(o), pa, len(ports), cells, x.ctypes.data_as(_DP), y.ctypes.data_as(_DP),
                       z.ctypes.data_as(_DP))
    return {"cells": (cells[0], cells[1], cells[2]), "mesh": {"x": x, "y": y, "z": z}}


def build_board_sim(cem, board, ports, **opts):
    """The finalized simulation of a board run (boardsim.ts:198-311):
    {"sim", "freqs", "ports", "cells", "t_stop", "mesh"}; the sim is a courant_amd.lib.Sim."""
    cem = _engine(cem)
    if board.cem is not cem and board.cem.path != cem.path:
        raise ValueError("board and simulation must come from the same library")
    plan = plan_board(board, ports, **opts)
    o = _options(**opts)
    pa, keep = _ports(ports)
    ids = (C.c_int * len(ports))()
    t_stop = C.c_double()
    h = cem.lib.cem_board_build(board.h, C.byref(o), pa, len(ports), ids, C.byref(t_stop))
    if not h:
        raise _err(cem)
    m = plan["mesh"]
    sim = L.Sim(cem, m["x"], m["y"], m["z"], handle=h)
    nf = int(math.floor(opts["fmax"] / opts["df"] + 0.5))
    sim.freqs = np.array([(i + 1) * opts["df"] for i in range(nf)])
    tau = math.sqrt(6.9) / (math.pi * 1.25 * opts["fmax"])
    sim.excitation_mode = (4.5 * tau, tau)
    return {"sim": sim, "freqs": sim.freqs, "ports": list(ids), "cells": plan["cells"], "t_stop": t_stop.value,
            "mesh": m}


# ---------------------------------------------------------------------------
# board runs (web/tools/run-board.ts:53-85)
# ---------------------------------------------------------------------------
def run_excitation(cem, board, ports, excite, decay=1e-5, tmax=5e-9, steps=None, **opts):
    """One run of a board with port `excite` driven: build, then step in chunks of 500
    until the field energy is below `decay` of its peak after the drive or the run passes
    `tmax` (run-board.ts:57-72), or exactly `steps` steps when given.  Returns the waves of
    both ports ({"a", "b", "z0"} each), the frequencies and the step count."""
    bs = build_board_sim(cem, board, ports, excite=excite, **opts)
    sim = bs["sim"]
    try:
        if steps:
            sim.step(int(steps))
        else:
            peak = 0.0
            while True:
                sim.step(500)
                W = sim.energy()
                peak = max(peak, W)
                t = sim.steps * sim.dt
                if (t > bs["t_stop"] and W < decay * peak) or t > tmax:
                    break
        waves = [sim.mode_port_waves(p) for p in bs["ports"][:2]]
        return {"waves": waves, "freqs": bs["freqs"], "steps": sim.steps, "cells": bs["cells"]}
    finally:
        sim.close()


def board_s_matrix(cem, runs):
    """S (nf x 2 x 2) renormalised to 50 ohm from one run (port 1 driven: S22 = S11 and S12
    = S21 assumed, a symmetric reciprocal line) or two (port 2 driven as well), each port
    referred to its line impedance of the first run (run-board.ts:74-83); FP64 on the GPU."""
    from . import twoport
    w1, w2 = runs[0]["waves"]
    nf = len(runs[0]["freqs"])
    S = np.empty((nf, 2, 2), complex)
    S[:, 0, 0] = w1["b"] / w1["a"]
    S[:, 1, 0] = w2["b"] / w1["a"]
    if len(runs) > 1:
        v1, v2 = runs[1]["waves"]
        S[:, 0, 1] = v1["b"] / v2["a"]
        S[:, 1, 1] = v2["b"] / v2["a"]
    else:
        S[:, 0, 1] = S[:, 1, 0]
        S[:, 1, 1] = S[:, 0, 0]
    return twoport.renormalise(cem, S, w1["z0"], w2["z0"], 50.0)


def run_board(cem, board, net=None, planes=(0.0, 0.0), hmin=20e-6, fmax=40e9, df=0.5e9, lossy=False, sheets=False,
              full2=False, stackup=None, decay=1e-5, tmax=5e-9, window=None):
    """The command-line board run of the reference (run-board.ts:36-85) as a function: the
    two-port net between the left and right port planes (m into the board), both planes
    extruded into feeds, mode ports, one run (or two with full2), S renormalised to 50 ohm.
    Defaults are the tool's: hmin 20 um, fmax 40 GHz, df 0.5 GHz, copper of its stackup
    thickness (sheets=True: zero-thickness), energy decay 1e-5, at most 5 ns per run.
    Returns {"f", "S", "z0", "net", "planes", "cells", "steps", "wall", "full"}."""
    import time
    if stackup is not None:
        board = board.apply_stackup(stackup)
    if net is None:
        nets = port_nets(board, planes[0], planes[1])
        if not nets:
            raise ValueError("no net crosses both port planes")
        net = nets[0]
    ports = [{"net": net, "edge": "xlo", "offset": planes[0]}, {"net": net, "edge": "xhi", "offset": planes[1]}]
    opts = dict(fmax=fmax, df=df, hmin=hmin, lossy=lossy, thick_copper=not sheets, y_window=window or 0.0)
    t0 = time.perf_counter()
    runs = [run_excitation(cem, board, ports, e, decay=decay, tmax=tmax, **opts) for e in range(2 if full2 else 1)]
    S = board_s_matrix(cem, runs)
    return {"f": runs[0]["freqs"], "S": S, "z0": runs[0]["waves"][0]["z0"], "net": net, "planes": tuple(planes),
            "cells": runs[0]["cells"], "steps": sum(r["steps"] for r in runs), "wall": time.perf_counter() - t0,
            "full": bool(full2)}


def board_touchstone(result, name, version):
    """Touchstone text of a run_board result, titled as the reference's tool titles it."""
    from . import touchstone
    def js_str(v):          # Number -> String as JavaScript prints it (shortest round trip)
        return str(int(v)) if float(v).is_integer() else repr(float(v))

    title = "%s, net %s, planes %s mm" % (name, result["net"], "/".join(js_str(p * 1e3) for p in result["planes"]))
    return touchstone.write_s2p(result["f"], result["S"], title, version, full=result["full"], renormalised=True)
