"""CPU tests of the board front-end (include/courant_board.h through courant_amd/board.py),
with the checks the reference's own tests of that path make and closed forms:

  * mesher rules on the GCPW cross-section (web/test/mesher.test.ts): neighbour ratio
    <= 1.3, no warnings, strictly increasing nodes, thirds-rule lines at every sheet edge
    and none on it, planes on grid lines, >= 6 substrate cells no larger than lambda/20,
    the estimator runs
  * closed forms: a uniform size field gives ceil(L/h) equal cells; a cap bounds every
    cell inside it; with_pml adds exact cells; the estimate's dt is the 3-D CFL limit of
    the smallest cells
  * KiCad import (web/test/import.test.ts:20-25, kicad.test.ts, launch.test.ts:21-22):
    layers, counts, the F.Cu y-edges with their metal sides, arc tracks, rotated and
    shaped pads, plated holes, an arc outline corner, port nets
  * the row rasteriser answers as the point test everywhere, exactly on copper edges
    included (kicad.test.ts:56-72)
  * the built-in GCPW board (web/src/scenarios/gcpwboard.ts; builtin.test.ts:21-24) and a
    mask added through the stackup (thick.test.ts:21)
  * plan and build: port and copper planes on grid lines, CPML cells, the strip as PEC
    edges through the board and the feed, the reference's error messages
  * a C11 caller (tests/c/board_build.c) builds a board through the header and gets the
    Python path's mesh and ports
"""
import math
import os

import numpy as np
import pytest

from courant_amd import board as B
from courant_amd import lib as L

HERE = os.path.dirname(os.path.abspath(__file__))
BOARDS = os.path.join(HERE, "golden", "boards")
RF_PORTS = [{"net": "RF", "edge": "xlo"}, {"net": "RF", "edge": "xhi"}]


def fixture(name):
    return B.import_kicad(open(os.path.join(BOARDS, name + ".kicad_pcb")).read())


def has(arr, x):
    return bool(np.any(np.abs(np.asarray(arr) - x) < 1e-12))


def js_fixed(x, d):
    """Number.prototype.toFixed for ordinary values (no negative zero)."""
    return "%.*f" % (d, x + 0.0 if abs(x) >= 0.5 * 10 ** -d else 0.0)


def mm(x, y):
    return (x * 1e-3, -y * 1e-3)     # KiCad (mm, y down) -> board model


def test_mesher_rules_on_gcpw_cross_section():
    h, w, s, fence, er, fmax = 254e-6, 500e-6, 200e-6, 300e-6, 3.66, 40e9
    a = w / 2
    b = a + s
    yf = b + fence
    yext = yf + 1e-3
    hmin = 20e-6
    sub = B.lambda_over_20(fmax, er)
    assert sub == pytest.approx(299792458.0 / (fmax * math.sqrt(er)) / 20, rel=1e-15)
    Y, ry = B.mesh_axis(-yext, yext, hmin, 100e-6, sheet_edges=[(-b, "lo"), (-a, "hi"), (a, "lo"), (b, "hi")],
                        planes=[-yf, 0, yf], spans=[(-b, -a, 8), (a, b, 8)], refine=[-b, -a, a, b])
    Z, rz = B.mesh_axis(0, h + 1.5e-3, hmin, 100e-6, planes=[h], spans=[(0, h, 6)], max_cell=[(0, h, sub)],
                        refine=[h])
    for m, r in ((Y, ry), (Z, rz)):
        assert r["max_ratio"] <= 1.3 + 1e-6
        assert r["warnings"] == []
        assert np.all(np.diff(m) > 0)
        assert r["cells"] == len(m) - 1
    for x, side in ((a, -1), (-a, 1), (b, 1), (-b, -1)):
        assert has(Y, x + side * hmin / 3) and has(Y, x - side * 2 * hmin / 3)
        assert not has(Y, x)
    assert has(Z, h)
    assert has(Y, 0) and has(Y, yf) and has(Y, -yf)
    assert np.count_nonzero((Z > 0) & (Z < h)) + 1 >= 6
    dz = np.diff(Z)
    assert np.all(dz[Z[1:] <= h] <= sub * (1 + 1e-9))
    X, _ = B.mesh_axis(0, 10e-3, 50e-6, 50e-6, max_cell=[(0, 10e-3, 50e-6)])
    est = B.estimate(X, Y, Z, 0.5e-9, "gpu")
    assert est["cells"] > 0 and est["dt"] > 0


def test_uniform_field_and_caps_closed_form():
    L_, h = 10e-3, 50e-6
    X, r = B.mesh_axis(0, L_, h, h)
    n = math.ceil(L_ / h - 1e-6)
    assert len(X) == n + 1 and r["cells"] == n
    assert np.max(np.abs(np.diff(X) - L_ / n)) <= 1e-9 * L_ / n
    assert X[0] == 0 and X[-1] == L_
    # a cap of 25 um on [4, 6] mm inside a 100 um field: every cell inside it is <= 25 um
    # and the cells grade out with the neighbour ratio
    X, r = B.mesh_axis(0, L_, 20e-6, 100e-6, max_cell=[(4e-3, 6e-3, 25e-6)], refine=[])
    d = np.diff(X)
    inside = (X[:-1] >= 4e-3) & (X[1:] <= 6e-3)
    assert inside.sum() >= 80 and np.all(d[inside] <= 25e-6 * (1 + 1e-9))
    assert d.max() <= 100e-6 * (1 + 1e-9) and r["max_ratio"] <= 1.3 + 1e-6


def test_line_merging_and_span_warning():
    hmin = 20e-6
    # two planes closer than hmin/4 at the same priority: one line, with a warning
    X, r = B.mesh_axis(0, 1e-3, hmin, 100e-6, planes=[0.5e-3, 0.5e-3 + 3e-6])
    assert has(X, 0.5e-3) and not has(X, 0.5e-3 + 3e-6)
    assert len(r["warnings"]) == 1 and "closer than hmin/4; merged" in r["warnings"][0]
    # a sheet edge's thirds-rule line within hmin/4 of a plane: the plane wins, silently
    e = 0.502e-3     # metal below e: lines at e - hmin/3 (4.7 um from the plane) and e + 2 hmin/3
    X, r = B.mesh_axis(0, 1e-3, hmin, 100e-6, planes=[0.5e-3], sheet_edges=[(e, "lo")])
    assert has(X, 0.5e-3) and not has(X, e - hmin / 3) and has(X, e + 2 * hmin / 3) and r["warnings"] == []
    # a span narrower than its cell count allows at hmin / 2 cells
    X, r = B.mesh_axis(0, 1e-3, hmin, 100e-6, spans=[(0.5e-3, 0.52e-3, 8)])
    assert len(r["warnings"]) == 1 and r["warnings"][0].endswith("wanted 8")


def test_mesher_rejects_bad_input():
    import ctypes as C
    cem = B._engine()
    f = B.AxisFeatures()
    f.lo, f.hi, f.nsheet, f.nrefine = 0.0, 1e-3, 2, -1          # a count without its array
    assert cem.lib.cem_mesh_axis(C.byref(f), 20e-6, 1e-4, 1.3, None, 0, None, None, 0) == -1
    assert "without its array" in cem.last_error()
    with pytest.raises(L.CourantError, match="at least one cell"):
        B.mesh_axis(0, 1e-3, 20e-6, 1e-4, spans=[(0, 1e-4, 0)])
    with pytest.raises(L.CourantError, match="bad features"):
        B.mesh_axis(1e-3, 0, 20e-6, 1e-4)


def test_with_pml_and_estimate_closed_form():
    assert list(B.with_pml([0.0, 1.0, 3.0], 2, 0.5)) == [-1.0, -0.5, 0.0, 1.0, 3.0, 3.5, 4.0]
    assert list(B.with_pml([0.0, 1.0], 2, 0.5, lo=False)) == [0.0, 1.0, 1.5, 2.0]
    assert list(B.with_pml([0.0, 1.0], 1, 0.5, hi=False)) == [-0.5, 0.0, 1.0]
    x = np.linspace(0, 2e-3, 41)            # 50 um
    y = np.r_[0, 10e-6, np.linspace(20e-6, 1e-3, 50)]   # smallest 10 um
    z = np.linspace(0, 1e-3, 41)            # 25 um
    e = B.estimate(x, y, z, 1e-9, "gpu", npml=8)
    dt = 0.99 / (L.C0 * math.sqrt(1 / 50e-6 ** 2 + 1 / 10e-6 ** 2 + 1 / 25e-6 ** 2))
    assert e["dt"] == pytest.approx(dt, rel=1e-12)
    assert e["steps"] == math.ceil(1e-9 / e["dt"])
    assert e["cells"] == 40 * 51 * 40 and e["points"] == 41 * 52 * 41
    assert e["cell_steps"] == e["cells"] * e["steps"]
    # FP32 E, H, ca, cb per point and 4 psi per axis on 2 npml planes
    psi = 4 * 2 * 8 * (52 * 41 + 41 * 41 + 41 * 52)
    assert e["bytes"] == 4 * (12 * e["points"] + psi)


def test_kicad_import_of_the_gcpw_fixture():
    b = fixture("gcpw")
    assert b.copper_layers == ["B.Cu", "F.Cu"]
    core = [l for l in b.layers if l["type"] == "dielectric"][0]
    assert abs(core["thickness"] - 254e-6) < 1e-12 and core["eps_r"] == 3.66
    assert len(b.vias) == 24 and len(b.tracks) == 1 and len(b.polys) == 3
    ye = " ".join(js_fixed(x * 1e3, 3) + m for x, m in b.copper_edges("F.Cu", 1))
    assert ye == "-6.000hi -3.450lo -3.250hi -2.750lo -2.550hi 0.000lo"
    # z bottom-up: copper sheets at the dielectric surfaces
    assert [(l["name"], l["z0"], l["z1"]) for l in b.layers] == [("B.Cu", 0, 0), ("dielectric 1", 0, 254e-6),
                                                                  ("F.Cu", 254e-6, 254e-6)]


def test_kicad_import_of_board_constructs():
    b = fixture("features")
    (x0, y0), (x1, y1) = b.outline
    assert abs(x0) < 1e-12 and abs(x1 - 30e-3) < 1e-9 and abs(y1) < 1e-9 and abs(y0 + 30e-3) < 1e-9
    arc = [t for t in b.tracks if t["net"] == 1]
    assert len(arc) >= 18
    at = lambda layer, x, y: b.copper_at(layer, mm(x, y))   # noqa: E731
    assert at("F.Cu", 25, 15) and at("F.Cu", 25.15, 15) and not at("F.Cu", 25.3, 15) and not at("F.Cu", 15, 15)
    # footprint at (10, 10) rotated 90: a 2 x 1 round-rect pad becomes 1 wide, 2 tall
    assert at("F.Cu", 10, 10.95) and at("F.Cu", 10.45, 10) and not at("F.Cu", 10.47, 10.97)
    assert not at("F.Cu", 10.6, 10)
    th = [v for v in b.vias if abs(v["at"][0] - 10e-3) < 1e-9 and abs(v["at"][1] + 7e-3) < 1e-9]
    assert th and abs(th[0]["drill"] - 0.6e-3) < 1e-12 and th[0]["net"] == 2
    assert at("F.Cu", 10.5, 7) and at("B.Cu", 10.5, 7)
    assert at("F.Cu", 10, 13.95) and not at("F.Cu", 10.45, 13.95)      # oval pad: a stadium
    assert at("F.Cu", 14, 10.9) and not at("F.Cu", 14.3, 10.9)         # custom pad primitive


@pytest.mark.parametrize("name", ["features", "gcpw"])
def test_row_rasteriser_equals_point_test(name):
    b = fixture(name)
    (x0, y0), (x1, y1) = b.outline
    xs = [x0 + (x1 - x0) * i / 300 for i in range(301)]
    ys = [y0 + (y1 - y0) * j / 200 for j in range(201)]
    for q in b.polys:           # every copper vertex coordinate, to hit edges exactly
        xs += list(q["pts"][:, 0])
        ys += list(q["pts"][:, 1])
    for t in b.tracks:
        xs += [t["a"][0] - t["width"] / 2, t["a"][0] + t["width"] / 2]
        ys += [t["a"][1] - t["width"] / 2, t["a"][1] + t["width"] / 2]
    bad = n = 0
    for layer in b.copper_layers:
        for y in ys:
            row = b.copper_row(layer, y, xs)
            pt = np.array([b.copper_at(layer, (x, y)) for x in xs])
            bad += int(np.count_nonzero(row != pt))
            n += len(xs)
    assert n > 50000 and bad == 0


def test_launch_fixture_import_and_port_nets():
    b = fixture("launch")
    assert any(l["mask"] for l in b.layers) and len(b.vias) > 90
    assert B.port_nets(b) == [] and B.port_nets(b, 3.5e-3, 3.5e-3) == ["RF"]


def test_builtin_gcpw_board():
    b = B.gcpw_board(10e-3, 3.66)
    rows = sorted({round(v["at"][1], 9) for v in b.vias})
    assert len(b.vias) == 24 and len(rows) == 2
    assert all(abs(v["at"][1]) - v["drill"] / 2 - 450e-6 > -1e-12 for v in b.vias)
    assert all(l["thickness"] == 35e-6 for l in b.layers if l["type"] == "copper")
    assert b.nets == {1: "RF", 2: "GND"} and B.port_nets(b) == ["RF"]


def test_apply_stackup_adds_a_mask():
    b = fixture("gcpw")
    rows = [{"name": "F.Mask", "type": "dielectric", "mask": True, "thickness": 20e-6, "eps_r": 3.5,
             "loss_tangent": 0.02}] + b.stackup()
    m = b.apply_stackup(rows)
    top = m.layers[-1]
    assert top["name"] == "F.Mask" and top["mask"] and top["z0"] == 254e-6 and top["z1"] == pytest.approx(274e-6)
    assert len(b.layers) == 3 and len(m.layers) == 4 and len(m.vias) == 24      # the original is unchanged
    with pytest.raises(ValueError, match="do not match"):
        b.apply_stackup([r for r in rows if r["name"] != "F.Cu"])


def interior_ratio(v, npml, lo=True, hi=True):
    d = np.diff(v)
    d = d[(npml if lo else 0):(len(d) - npml if hi else len(d))]
    return float(np.max(np.maximum(d[1:] / d[:-1], d[:-1] / d[1:])))


def test_plan_of_the_gcpw_fixture():
    b = fixture("gcpw")
    p = B.plan_board(b, RF_PORTS, fmax=40e9, df=1e9, hmin=40e-6)
    x, y, z = p["mesh"]["x"], p["mesh"]["y"], p["mesh"]["z"]
    assert p["cells"] == (len(x) - 1, len(y) - 1, len(z) - 1)
    hp = min(B.lambda_over_20(40e9, 1), 200e-6) / 2
    # 8 CPML cells of hp at both x ends (ports on both) and both y ends, only at the top in z
    for v, lo in ((x, True), (y, True), (z, False)):
        assert np.allclose(np.diff(v)[-8:], hp, rtol=0, atol=1e-15)
        if lo:
            assert np.allclose(np.diff(v)[:8], hp, rtol=0, atol=1e-15)
    assert z[0] == 0
    # port planes at the board edges, 4 mm feeds
    assert has(x, 0.0) and has(x, 10e-3) and abs(x[8] + 4e-3) < 1e-15 and abs(x[-9] - 14e-3) < 1e-15
    # board edges and copper planes on grid lines; the strip's thirds-rule lines
    assert has(y, -6e-3) and has(y, 0.0) and has(z, 254e-6)
    for e, m in b.copper_edges("F.Cu", 1):
        if -6e-3 + 1e-9 < e < -1e-9:
            s = 1 if m == "hi" else -1
            assert has(y, e + s * 40e-6 / 3) and has(y, e - s * 2 * 40e-6 / 3) and not has(y, e)
    # this fixture's ground pours run to the board edges: the thirds rule then puts a line
    # hmin/3 inside the outline (the effect web/src/scenarios/gcpwboard.ts:8-10 describes)
    assert has(y, -6e-3 + 40e-6 / 3) and has(y, -40e-6 / 3)
    for v, lo, hi in ((x, True, True), (z, False, True)):
        assert interior_ratio(v, 8, lo, hi) <= 1.3 + 1e-6
    # the built-in board keeps its copper 0.1 mm inside the outline: graded everywhere
    g = B.plan_board(B.gcpw_board(10e-3, 3.66), RF_PORTS, fmax=40e9, df=1e9, hmin=20e-6, thick_copper=True)
    gx, gy, gz = g["mesh"]["x"], g["mesh"]["y"], g["mesh"]["z"]
    for v, lo in ((gx, True), (gy, True), (gz, False)):
        assert interior_ratio(v, 8, lo, True) <= 1.3 + 1e-6
    assert has(gz, 254e-6) and has(gz, 254e-6 + 35e-6)      # thick copper: its top face on a grid line
    # the plan is what the builder uses
    cem = L.Courant()
    bs = B.build_board_sim(cem, b, RF_PORTS, fmax=40e9, df=1e9, hmin=40e-6)
    try:
        assert bs["cells"] == p["cells"] and bs["ports"] == [0, 1]
        tau = math.sqrt(6.9) / (math.pi * 1.25 * 40e9)
        assert bs["t_stop"] == pytest.approx(10.5 * tau, rel=1e-15)
        assert list(bs["freqs"]) == [(i + 1) * 1e9 for i in range(40)]
        # the strip is PEC on the F.Cu plane through the board and the feeds (the
        # cross-section is extruded), the gap is not
        sim = bs["sim"]
        nx, ny, nz = p["cells"]
        cb = np.ctypeslib.as_array(cem.lib.cem_plan_array(sim.h, 1, 0), shape=(sim.N,)).reshape(nx + 1, ny + 1, nz + 1)
        k = int(np.argmin(np.abs(z - 254e-6)))
        jc, jg = int(np.argmin(np.abs(y + 3.0e-3))), int(np.argmin(np.abs(y + 3.35e-3)))
        for xq in (-2e-3, 1e-3, 5e-3, 9e-3, 12e-3):
            i = int(np.argmin(np.abs(x - xq)))
            assert cb[i, jc, k] == 0 and cb[i, jg, k] > 0
    finally:
        bs["sim"].close()


def test_board_errors_follow_the_reference():
    cem = L.Courant()
    b = B.Board.empty(cem)
    b.add_layer("B.Cu", "copper", 35e-6).add_layer("core", "dielectric", 254e-6, eps_r=3.66)
    b.add_layer("F.Cu", "copper", 35e-6).set_outline((0, -1e-3), (5e-3, 1e-3)).set_net(1, "RF")
    b.add_track((0, 0), (5e-3, 0), 0.3e-3, "F.Cu", 1)
    b.add_poly([(0, -1e-3), (1e-3, -1e-3), (1e-3, 1e-3), (0, 1e-3)], "B.Cu", 0)   # a partial floor
    with pytest.raises(L.CourantError, match="needs a solid bottom ground plane"):
        B.plan_board(b, RF_PORTS, fmax=40e9, df=1e9, hmin=40e-6)
    g = B.gcpw_board(10e-3, 3.66)
    with pytest.raises(L.CourantError, match="port planes leave less than 0.1 mm of board"):
        B.plan_board(g, [{"net": "RF", "edge": "xlo", "offset": 5e-3}, {"net": "RF", "edge": "xhi", "offset": 5e-3}],
                     fmax=40e9, df=1e9, hmin=40e-6)
    with pytest.raises(L.CourantError, match="net XYZ not found"):
        B.build_board_sim(cem, g, [{"net": "XYZ", "edge": "xlo"}], fmax=40e9, df=1e9, hmin=40e-6)
    with pytest.raises(L.CourantError, match="no x-directed track of net RF crosses the xlo port plane"):
        B.build_board_sim(cem, fixture("launch"), RF_PORTS, fmax=40e9, df=1e9, hmin=60e-6, y_window=3e-3)
    with pytest.raises(L.CourantError, match="not a kicad_pcb file"):
        B.import_kicad("(something_else)")
    with pytest.raises(L.CourantError, match="outline"):
        B.import_kicad("(kicad_pcb (net 0 \"\"))")
    with pytest.raises(L.CourantError, match="nested too deeply"):
        B.import_kicad("(" * 300 + ")" * 300)
    with pytest.raises(L.CourantError, match="unbalanced"):
        B.import_kicad("(kicad_pcb (net 0 \"\")")
    with pytest.raises(L.CourantError, match="hmin too small"):
        B.mesh_axis(0, 1.0, 1e-9, 1e-3)


def test_c_caller_builds_a_board(tmp_path):
    """tests/c/board_build.c, compiled as C11 against include/courant_board.h and the
    library: the mesh and ports it gets are the Python path's."""
    import subprocess
    root = os.path.dirname(HERE)
    exe = str(tmp_path / "board_build")
    lib_dir = os.path.join(root, "courant_amd")
    subprocess.run(["gcc", "-std=c11", "-O1", "-Wall", "-Werror", os.path.join(HERE, "c", "board_build.c"),
                    "-I" + os.path.join(root, "include"), "-L" + lib_dir, "-lcourant", "-Wl,-rpath," + lib_dir, "-lm",
                    "-o", exe], check=True)
    out = subprocess.run([exe, os.path.join(BOARDS, "msl.kicad_pcb")], stdout=subprocess.PIPE, text=True,
                         check=True, timeout=120).stdout.split()
    p = B.plan_board(fixture("msl"), RF_PORTS, fmax=40e9, df=1e9, hmin=40e-6)
    assert tuple(map(int, out[1:4])) == p["cells"] == tuple(map(int, out[5:8]))
    assert out[9:11] == ["0", "1"] and int(out[14]) == 40
    assert float(out[12]) == pytest.approx(10.5 * math.sqrt(6.9) / (math.pi * 1.25 * 40e9), rel=1e-15)
