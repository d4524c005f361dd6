"""CPU tests of the C-ABI boundary: the library loads, exports every entry point
include/courant.h declares, and its error behaviour matches the reference's
(solver/src/sim.c:40-117, 244-340)."""
import ctypes
import os
import re

import numpy as np
import pytest

from courant_amd import lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"CEM_API[^;]*?\b(cem_\w+)\s*\(", text, re.S)))


def exported(path):
    lib = ctypes.CDLL(path)
    return lib


def test_header_declares_91_entry_points():
    names = declared("courant.h")
    assert len(names) == 91, names


def test_library_exports_every_declared_symbol():
    lib = exported(L.DEFAULT_LIB)
    missing = [n for n in declared("courant.h") + declared("courant_amd.h") + declared("courant_board.h")
               if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    from courant_amd import board
    assert set(declared("courant.h")) == set(L.SIGNATURES)
    assert set(declared("courant_amd.h")) == set(L.AMD_SIGNATURES)
    assert set(declared("courant_board.h")) == set(board.SIGNATURES)


def test_constants_and_version(cem):
    assert cem.version().startswith("20261017.0")
    assert L.C0 == 299792458.0 and L.MU0 == 1.25663706212e-6 and L.EPS0 == 8.8541878128e-12


def test_mode_port_def_layout():
    # 15 ints then 3 doubles at offset 64 on LP64 (courant.h:198-208)
    assert ctypes.sizeof(L.ModePortDef) == 88
    assert L.ModePortDef.x_ref.offset == 64


def test_error_paths(cem):
    lib = cem.lib
    x = np.linspace(0, 1e-3, 2)
    assert not lib.cem_create(1, 4, 4, x.ctypes.data_as(L._DP), x.ctypes.data_as(L._DP), x.ctypes.data_as(L._DP))
    assert "at least 2 cells" in cem.last_error()
    bad = np.array([0, 1e-3, 0.5e-3, 2e-3])
    with pytest.raises(L.CourantError, match="strictly increasing"):
        cem.create_sim(bad, np.linspace(0, 1e-3, 4), np.linspace(0, 1e-3, 4))
    s = cem.create_sim(np.linspace(0, 1e-3, 11), np.linspace(0, 1e-3, 11), np.linspace(0, 1e-3, 11))
    with pytest.raises(L.CourantError, match="bad face"):
        s.set_boundary(6, L.CPML, 2)
    with pytest.raises(L.CourantError, match="bad CPML thickness"):
        s.set_boundary(L.XLO, L.CPML, 5)
    with pytest.raises(L.CourantError, match="bad port definition"):
        s.add_lumped_port(L.Z, 5, 5, 3, 3, 1, 50)
    with pytest.raises(L.CourantError, match="bad probe"):
        s.add_vprobe(L.Z, 5, 50, 0, 3, 1)
    with pytest.raises(L.CourantError, match="bad current probe"):
        s.add_iprobe(L.X, 5, 3, 3, 1, 2)
    with pytest.raises(L.CourantError, match="bad cell box"):
        s.pec_cells(8, 0, 0, np.ones((4, 1, 1)))
    with pytest.raises(L.CourantError, match="set Debye poles first"):
        s.material_box_debye(0, 1, 0, 1, 0, 1, 2.0, 0.0, [0.1])
    with pytest.raises(L.CourantError, match="bad conductor-loss"):
        s.set_conductor_loss(-1, 0, 0, 0, 1e9, 2e9)
    s.finalize(0.99)
    with pytest.raises(L.CourantError, match="already finalized"):
        s.material_box(0, 1, 0, 1, 0, 1, 2.0)
    with pytest.raises(L.CourantError, match="already finalized"):
        s.finalize(0.99)


def test_port_on_pec_edge_is_refused(cem):
    g = np.linspace(0, 1e-3, 11)
    s = cem.create_sim(g, g, g)
    s.pec_box(0, 1, 0, 1, 0, 1)
    s.add_lumped_port(L.Z, 5, 5, 3, 5, 1, 50)
    with pytest.raises(L.CourantError, match="port edge is PEC"):
        s.finalize(0.99)


def test_error_string_is_per_thread(cem):
    import threading
    seen = {}

    def worker():
        x = np.linspace(0, 1e-3, 2)
        cem.lib.cem_create(1, 1, 1, x.ctypes.data_as(L._DP), x.ctypes.data_as(L._DP), x.ctypes.data_as(L._DP))
        seen["t"] = cem.last_error()

    cem.lib.cem_set_boundary  # main thread keeps its own message
    g = np.linspace(0, 1e-3, 11)
    s = cem.create_sim(g, g, g)
    with pytest.raises(L.CourantError):
        s.set_boundary(9, L.CPML, 1)
    t = threading.Thread(target=worker)
    t.start()
    t.join()
    assert "at least 2 cells" in seen["t"]
    assert cem.last_error() == "bad face"
