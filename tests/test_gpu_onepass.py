"""The one-pass step kernels (H and E updates of a step in one sweep, field sets alternating:
k_step with LDS tiles, COURANT_STEP=onepass; k_step_rows, which recomputes row j - 1's H per
lane, COURANT_STEP=rows) against the two separate sweeps: bit-identical fields, psi and Debye
state (the checkpoint bytes), port and probe spectra, after every stretch of a schedule
that mixes graph-replayed chunks, single steps, host reads and a host write; on scenes with
CPML on five faces, fused and post-kernel Debye currents, SIBC, lumped and mode ports, the
(ca, cb) class table and the full coefficient arrays, tiles that do not divide the rows and
rows split into k tiles."""
import numpy as np
import pytest

from courant_amd import lib as L
from courant_amd import scenes as S

pytestmark = pytest.mark.gpu

MODE_MINI = dict(hmin=40e-6, hmax=100e-6, dx=100e-6, length=1.5e-3, feed=1.2e-3, mon=0.6e-3, yair=0.3e-3,
                 zair=0.4e-3, ports="mode", lossy=True, tcu=35e-6, mask=25e-6, dc=False, df=4e9)


def print_case(name):
    print(f"\n[gpu case] {name}", flush=True)


def fields(sim):
    return [np.array(sim.field(c)) for c in range(6)]


def tall_cavity(cem, nz=200):
    """Debye-filled box of 30 x 22 x nz cells with CPML on all faces.  nz = 200: z pitch
    204, so a tile holds two whole rows of k and 23 rows leave a partial last tile;
    nz = 400: z pitch 404, so the rows split into two k tiles of three rows."""
    nx, ny = 30, 22
    x = [1e-4 * i for i in range(nx + 1)]
    y = [1e-4 * j for j in range(ny + 1)]
    z = [2e-5 * k for k in range(nz + 1)]
    s = cem.create_sim(x, y, z)
    for f in range(6):
        s.set_boundary(f, L.CPML, 6)
    s.set_debye_poles([2e9, 30e9])
    s.material_box_debye(x[8], x[22], y[5], y[16], z[40], z[160], 3.9, 0.0, [0.05, 0.03])
    p = s.add_lumped_port(L.Z, 15, 10, 90, 110, 1, 50)
    s.set_gaussian(p, 1.0, 40e-12, 10e-12)
    s.set_frequencies([5e9, 20e9])
    s.add_vprobe(L.Z, 12, 9, 80, 120, 1)
    s.add_iprobe(L.X, 12, 7, 11, 90, 100)
    s.finalize(0.95)
    return s


SCENES = {
    "debye_sibc": lambda cem: S.checkpoint_scene(cem),
    "mode_lossy": lambda cem: S.gcpw(cem, **MODE_MINI)[0],
    "tall_cavity": tall_cavity,
    "deep_cavity": lambda cem: tall_cavity(cem, nz=400),
}


def compare(a, b, tag):
    for c, (fa, fb) in enumerate(zip(fields(a), fields(b))):
        assert np.array_equal(fa, fb), (tag, c)
    assert np.array_equal(a.checkpoint(), b.checkpoint()), tag
    nport = a.lib.cem_plan_nport(a.h)
    for p in range(nport):
        assert np.array_equal(a.port_v(p), b.port_v(p)), (tag, p)
        assert np.array_equal(a.port_i(p), b.port_i(p)), (tag, p)
    assert np.array_equal(a.source_spectrum(), b.source_spectrum()), tag


@pytest.mark.parametrize("mode", ["onepass", "rows"])
@pytest.mark.parametrize("scene", sorted(SCENES))
@pytest.mark.parametrize("variant", ["classes", "arrays", "debye_list"])
def test_onepass_step_matches_separate_sweeps(cem, gpu_available, monkeypatch, scene, variant, mode):
    if variant != "classes" and scene != "debye_sibc":
        pytest.skip("coefficient arrays and post-kernel Debye run on the CPML + Debye + SIBC scene")
    print_case(f"one-pass step ({mode}) vs separate sweeps: {scene}, {variant}")
    if variant == "arrays":
        monkeypatch.setenv("COURANT_COEF_ARRAYS", "1")
    if variant == "debye_list":
        monkeypatch.setenv("COURANT_DEBYE_LIST", "1")
    monkeypatch.setenv("COURANT_STEP", mode)
    a = SCENES[scene](cem)
    a.energy()          # set-up reads the switches when the sim first runs on the GPU
    monkeypatch.delenv("COURANT_STEP")
    b = SCENES[scene](cem)
    b.energy()
    rows, planes = a.step_shape()
    assert rows >= (2 if mode == "onepass" else 1) and planes >= 1, (rows, planes)
    assert b.step_shape() == (0, 0)
    for n in (200, 37, 64, 1, 450):
        a.step(n)
        b.step(n)
        compare(a, b, n)
    # a host write of E between steps lands in the current field set of both
    e = a.field(L.Z)
    e[3, 4, 5] += 1e-3
    e = b.field(L.Z)
    e[3, 4, 5] += 1e-3
    a.step(65)
    b.step(65)
    compare(a, b, "after host write")


@pytest.mark.parametrize("mode", ["onepass", "rows"])
def test_onepass_launch_shape_is_bit_identical(cem, gpu_available, monkeypatch, mode):
    """The autotuner picks the x planes per workgroup of the one-pass step on the first steps;
    a run with a fixed, different shape gives bit-identical fields and spectra."""
    print_case(f"one-pass step ({mode}): autotuned vs fixed planes per workgroup")
    monkeypatch.setenv("COURANT_STEP", mode)
    a = S.checkpoint_scene(cem)
    a.step(300)
    rows, planes = a.step_shape()
    assert rows >= (2 if mode == "onepass" else 1) and planes in (4, 8, 16)
    monkeypatch.setenv("COURANT_ISEG_S", "3" if planes != 3 else "5")
    b = S.checkpoint_scene(cem)
    b.step(300)
    monkeypatch.delenv("COURANT_ISEG_S")
    monkeypatch.delenv("COURANT_STEP")
    assert b.step_shape()[1] != planes
    compare(a, b, "fixed shape")
