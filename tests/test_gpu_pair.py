"""Both excitations of a 2-port as right-hand sides of shared launches
(cem_amd_clone_excitation, cem_amd_step_pair): the clone copies the plan instead of
running set-up again, and a pair step runs the H and E updates of both excitations as one
launch each.  Each simulation of a pair must equal the same excitation built and stepped
alone bit for bit (fields, Debye / SIBC / CPML state, port spectra), over schedules that
mix graph-replayed chunks and single steps, on a lossy lumped-port board (Debye + SIBC +
CPML) and a lossy mode-port GCPW."""
import numpy as np
import pytest

from courant_amd import lib as L
from courant_amd import scenes as S
from courant_amd import serve

pytestmark = pytest.mark.gpu

MODE_MINI = dict(hmin=40e-6, hmax=100e-6, dx=100e-6, length=1.5e-3, feed=1.2e-3, mon=0.6e-3, yair=0.3e-3,
                 zair=0.4e-3, ports="mode", lossy=True, tcu=35e-6, mask=25e-6, dc=False, df=4e9)
BOARD = dict(serve.random_board(4001), nx=84, ny=72, nz=60)


def print_case(name):
    print(f"\n[gpu case] {name}", flush=True)


def same(a, b, ports, tag):
    for c in range(6):
        assert np.array_equal(np.array(a.field(c)), np.array(b.field(c))), (tag, c)
    assert np.array_equal(a.checkpoint(), b.checkpoint()), tag
    for p in ports:
        assert np.array_equal(a.port_v(p), b.port_v(p)), (tag, p)
        assert np.array_equal(a.port_i(p), b.port_i(p)), (tag, p)
    assert np.array_equal(a.source_spectrum(), b.source_spectrum()), tag


@pytest.mark.parametrize("order", [None, "plain", "xcd", "plain4", "reversed"])
def test_pair_matches_each_excitation_alone_lumped(cem, gpu_available, monkeypatch, order):
    print_case(f"pair step vs alone: lossy lumped-port board (Debye + SIBC + CPML), tile order {order}")
    if order == "reversed":   # the H sweep from its last tile: the right-hand sides swap parity
        monkeypatch.setenv("COURANT_H_REVERSE", "1")
    elif order:   # the tile orders number the two right-hand sides' tiles differently
        monkeypatch.setenv("COURANT_TILE_ORDER", order)
    a, ports = serve.build(cem, BOARD, 0)
    b = a.clone_excitation(ports[1])
    a1, _ = serve.build(cem, BOARD, 0)
    b1, _ = serve.build(cem, BOARD, 1)
    try:
        assert b.dt == a.dt and b.steps == 0
        done = 0
        for n in (200, 37, 64):   # graph chunks, a remainder, a chunk that starts unaligned
            a.step_pair(b, n)
            a1.step_device(n)
            b1.step_device(n)
            done += n
            assert a.steps == b.steps == done
            same(a, a1, ports, ("a", done))
            same(b, b1, ports, ("b", done))
        # the clone drives the other port: its port-2 source spectrum is the driven one
        assert np.abs(b.port_v(ports[1])).max() > 0 and np.abs(a.port_v(ports[0])).max() > 0
    finally:
        for s in (a, b, a1, b1):
            s.close()


def test_pair_matches_each_excitation_alone_mode_ports(cem, gpu_available):
    print_case("pair step vs alone: lossy mode-port GCPW")
    a, info = S.gcpw(cem, **MODE_MINI)
    b = a.clone_excitation(1)
    a1, _ = S.gcpw(cem, **MODE_MINI)
    b1, _ = S.gcpw(cem, **dict(MODE_MINI, excite=2))
    try:
        for n in (130, 64):
            a.step_pair(b, n)
            a1.step_device(n)
            b1.step_device(n)
            same(a, a1, range(a.lib.cem_plan_nport(a.h)), ("a", n))
            same(b, b1, range(b.lib.cem_plan_nport(b.h)), ("b", n))
    finally:
        for s in (a, b, a1, b1):
            s.close()


def test_pair_refuses_other_plans_and_steps(cem, gpu_available):
    print_case("pair step: plan and step checks")
    a, ports = serve.build(cem, BOARD, 0)
    c, _ = serve.build(cem, BOARD, 1)
    b = a.clone_excitation(ports[1])
    try:
        with pytest.raises(L.CourantError, match="clone"):
            a.step_pair(c, 4)
        b.step_device(3)
        with pytest.raises(L.CourantError, match="same step"):
            a.step_pair(b, 4)
        with pytest.raises(L.CourantError, match="bad port"):
            a.clone_excitation(7)
    finally:
        for s in (a, b, c):
            s.close()
