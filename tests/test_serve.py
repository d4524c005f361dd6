"""Batched and served solves: board generators, S-matrix assembly, admission of the request
loop (CPU), and the served board's fixture against the compiled reference core."""
import os
import threading
import time

import numpy as np
import pytest

from courant_amd import lib as L
from courant_amd import serve

HERE = os.path.dirname(os.path.abspath(__file__))


def test_batched_boards_are_seeded_and_in_range():
    boards = [serve.random_board(4000 + b) for b in range(64)]
    assert boards[5] == serve.random_board(4005)
    for b in boards:
        assert 80 <= b["nx"] <= 256 and 80 <= b["ny"] <= 192 and 60 <= b["nz"] <= 96
        assert 0.1e-3 <= b["width"] <= 0.5e-3 and 3.0 <= b["er"] <= 4.5 and 2 <= b["npole"] <= 6
        assert b["kind"] in ("gcpw", "microstrip") and b["nport"] == 2
    assert {b["kind"] for b in boards} == {"gcpw", "microstrip"}
    assert len(serve.jobs_of(boards)) == 128


def test_served_requests_follow_the_lognormal_draw():
    reqs = [serve.random_request(5000 + r) for r in range(400)]
    c = np.array([serve.cells(b) for b in reqs], float)
    # nx, ny are rounded from the drawn cell count: within a few percent of [0.5 M, 20 M]
    assert c.min() >= 0.45e6 and c.max() <= 21e6
    assert abs(np.log(np.median(c)) - np.log(3e6)) < 0.25
    assert 0.4 < np.mean([b["lossy"] for b in reqs]) < 0.6
    assert 0.4 < np.mean([b["nport"] == 2 for b in reqs]) < 0.6


def test_estimate_grows_with_the_board():
    b = serve.random_board(4000)
    big = dict(b, nx=2 * b["nx"])
    assert serve.estimate_bytes(big) > serve.estimate_bytes(b)
    assert serve.estimate_bytes(dict(b, lossy=False)) < serve.estimate_bytes(b)
    # at least the six FP32 field arrays
    assert serve.estimate_bytes(b) > 24 * (b["nx"] + 1) * (b["ny"] + 1) * (b["nz"] + 1)


def test_s_matrix_recovers_a_known_two_port():
    """Port V/I made from a known S and known incident waves (V = sqrt Z (a + b),
    I = (a - b) / sqrt Z, b = S a) give S back; one excitation gives b / a1."""
    rng = np.random.default_rng(1)
    nf, z = 7, 50.0
    S = rng.normal(size=(nf, 2, 2)) + 1j * rng.normal(size=(nf, 2, 2))
    A = rng.normal(size=(nf, 2, 2)) + 1j * rng.normal(size=(nf, 2, 2))
    vis = []
    for q in range(2):
        a = A[:, :, q]
        b = np.einsum("fij,fj->fi", S, a)
        v = np.sqrt(z) * (a + b)
        i = (a - b) / np.sqrt(z)
        vis.append((v.T, i.T))
    np.testing.assert_allclose(serve.s_matrix(vis, z), S, rtol=1e-12, atol=1e-12)
    one = serve.s_matrix(vis[:1], z)
    s11, s21 = L.s_params(vis[0][0][0], vis[0][1][0], vis[0][0][1], vis[0][1][1], z)
    np.testing.assert_allclose(one[:, 0, 0], s11, rtol=1e-13)
    np.testing.assert_allclose(one[:, 1, 0], s21, rtol=1e-13)


def _fake_solver(monkeypatch, hold=0.05):
    """Replace the device solve by a timed stand-in that records the admitted footprint."""
    state = {"in_use": 0, "peak": 0, "order": []}
    lock = threading.Lock()

    def fake(cem, b, e):
        need = serve.estimate_bytes(b)
        with lock:
            state["in_use"] += need
            state["peak"] = max(state["peak"], state["in_use"])
            state["order"].append((b["seed"], e))
        time.sleep(hold)
        with lock:
            state["in_use"] -= need
        nf = b["nf"]
        # independent columns: port 2 sees 2x (excitation 1) or 3x (excitation 2) port 1
        v = np.ones((2, nf), complex) * np.array([[1.0], [2.0 + e]]) * (1 + b["seed"] + e)
        return (v, v / 60.0), 1

    monkeypatch.setattr(serve, "solve", fake)
    return state


def test_admission_keeps_the_footprint_under_budget_in_arrival_order(monkeypatch):
    state = _fake_solver(monkeypatch)
    boards = [serve.random_board(4000 + b) for b in range(6)]
    budget = 2.2 * max(serve.estimate_bytes(b) for b in boards)
    srv = serve.Server(None, workers=6, budget_bytes=budget)
    try:
        handles = [srv.submit(b) for b in boards]
        out = [h() for h in handles]
    finally:
        srv.close()
    assert state["peak"] <= budget
    # jobs start first come, first served
    assert state["order"] == [(b["seed"], e) for b in boards for e in (0, 1)]
    assert all(o[0].shape == (boards[0]["nf"], 2, 2) and o[1] > 0 for o in out)
    assert srv.in_use == 0


def test_oversize_board_is_refused(monkeypatch):
    _fake_solver(monkeypatch)
    b = serve.random_board(4000)
    srv = serve.Server(None, workers=2, budget_bytes=serve.estimate_bytes(b) - 1)
    try:
        with pytest.raises(ValueError):
            srv.submit(b)
    finally:
        srv.close()


def test_serve_returns_every_board(monkeypatch):
    _fake_solver(monkeypatch, hold=0.01)
    boards = [serve.random_request(5000 + r) for r in range(8)]
    lat, wall, S, upd = serve.serve(None, boards, rate=200.0, workers=4, budget_bytes=1 << 40)
    assert len(lat) == len(S) == 8 and np.all(lat > 0) and wall > 0
    for b, s in zip(boards, S):
        assert s.shape == (b["nf"], 2, len(serve.excitations(b)))


def test_served_board_fixture_reproduces_on_the_reference(ref):
    """The served board's committed S matrix is what the f64 reference core computes."""
    from golden.make_golden import SERVED_MINI
    g = np.load(os.path.join(HERE, "golden", "served_mini.npz"))
    S, _, _ = serve.run_batch(ref, [SERVED_MINI], 1)
    np.testing.assert_array_equal(S[0], g["S"])
    # a passive, reciprocal 2-port: S12 = S21 (the board is mirror-symmetric in x) and
    # no gain at any frequency
    np.testing.assert_allclose(g["S"][:, 0, 1], g["S"][:, 1, 0], rtol=1e-9)
    assert np.all(np.abs(g["S"]) < 1.0 + 1e-6)


def board_fixture(name):
    from courant_amd import board as B
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "boards", name + ".kicad_pcb")
    return B.import_kicad(open(path).read())


def test_board_request_plans_its_mesh():
    from courant_amd import board as B
    b = board_fixture("msl")
    r = serve.board_request(b, hmin=40e-6, df=1e9, sheets=True)
    p = B.plan_board(b, r["ports"], **r["opts"])
    assert (r["nx"], r["ny"], r["nz"]) == p["cells"] and serve.cells(r) == np.prod(p["cells"])
    assert r["ports"][0]["net"] == "RF" and serve.excitations(r) == (0, 1)
    lossy = serve.board_request(b, hmin=40e-6, df=1e9, sheets=True, lossy=True)
    n = (r["nx"] + 1) * (r["ny"] + 1) * (r["nz"] + 1)
    assert 27 * n < serve.estimate_bytes(r) < serve.estimate_bytes(lossy)
    assert serve.excitations(serve.board_request(b, hmin=40e-6, df=1e9, full2=False)) == (0,)
    with pytest.raises(ValueError, match="no net crosses"):     # the launch pads stop the line at the edges
        serve.board_request(board_fixture("launch"))


def test_board_s_matrix_assembly(ref):
    """Wave ratios of one or two runs, renormalised to 50 ohm through the reference core's
    cem_s2_renormalize: the identity when both lines are 50 ohm, the symmetric fill of a
    single run (S22 = S11, S12 = S21) as the reference's board tool assumes, and a matched
    25 ohm line seen from 50 ohm."""
    from courant_amd import board as B
    nf = 5
    rng = np.random.default_rng(7)
    a1, a2 = rng.normal(size=nf) + 1j * rng.normal(size=nf), rng.normal(size=nf) + 1j * rng.normal(size=nf)
    S = 0.3 * (rng.normal(size=(nf, 2, 2)) + 1j * rng.normal(size=(nf, 2, 2)))
    f = np.arange(1, nf + 1) * 1e9
    z50 = np.full(nf, 50.0)
    run1 = {"freqs": f, "waves": [{"a": a1, "b": S[:, 0, 0] * a1, "z0": z50}, {"a": 0 * a1, "b": S[:, 1, 0] * a1, "z0": z50}]}
    run2 = {"freqs": f, "waves": [{"a": 0 * a2, "b": S[:, 0, 1] * a2, "z0": z50}, {"a": a2, "b": S[:, 1, 1] * a2, "z0": z50}]}
    assert np.allclose(B.board_s_matrix(ref, [run1, run2]), S, rtol=0, atol=1e-13)
    one = B.board_s_matrix(ref, [run1])
    for got, want in ((one[:, 0, 0], S[:, 0, 0]), (one[:, 1, 1], S[:, 0, 0]), (one[:, 0, 1], S[:, 1, 0]),
                      (one[:, 1, 0], S[:, 1, 0])):
        assert np.allclose(got, want, rtol=0, atol=1e-13)
    # a matched, lossless 25 ohm line between 50 ohm references: Gamma = (25 - 50) / (25 + 50)
    # at the input, and S11 = Gamma (1 - T^2) / (1 - Gamma^2 T^2), S21 = T (1 - Gamma^2) / (1 - Gamma^2 T^2)
    z25 = np.full(nf, 25.0)
    T = np.exp(-0.3j * np.arange(1, nf + 1))     # T = 1 would make I - S singular
    line = {"freqs": f, "waves": [{"a": a1, "b": 0 * a1, "z0": z25}, {"a": 0 * a1, "b": T * a1, "z0": z25}]}
    r = B.board_s_matrix(ref, [line])
    g = -1 / 3
    d = 1 - g * g * T * T
    assert np.allclose(r[:, 0, 0], g * (1 - T * T) / d, rtol=0, atol=1e-13)
    assert np.allclose(r[:, 1, 0], T * (1 - g * g) / d, rtol=0, atol=1e-13)
