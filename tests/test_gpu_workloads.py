"""The batched, served and large-board workloads on the GPU at their drawn sizes, against
fixtures written from the f64 reference core by tests/golden/make_fullsize_golden.py:

  * three seeded batched boards (serve.random_board(4001..4003): 1.78 M, 1.16 M and
    0.90 M cells, Debye 2-6 poles + SIBC copper, full runs of both excitations) solved
    concurrently: every entry of the 2 x 2 S matrix within 1e-4 dB and 1e-3 deg of the core's
    wherever |S| > -40 dB, and |dS| <= 1e-4 everywhere (the tolerance of
    web/test/ports.test.ts:62-64 tightened to the project's)
  * a served Poisson stream of four seeded request boards at their drawn sizes (0.83 M to
    1.43 M cells, lossy and lossless, 1- and 2-port) through the admission loop: each S
    matrix bit-identical to the same board solved alone, and the two boards with fixtures
    within the same tolerance of the core
  * the 38.8 M-point four-layer board after 600 steps: the port voltage and current spectra
    (801 frequencies) within 1e-4 of their largest magnitude, the field energy within 1e-5
"""
import os

import numpy as np
import pytest

from courant_amd import scenes as S
from courant_amd import serve

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def print_case(name):
    print(f"\n[gpu case] {name}", flush=True)


def s_errors(a, b):
    """(max dB error, max degree error) where |a| > -40 dB, and max |a - b| over all entries."""
    big = np.abs(a) > 10 ** (-40 / 20)
    ddb = np.abs(20 * np.log10(np.abs(b[big])) - 20 * np.log10(np.abs(a[big])))
    ddeg = np.abs(np.angle(b[big] / a[big], deg=True))
    return ddb.max(), ddeg.max(), np.abs(a - b).max()


def test_batched_boards_at_drawn_sizes_match_reference(cem, gpu_available):
    g = np.load(os.path.join(GOLDEN, "batched_cfg4.npz"))
    seeds = [int(x) for x in g["seeds"]]
    boards = [serve.random_board(sd) for sd in seeds]
    print_case("batched boards %s (%s cells), both excitations, 3 workers"
               % (seeds, [serve.cells(b) for b in boards]))
    got, wall, _ = serve.run_batch(cem, boards, 3)
    print("  %.2f s" % wall)
    for sd, b in zip(seeds, got):
        db, deg, ds = s_errors(g[f"S{sd}"], b)
        print("  board %d: |S| %.2e dB, %.2e deg, |dS| %.2e" % (sd, db, deg, ds))
        assert db <= 1e-4 and deg <= 1e-3 and ds <= 1e-4


def test_served_stream_at_drawn_sizes(cem, gpu_available):
    g = np.load(os.path.join(GOLDEN, "served_cfg5.npz"))
    seeds = [5033, 5002, 5037, 5022]
    reqs = [serve.random_request(sd) for sd in seeds]
    assert all(serve.cells(r) >= 0.5e6 for r in reqs)
    print_case("served stream %s (%s cells), alone" % (seeds, [serve.cells(r) for r in reqs]))
    alone = [serve.run_batch(cem, [r], 1)[0][0] for r in reqs]
    print_case("served stream, Poisson arrivals at 20 boards/s, 3 workers")
    lat, wall, got, _ = serve.serve(cem, reqs, rate=20.0, workers=3, budget_bytes=200e9, seed=5000)
    print("  %.2f s, p50 %.3f s, max %.3f s" % (wall, np.median(lat), np.max(lat)))
    for x, y in zip(alone, got):
        assert np.array_equal(x, y)
    for sd in (int(x) for x in g["seeds"]):
        db, deg, ds = s_errors(g[f"S{sd}"], alone[seeds.index(sd)])
        print("  board %d vs f64 core: |S| %.2e dB, %.2e deg, |dS| %.2e" % (sd, db, deg, ds))
        assert db <= 1e-4 and deg <= 1e-3 and ds <= 1e-4


def test_four_layer_board_600_steps_matches_reference(cem, gpu_available):
    g = np.load(os.path.join(GOLDEN, "board4_600.npz"))
    steps = int(g["steps"])
    print_case("four-layer board, 38.8 M points, %d steps: port spectra against the f64 core" % steps)
    b, info = S.board4(cem)
    try:
        b.step(steps)
        p1, p2 = info["ports"]
        for name, got in (("v1", b.port_v(p1)), ("i1", b.port_i(p1)), ("v2", b.port_v(p2)), ("i2", b.port_i(p2))):
            want = g[name]
            err = np.max(np.abs(got - want)) / np.max(np.abs(want))
            print("  %s: %.2e of its largest magnitude" % (name, err))
            assert err <= 1e-4
        de = abs(b.energy() - float(g["energy"])) / float(g["energy"])
        print("  energy: %.2e" % de)
        assert de <= 1e-5
    finally:
        b.close()
