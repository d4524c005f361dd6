"""Batched and served solves on the GPU.

  * a served board's full 2 x 2 S matrix against the f64 reference core's (committed
    fixture): |S| within 1e-4 dB and the angle within 1e-3 deg wherever |S| > -40 dB,
    and |dS| <= 1e-4 everywhere
  * every board's S matrix is bit-identical alone and in a concurrent mix (the kernels
    have no cross-simulation state and no order-dependent reductions)
"""
import os

import numpy as np
import pytest

from courant_amd import serve

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def print_case(name):
    print(f"\n[gpu case] {name}", flush=True)


def _mix():
    """Three small boards of different shape and physics (lossy 2-port GCPW, lossless
    1-port microstrip, lossy 2-port microstrip)."""
    from golden.make_golden import SERVED_MINI
    b = SERVED_MINI
    c = dict(serve.random_request(5001), nx=72, ny=40, nz=30, h_cells=6, lossy=False, nport=1,
             kind="microstrip", nf=21, steps=1200)
    d = dict(serve.random_board(4003), nx=56, ny=52, nz=34, h_cells=6, npole=2, kind="microstrip",
             nf=21, steps=1500)
    return [b, c, d]


def test_served_board_matches_reference(cem, gpu_available):
    print_case("served board 64x48x36, both excitations")
    from golden.make_golden import SERVED_MINI
    g = np.load(os.path.join(HERE, "golden", "served_mini.npz"))
    S, _, _ = serve.run_batch(cem, [SERVED_MINI], 1)
    a, b = g["S"], S[0]
    big = np.abs(a) > 10 ** (-40 / 20)
    ddb = np.abs(20 * np.log10(np.abs(b[big])) - 20 * np.log10(np.abs(a[big])))
    ddeg = np.abs(np.angle(b[big] / a[big], deg=True))
    print("max |S| err %.2e dB, %.2e deg, |dS| %.2e" % (ddb.max(), ddeg.max(), np.abs(a - b).max()))
    assert ddb.max() <= 1e-4
    assert ddeg.max() <= 1e-3
    assert np.abs(a - b).max() <= 1e-4


def test_results_identical_alone_and_in_the_mix(cem, gpu_available):
    boards = _mix()
    print_case("alone, one at a time")
    alone = [serve.run_batch(cem, [b], 1)[0][0] for b in boards]
    print_case("batched on 4 threads")
    batched, _, _ = serve.run_batch(cem, boards + boards[:1], 4)
    for x, y in zip(alone + alone[:1], batched):
        assert np.array_equal(x, y)
    print_case("served, Poisson arrivals, budget for two excitations at a time")
    budget = 2 * max(serve.estimate_bytes(b) for b in boards)
    lat, wall, served, _ = serve.serve(cem, boards[::-1] + boards, rate=50.0, workers=3,
                                       budget_bytes=budget, seed=5000)
    for x, y in zip(alone[::-1] + alone, served):
        assert np.array_equal(x, y)
    print("served %d boards in %.2f s, p50 %.3f s, max %.3f s" % (len(lat), wall, np.median(lat), lat.max()))
