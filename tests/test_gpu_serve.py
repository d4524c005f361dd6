"""Batched and served solves on the GPU.

  * a served board's full 2 x 2 S matrix against the f64 reference core's (committed
    fixture): |S| within 1e-4 dB and the angle within 1e-3 deg wherever |S| > -40 dB,
    and |dS| <= 1e-4 everywhere
  * every board's S matrix is bit-identical alone and in a concurrent mix (the kernels
    have no cross-simulation state and no order-dependent reductions)
  * a full mode-port 2-port (both excitations, renormalised to 50 ohm on the GPU) against
    the f64 reference core: S21, S12 within 1e-4 dB / 1e-3 deg, |dS11|, |dS22| <= 1e-4, and
    its Touchstone file against the core's under the "gpu" rule of tools/compare_s2p.py
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


def test_pairs_identical_to_single_excitations(cem, gpu_available):
    """2-port boards solved as pairs (both excitations as right-hand sides of shared
    launches, one set-up) give the S matrices of their excitations solved one at a time,
    bit for bit, batched and served."""
    boards = _mix()
    alone = [serve.run_batch(cem, [b], 1)[0][0] for b in boards]
    print_case("pairs, batched on 3 threads")
    paired, _, _ = serve.run_batch(cem, boards, 3, pairs=True)
    for x, y in zip(alone, paired):
        assert np.array_equal(x, y)
    print_case("pairs, served, Poisson arrivals")
    budget = 2 * max(2 * serve.estimate_bytes(b) for b in boards)
    lat, wall, served, _ = serve.serve(cem, boards + boards[:1], rate=50.0, workers=3, budget_bytes=budget,
                                       seed=5001, pairs=True)
    for x, y in zip(alone + alone[:1], served):
        assert np.array_equal(x, y)


def test_full_two_port_touchstone_matches_reference(cem, ref, gpu_available):
    print_case("full 2-port, mode ports, lossy gcpw, 3000 steps per excitation")
    from courant_amd import scenes as S
    from courant_amd import touchstone, twoport
    from test_gpu_parity import MODE_MINI
    a = twoport.run(ref, lambda e: S.gcpw(ref, **dict(MODE_MINI, excite=e)), 3000)
    b = twoport.run(cem, lambda e: S.gcpw(cem, **dict(MODE_MINI, excite=e)), 3000)
    # transmission (S21, S12) in dB and degrees, reflection (S11, S22) as |dS|: the
    # tolerances of the rest of the suite
    ta, tb = a["S"][:, [1, 0], [0, 1]], b["S"][:, [1, 0], [0, 1]]
    ddb = np.abs(20 * np.log10(np.abs(tb)) - 20 * np.log10(np.abs(ta)))
    ddeg = np.abs(np.angle(tb / ta, deg=True))
    drefl = np.abs(b["S"][:, [0, 1], [0, 1]] - a["S"][:, [0, 1], [0, 1]])
    print("max |S21|, |S12| err %.2e dB, %.2e deg; |dS11|, |dS22| %.2e" % (ddb.max(), ddeg.max(), drefl.max()))
    assert ddb.max() <= 1e-4 and ddeg.max() <= 1e-3 and drefl.max() <= 1e-4
    # the line is mirror-symmetric: reciprocal and symmetric up to rounding
    assert np.max(np.abs(b["S"][:, 0, 1] - b["S"][:, 1, 0])) <= 1e-5 * np.max(np.abs(b["S"][:, 1, 0]))
    assert np.max(np.abs(b["S"][:, 0, 0] - b["S"][:, 1, 1])) <= 1e-5 * np.max(np.abs(b["S"][:, 0, 0]))
    fa = touchstone.read_s2p(twoport.to_touchstone(a, "mode-port gcpw", ref.version()))
    fb = touchstone.read_s2p(twoport.to_touchstone(b, "mode-port gcpw", cem.version()))
    res = touchstone.compare(fb, fa, "gpu")
    for c in res["checks"]:
        print("  %s: %.3g %s" % (c["name"], c["value"], c["unit"]))
    assert res["pass"]
