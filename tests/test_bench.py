"""The bench contract on the CPU: one GPU per run, and the byte models the roofline is
priced with, on the four-layer board's own plan (built by the host set-up here)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and not r.stdout


@pytest.fixture(scope="module")
def board_counts(cem):
    from courant_amd import scenes
    sim, info = scenes.board4(cem)
    try:
        yield bench.plan_counts(sim), info
    finally:
        sim.close()


def test_survey_byte_model_of_the_board(board_counts):
    """72 B per grid point (E, H read and written; ca, cb read), 8 B per psi entry, 12 B per
    pole of a dispersive edge, (8K + 36) B per SIBC edge: 4.394 GB a step on the 600 x 400 x
    160 board (the figure the headline fraction of 8 TB/s is priced on)."""
    pc, info = board_counts
    assert tuple(info["cells"]) == (600, 400, 160)
    assert pc["N"] == 601 * 401 * 161 and pc["npole"] == 6 and pc["ndisp"] == 19287641
    assert pc["nsibc"] == 40515 and pc["nterm"] > 0
    assert bench.step_bytes(pc) == 4393731100


def test_kernel_byte_models_follow_the_class_storage(board_counts):
    """The E update's bytes drop by 3 B a point when it reads class bytes instead of class
    words, and by 24 - 4 when classes replace the ca, cb arrays; H and post do not change."""
    pc, _ = board_counts
    one, four, arrays = (bench.kernel_bytes(pc, c) for c in (1, 4, 24))
    assert four["E"] - one["E"] == 3 * pc["N"]
    assert arrays["E"] - four["E"] == 20 * pc["N"]
    assert one["H"] == four["H"] == arrays["H"] and one["post"] == four["post"]
