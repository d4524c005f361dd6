"""Touchstone output and comparison (courant_amd/touchstone.py) and the 2-port assembly
(courant_amd/twoport.py), on the CPU."""
import numpy as np
import pytest

from courant_amd import touchstone as T
from courant_amd import twoport


def test_numbers_print_as_javascript_toexponential():
    # values as Number.prototype.toExponential prints them
    assert T.js_exp(1e9, 6) == "1.000000e+9"
    assert T.js_exp(40e9, 6) == "4.000000e+10"
    assert T.js_exp(0.0375, 9) == "3.750000000e-2"
    assert T.js_exp(-0.0, 9) == "0.000000000e+0"
    assert T.js_exp(123456.789, 6) == "1.234568e+5"
    assert T.js_exp(-2.5e-12, 9) == "-2.500000000e-12"


def test_write_read_round_trip_and_header():
    rng = np.random.default_rng(3)
    f = np.linspace(0.1e9, 40e9, 17)
    S = 0.4 * (rng.normal(size=(17, 2, 2)) + 1j * rng.normal(size=(17, 2, 2)))
    text = T.write_s2p(f, S, "line", "20261017.0", full=True, date="2026-01-01T00:00:00.000Z")
    head = text.splitlines()[:3]
    assert head == ["! Courant EM 20261017.0 — line", "! 2026-01-01T00:00:00.000Z; reference 50; both ports excited",
                    "# HZ S RI R 50"]
    f2, S2, R = T.read_s2p(text)
    assert R == 50.0 and np.array_equal(f2, f)
    assert np.max(np.abs(S2 - S)) <= 1e-9 * np.max(np.abs(S))
    half = T.write_s2p(f, S, "line", "v", full=False)
    _, S3, _ = T.read_s2p(half)
    assert np.allclose(S3[:, 0, 1], S3[:, 1, 0]) and np.allclose(S3[:, 1, 1], S3[:, 0, 0])
    assert "S22 = S11 and S12 = S21 assumed" in half.splitlines()[1]


def test_reader_accepts_ma_db_and_units():
    text = "# GHZ S MA R 50\n1.0 0.5 90 0.25 -45 0.25 -45 0.5 90\n"
    f, S, R = T.read_s2p(text)
    assert f[0] == 1e9 and abs(S[0, 0, 0] - 0.5j) < 1e-15
    text = "# MHZ S DB R 75\n! comment\n1000 -6.0205999132796239 0 0 0 0 0 -20 180\n"
    f, S, R = T.read_s2p(text)
    assert f[0] == 1e9 and R == 75.0 and abs(S[0, 0, 0] - 0.5) < 1e-15 and abs(S[0, 1, 1] + 0.1) < 1e-15


def test_common_reference_renormalisation_matches_the_core(ref):
    """renorm(S, R) with both ports at R equals the core's cem_s2_renormalize(S, R, R, 50)."""
    rng = np.random.default_rng(5)
    S = 0.5 * (rng.normal(size=(9, 2, 2)) + 1j * rng.normal(size=(9, 2, 2)))
    got = T.renorm(S, 75.0)
    for k in range(9):
        want = np.array(ref.renormalize2(list(S[k].ravel()), 75.0, 75.0, 50.0)).reshape(2, 2)
        np.testing.assert_allclose(got[k], want, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(twoport.renormalise(ref, S, np.full(9, 75.0), np.full(9, 75.0), 50.0), got,
                               rtol=1e-12, atol=1e-14)


def test_comparison_rules():
    f = np.linspace(1e9, 40e9, 60)
    w = 2 * np.pi * f
    s21 = 0.9 * np.exp(-1j * w * 50e-12)
    s11 = 0.2 * np.exp(-1j * w * 20e-12) * np.cos(w * 30e-12)
    S = np.empty((60, 2, 2), complex)
    S[:, 0, 0] = S[:, 1, 1] = s11
    S[:, 0, 1] = S[:, 1, 0] = s21
    same = T.compare((f, S, 50.0), (f, S.copy(), 50.0), "gpu")
    # resampling goes through dB and unwrapped phase, which rounds
    assert same["pass"] and all(c["value"] <= 1e-9 for c in same["checks"])
    off = S.copy()
    off[:, 1, 0] *= 10 ** (0.02 / 20)        # |S21| 0.02 dB off
    assert not T.compare((f, off, 50.0), (f, S, 50.0), "gpu")["pass"]
    assert T.compare((f, off, 50.0), (f, S, 50.0), "openems")["pass"]
    rot = S.copy()
    rot[:, 0, 1] *= np.exp(1j * np.radians(0.5))   # 0.5 deg on S12
    res = T.compare((f, rot, 50.0), (f, S, 50.0), "gpu")
    assert not res["pass"] and any(c["name"] == "phase S12 max difference" and not c["pass"] for c in res["checks"])
    # the same S at another reference impedance compares equal after renormalisation
    assert T.compare((f, T.renorm(S, 50.0, 75.0), 75.0), (f, S, 50.0), "gpu")["pass"]
