"""GPU runs at the full sizes of the benchmark configurations, checked through properties
that do not depend on size, and against the f64 reference core over first steps it
finishes in seconds:

  * the lossless GCPW with lumped ports (120 x 80 x 60 cells, 20 000 steps, 401
    frequencies): the source spectrum equals the Gaussian drive's closed-form transform to
    1e-8 of its peak; the line conserves power (|S11|^2 + |S21|^2 within 2 % of 1) from 1 to
    40 GHz
  * the lossy GCPW with mode ports (300 x 152 x 99 cells; 10 output frequencies, so that the
    ports' 2D mode sweep takes seconds, not the minutes of 400): the modal source spectrum equals
    the closed-form transform of the differentiated Gaussian tau d/dt exp(-x^2)
    (modeport.c:124-128), j w tau^2 sqrt(pi) exp(-(w tau / 2)^2) exp(-j w t0), to 1e-7 of its
    peak; after the run it is passive with a loss that grows with frequency; the same board
    with lumped ports matches the f64 core over 200 steps (every field within 1e-4 of its
    largest value, energy within 1e-6)
  * the four-layer board (38.8 M points): every field, the energy and both ports' spectra
    against the f64 core after 200 steps of a drive started at its peak
"""
import numpy as np
import pytest

from courant_amd import scenes as S

pytestmark = pytest.mark.gpu


def print_case(name):
    print(f"\n[gpu case] {name}", flush=True)


def fields(s):
    return [np.array(s.field(c)) for c in range(6)]


def compare_fields(a, b):
    """Largest deviation of each field from the reference, relative to the largest value of
    all six, and the energy's relative deviation."""
    fa, fb = fields(a), fields(b)
    scale = max(np.max(np.abs(f)) for f in fa)
    dev = max(np.max(np.abs(x - y)) for x, y in zip(fa, fb)) / scale
    return dev, abs(b.energy() - a.energy()) / a.energy()


def gaussian_transform(amp, t0, tau, freqs):
    w = 2 * np.pi * np.asarray(freqs)
    return amp * tau * np.sqrt(np.pi) * np.exp(-(w * tau / 2) ** 2) * np.exp(-1j * w * t0)


def test_lossless_gcpw_full_size(cem, gpu_available):
    print_case("lossless GCPW, lumped ports, 120 x 80 x 60 cells, 20000 steps")
    s, info = S.gcpw(cem, **S.CFG1)
    try:
        assert tuple(info["cells"]) == (120, 80, 60) and len(info["freqs"]) == 401
        s.step(20000)
        amp, t0, tau = s.excitation_lumped
        want = gaussian_transform(amp, t0, tau, info["freqs"])
        err = np.max(np.abs(s.source_spectrum() - want)) / np.max(np.abs(want))
        s11, s21 = S.gcpw_sparams(s, info)
        band = (info["freqs"] >= 1e9) & (info["freqs"] <= 40e9)
        power = np.abs(s11[band]) ** 2 + np.abs(s21[band]) ** 2
        print("source spectrum %.2e of the peak; |S11| max %.1f dB; power balance %.4f..%.4f" %
              (err, 20 * np.log10(np.abs(s11[band]).max()), power.min(), power.max()))
        assert err <= 1e-8
        assert np.all(np.abs(power - 1) < 0.02)
    finally:
        s.close()


def test_lossy_mode_port_gcpw_full_size(cem, gpu_available):
    print_case("lossy GCPW, mode ports, 300 x 152 x 99 cells: source spectrum and S after 20000 steps")
    s, info = S.gcpw(cem, **dict(S.CFG2, df=4e9))
    try:
        assert tuple(info["cells"]) == (300, 152, 99)
        t0, tau = s.excitation_mode
        s.step(int(np.ceil((t0 + 7 * tau) / s.dt)))
        w = 2 * np.pi * info["freqs"]
        want = 1j * w * tau * gaussian_transform(1.0, t0, tau, info["freqs"])
        err = np.max(np.abs(s.source_spectrum() - want)) / np.max(np.abs(want))
        print("modal source spectrum %.2e of the peak" % err)
        assert err <= 1e-7
        s.step(20000 - s.steps)
        print("20000 steps done; port mode sweeps", flush=True)
        p1, p2 = info["ports"]
        u1, u2 = s.mode_port_waves(p1), s.mode_port_waves(p2)
        s11, s21 = np.abs(u1["b"] / u1["a"]), np.abs(u2["b"] / u1["a"])
        f = info["freqs"]
        lo, hi = np.argmin(np.abs(f - 5e9)), np.argmin(np.abs(f - 40e9))
        print("|S21| %.4f at 5 GHz, %.4f at 40 GHz; |S11|^2 + |S21|^2 max %.4f" %
              (s21[lo], s21[hi], np.max(s11 ** 2 + s21 ** 2)))
        assert np.all(s11 ** 2 + s21 ** 2 < 1 + 1e-3)
        assert s21[hi] < s21[lo] < 1
    finally:
        s.close()


def test_lossy_gcpw_full_size_matches_reference(cem, ref, gpu_available):
    print_case("lossy GCPW at full size with lumped ports: 200 steps against the f64 core")
    p = dict(S.CFG2, ports="lumped", probes=False)
    a, _ = S.gcpw(ref, **p)
    b, _ = S.gcpw(cem, **p)
    try:
        a.step(200)
        b.step(200)
        dev, de = compare_fields(a, b)
        print("cells %s: field deviation %.2e of the largest, energy %.2e" % (b.n, dev, de))
        assert dev <= 1e-4 and de <= 1e-6
    finally:
        a.close()
        b.close()


def test_four_layer_board_full_size_matches_reference(cem, ref, gpu_available):
    """The headline board at full size against the f64 core over 200 steps of a drive started
    at its peak: every field within 1e-4 of its largest value, the energy within 1e-6, and both
    ports' voltage and current spectra (801 frequencies, the DFT of those 200 steps) within
    1e-4 of their largest magnitude.  The f64 core takes about a minute on the host cores; it
    runs in stretches of 50 steps that each print a line."""
    print_case("four-layer board, 38.8 M points: 200 steps against the f64 core, fields and port spectra")
    a, ia = S.board4(ref)
    b, ib = S.board4(cem)
    try:
        tau = ia["tau"]
        for s, info in ((a, ia), (b, ib)):
            s.set_gaussian(info["ports"][0], 1.0, 0.0, tau)     # the drive at its peak from the start
        for done in range(50, 201, 50):
            a.step(50)
            print(f"f64 core: {done} steps", flush=True)
        b.step(200)
        dev, de = compare_fields(a, b)
        print("cells %s: field deviation %.2e of the largest, energy %.2e" % (b.n, dev, de))
        assert dev <= 1e-4 and de <= 1e-6
        worst = 0.0
        for p in ia["ports"]:
            for get in ("port_v", "port_i"):
                x, y = np.asarray(getattr(a, get)(p)), np.asarray(getattr(b, get)(p))
                assert np.all(np.isfinite(y)) and np.max(np.abs(x)) > 0
                worst = max(worst, float(np.max(np.abs(y - x)) / np.max(np.abs(x))))
        print("port spectra: largest deviation %.2e of the largest magnitude" % worst)
        assert worst <= 1e-4
    finally:
        a.close()
        b.close()
