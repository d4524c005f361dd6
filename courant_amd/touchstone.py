"""Touchstone v1 2-port files: the writer of the reference's export
(web/src/app/touchstone.ts:1-23), a reader of any v1 2-port file, and the comparison rules
of its validation tool (tools/compare_s2p.py).

write_s2p: RI format, frequencies in Hz, reference 50 ohm; all four S-parameters when both
ports were excited, otherwise S22 = S11 and S12 = S21 are assumed and the header says so.
Numbers are printed as JavaScript's toExponential prints them (6 digits for the frequency,
9 for each component, exponent without leading zeros), so files match the reference's
byte for byte apart from the date line.
"""
import datetime
import math

import numpy as np

RULES = {   # tolerances of tools/compare_s2p.py:3-10
    "openems": {"s21_db": 0.05, "s21_deg": 1.0, "s11_res": 0.01, "s11_db": 1.0, "fmax": 40e9},
    "hfss": {"s21_db": 0.15, "s21_deg": 3.0, "s11_res": 0.01, "s11_db": 1.0, "fmax": 40e9},
    "measured": {"s21_db": 0.3, "s21_deg": None, "s11_res": None, "s11_db": None, "fmax": 20e9},
    "gpu": {"s21_db": 0.01, "s21_deg": 0.1, "s11_res": None, "s11_db": None, "fmax": 40e9},
}


def js_exp(v, digits):
    """Number.prototype.toExponential(digits): d.ddde+N / e-N, no sign on zero."""
    v = float(v)
    if v == 0:
        return f"{0:.{digits}f}e+0"
    if not math.isfinite(v):
        return "NaN" if math.isnan(v) else ("Infinity" if v > 0 else "-Infinity")
    mant, exp = f"{v:.{digits}e}".split("e")
    e = int(exp)
    return f"{mant}e{'+' if e >= 0 else '-'}{abs(e)}"


def write_s2p(f, S, title, version, full=True, renormalised=True, date=None):
    """Touchstone text of S (nf x 2 x 2, [[S11, S12], [S21, S22]]) at frequencies f (Hz)."""
    date = date or datetime.datetime.now(datetime.timezone.utc).strftime("%Y-%m-%dT%H:%M:%S.%f")[:-3] + "Z"
    ref = "50" if renormalised else "line Z0 (not renormalised)"
    excited = "both ports excited" if full else "port 1 excited, S22 = S11 and S12 = S21 assumed"
    lines = [f"! Courant EM {version} — {title}", f"! {date}; reference {ref}; {excited}", "# HZ S RI R 50"]
    S = np.asarray(S, complex)
    for k, fk in enumerate(f):
        s11, s21 = S[k, 0, 0], S[k, 1, 0]
        s12, s22 = (S[k, 0, 1], S[k, 1, 1]) if full else (s21, s11)
        vals = [fk, s11.real, s11.imag, s21.real, s21.imag, s12.real, s12.imag, s22.real, s22.imag]
        lines.append(" ".join(js_exp(v, 9 if i else 6) for i, v in enumerate(vals)))
    return "\n".join(lines) + "\n"


def read_s2p(text):
    """(f in Hz, S nf x 2 x 2, reference R) of a Touchstone v1 2-port text: option line
    '# <unit> S <RI|MA|DB> R <ohm>' (defaults GHz, MA, 50), '!' comments, nine numbers per
    frequency in the order f S11 S21 S12 S22, possibly wrapped over lines."""
    scale = {"hz": 1.0, "khz": 1e3, "mhz": 1e6, "ghz": 1e9}
    opt = {"unit": 1e9, "fmt": "ma", "R": 50.0}
    nums = []
    for raw in text.splitlines():
        body = raw.partition("!")[0].strip()
        if body.startswith("#"):
            words = body[1:].lower().split()
            for k, w in enumerate(words):
                if w in scale:
                    opt["unit"] = scale[w]
                elif w in ("ri", "ma", "db"):
                    opt["fmt"] = w
                elif w == "r" and k + 1 < len(words):
                    opt["R"] = float(words[k + 1])
        elif body:
            nums += [float(x) for x in body.split()]
    rec = np.asarray(nums).reshape(-1, 9)
    first, second = rec[:, 1::2], rec[:, 2::2]          # 4 pairs per row: S11 S21 S12 S22
    if opt["fmt"] == "ri":
        val = first + 1j * second
    else:
        mag = 10.0 ** (first / 20.0) if opt["fmt"] == "db" else first
        val = mag * np.exp(1j * np.deg2rad(second))
    S = np.stack([np.stack([val[:, 0], val[:, 2]], -1), np.stack([val[:, 1], val[:, 3]], -1)], 1)
    return rec[:, 0] * opt["unit"], S, opt["R"]


def renorm(S, R, R0=50.0):
    """S (nf x 2 x 2) from a common port reference R to R0 through the impedance matrix
    Z = R (I + S)(I - S)^-1: S' = (Z - R0 I)(Z + R0 I)^-1."""
    if abs(R - R0) < 1e-12:
        return S
    eye = np.eye(2)
    Z = R * (eye + S) @ np.linalg.inv(eye - S)
    return (Z - R0 * eye) @ np.linalg.inv(Z + R0 * eye)


def _mag_db(x):
    return 20.0 * np.log10(np.maximum(np.abs(x), 1e-30))


def _resample(f_new, f, x):
    """x at f_new, interpolating dB magnitude and unwrapped phase linearly."""
    mag = np.interp(f_new, f, _mag_db(x))
    ph = np.interp(f_new, f, np.unwrap(np.angle(x)))
    return 10.0 ** (mag / 20.0) * np.exp(1j * ph)


def _minima_below(f, x, floor_db=-10.0):
    """Frequencies of the interior local minima of |x| below floor_db."""
    d = _mag_db(x)
    k = np.arange(1, len(d) - 1)
    keep = (d[k] < floor_db) & (d[k] <= d[k - 1]) & (d[k] <= d[k + 1])
    return f[k[keep]]


def compare(a, b, rule, fmax=None):
    """File a against file b (each (f, S, R)) under one rule of RULES, on a's frequencies
    inside the overlap (b resampled): the worst |S21|, |S12| difference in dB and phase
    difference in degrees; the worst |S11|, |S22| difference where either file is above
    -30 dB; the worst relative distance from each dip of a below -10 dB to the nearest dip
    of b.  Returns {"rule", "points", "checks": [...], "pass"}."""
    fa, Sa, Ra = a
    fb, Sb, Rb = b
    Sa, Sb = renorm(Sa, Ra), renorm(Sb, Rb)
    tol = RULES[rule]
    hi = min(fa.max(), fb.max(), fmax or tol["fmax"])
    use = (fa >= fb.min()) & (fa <= hi)
    f = fa[use]
    checks = []

    def record(name, value, limit, unit):
        checks.append({"name": name, "value": value, "tol": limit, "unit": unit,
                       "pass": None if limit is None else bool(value <= limit)})

    for r, c in ((1, 0), (0, 1)):
        ours, theirs = Sa[use, r, c], _resample(f, fb, Sb[:, r, c])
        tag = f"S{r + 1}{c + 1}"
        record(f"|{tag}| max difference", float(np.abs(_mag_db(ours) - _mag_db(theirs)).max()), tol["s21_db"], "dB")
        record(f"phase {tag} max difference", float(np.abs(np.degrees(np.angle(ours / theirs))).max()),
               tol["s21_deg"], "deg")
    for k in (0, 1):
        ours, theirs = Sa[use, k, k], _resample(f, fb, Sb[:, k, k])
        tag = f"S{k + 1}{k + 1}"
        loud = (_mag_db(ours) > -30) | (_mag_db(theirs) > -30)
        diff = np.abs(_mag_db(ours) - _mag_db(theirs))
        record(f"|{tag}| max difference where above -30 dB", float(diff[loud].max()) if loud.any() else 0.0,
               tol["s11_db"], "dB")
        da, dbb = _minima_below(f, ours), _minima_below(f, theirs)
        if len(da) and len(dbb):
            worst = max(float(np.abs(dbb - x).min() / x) for x in da)
            record(f"{tag} dip frequencies (nearest match)", worst, tol["s11_res"], "relative")
    return {"rule": rule, "points": int(len(f)), "checks": checks,
            "pass": all(ch["pass"] is not False for ch in checks)}
