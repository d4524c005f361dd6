"""Full 2-port S-parameters of a board from its two excitations, the assembly of the
reference's command-line board run (web/tools/run-board.ts:61-85), and their Touchstone file.

Lumped ports: S = B A^-1 from the port voltages and currents of both runs
(courant_amd.serve.s_matrix).  Mode ports: from the mode-port waves, with port 1 driven
S11 = b1 / a1 and S21 = b2 / a1, with port 2 driven S12 = b1 / a2 and S22 = b2 / a2, each
port referred to its own line impedance Z0(f); the matrix is then renormalised to 50 ohm
in FP64, on the GPU (cem_amd_s2_renormalize) when the library has it, else through
cem_s2_renormalize per frequency.
"""
import numpy as np

from . import serve
from . import touchstone


def run(cem, make, steps):
    """Both excitations: make(excite) builds and finalizes the board with port `excite` (1 or
    2) driven and returns (sim, info) with info["ports"] = (p1, p2); each runs `steps` steps.
    Returns a dict with f, S (nf x 2 x 2, 50 ohm), z0 (2 x nf, mode ports) and the runs'
    sims closed."""
    sims = []
    try:
        for excite in (1, 2):
            s, info = make(excite)
            sims.append((s, info))
            s.step(steps)
        s1, info = sims[0]
        s2, _ = sims[1]
        p1, p2 = info["ports"]
        if info["params"]["ports"] == "lumped":
            vis = [(np.array([s.port_v(p1), s.port_v(p2)]), np.array([s.port_i(p1), s.port_i(p2)])) for s, _ in sims]
            return {"f": info["freqs"], "S": serve.s_matrix(vis, 50.0), "z0": None}
        u1, u2 = s1.mode_port_waves(p1), s1.mode_port_waves(p2)   # port 1 driven
        v1, v2 = s2.mode_port_waves(p1), s2.mode_port_waves(p2)   # port 2 driven
        S = np.empty((len(info["freqs"]), 2, 2), complex)
        S[:, 0, 0] = u1["b"] / u1["a"]
        S[:, 1, 0] = u2["b"] / u1["a"]
        S[:, 0, 1] = v1["b"] / v2["a"]
        S[:, 1, 1] = v2["b"] / v2["a"]
        z1, z2 = u1["z0"], u2["z0"]
        return {"f": info["freqs"], "S": renormalise(cem, S, z1, z2, 50.0), "z0": np.array([z1, z2])}
    finally:
        for s, _ in sims:
            s.close()


def renormalise(cem, S, z1, z2, zref):
    """S referred to (z1[f], z2[f]) -> zref, FP64."""
    if getattr(cem, "has_amd", False):
        return cem.renormalize_batch(S, z1, z2, zref)
    out = np.empty_like(S)
    for k in range(len(S)):
        out[k] = np.array(cem.renormalize2(list(S[k].ravel()), z1[k], z2[k], zref)).reshape(2, 2)
    return out


def to_touchstone(result, title, version):
    """Touchstone text of a run() result (both ports excited, 50 ohm)."""
    return touchstone.write_s2p(result["f"], result["S"], title, version, full=True, renormalised=True)
