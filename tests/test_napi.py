"""The Node addon (courant_amd/courant_napi.node): the courant.h ABI presented as the
wasm-shaped `exports` object web/src/solver/courant.ts:19-85 reads, driven by
tests/napi/harness.js the way that wrapper drives the wasm module.

  * CPU: every function of include/courant.h is exported, with malloc / free and a
    memory.buffer arena; strings, plan arrays, error returns and the allocator behave
    as through the C ABI (the same values as the ctypes path, bit for bit)
  * CPU: each Node environment gets its own instance: two worker threads that load the
    addon while the main thread holds a plan and data in its arena each start from a
    fresh arena, report what the main thread reports, and leave the main thread's data
    and plan untouched
  * GPU: a lumped-port line stepped in chunks of 500 through the addon, with a field
    written through a held view before stepping and read through it after, gives
    bit-identical spectra, energy and field values to the same calls made through ctypes
  * GPU: the command-line board run (web/tools/run-board.ts:53-85) through the addon on the
    lossy mode-port GCPW: both excitations (their set-up calls recorded from the Python
    scene builder and replayed), stepped in chunks of 500 to the energy-decay stop, the
    mode-port waves of both ports, the 2 x 2 S matrix, its per-frequency 50-ohm
    renormalisation (cem_s2_renormalize) and the Touchstone text: every return value,
    energy, wave ratio and renormalised entry bit-identical to the ctypes path, and the
    Touchstone text identical to the Python writer's apart from the date line
"""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from courant_amd import lib as L

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
ADDON = os.path.join(ROOT, "courant_amd", "courant_napi.node")
HARNESS = os.path.join(HERE, "napi", "harness.js")
NODE = shutil.which("node")

needs_addon = pytest.mark.skipif(not NODE or not os.path.exists(ADDON), reason="node or courant_napi.node missing")


def run_harness(mode):
    r = subprocess.run([NODE, HARNESS, mode, ADDON], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       timeout=240)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout.strip().splitlines()[-1])


def build_line(cem):
    """The harness's line, through the ctypes binding (same calls, same order)."""
    x = [i * 1e-4 for i in range(61)]
    y = [j * 1e-4 for j in range(41)]
    z = [k * 5e-5 for k in range(31)]
    s = cem.create_sim(x, y, z)
    for f in range(6):
        s.set_boundary(f, 1, 6)
    s.material_box(-1, 1, -1, 1, z[6], z[12], 4.3, 0)
    s.pec_sheet(2, z[6], -1, 1, -1, 1)
    s.pec_sheet(2, z[12], x[10], x[50], y[18], y[22])
    p1 = s.add_lumped_port(2, 10, 20, 6, 12, -1, 50)
    p2 = s.add_lumped_port(2, 50, 20, 6, 12, -1, 50)
    s.set_gaussian(p1, 1, 60e-12, 15e-12)
    s.set_frequencies([1e9 * (i + 1) for i in range(40)])
    s.finalize(0.99)
    return s, p1, p2


def interleave(z):
    return np.stack([z.real, z.imag], -1).ravel()


@needs_addon
def test_addon_exports_the_whole_abi(cem):
    from test_abi import declared
    out = run_harness("exports")
    assert out["names"] == declared(os.path.join(ROOT, "include", "courant.h"))
    assert len(out["names"]) == 91 and out["malloc"] and out["arena_bytes"] >= 16 << 20
    assert out["version"] == cem.version()
    s, _, _ = build_line(cem)
    assert out["dt"] == s.dt and out["nfreq"] == s.nf
    cb = np.ctypeslib.as_array(cem.lib.cem_plan_array(s.h, 1, 2), shape=(s.N,))
    assert out["cb_sum"] == sum(cb.tolist())
    ide = np.ctypeslib.as_array(cem.lib.cem_plan_array(s.h, 2, 1), shape=(41,))
    assert out["ide_y"] == ide.tolist()
    assert out["bad_port"] == -1 and out["error"] == "already finalized"
    assert out["bad_ptr"] == [True, True, True, True]
    assert out["alloc"] == {"distinct": True, "reused": True}


@needs_addon
def test_addon_state_is_per_worker_thread():
    out = run_harness("workers")
    assert out["workers"] == 2 and out["workers_equal_main"]
    assert out["worker_first_blocks"] == [16, 16]
    assert out["pattern_ok"] and out["plan_ok"]


@needs_addon
@pytest.mark.gpu
def test_addon_line_matches_the_ctypes_path(cem, gpu_available):
    print("\n[gpu case] node addon: lumped line, 2000 steps", flush=True)
    out = run_harness("line")
    s, p1, p2 = build_line(cem)
    f = s.field(2).reshape(-1)
    f[s.index(30, 10, 9)] = 1e-3
    for _ in range(4):
        s.step(500)
    assert out["steps"] == s.steps == 2000
    assert out["energy"] == s.energy()
    for key, got in (("v1", s.port_v(p1)), ("i1", s.port_i(p1)), ("v2", s.port_v(p2)), ("i2", s.port_i(p2))):
        assert np.array_equal(np.array(out[key]), interleave(got)), key
    assert out["ez_probe"] == s.field(2).reshape(-1)[s.index(30, 20, 9)]
    assert np.max(np.abs(out["v2"])) > 0


MODE_MINI = dict(hmin=40e-6, hmax=100e-6, dx=100e-6, length=1.5e-3, feed=1.2e-3, mon=0.6e-3, yair=0.3e-3,
                 zair=0.4e-3, ports="mode", lossy=True, tcu=35e-6, mask=25e-6, dc=False, df=4e9)


def cdiv(x, y):
    """run-board.ts's complex division, elementwise on complex arrays."""
    d = y.real * y.real + y.imag * y.imag
    return (x.real * y.real + x.imag * y.imag) / d + 1j * ((x.imag * y.real - x.real * y.imag) / d)


@needs_addon
@pytest.mark.gpu
def test_addon_board_run_matches_the_ctypes_path(cem, gpu_available, tmp_path):
    from courant_amd import scenes as S
    from courant_amd import touchstone
    from napi_trace import Recorder
    ver = subprocess.run([NODE, "--version"], stdout=subprocess.PIPE, text=True).stdout.strip()
    print(f"\n[gpu case] node addon ({ver}): board run, lossy mode-port GCPW, both excitations", flush=True)
    rec_cem = L.Courant()
    rec = Recorder(rec_cem.lib)
    rec_cem.lib = rec
    decay = 1e-5
    runs, waves, energies, steps = [], [], [], []
    for excite in (1, 2):
        rec.calls, rec.handles, rec.active = [], {}, True
        s, info = S.gcpw(rec_cem, **dict(MODE_MINI, excite=excite))
        rec.active = False
        tstop, tmax = 9 * info["tau"], 4000 * s.dt
        peak, ws = 0.0, []
        while True:
            s.lib.cem_step(s.h, 500)
            w = s.lib.cem_energy(s.h)
            ws.append(w)
            peak = max(peak, w)
            n = s.lib.cem_step_count(s.h)
            if (n * s.dt > tstop and w < decay * peak) or n * s.dt > tmax:
                break
        p1, p2 = info["ports"]
        waves.append((s.mode_port_waves(p1), s.mode_port_waves(p2)))
        energies.append(ws)
        steps.append(n)
        runs.append({"calls": rec.calls, "ports": [p1, p2]})
        s.close()
    f = [float(v) for v in info["freqs"]]
    title = "lossy mode-port GCPW, planes at the port references"
    trace = {"runs": runs, "tstop": tstop, "tmax": tmax, "decay": decay, "f": f, "title": title}
    path = tmp_path / "trace.json"
    path.write_text(json.dumps(trace))
    r = subprocess.run([NODE, HARNESS, "board", ADDON, str(path)], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    for k, run in enumerate(runs):
        got = out["runs"][k]
        assert got["results"] == [c[2] for c in run["calls"]], k
        assert got["outputs"] == [c[3] for c in run["calls"]], k
        assert got["energies"] == energies[k] and got["steps"] == steps[k], k
    (u1, u2), (v1, v2) = waves
    Sm = {"s11": cdiv(u1["b"], u1["a"]), "s21": cdiv(u2["b"], u1["a"]),
          "s12": cdiv(v1["b"], v2["a"]), "s22": cdiv(v2["b"], v2["a"])}
    for key, val in Sm.items():
        assert np.array_equal(np.array(out[key]), interleave(val)), key
    assert out["z1"] == u1["z0"].tolist() and out["z2"] == u2["z0"].tolist()
    s50 = np.array([cem.renormalize2([Sm["s11"][i], Sm["s12"][i], Sm["s21"][i], Sm["s22"][i]],
                                     u1["z0"][i], u2["z0"][i], 50.0) for i in range(len(f))])
    assert np.array_equal(np.array(out["s50"]), np.stack([s50.real, s50.imag], -1).reshape(len(f), 8))
    py = touchstone.write_s2p(f, s50.reshape(-1, 2, 2), title, cem.version(), full=True, renormalised=True)
    js, pl = out["touchstone"].split("\n"), py.split("\n")
    assert len(js) == len(pl) == len(f) + 4
    assert js[0] == pl[0] and js[2:] == pl[2:]
    assert js[1].endswith("; reference 50; both ports excited")
    assert steps[0] > 500 and np.all(np.abs(s50) < 1.5)
