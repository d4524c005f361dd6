"""Benchmark of the FDTD time stepper on one MI355X.

Workload (default): the large single board, 600 x 400 x 160 non-uniform cells
(38,801,161 grid points): a synthetic four-layer stackup with 5 seeded meanders per
copper layer, 40 PEC vias, a 6-pole Djordjevic-Sarkar dielectric on all three layers
and Huray SIBC on every conductor, CPML on all faces, lumped ports, 801 DFT
frequencies; FP32 fields, FP64 DFT.  A "step" is one full leapfrog step of that board
(H, E with CPML and the Debye currents, the post-phase SIBC / source edges, the probe
gather and the running DFT), with every input resident in HBM.  The timed steps run as
the engine runs them (chunks replayed from a captured HIP graph) and end with the DFT
of their samples; per-kernel durations come from a second, event-bracketed pass of
direct launches of the same kernels.

Prints one JSON line; `value` is Mcell-updates/s = nx*ny*nz*steps / wall.  The line
also carries the roofline of the dominant kernel and of the whole step, the reference
core's f64 rate on the host cores (cpu_baseline), the lossy mode-port GCPW (4.6 M
points, MALL-resident) as `cfg2`, `batched` (seeded 2-port boards, both excitations,
one at a time versus concurrent) and `served` (a Poisson stream of seeded boards at
several arrival rates through the admission loop: solves/s, p50/p99 time to
S-parameters, against the same boards run one at a time); `--no-batch` leaves the
last two out.
"""
import argparse
import csv
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Mcell-updates/s & %HBM peak; batched solves/s + p99 latency at <=1e-4 dB"
WORKLOADS = {
    "board4": "four-layer board, 20 meanders + 40 vias, Debye 6 poles + Huray SIBC, lumped ports",
    "cfg2": "lossy GCPW, mode ports, Debye 6 poles + Huray SIBC",
    "cfg1": "lossless GCPW, lumped ports, CPML on six faces",
}


def make_workload(cem, name, **over):
    from courant_amd import scenes
    if name == "board4":
        return scenes.board4(cem, **over)
    return scenes.gcpw(cem, **dict(scenes.CFG2 if name == "cfg2" else scenes.CFG1, **over))


def plan_counts(sim):
    L = sim.lib
    psiE = sum(L.cem_state_len(sim.h, 0, c, w) for c in range(3) for w in range(2))
    psiH = sum(L.cem_state_len(sim.h, 1, c, w) for c in range(3) for w in range(2))
    nd, npole, ns = L.cem_plan_ndisp(sim.h), L.cem_plan_npole(sim.h), L.cem_plan_nsibc(sim.h)
    nterm = L.cem_state_len(sim.h, 4, 0, 0) // ns if ns else 0
    return dict(N=sim.N, psiE=psiE, psiH=psiH, ndisp=nd, npole=npole, nsibc=ns, nterm=nterm)


def step_bytes(pc):
    """Algorithmic bytes per step of the survey's FP32 byte model, each array counted once
    per access: B_step = 72 N + 8 N_psi + 12 P N_disp + (8K + 36) N_sibc (E, H read+write
    and ca, cb read per point; psi read+write; J read+write and beta read per pole of a
    dispersive edge; SIBC states, index, 4 H faces and weights per surface edge)."""
    return (72 * pc["N"] + 8 * (pc["psiE"] + pc["psiH"]) + 12 * pc["npole"] * pc["ndisp"]
            + (8 * pc["nterm"] + 36) * pc["nsibc"])


def kernel_bytes(pc, cbytes=4):
    """Bytes each kernel must move in this engine's storage, per launch:
      H    : E read (12 B/pt) + H read/write (24 B/pt) + psi_H read/write (8 B/entry)
      E    : E read/write (24) + H read (12) + coefficient classes (cbytes: a class byte
             when the grid has at most 256 distinct class triples, else a word of three 8-bit
             class numbers, instead of 24 B of ca, cb) per point + psi_E (8 B/entry) + the Debye
             currents it updates, read/write (8 B per pole per dispersive edge; beta / dt
             comes from a per-material table)
      step : the one-pass step (H and E of a step in one sweep): E and H read once and
             written once (48 B/pt) + classes (4) + psi_H and psi_E read/write (8 B/entry)
             + the Debye currents (8 B per pole per dispersive edge)
      post : SIBC states read/write (8 B per term) + index, 4 H faces and weights (36 B)
             per surface edge (and the few dispersive edges that are also surface or
             source edges, not counted)
    """
    h = 36 * pc["N"] + 8 * pc["psiH"]
    e = (36 + cbytes) * pc["N"] + 8 * pc["psiE"] + 8 * pc["npole"] * pc["ndisp"]
    step = 52 * pc["N"] + 8 * (pc["psiH"] + pc["psiE"]) + 8 * pc["npole"] * pc["ndisp"]
    post = (8 * pc["nterm"] + 36) * pc["nsibc"]
    return {"H": h, "E": e, "step": step, "post": post}


def cpu_baseline(workload, target_s=12.0):
    """Reference f64 core (compiled from the reference sources, OpenMP) on the host cores,
    on the same board, timed over a bounded number of steps; then the GPU engine on the
    same board for the same steps, and the relative difference of the field energies."""
    path = os.path.join(ROOT, "oracle", "_ref", "libcourant_ref.so")
    if not os.path.exists(path):
        return None
    from courant_amd import lib, scenes
    ref = lib.Courant(path, extensions=False)
    over = {} if workload == "board4" else dict(ports="lumped", probes=False)
    t = time.perf_counter()
    s, info = make_workload(ref, workload, **over)
    setup = time.perf_counter() - t
    cells = scenes.cells(info)
    s.step(2)
    t = time.perf_counter()
    s.step(2)
    per = (time.perf_counter() - t) / 2
    n = int(max(4, min(400, target_s / max(per, 1e-6))))
    t = time.perf_counter()
    s.step(n)
    wall = time.perf_counter() - t
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    w_ref = s.energy()
    s.close()
    gpu, _ = make_workload(lib.Courant(), workload, **over)
    gpu.step(4 + n)
    w_gpu = gpu.energy()
    gpu.close()
    sample = ("reference f64 OpenMP core, same board" + ("" if workload == "board4" else " with lumped ports")
              + f", {n} steps after 4 ({wall:.1f} s; set-up {setup:.0f} s untimed)")
    return {"value": round(cells * n / wall / 1e6, 2), "unit": "Mcell-updates/s", "cores": cores,
            "kind": "reference", "sample": sample,
            "energy_rel_err_gpu_vs_f64": float(f"{abs(w_gpu - w_ref) / w_ref:.3e}"), "energy_steps": 4 + n}


def _pmc_pass(counter, args_tail, tmp, tag, env=None):
    """One rocprofv3 counter pass over a child run of this script; mean counter value per
    dispatch, by kernel name (template arguments and parameter lists dropped)."""
    exe = shutil.which("rocprofv3")
    if not exe:
        return None
    d = os.path.join(tmp, f"{tag}_{counter}")
    cmd = ["timeout", "-s", "KILL", "240", exe, "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "pmc",
           "--", sys.executable, os.path.abspath(__file__)] + args_tail
    r = subprocess.run(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, cwd=tmp, env=env)
    path = os.path.join(d, "pmc_counter_collection.csv")
    if r.returncode != 0 or not os.path.exists(path):
        return None
    acc = {}
    for row in csv.DictReader(open(path)):
        name = row["Kernel_Name"].split("(")[0].replace("void ", "").split("<")[0]
        acc.setdefault(name, []).append(float(row["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in acc.items()}


def pmc_traffic(workload, shape, lanes, order, hrev):
    """HBM bytes per dispatch of each stepper kernel: FETCH_SIZE and WRITE_SIZE (KiB per
    dispatch, one counter per rocprofv3 run), each scaled by what the same counter reads for
    the plain copy kernel moving 5e8 B in and 5e8 B out per dispatch.  The counted child runs
    use the launch shape the timed run was tuned to (planes and lanes per thread, tile order)."""
    env = dict(os.environ, COURANT_ISEG_H=str(shape[0]), COURANT_ISEG_E=str(shape[1]),
               COURANT_LANES_H=str(lanes[0]), COURANT_TILE_ORDER=order, COURANT_H_REVERSE=str(int(hrev)))
    if shape[3]:
        env["COURANT_ISEG_S"] = str(shape[3])
    tmp = tempfile.mkdtemp(prefix="courant_pmc_", dir="/tmp")
    try:
        out = {}
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            run = _pmc_pass(counter, ["--kernel-pass", "--workload", workload, "--steps", "8", "--warmup", "2"],
                            tmp, "step", env)
            cal = _pmc_pass(counter, ["--copy-only", "0.5"], tmp, "copy")
            if not run or not cal or not cal.get("cem::k_copy"):
                return None
            scale = 0.5e9 / 1024 / cal["cem::k_copy"]
            out[counter] = {"scale": scale, "kib": run}
        kernels = {}
        for k in ("cem::k_step", "cem::k_update_h", "cem::k_update_e", "cem::k_post", "cem::k_gather"):
            f = out["FETCH_SIZE"]["kib"].get(k)
            w = out["WRITE_SIZE"]["kib"].get(k)
            if f is not None and w is not None:
                kernels[k] = (f * out["FETCH_SIZE"]["scale"] + w * out["WRITE_SIZE"]["scale"]) * 1024
        return {"kernels": kernels, "fetch_scale": round(out["FETCH_SIZE"]["scale"], 3),
                "write_scale": round(out["WRITE_SIZE"]["scale"], 3)}
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def timed_run(sim, steps, warmup):
    """Autotune (set-up), W warm-up steps, then exactly `steps` steps and the DFT of their
    samples, bracketed by synchronisations.  Returns the wall seconds of the timed steps."""
    sim.tune()
    sim.step_device(warmup)
    sim.flush_dft()
    sim.synchronize()
    t0 = time.perf_counter()
    sim.step_device(steps)
    sim.flush_dft()
    sim.synchronize()
    return time.perf_counter() - t0


def extra_workload(cem, name, steps):
    """The lossy mode-port GCPW (its ~185 MB of state sits in the 256 MB MALL, so its rate
    is not an HBM figure): Mcell-updates/s over `steps` graph-replayed steps."""
    from courant_amd import scenes
    sim, info = make_workload(cem, name)
    try:
        wall = timed_run(sim, steps, 128)
        return {"workload": WORKLOADS[name], "cells": list(info["cells"]),
                "value": round(scenes.cells(info) * steps / wall / 1e6, 1), "unit": "Mcell-updates/s",
                "ms_per_step": round(wall / steps * 1e3, 5), "steps": steps,
                "note": "state fits the 256 MB MALL (Infinity Cache): not an HBM rate"}
    finally:
        sim.close()


def batched_and_served(cem, nboards, workers, nreq, rates, budget_gb):
    """Batched: the first `nboards` of the 64 seeded GCPW / microstrip boards (seed 4000 + b),
    both excitations each, one at a time and then on `workers` concurrent streams; the two
    runs must give bit-identical S matrices.  Served: `nreq` seeded request boards (seed
    5000 + r) arriving as a Poisson stream at each of `rates` boards/s, admitted by estimated
    footprint under `budget_gb`, against the same boards run one at a time."""
    from courant_amd import serve
    serve.run_batch(cem, [dict(serve.random_board(4000), nx=80, ny=80, nz=60, steps=64)], 1)   # context warm-up
    batched = served = None
    if nboards > 0:
        batched = _batched(cem, serve, nboards, workers)
    if nreq > 0 and rates:
        served = _served(cem, serve, nreq, rates, workers, budget_gb)
    return batched, served


def _batched(cem, serve, nboards, workers):
    """Plain one-at-a-time runs (each excitation built, stepped and read alone, in turn)
    against the batched form: each board's two excitations as a pair (one set-up, shared
    H / E launches), `workers` boards at once; and, for reference, the excitations as
    independent simulations on `workers` threads."""
    boards = [serve.random_board(4000 + b) for b in range(nboards)]
    s1, w1, u1 = serve.run_batch(cem, boards, 1)
    print(f"batched: one at a time {w1:.2f} s", file=sys.stderr, flush=True)
    sc, wc, uc = serve.run_batch(cem, boards, workers)
    print(f"batched: {workers} workers, single excitations {wc:.2f} s", file=sys.stderr, flush=True)
    sp, wp, up = serve.run_batch(cem, boards, workers, pairs=True)
    print(f"batched: {workers} workers, pairs {wp:.2f} s", file=sys.stderr, flush=True)
    nsol = len(serve.jobs_of(boards))
    return {
        "workload": "seeded GCPW/microstrip boards 80..256 x 80..192 x 60..96, Debye 2-6 poles + SIBC, "
                    "both excitations per board (seed 4000+b)",
        "boards": nboards, "solves": nsol, "workers": workers, "form": "pairs (both excitations as right-hand "
        "sides of shared H/E launches, one set-up), boards concurrent",
        "solves_per_s": round(nsol / wp, 3), "one_at_a_time_solves_per_s": round(nsol / w1, 3),
        "speedup": round(w1 / wp, 3), "mcell_updates_per_s": round(up / wp / 1e6, 1),
        "one_at_a_time_mcell_updates_per_s": round(u1 / w1 / 1e6, 1),
        "single_excitations_concurrent_solves_per_s": round(nsol / wc, 3),
        "bitwise_same_as_one_at_a_time": all(np.array_equal(a, b) for a, b in zip(s1, sp))
        and all(np.array_equal(a, b) for a, b in zip(s1, sc)),
    }


def _served(cem, serve, nreq, rates, workers, budget_gb):
    reqs = [serve.random_request(5000 + r) for r in range(nreq)]
    nsol = len(serve.jobs_of(reqs))
    t0 = time.perf_counter()
    seq = [serve.run_batch(cem, [b], 1)[0][0] for b in reqs]
    wseq = time.perf_counter() - t0
    print(f"served: one at a time {wseq:.2f} s", file=sys.stderr, flush=True)
    sweep = []
    for rate in rates:
        lat, wall, got, upd = serve.serve(cem, reqs, rate, workers, budget_gb * 1e9, pairs=True)
        print(f"served: {rate} boards/s {wall:.2f} s", file=sys.stderr, flush=True)
        sweep.append({
            "rate_per_s": rate, "solves_per_s": round(nsol / wall, 3), "speedup": round(wseq / wall, 3),
            "mcell_updates_per_s": round(upd / wall / 1e6, 1),
            "p50_s": round(float(np.percentile(lat, 50)), 3), "p99_s": round(float(np.percentile(lat, 99)), 3),
            "bitwise_same_as_one_at_a_time": all(np.array_equal(a, b) for a, b in zip(seq, got)),
        })
    return {
        "workload": "Poisson arrivals of seeded boards, cells log-normal (median 3 M, sigma 0.8) in "
                    "[0.5 M, 20 M], half lossy, half 2-port (seed 5000+r)",
        "boards": nreq, "solves": nsol, "workers": workers, "budget_gb": budget_gb,
        "form": "admission by footprint; 2-port boards as pairs (both excitations in shared H/E launches)",
        "mean_mcells": round(float(np.mean([serve.cells(b) for b in reqs])) / 1e6, 2),
        "one_at_a_time_solves_per_s": round(nsol / wseq, 3), "rates": sweep,
    }


def mode_sweep(cem, df):
    """Wall time of the eigenmode sweep of both ports of the lossy GCPW (mode ports,
    300 x 152 x 99 cells) at output frequencies spaced df: the sweep runs on the first request
    for a port's waves, after a few steps."""
    from courant_amd import scenes as S
    sim, info = S.gcpw(cem, **dict(S.CFG2, df=df))
    try:
        sim.step(64)
        t = time.perf_counter()
        for port in info["ports"]:   # the first request solves every port; each must have solved
            sim.mode_port_waves(port)
        wall = time.perf_counter() - t
        return {"frequencies": len(info["freqs"]), "ports": 2, "wall_s": round(wall, 2)}
    finally:
        sim.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=512)
    ap.add_argument("--warmup", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-pmc", action="store_true", help="skip the rocprofv3 passes that measure HBM traffic")
    ap.add_argument("--kernel-pass", action="store_true",
                    help="only set up the workload and run warmup + steps (the child run of a counter pass)")
    ap.add_argument("--workload", choices=tuple(WORKLOADS), default="board4",
                    help="board4: four-layer board, 600x400x160 (38.8 M grid points; default); cfg2: lossy "
                         "mode-port GCPW, 4.6 M points; cfg1: lumped GCPW, 120x80x60")
    ap.add_argument("--copy-only", type=float, default=0,
                    help="only run the stream-copy kernel on this many GB (byte-counter calibration)")
    ap.add_argument("--no-extra", action="store_true", help="skip the cfg2 line and the mode sweep")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched and served measurements")
    ap.add_argument("--batch-boards", type=int, default=8, help="boards of the batched set (of 64; 0: none)")
    ap.add_argument("--serve-requests", type=int, default=10, help="boards of the served stream (0: none)")
    ap.add_argument("--serve-rates", type=str, default="1,5,20", help="Poisson arrival rates, boards/s")
    ap.add_argument("--workers", type=int, default=4, help="concurrent solves (host threads, one stream each)")
    ap.add_argument("--budget-gb", type=float, default=200.0, help="device footprint admitted at once")
    ap.add_argument("--mode-sweep", type=float, default=1e8, metavar="DF",
                    help="time the lossy GCPW's port-mode sweep at output frequencies spaced DF Hz "
                         "(default 1e8: 400 frequencies; 0 or --no-extra: skip)")
    args = ap.parse_args()
    if args.copy_only:
        from courant_amd import lib
        gbs = lib.Courant().lib.cem_amd_copy_bandwidth(int(args.copy_only * 1e9), 5)
        print(json.dumps({"copy_gb": args.copy_only, "copy_gbs": round(gbs, 1), "launches": 5}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this engine runs one board per GPU and --gpus {args.gpus} is not "
              "supported (multi-GPU slab sharding is design-only)", file=sys.stderr)
        sys.exit(2)

    from courant_amd import lib, scenes
    cem = lib.Courant()
    t0 = time.perf_counter()
    sim, info = make_workload(cem, args.workload)
    setup_s = time.perf_counter() - t0
    nx, ny, nz = info["cells"]
    cells = nx * ny * nz
    if args.kernel_pass:
        sim.step_device(args.warmup)
        sim.synchronize()
        sim.set_profiling(True)   # direct launches: one dispatch per kernel and step
        sim.step_device(args.steps)
        sim.synchronize()
        sim.close()
        return
    # timed region: the production path (runs of steps replayed from a captured graph)
    wall = timed_run(sim, args.steps, args.warmup)
    # per-kernel durations: a separate pass of direct launches bracketed by HIP events
    # on the engine's stream (same kernels and arguments as the graph)
    nprof = max(64, min(args.steps, 512))
    sim.set_profiling(True)
    t1 = time.perf_counter()
    sim.step_device(nprof)
    sim.flush_dft()
    sim.synchronize()
    wall_direct = time.perf_counter() - t1
    stats = {k: sim.kernel_stats(i) for i, k in enumerate(("H", "E", "gather", "dft", "post", "step"))}
    sim.set_profiling(False)
    pc = plan_counts(sim)
    kb = kernel_bytes(pc, sim.class_bytes())
    b8d = step_bytes(pc)
    shape = sim.launch_shape()
    lanes = sim.lane_widths()
    order = sim.tile_order()
    sim_hrev = sim.h_reversed()
    rows, planes = sim.step_shape()
    shape = (shape[0], shape[1], shape[2], planes)
    copy_gbs = cem.lib.cem_amd_copy_bandwidth(2 << 30, 10)
    value = cells * args.steps / wall / 1e6
    avg = {k: v[0] / max(v[1], 1) for k, v in stats.items()}
    dom = max(("step", "H", "E", "post"), key=lambda k: stats[k][0])
    achieved = kb[dom] / (avg[dom] * 1e-3) / 1e9
    t_step = wall / args.steps
    moved = (kb["step"] if rows else kb["H"] + kb["E"]) + kb["post"]
    sim.step(1)  # refresh spectra through the public path: probes must be finite
    spec = sim.vprobe(0) if cem.lib.cem_plan_nvprobe(sim.h) > 0 else sim.port_v(0)
    ok = bool(np.all(np.isfinite(spec)))
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_step * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": WORKLOADS[args.workload], "cells": [nx, ny, nz],
                   "grid_points": pc["N"], "nf": int(len(info["freqs"])), "ndisp": pc["ndisp"],
                   "npole": pc["npole"], "nsibc": pc["nsibc"], "dft": "f64", "setup_s": round(setup_s, 2),
                   "finite": ok},
        "roofline": {"bound": "hbm",
                     "kernel": {"step": "k_step", "H": "k_update_h", "E": "k_update_e", "post": "k_post"}[dom],
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "bytes_per_launch": kb[dom], "avg_launch_ms": round(avg[dom], 5),
                     "step": {"bytes_survey_model": b8d,
                              "frac_survey_model": round(b8d / t_step / 1e9 / HBM_PEAK_GBS, 4),
                              "bytes_moved_model": moved,
                              "frac_moved_model": round(moved / t_step / 1e9 / HBM_PEAK_GBS, 4)},
                     "measured_copy_gbs": round(copy_gbs, 1),
                     "frac_of_copy": round(achieved / copy_gbs, 4) if copy_gbs > 0 else None},
        "kernels_ms": {k: round(v, 5) for k, v in avg.items() if stats[k][1]},
        "kernel_bytes_mb": {k: round(v / 1e6, 1) for k, v in kb.items() if stats[k][1]},
        "launch": {"timed": "graph replay, DFT of the timed steps included",
                   "ms_per_step_direct_launches": round(wall_direct / nprof * 1e3, 5),
                   "profiled_steps": nprof,
                   "one_pass_step": {"rows_per_workgroup": rows, "xplanes_per_workgroup": planes} if rows else None,
                   "iplanes_per_thread_split": None if rows else {"H": shape[0], "E": shape[1]},
                   "kpoints_per_thread_split": None if rows else {"H": lanes[0], "E": lanes[1]},
                   "tile_order_split": None if rows else order,
                   "h_sweep_reversed": None if rows else sim_hrev,
                   "class_bytes_per_point": sim.class_bytes(),
                   "autotuned": shape[2]},
    }
    sim.close()
    if not args.no_pmc:
        tr = pmc_traffic(args.workload, shape, lanes, order, sim_hrev)
        if tr:
            kname = "cem::" + out["roofline"]["kernel"]
            out["roofline"]["traffic"] = round(tr["kernels"].get(kname, 0.0))
            out["roofline"]["traffic_by_kernel"] = {k: round(v) for k, v in tr["kernels"].items()}
            out["roofline"]["pmc_scale"] = {"FETCH_SIZE": tr["fetch_scale"], "WRITE_SIZE": tr["write_scale"]}
            tsum = sum(tr["kernels"].values())
            out["roofline"]["step"]["bytes_counters"] = round(tsum)
            out["roofline"]["step"]["frac_counters"] = round(tsum / t_step / 1e9 / HBM_PEAK_GBS, 4)
    if not args.no_extra and args.workload != "cfg2":
        out["cfg2"] = extra_workload(cem, "cfg2", 512)
    if not args.no_batch:
        rates = [float(r) for r in args.serve_rates.split(",") if r.strip()]
        b, v = batched_and_served(cem, args.batch_boards, args.workers, args.serve_requests, rates, args.budget_gb)
        if b:
            out["batched"] = b
        if v:
            out["served"] = v
    if args.mode_sweep > 0 and not args.no_extra:
        out["mode_sweep"] = mode_sweep(cem, args.mode_sweep)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.workload)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
