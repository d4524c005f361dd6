"""Batched and served board solves on one GPU.

Every simulation owns a HIP stream and replays its own graph-captured step blocks, so
independent solves driven from concurrent host threads overlap on the device; the C ABI
releases the interpreter lock for the duration of each call.

* random_board(seed)        the batched workload's board: seeded GCPW or microstrip 2-port,
                            grid 80..256 x 80..192 x 60..96 cells, trace width 0.1-0.5 mm,
                            eps_r 3.0-4.5, 2-6 Debye poles, SIBC copper
* random_request(seed)      the served workload's board: cells log-normal (median 3 M,
                            sigma 0.8) in [0.5 M, 20 M], lossless or lossy, 1 or 2 ports
* solve(cem, b, excite)     one right-hand side: build, step, port V/I spectra
* solve_pair(cem, b)        both right-hand sides of a 2-port board: one set-up, the second
                            excitation a clone of the first's plan, the two stepped in shared
                            launches (one H and one E launch per step for both)
* s_matrix(...)             S = B A^-1 from the excitations' port waves (per frequency)
* run_batch(...)            many boards at once on a pool of host threads
* Server / serve(...)       a request loop: boards arrive as a Poisson stream, each
                            excitation is admitted while its estimated device footprint fits
                            the budget, runs concurrently with the others, and every board is
                            timed from arrival to S-parameters

Boards are built on a uniform 50 x 50 x 25 um grid with 8 CPML cells on every face, a ground
sheet under an 8-cell substrate, the trace (and for GCPW the coplanar grounds) as PEC sheets
on top, and 50 ohm lumped ports 6 cells inside the CPML at both ends of the trace.  The S
matrix follows the two-excitation assembly of the reference's port tests
(web/src/solver/ports.test.ts:78-81): S = B A^-1 with a = (V + Z I) / 2 sqrt Z and
b = (V - Z I) / 2 sqrt Z (gcpw_line.c:379-391).
"""
import math
import threading
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import lib as L

DX, DY, DZ = 50e-6, 50e-6, 25e-6
NPML = 8
PORT_INSET = 6
C0 = 299792458.0


def random_board(seed):
    """A batched-workload board (seed 4000 + b for board b of 64)."""
    rng = np.random.default_rng(seed)
    return dict(
        seed=int(seed),
        nx=int(rng.integers(80, 257)), ny=int(rng.integers(80, 193)), nz=int(rng.integers(60, 97)),
        width=float(rng.uniform(0.1e-3, 0.5e-3)), er=float(rng.uniform(3.0, 4.5)),
        npole=int(rng.integers(2, 7)), kind="gcpw" if rng.random() < 0.5 else "microstrip",
        lossy=True, nport=2, h_cells=8, gap=0.2e-3, nf=101, fmax=40e9,
    )


def random_request(seed):
    """A served-workload board: cells log-normal (median 3e6, sigma 0.8) truncated to
    [0.5e6, 20e6]; nz uniform 60..96 and nx = 1.4 ny; half lossless (plain dielectric, PEC)
    and half lossy (Debye substrate, SIBC copper); half 1-port and half 2-port."""
    rng = np.random.default_rng(seed)
    while True:
        cells = float(np.exp(rng.normal(math.log(3e6), 0.8)))
        if 0.5e6 <= cells <= 20e6:
            break
    nz = int(rng.integers(60, 97))
    ny = max(40, int(round(math.sqrt(cells / nz / 1.4))))
    nx = max(60, int(round(cells / nz / ny)))
    return dict(
        seed=int(seed), nx=nx, ny=ny, nz=nz,
        width=float(rng.uniform(0.1e-3, 0.5e-3)), er=float(rng.uniform(3.0, 4.5)),
        npole=int(rng.integers(2, 7)), kind="gcpw" if rng.random() < 0.5 else "microstrip",
        lossy=bool(rng.random() < 0.5), nport=1 if rng.random() < 0.5 else 2,
        h_cells=8, gap=0.2e-3, nf=101, fmax=40e9,
    )


def cells(b):
    return b["nx"] * b["ny"] * b["nz"]


def board_request(board, net=None, planes=(0.0, 0.0), hmin=20e-6, fmax=40e9, df=0.5e9, lossy=False, sheets=False,
                  full2=True, decay=1e-5, tmax=5e-9, window=None, steps=None):
    """A served request for an imported board (courant_amd.board.Board): the two-port net
    between the port planes, as courant_amd.board.run_board runs it.  Its mesh is planned at
    once, for the admission estimate."""
    from . import board as BD
    if net is None:
        nets = BD.port_nets(board, planes[0], planes[1])
        if not nets:
            raise ValueError("no net crosses both port planes")
        net = nets[0]
    ports = [{"net": net, "edge": "xlo", "offset": planes[0]}, {"net": net, "edge": "xhi", "offset": planes[1]}]
    opts = dict(fmax=fmax, df=df, hmin=hmin, lossy=lossy, thick_copper=not sheets, y_window=window or 0.0)
    plan = BD.plan_board(board, ports, **opts)
    nx, ny, nz = plan["cells"]
    return {"kind": "board", "board": board, "ports": ports, "opts": opts, "nport": 2 if full2 else 1,
            "decay": decay, "tmax": tmax, "steps": steps, "nx": nx, "ny": ny, "nz": nz, "lossy": lossy}


def estimate_bytes(b):
    """Device footprint of one excitation, the admission test of the request loop.  Grows
    from the reference's cell estimate (web/src/mesh/mesher.ts:143-156) by what it leaves
    out: six FP32 field arrays and one byte of coefficient index per component (27 B per
    grid point), the CPML psi planes (two per face, 8 B per plane point), the Debye states
    of the substrate edges (4 B per pole plus 16 B of index, table row and old value), the
    SIBC states (12 terms of FP32 plus 36 B of indices and weights per conductor face), and
    a fixed 16 MiB for DFT series, probes and tables."""
    nx, ny, nz = b["nx"], b["ny"], b["nz"]
    n = (nx + 1) * (ny + 1) * (nz + 1)
    psi = 2 * NPML * 8 * 2 * ((nx + 1) * (ny + 1) + (ny + 1) * (nz + 1) + (nx + 1) * (nz + 1))
    lossy = b.get("lossy", True)
    if b.get("kind") == "board":
        # an imported board: Debye states on up to every edge (8 poles), SIBC on up to
        # three edges per point of every z plane that holds copper, bounded by the grid
        ndisp = 3 * n if lossy else 0
        nsibc = min(3 * n, 3 * (nx + 1) * (ny + 1) * 8) if lossy else 0
        return 27 * n + psi + ndisp * (4 * 8 + 16) + nsibc * (12 * 4 + 36) + (16 << 20)
    ndisp = 3 * (nx + 1) * (ny + 1) * (b["h_cells"] + 1)
    nsibc = 3 * (nx + 1) * (ny + 1)
    return (27 * n + psi + (ndisp * (4 * b["npole"] + 16) if lossy else 0)
            + (nsibc * (12 * 4 + 36) if lossy else 0) + (16 << 20))


def build(cem, b, excite):
    """One excitation of board b (`excite` 0 or 1 names the driven port); returns (sim, ports)."""
    nx, ny, nz = b["nx"], b["ny"], b["nz"]
    X = (np.arange(nx + 1) - nx / 2) * DX
    Y = (np.arange(ny + 1) - ny / 2) * DY
    Z = (np.arange(nz + 1) - (NPML + 2)) * DZ
    s = cem.create_sim(X, Y, Z)
    try:
        for f in range(6):
            s.set_boundary(f, L.CPML, NPML)
        k0 = NPML + 2
        kh = k0 + b["h_cells"]
        xlo, xhi, ylo, yhi = X[0] - 1, X[-1] + 1, Y[0] - 1, Y[-1] + 1
        fmax = b["fmax"]
        freqs = np.linspace(fmax / b["nf"], fmax, b["nf"])
        if b.get("lossy", True):
            m = cem.fit_djordjevic_sarkar(b["er"], 0.02, 1e9, freqs[0], fmax, b["npole"])
            s.set_debye_poles(m["fp"])
            s.material_box_debye(xlo, xhi, ylo, yhi, Z[k0], Z[kh], m["eps_inf"], 0.0, m["deps"])
            s.set_conductor_loss(5.8e7, 0, 0, 0, freqs[0], fmax)
        else:
            s.material_box(xlo, xhi, ylo, yhi, Z[k0], Z[kh], b["er"])
        s.pec_sheet(L.Z, Z[k0], xlo, xhi, ylo, yhi)
        i1, i2 = NPML + PORT_INSET, nx - NPML - PORT_INSET
        w = b["width"]
        s.pec_sheet(L.Z, Z[kh], X[i1], X[i2], -w / 2, w / 2)
        if b["kind"] == "gcpw":
            s.pec_sheet(L.Z, Z[kh], xlo, xhi, ylo, -w / 2 - b["gap"])
            s.pec_sheet(L.Z, Z[kh], xlo, xhi, w / 2 + b["gap"], yhi)
        j0 = ny // 2
        p1 = s.add_lumped_port(L.Z, i1, j0, k0, kh, -1, 50.0)
        p2 = s.add_lumped_port(L.Z, i2, j0, k0, kh, -1, 50.0)
        tau = pulse_tau(fmax)
        s.set_gaussian((p1, p2)[excite], 1.0, 4.5 * tau, tau)
        s.set_frequencies(freqs)
        s.finalize(0.99)
    except Exception:
        s.close()
        raise
    return s, (p1, p2)


def pulse_tau(fmax):
    return math.sqrt(6.9) / (math.pi * 1.25 * fmax)


def steps_for(b, dt):
    """Fixed run length: the pulse (peak at 4.5 tau, then 2 tau of tail) plus five transits of
    the trace at the substrate's phase velocity; a board may name its own count ("steps")."""
    if b.get("steps"):
        return int(b["steps"])
    tau = pulse_tau(b["fmax"])
    length = (b["nx"] - 2 * (NPML + PORT_INSET)) * DX
    t_stop = 6.5 * tau + 5 * length * math.sqrt(b["er"]) / C0
    return int(math.ceil(t_stop / dt))


def solve(cem, b, excite):
    """One right-hand side: build, step, return (V, I) port spectra (2 x nf each) and steps;
    for an imported board, the run of courant_amd.board.run_excitation and its steps."""
    if b.get("kind") == "board":
        from . import board as BD
        r = BD.run_excitation(cem, b["board"], b["ports"], excite, decay=b["decay"], tmax=b["tmax"],
                              steps=b.get("steps"), **b["opts"])
        return r, r["steps"]
    s, ports = build(cem, b, excite)
    try:
        n = steps_for(b, s.dt)
        s.step(n)
        v = np.array([s.port_v(p) for p in ports])
        i = np.array([s.port_i(p) for p in ports])
    finally:
        s.close()
    return (v, i), n


def solve_pair(cem, b):
    """Both right-hand sides of a 2-port lumped board: the board is built once (port 1
    driven), port 2's excitation is a clone of that plan (Sim.clone_excitation: no second
    set-up), and the two advance together (Sim.step_pair: each step's H and E updates of
    both in one launch each).  Returns ([(V, I) of each excitation], steps); the spectra
    equal those of solve(cem, b, 0) and solve(cem, b, 1) bit for bit."""
    s, ports = build(cem, b, 0)
    try:
        c = s.clone_excitation(ports[1])
        try:
            n = steps_for(b, s.dt)
            s.step_pair(c, n)
            out = [(np.array([x.port_v(p) for p in ports]), np.array([x.port_i(p) for p in ports])) for x in (s, c)]
        finally:
            c.close()
    finally:
        s.close()
    return out, n


MALL_BYTES = 256 << 20   # the MI355X's last-level cache (Infinity Cache)


def pairable(b):
    """A 2-port lumped board whose excitations pay off as a pair.  A pair steps two field
    sets through every launch, so its working set is twice an excitation's: a board whose
    excitation fits the 256 MiB last-level cache but whose pair does not loses that
    residency (a 4.7 M-cell board: a pair step took 2.8x an excitation's step, a 38.8
    M-point board 1.98x).  Pairs are formed when both excitations fit the cache together
    (with a quarter to spare) or when one alone does not fit it anyway."""
    if b.get("nport", 2) != 2 or b.get("kind") == "board":
        return False
    need = estimate_bytes(b)
    return 2 * need <= MALL_BYTES * 3 // 4 or need >= MALL_BYTES


def excitations(b):
    return (0, 1) if b.get("nport", 2) == 2 else (0,)


def s_matrix(vis, z0=50.0):
    """S per frequency from the excitations' port spectra.  Two excitations: the full 2 x 2
    S = B A^-1 (ports.test.ts:78-81).  One excitation (port 1 driven): S11 and S21 only,
    as b / a1 (gcpw_line.c:379-391), returned as an nf x 2 x 1 array."""
    sq = math.sqrt(z0)
    nf = vis[0][0].shape[1]
    A = np.empty((nf, 2, len(vis)), complex)
    B = np.empty((nf, 2, len(vis)), complex)
    for q, (v, i) in enumerate(vis):
        A[:, :, q] = ((v + z0 * i) / (2 * sq)).T
        B[:, :, q] = ((v - z0 * i) / (2 * sq)).T
    if len(vis) == 1:
        return B / A[:, 0:1, :]
    return B @ np.linalg.inv(A)


def assemble(cem, b, results):
    """S of board b from its excitations' results: B A^-1 from the lumped ports' spectra,
    or the mode-port wave ratios renormalised to 50 ohm for an imported board."""
    if b.get("kind") == "board":
        from . import board as BD
        return BD.board_s_matrix(cem, results)
    return s_matrix(results)


def jobs_of(boards):
    return [(k, b, e) for k, b in enumerate(boards) for e in excitations(b)]


def run_batch(cem, boards, workers, pairs=False):
    """Every excitation of `boards` on `workers` host threads (1 = plain one-at-a-time).
    pairs: the two excitations of each 2-port board run as one pair (solve_pair).
    Returns (S per board, wall seconds, cell updates)."""
    jobs = []
    for k, b in enumerate(boards):
        if pairs and pairable(b):
            jobs.append((k, b, None))
        else:
            jobs += [(k, b, e) for e in excitations(b)]

    def one(j):
        _, b, e = j
        if e is None:
            vis, n = solve_pair(cem, b)
            return vis, n, 2
        vi, n = solve(cem, b, e)
        return [vi], n, 1

    t0 = time.perf_counter()
    if workers <= 1:
        res = [one(j) for j in jobs]
    else:
        with ThreadPoolExecutor(max_workers=workers) as ex:
            res = list(ex.map(one, jobs))
    wall = time.perf_counter() - t0
    per = [[] for _ in boards]
    upd = 0
    for (k, b, _), (vis, n, nrhs) in zip(jobs, res):
        per[k] += vis
        upd += cells(b) * n * nrhs
    return [assemble(cem, b, p) for b, p in zip(boards, per)], wall, upd


class Server:
    """Request loop with continuous admission.  Each excitation of a submitted board is a
    job; a job starts as soon as a worker is free and its estimated footprint fits the
    remaining device budget, first come first served (a job that does not fit waits while
    running ones finish and give their memory back).  A board's future resolves with its S
    matrix and the time from its arrival to the S matrix."""

    def __init__(self, cem, workers, budget_bytes, pairs=False):
        self.cem, self.budget, self.pairs = cem, budget_bytes, pairs
        self.pool = ThreadPoolExecutor(max_workers=workers)
        self.cv = threading.Condition()
        self.in_use = 0
        self.queue = []          # job tickets in arrival order
        self.ticket = 0
        self.updates = 0

    def _admit(self, need):
        with self.cv:
            t = self.ticket
            self.ticket += 1
            self.queue.append(t)
            while self.queue[0] != t or self.in_use + need > self.budget:
                self.cv.wait()
            self.queue.pop(0)
            self.in_use += need
            self.cv.notify_all()

    def _release(self, need, upd):
        with self.cv:
            self.in_use -= need
            self.updates += upd
            self.cv.notify_all()

    def _job(self, b, e):
        """One excitation (e), or both of a 2-port board as a pair (e None; twice the
        footprint).  Returns ([(V, I) per excitation], completion time)."""
        nrhs = 2 if e is None else 1
        need = nrhs * estimate_bytes(b)
        self._admit(need)
        vis, n = None, 0
        try:
            if e is None:
                vis, n = solve_pair(self.cem, b)
            else:
                vi, n = solve(self.cem, b, e)
                vis = [vi]
        finally:
            self._release(need, nrhs * cells(b) * n if vis is not None else 0)
        return vis, time.perf_counter()

    def submit(self, b):
        """Queue every excitation of board b; returns a callable that blocks for
        (S, seconds from submission to S)."""
        if estimate_bytes(b) > self.budget:
            raise ValueError("board larger than the device budget")
        t_arr = time.perf_counter()
        if self.pairs and pairable(b) and 2 * estimate_bytes(b) <= self.budget:
            futs = [self.pool.submit(self._job, b, None)]
        else:
            futs = [self.pool.submit(self._job, b, e) for e in excitations(b)]
        done = {}

        def result():
            if not done:
                res = [f.result() for f in futs]
                vis = [v for r in res for v in r[0]]
                done["r"] = (assemble(self.cem, b, vis), max(r[1] for r in res) - t_arr)
            return done["r"]

        return result

    def close(self):
        self.pool.shutdown(wait=True)


def serve(cem, boards, rate, workers, budget_bytes, seed=5000, pairs=False):
    """Poisson arrivals (`rate` boards per second) of `boards`; pairs: 2-port boards run
    both excitations as one pair.  Returns (latencies s, wall s, S per board, cell updates)."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, len(boards))
    srv = Server(cem, workers, budget_bytes, pairs=pairs)
    t0 = time.perf_counter()
    handles = []
    t_next = t0
    try:
        for b, gap in zip(boards, gaps):
            t_next += gap
            now = time.perf_counter()
            if t_next > now:
                time.sleep(t_next - now)
            handles.append(srv.submit(b))
        out = [h() for h in handles]
    finally:
        srv.close()
    wall = time.perf_counter() - t0
    return np.array([o[1] for o in out]), wall, [o[0] for o in out], srv.updates
