"""Scene builders that drive any courant.h library through courant_amd.lib.

* parity_scene     - the all-features deterministic scene of solver/apps/parity.c:11-31
* checkpoint_scene - the stateful scene of solver/tests/test_checkpoint.c:221-243
* gcpw             - grounded coplanar waveguide in the style of
                     solver/apps/gcpw_line.c:125-292 and web/src/scenarios/gcpw.ts:94-166,
                     parameterised so the workloads below can be built:
    cfg1  "gcpw_line": ~120x80x60, lossless eps_r 4.3, 0.2 mm substrate, trace
          0.35 mm, gaps 0.2 mm, zero-thickness sheets, ground sheet below, CPML 8
          on all six faces, two lumped ports, 401 frequencies 0-40 GHz
    cfg2  lossy GCPW with mode ports: ~300x150x100 (4.5 M cells), 35 um copper
          meshed with 3 cells, SIBC sigma 5.8e7 with Huray roughness, 6-pole
          Djordjevic-Sarkar substrate (4.1, 0.02 at 1 GHz), 25 um solder mask,
          2D eigenmode ports, 401 frequencies 0.1-40 GHz
* board4           - large single board: four copper layers, three Debye dielectrics,
                     seeded meanders and vias on 600x400x160 non-uniform cells
"""
import math

import numpy as np

from . import lib as L


def graded(fix, ref, hmin, hmax, ratio):
    """Graded 1D mesh: fixed lines kept; spacing hmin at refinement points, growing by
    `ratio` per cell up to hmax; each segment rescaled so its last gap matches its
    neighbour (the rule of gcpw_line.c:41-78)."""
    pts = sorted(fix)
    out = [pts[0]]

    def size_at(x):
        d = min((abs(x - r) for r in ref), default=1e300)
        return min(hmax, hmin + (ratio - 1) * d)

    for a, b in zip(pts[:-1], pts[1:]):
        if b - a < 1e-15:
            continue
        seg = []
        x = a
        while True:
            h = size_at(x)
            h = min(h, size_at(x + h))
            if x + h >= b - 0.5 * h:
                break
            x += h
            seg.append(x)
        if seg:
            last = seg[-1]
            prev = last - seg[-2] if len(seg) > 1 else last - a
            k = (b - a) / ((last - a) + prev)
            out.extend(a + (v - a) * k for v in seg)
        out.append(b)
    return out


def nearest(a, x):
    a = np.asarray(a)
    return int(np.argmin(np.abs(a - x)))


# ---------------------------------------------------------------------------
def parity_scene(cem, steps=None):
    """solver/apps/parity.c:11-31: non-uniform grid, CPML on 5 faces, lossy dielectric,
    PEC box and sheet, two lumped ports, V and I probes; returns (sim, info)."""
    NX, NY, NZ = 24, 18, 14
    x = [1e-4 * i + 2e-6 * i * i for i in range(NX + 1)]
    y = [1e-4 * (j - NY // 2) * (1 + 0.01 * abs(j - NY / 2.0)) for j in range(NY + 1)]
    z = [5e-5 * k + 1e-6 * k * k for k in range(NZ + 1)]
    s = cem.create_sim(x, y, z)
    for f in range(6):
        if f != L.ZLO:
            s.set_boundary(f, L.CPML, 4)
    s.material_box(-1, 1, -1, 1, 0, z[5], 3.4, 0.02)
    s.pec_box(-1, 1, y[NY - 6], 1, 0, z[5])
    s.pec_sheet(L.Z, z[5], x[5], x[19], y[8], y[10])
    p1 = s.add_lumped_port(L.Z, 5, 9, 0, 5, -1, 50)
    p2 = s.add_lumped_port(L.Z, 19, 9, 0, 5, -1, 50)
    s.set_gaussian(p1, 1.0, 60e-12, 15e-12)
    s.set_frequencies([5e9, 20e9, 40e9])
    v = s.add_vprobe(L.Z, 12, 9, 0, 5, -1)
    ip = s.add_iprobe(L.X, 12, 7, 11, 4, 5)
    s.finalize(0.95)
    if steps:
        s.step(steps)
    return s, {"ports": (p1, p2), "vprobe": v, "iprobe": ip}


def checkpoint_scene(cem):
    """solver/tests/test_checkpoint.c:221-243: CPML, 3 Debye poles, SIBC, two lumped ports, probes."""
    NX, NY, NZ = 24, 18, 14
    x = [1e-4 * i + 2e-6 * i * i for i in range(NX + 1)]
    y = [1e-4 * (j - NY // 2) for j in range(NY + 1)]
    z = [5e-5 * k + 1e-6 * k * k for k in range(NZ + 1)]
    s = cem.create_sim(x, y, z)
    for f in range(6):
        if f != L.ZLO:
            s.set_boundary(f, L.CPML, 4)
    s.set_debye_poles([1e9, 10e9, 100e9])
    s.material_box_debye(-1, 1, -1, 1, 0, z[5], 3.4, 0.0, [0.02, 0.03, 0.02])
    s.set_conductor_loss(5.8e7, 0, 0, 0, 1e9, 40e9)
    s.pec_sheet(L.Z, z[5], x[5], x[19], y[8], y[10])
    p1 = s.add_lumped_port(L.Z, 5, 9, 0, 5, -1, 50)
    s.add_lumped_port(L.Z, 19, 9, 0, 5, -1, 50)
    s.set_gaussian(p1, 1.0, 60e-12, 15e-12)
    s.set_frequencies([5e9, 20e9, 40e9])
    s.add_vprobe(L.Z, 12, 9, 0, 5, -1)
    s.add_iprobe(L.X, 12, 7, 11, 4, 5)
    s.finalize(0.95)
    return s


# ---------------------------------------------------------------------------
GCPW_DEFAULTS = dict(
    h=0.2e-3, w=0.35e-3, gap=0.2e-3, fence=0.3e-3, er=4.3,
    hmin=30e-6, hmax=120e-6, ratio=1.3, dx=100e-6, length=9.4e-3, ext=0.5e-3,
    yair=1.0e-3, zair=1.2e-3, zbelow=0.0, npml=8, pml_below=False,
    ports="lumped", fmax=40e9, df=0.1e9, dc=True, cfl=0.99,
    lossy=False, tcu=0.0, mask=0.0, mask_er=3.5, ds=(4.1, 0.02, 1e9), npole=6,
    sigma_cu=5.8e7, rough=2, rough_p1=0.5e-6, rough_p2=2.0 / (4 * math.pi * 0.25e-12),
    feed=4.0e-3, mon=2.5e-3, neff0=1.65, probes=True, excite=1,
)

# cfg1: lumped-port GCPW on 120x80x60 with CPML on all six faces (ground sheet above a bottom air layer)
CFG1 = dict(hmin=20e-6, hmax=50e-6, dx=100e-6, length=9.4e-3, yair=0.6e-3, zair=1.2e-3, zbelow=0.6e-3,
            pml_below=True, ports="lumped", dc=True)
# cfg2: lossy GCPW with mode ports on ~300x150x100
CFG2 = dict(hmin=10e-6, hmax=20e-6, dx=30e-6, length=3.6e-3, feed=2.45e-3, mon=1.5e-3, yair=0.6e-3, zair=1.5e-3,
            zbelow=0.0, pml_below=False, ports="mode", lossy=True, tcu=35e-6, mask=25e-6, dc=False)


def gcpw_mesh(p):
    """Node coordinates (x, y, z) and key indices of the GCPW scene."""
    h, w, gap, fence = p["h"], p["w"], p["gap"], p["fence"]
    a, b = w / 2, w / 2 + gap
    yf = b + fence
    yext = yf + p["yair"]
    tcu, msk = p["tcu"], p["mask"]
    ztop = h + tcu + msk + p["zair"]
    hmin, hmax, npml = p["hmin"], p["hmax"], p["npml"]
    t1, t2 = hmin / 3, 2 * hmin / 3
    if tcu > 0:   # thick copper: grid lines on the copper faces
        Y = graded([-yext, -yf, -b, -a, 0, a, b, yf, yext], [-b, -a, a, b], hmin, hmax, p["ratio"])
    else:         # thirds rule at zero-thickness edges (docs/design/0002-edge-meshing.md)
        Y = graded([-yext, -yf, -b - t1, -b + t2, -a - t2, -a + t1, 0, a - t1, a + t2, b - t2, b + t1, yf, yext],
                   [-b, -a, a, b], hmin, hmax, p["ratio"])
    zfix = [0.0, h, ztop]
    zref = [h]
    if tcu > 0:
        zfix += [h + tcu / 3, h + 2 * tcu / 3, h + tcu]
        zref += [h + tcu]
    if msk > 0:
        zfix += [h + tcu + msk]
    if p["zbelow"] > 0:
        zfix += [-p["zbelow"]]
    Z = graded(zfix, zref, hmin, hmax, p["ratio"])
    Y2 = [Y[0] - i * hmax for i in range(npml, 0, -1)] + Y + [Y[-1] + i * hmax for i in range(1, npml + 1)]
    Z = Z + [ztop + i * hmax for i in range(1, npml + 1)]
    if p["pml_below"]:
        Z = [Z[0] - i * hmax for i in range(npml, 0, -1)] + Z
    mode = p["ports"] == "mode"
    dx = p["dx"]
    nlen = int(round(p["length"] / dx))
    nxt = int(round((p["feed"] if mode else p["ext"]) / dx))
    X = [i * dx for i in range(-nxt - npml, nlen + nxt + npml + 1)]
    return np.array(X), np.array(Y2), np.array(Z), dict(a=a, b=b, yf=yf, nlen=nlen, nxt=nxt)


def gcpw(cem, **over):
    """Build and finalize the GCPW scene; returns (sim, info)."""
    p = dict(GCPW_DEFAULTS)
    p.update(over)
    X, Y, Z, g = gcpw_mesh(p)
    a, b, yf = g["a"], g["b"], g["yf"]
    h, tcu, msk, npml = p["h"], p["tcu"], p["mask"], p["npml"]
    nx, ny, nz = len(X) - 1, len(Y) - 1, len(Z) - 1
    ip1 = g["nxt"] + npml
    ip2 = ip1 + g["nlen"]
    j0 = nearest(Y, 0.0)
    kg = nearest(Z, 0.0)          # ground plane node
    kh = nearest(Z, h)            # substrate top
    kt = nearest(Z, h + tcu)      # copper top
    s = cem.create_sim(X, Y, Z)
    xlo, xhi, ylo, yhi = X[0] - 1, X[-1] + 1, Y[0] - 1, Y[-1] + 1
    zlo = Z[0] - 1
    for f in (L.XLO, L.XHI, L.YLO, L.YHI, L.ZHI):
        s.set_boundary(f, L.CPML, npml)
    if p["pml_below"]:
        s.set_boundary(L.ZLO, L.CPML, npml)
    fmax, df = p["fmax"], p["df"]
    fmin_fit = df
    if p["lossy"]:
        dk, tand, fref = p["ds"]
        m = cem.fit_djordjevic_sarkar(dk, tand, fref, fmin_fit, fmax, p["npole"])
        s.set_debye_poles(m["fp"])
        s.material_box_debye(xlo, xhi, ylo, yhi, 0, h, m["eps_inf"], 0.0, m["deps"])
    else:
        s.material_box(xlo, xhi, ylo, yhi, 0, h, p["er"], 0.0)
    if msk > 0:
        s.material_box(xlo, xhi, ylo, yhi, h + tcu, h + tcu + msk, p["mask_er"], 0.0)
    # ground: sheet at z = 0 over the whole board (the domain floor when ZLO is PEC)
    s.pec_sheet(L.Z, 0.0, xlo, xhi, ylo, yhi)
    # via fences: solid metal for |y| > yf through the substrate
    s.pec_box(xlo, xhi, ylo, -yf, 0, h)
    s.pec_box(xlo, xhi, yf, yhi, 0, h)
    mode = p["ports"] == "mode"
    sx0, sx1 = (xlo, xhi) if mode else (X[ip1], X[ip2])
    if tcu > 0:
        s.pec_box(xlo, xhi, ylo, -b, h, h + tcu)
        s.pec_box(xlo, xhi, b, yhi, h, h + tcu)
        s.pec_box(sx0, sx1, -a, a, h, h + tcu)
    else:
        s.pec_sheet(L.Z, h, xlo, xhi, ylo, -b)
        s.pec_sheet(L.Z, h, xlo, xhi, b, yhi)
        s.pec_sheet(L.Z, h, sx0, sx1, -a, a)
    if p["lossy"]:
        s.set_conductor_loss(p["sigma_cu"], p["rough"], p["rough_p1"], p["rough_p2"], fmin_fit, fmax)
    tau = math.sqrt(6.9) / (math.pi * 1.25 * fmax)
    info = dict(ip1=ip1, ip2=ip2, j0=j0, kg=kg, kh=kh, kt=kt, cells=(nx, ny, nz), tau=tau, params=p)
    if not mode:
        p1 = s.add_lumped_port(L.Z, ip1, j0, kg, kh, -1, 50.0)
        p2 = s.add_lumped_port(L.Z, ip2, j0, kg, kh, -1, 50.0)
        s.set_gaussian(p1 if p["excite"] == 1 else p2, 1.0, 4.5 * tau, tau)
    else:
        jl, jr = nearest(Y, -(a + b) / 2), nearest(Y, (a + b) / 2)
        mon = int(round(p["mon"] / p["dx"]))
        common = dict(j0=0, j1=ny, k0=0, k1=nz, vj=j0, vk0=kg, vk1=kh, vsign=-1, ja=jl, jb=jr - 1, ka=kh - 1,
                      kb=kt, f0=0.5 * fmax, neff0=p["neff0"])
        s1, s2 = npml + 2, nx - npml - 2
        p1 = s.add_mode_port(dir=1, i_src=s1, i_mon=s1 + mon, x_ref=X[ip1], **common)
        p2 = s.add_mode_port(dir=-1, i_src=s2, i_mon=s2 - mon - 1, x_ref=X[ip2], **common)
        s.set_mode_excitation(p1 if p["excite"] == 1 else p2, 4.5 * tau, tau)
    info["ports"] = (p1, p2)
    nf = int(round(fmax / df))
    if p["dc"]:
        freqs = [i * df for i in range(nf + 1)]
    else:
        freqs = [(i + 1) * df for i in range(nf)]
    s.set_frequencies(freqs)
    info["freqs"] = np.array(freqs)
    probes = []
    if p["probes"] and not mode:
        jl, jr = nearest(Y, -(a + b) / 2), nearest(Y, (a + b) / 2)
        step = max(1, int(round(1e-3 / p["dx"])))
        m = ip1 + 2 * step
        while m <= ip2 - 2 * step and len(probes) < 20:
            probes.append((s.add_vprobe(L.Z, m, j0, kg, kh, -1), s.add_iprobe(L.X, m, jl, jr - 1, kh - 1, kt)))
            m += step
    info["probes"] = probes
    s.finalize(p["cfl"])
    return s, info


def gcpw_sparams(s, info):
    """Port-1-excited S11, S21 of a GCPW scene (lumped: V/I; mode: wave amplitudes)."""
    p1, p2 = info["ports"]
    if info["params"]["ports"] == "lumped":
        return L.s_params(s.port_v(p1), s.port_i(p1), s.port_v(p2), s.port_i(p2), 50.0)
    w1, w2 = s.mode_port_waves(p1), s.mode_port_waves(p2)
    return w1["b"] / w1["a"], w2["b"] / w1["a"]


def cells(info):
    nx, ny, nz = info["cells"]
    return nx * ny * nz


# ---------------------------------------------------------------------------
# Large single board: 4-layer stackup on 600 x 400 x 160 cells (38.8 M grid points)
BOARD4 = dict(
    nx=600, ny=400, dx=0.1e-3, dy=0.1e-3, npml=8,
    layers=(0.2e-3, 0.3e-3, 0.2e-3),    # dielectric thicknesses between copper layers 1-2, 2-3, 3-4
    zair=3.0e-3, hd=25e-6, nz=160,
    meanders=5, width=(0.1e-3, 0.4e-3), length=(10e-3, 40e-3), seed_layer=3000,
    vias=40, via_r=0.15e-3, seed_via=3100,
    ds=(4.1, 0.02, 1e9), npole=6, sigma_cu=5.8e7, rough=2, rough_p1=0.5e-6,
    rough_p2=2.0 / (4 * math.pi * 0.25e-12), fmax=40e9, nf=801, cfl=0.99,
)


def _z_stack(p):
    """z lines: uniform cells of hd through the dielectric stack (copper planes at the layer
    interfaces), air on both sides growing geometrically from hd to fill zair, the growth
    ratio chosen so the domain has exactly p["nz"] cells with npml CPML cells per side at
    the outermost air spacing."""
    hd, npml = p["hd"], p["npml"]
    zc = [0.0]
    zd = [0.0]
    for t in p["layers"]:
        m = max(1, int(round(t / hd)))
        zd += [zd[-1] + t * (i + 1) / m for i in range(m)]
        zc.append(zd[-1])
    nair = p["nz"] - 2 * npml - (len(zd) - 1)
    if nair < 4:
        raise ValueError("board4: not enough cells for the air layers")

    def side(n):
        lo, hi = 1.0, 2.0   # geometric growth ratio with hd * (r^n - 1) / (r - 1) = zair
        if hd * n >= p["zair"]:
            return [hd * p["zair"] / (hd * n)] * n
        for _ in range(200):
            r = 0.5 * (lo + hi)
            if hd * (r ** n - 1) / (r - 1) > p["zair"]:
                hi = r
            else:
                lo = r
        r = 0.5 * (lo + hi)
        h = [hd * r ** i for i in range(n)]
        f = p["zair"] / sum(h)
        return [v * f for v in h]

    lo_h, hi_h = side(nair // 2), side(nair - nair // 2)
    Z = [0.0]
    for h in lo_h:
        Z.insert(0, Z[0] - h)
    Z += zd[1:]
    for h in hi_h:
        Z.append(Z[-1] + h)
    Z = [Z[0] - lo_h[-1] * i for i in range(npml, 0, -1)] + Z + [Z[-1] + hi_h[-1] * i for i in range(1, npml + 1)]
    return np.array(Z), zc


def _meander(rng, xlo, xhi, ylo, yhi, w, length):
    """Axis-aligned serpentine of total centre-line length `length`, width w: a list of
    rectangles (x0, x1, y0, y1)."""
    amp = rng.uniform(1.0e-3, 3.0e-3)
    pitch = 4 * w
    x = rng.uniform(xlo + 1e-3, xhi - amp - 1e-3)
    y = rng.uniform(ylo + 1e-3, yhi - 1e-3)
    rects, left, dirn, pts = [], length, 1, [(x, y)]
    horiz = True
    while left > 0:
        if horiz:
            run = min(amp, left)
            nxp = x + dirn * run
            rects.append((min(x, nxp) - w / 2, max(x, nxp) + w / 2, y - w / 2, y + w / 2))
            x = nxp
            dirn = -dirn
        else:
            run = min(pitch, left)
            if y + run > yhi - 1e-3:
                run = -run
            nyp = y + run
            rects.append((x - w / 2, x + w / 2, min(y, nyp) - w / 2, max(y, nyp) + w / 2))
            y = nyp
        left -= abs(run)
        horiz = not horiz
        pts.append((x, y))
    return rects, pts


def board4(cem, **over):
    """Large single board: copper layers 1-4 at the dielectric interfaces, each with seeded
    meander traces (PEC edges on the layer plane), PEC via cylinders through the stack,
    Djordjevic-Sarkar Debye dielectrics, SIBC on every conductor, two lumped ports on the
    first meander of layer 1, CPML on all faces."""
    p = dict(BOARD4)
    p.update(over)
    npml = p["npml"]
    nxi, nyi = p["nx"] - 2 * npml, p["ny"] - 2 * npml
    X = np.array([(i - npml) * p["dx"] for i in range(p["nx"] + 1)])
    Y = np.array([(j - npml) * p["dy"] for j in range(p["ny"] + 1)])
    Z, zc = _z_stack(p)
    nx, ny, nz = len(X) - 1, len(Y) - 1, len(Z) - 1
    s = cem.create_sim(X, Y, Z)
    for f in range(6):
        s.set_boundary(f, L.CPML, npml)
    xlo, xhi, ylo, yhi = X[0] - 1, X[-1] + 1, Y[0] - 1, Y[-1] + 1
    bx0, bx1, by0, by1 = 0.0, nxi * p["dx"], 0.0, nyi * p["dy"]
    fmax = p["fmax"]
    freqs = np.linspace(fmax / p["nf"], fmax, p["nf"])
    dk, tand, fref = p["ds"]
    m = cem.fit_djordjevic_sarkar(dk, tand, fref, freqs[0], fmax, p["npole"])
    s.set_debye_poles(m["fp"])
    s.material_box_debye(bx0, bx1, by0, by1, zc[0], zc[-1], m["eps_inf"], 0.0, m["deps"])
    # meanders on each copper plane
    XX, YY = np.meshgrid(X, Y, indexing="ij")      # nodes (nx+1, ny+1)
    ks = [nearest(Z, z) for z in zc]
    first = None
    for lay, k in enumerate(ks):
        rng = np.random.default_rng(p["seed_layer"] + lay)
        inside = np.zeros((nx + 1, ny + 1), dtype=bool)
        for t in range(p["meanders"]):
            w = rng.uniform(*p["width"])
            rects, pts = _meander(rng, bx0, bx1, by0, by1, w, rng.uniform(*p["length"]))
            for (x0, x1, y0, y1) in rects:
                inside |= (XX >= x0 - 1e-12) & (XX <= x1 + 1e-12) & (YY >= y0 - 1e-12) & (YY <= y1 + 1e-12)
            if lay == 0 and t == 0:
                first = pts
        mu = inside[:-1, :] & inside[1:, :]      # x edges (nx, ny+1)
        mv = inside[:, :-1] & inside[:, 1:]      # y edges (nx+1, ny)
        s.pec_edges_plane(L.Z, k, mu.astype(np.uint8), mv.astype(np.uint8))
    # vias: PEC cells whose centres lie within the cylinder, through the dielectric stack
    rng = np.random.default_rng(p["seed_via"])
    xc = 0.5 * (X[:-1] + X[1:])
    yc = 0.5 * (Y[:-1] + Y[1:])
    k0, k1 = ks[0], ks[-1]
    r = p["via_r"]
    for _ in range(p["vias"]):
        vx, vy = rng.uniform(bx0 + 1e-3, bx1 - 1e-3), rng.uniform(by0 + 1e-3, by1 - 1e-3)
        i0, i1 = int(np.searchsorted(xc, vx - r)), int(np.searchsorted(xc, vx + r))
        j0, j1 = int(np.searchsorted(yc, vy - r)), int(np.searchsorted(yc, vy + r))
        if i1 <= i0 or j1 <= j0:
            continue
        cx, cy = np.meshgrid(xc[i0:i1], yc[j0:j1], indexing="ij")
        disk = (cx - vx) ** 2 + (cy - vy) ** 2 <= r * r
        mask = np.repeat(disk[:, :, None], k1 - k0, axis=2).astype(np.uint8)
        s.pec_cells(i0, j0, k0, mask)
    s.set_conductor_loss(p["sigma_cu"], p["rough"], p["rough_p1"], p["rough_p2"], freqs[0], fmax)
    # lumped ports from layer 1 down to layer 2 at the two ends of the first meander
    (xa, ya), (xb, yb) = first[0], first[-1]
    ia, ja, ib, jb = nearest(X, xa), nearest(Y, ya), nearest(X, xb), nearest(Y, yb)
    p1 = s.add_lumped_port(L.Z, ia, ja, ks[0], ks[1], -1, 50.0)
    p2 = s.add_lumped_port(L.Z, ib, jb, ks[0], ks[1], -1, 50.0)
    tau = math.sqrt(6.9) / (math.pi * 1.25 * fmax)
    s.set_gaussian(p1, 1.0, 4.5 * tau, tau)
    s.set_frequencies(freqs)
    s.finalize(p["cfl"])
    info = dict(cells=(nx, ny, nz), ports=(p1, p2), freqs=freqs, params=p, layers=ks, tau=tau)
    return s, info
