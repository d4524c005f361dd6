// Host-side set-up of a simulation: creation, geometry rasterisers, ports and
// probes, finalize (time step, CPML profiles, edge coefficients, lumped ports,
// Debye edge list), plan export, checkpoints and the C entry points that wrap
// the device engine.
//
// Semantics follow solver/src/sim.c and solver/src/export.c of the reference
// (line numbers cited per function).  Floating-point set-up arithmetic keeps the
// reference's operation order and is compiled without contraction, so the
// exported plan (coefficients and every integer list) is bit-identical to the
// reference's for the same calls.
#include "courant_amd.h"
#include "sim.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <atomic>
#include <new>
#include <thread>
#include <vector>

using namespace cem;

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

// ---- errors (sim.c:16-24; per thread here because serving runs many sims) ----
static thread_local char g_err[256] = "";

void cem::set_error(const char *m) {
    std::snprintf(g_err, sizeof g_err, "%s", m);
}

extern "C" const char *cem_last_error(void) { return g_err; }
extern "C" const char *cem_version(void) { return "20261017.0-mi355x"; }

#define REQUIRE_SETUP(s)                                  \
    do {                                                  \
        if ((s)->finalized) {                             \
            set_error("already finalized");               \
            return -1;                                    \
        }                                                 \
    } while (0)

// ---- trivial accessors (sim.c:26-33) ----
extern "C" size_t cem_index(const cem_sim *s, int i, int j, int k) { return IDX(s, i, j, k); }
extern "C" void cem_dims(const cem_sim *s, int *nx, int *ny, int *nz) {
    *nx = s->n[0];
    *ny = s->n[1];
    *nz = s->n[2];
}
extern "C" double cem_dt(const cem_sim *s) { return s->dt; }
extern "C" long cem_step_count(const cem_sim *s) { return s->step; }
extern "C" int cem_nfreq(const cem_sim *s) { return s->nf; }

// ---- create / destroy (sim.c:40-107) ----
extern "C" cem_sim *cem_create(int nx, int ny, int nz, const double *x, const double *y, const double *z) {
    if (nx < 2 || ny < 2 || nz < 2) {
        set_error("grid needs at least 2 cells per axis");
        return nullptr;
    }
    cem_sim *s = new (std::nothrow) cem_sim();
    if (!s) {
        set_error("out of memory");
        return nullptr;
    }
    try {
        s->n[0] = nx;
        s->n[1] = ny;
        s->n[2] = nz;
        const double *src[3] = {x, y, z};
        for (int a = 0; a < 3; a++) {
            int n = s->n[a];
            s->node[a].assign(src[a], src[a] + n + 1);
            s->d[a].assign(n, 0.0);
            s->dd[a].assign(n + 1, 0.0);
            for (int i = 0; i < n; i++) {
                s->d[a][i] = src[a][i + 1] - src[a][i];
                if (!(s->d[a][i] > 0)) {
                    set_error("node coordinates must be strictly increasing");
                    delete s;
                    return nullptr;
                }
            }
            for (int i = 0; i <= n; i++) {
                double lo = i > 0 ? s->d[a][i - 1] : 0.0, hi = i < n ? s->d[a][i] : 0.0;
                s->dd[a][i] = 0.5 * (lo + hi);
            }
        }
        s->sj = (size_t)nz + 1;
        s->si = s->sj * ((size_t)ny + 1);
        s->N = s->si * ((size_t)nx + 1);
        for (int c = 0; c < 3; c++) {
            s->ca[c].assign(s->N, 0.0);
            s->cb[c].assign(s->N, 0.0);
            s->edge_pec[c].assign(s->N, 0);
        }
        s->eps_r.assign(s->N, 1.0);
        s->sigma.assign(s->N, 0.0);
        s->cell_pec.assign(s->N, 0);
    } catch (const std::bad_alloc &) {
        set_error("out of memory");
        delete s;
        return nullptr;
    }
    return s;
}

extern "C" void cem_destroy(cem_sim *s) {
    if (!s) return;
    dev_destroy(s);
    modeports_free(s);
    delete s;
}

// ---- boundary (sim.c:109-117) ----
extern "C" int cem_set_boundary(cem_sim *s, int face, int type, int ncells) {
    REQUIRE_SETUP(s);
    if (face < 0 || face > 5) {
        set_error("bad face");
        return -1;
    }
    int a = face / 2;
    if (type == CEM_BC_CPML && (ncells < 1 || ncells * 2 >= s->n[a])) {
        set_error("bad CPML thickness");
        return -1;
    }
    s->bc[face] = type;
    s->npml[face] = type == CEM_BC_CPML ? ncells : 0;
    return 0;
}

// ---- geometry (sim.c:121-240) ----
static int nearest_node(const cem_sim *s, int a, double pos) {
    int best = 0;
    double bd = std::fabs(s->node[a][0] - pos);
    for (int i = 1; i <= s->n[a]; i++) {
        double d = std::fabs(s->node[a][i] - pos);
        if (d < bd) {
            bd = d;
            best = i;
        }
    }
    return best;
}

// cells whose centre lies in [lo, hi]; returns 0 when none
static int cell_range(const cem_sim *s, int a, double lo, double hi, int *i0, int *i1) {
    *i0 = s->n[a];
    *i1 = -1;
    for (int i = 0; i < s->n[a]; i++) {
        double c = 0.5 * (s->node[a][i] + s->node[a][i + 1]);
        if (c >= lo && c <= hi) {
            if (i < *i0) *i0 = i;
            if (i > *i1) *i1 = i;
        }
    }
    return *i1 >= *i0;
}

static int box_cells(const cem_sim *s, double x0, double x1, double y0, double y1, double z0, double z1,
                     int r[3][2]) {
    double lo[3] = {x0, y0, z0}, hi[3] = {x1, y1, z1};
    for (int a = 0; a < 3; a++)
        if (!cell_range(s, a, lo[a], hi[a], &r[a][0], &r[a][1])) return 0;
    return 1;
}

extern "C" int cem_material_box(cem_sim *s, double x0, double x1, double y0, double y1, double z0, double z1,
                                double eps_r, double sigma) {
    REQUIRE_SETUP(s);
    int r[3][2];
    if (!box_cells(s, x0, x1, y0, y1, z0, z1, r)) return 0;
    for (int i = r[0][0]; i <= r[0][1]; i++)
        for (int j = r[1][0]; j <= r[1][1]; j++)
            for (int k = r[2][0]; k <= r[2][1]; k++) {
                size_t q = IDX(s, i, j, k);
                s->eps_r[q] = eps_r;
                s->sigma[q] = sigma;
            }
    return 0;
}

extern "C" int cem_pec_box(cem_sim *s, double x0, double x1, double y0, double y1, double z0, double z1) {
    REQUIRE_SETUP(s);
    int r[3][2];
    if (!box_cells(s, x0, x1, y0, y1, z0, z1, r)) return 0;
    for (int i = r[0][0]; i <= r[0][1]; i++)
        for (int j = r[1][0]; j <= r[1][1]; j++)
            for (int k = r[2][0]; k <= r[2][1]; k++) s->cell_pec[IDX(s, i, j, k)] = 1;
    return 0;
}

static size_t at3(const cem_sim *s, int axis, int m, int ua, int u, int va, int v) {
    int ijk[3];
    ijk[axis] = m;
    ijk[ua] = u;
    ijk[va] = v;
    return IDX(s, ijk[0], ijk[1], ijk[2]);
}

extern "C" int cem_pec_sheet(cem_sim *s, int axis, double pos, double a0, double a1, double b0, double b1) {
    REQUIRE_SETUP(s);
    if (axis < 0 || axis > 2) {
        set_error("bad axis");
        return -1;
    }
    int m = nearest_node(s, axis, pos);
    int ua = (axis + 1) % 3, va = (axis + 2) % 3;
    const cem::dvec &U = s->node[ua], &V = s->node[va];
    double tol = 1e-9 * (U[s->n[ua]] - U[0] + V[s->n[va]] - V[0]);
    for (int u = 0; u < s->n[ua]; u++) {          // edges along ua
        if (U[u] < a0 - tol || U[u + 1] > a1 + tol) continue;
        for (int v = 0; v <= s->n[va]; v++) {
            if (V[v] < b0 - tol || V[v] > b1 + tol) continue;
            s->edge_pec[ua][at3(s, axis, m, ua, u, va, v)] = 1;
        }
    }
    for (int u = 0; u <= s->n[ua]; u++) {         // edges along va
        if (U[u] < a0 - tol || U[u] > a1 + tol) continue;
        for (int v = 0; v < s->n[va]; v++) {
            if (V[v] < b0 - tol || V[v + 1] > b1 + tol) continue;
            s->edge_pec[va][at3(s, axis, m, ua, u, va, v)] = 1;
        }
    }
    return 0;
}

extern "C" int cem_pec_edges_plane(cem_sim *s, int axis, int m, const unsigned char *mask_u,
                                   const unsigned char *mask_v) {
    REQUIRE_SETUP(s);
    if (axis < 0 || axis > 2 || m < 0 || m > s->n[axis]) {
        set_error("bad plane");
        return -1;
    }
    int ua = (axis + 1) % 3, va = (axis + 2) % 3;
    int nu = s->n[ua], nv = s->n[va];
    if (mask_u)
        for (int u = 0; u < nu; u++)
            for (int v = 0; v <= nv; v++)
                if (mask_u[(size_t)u * (nv + 1) + v]) s->edge_pec[ua][at3(s, axis, m, ua, u, va, v)] = 1;
    if (mask_v)
        for (int u = 0; u <= nu; u++)
            for (int v = 0; v < nv; v++)
                if (mask_v[(size_t)u * nv + v]) s->edge_pec[va][at3(s, axis, m, ua, u, va, v)] = 1;
    return 0;
}

extern "C" int cem_pec_cells(cem_sim *s, int i0, int j0, int k0, int ni, int nj, int nk, const unsigned char *mask) {
    REQUIRE_SETUP(s);
    if (i0 < 0 || j0 < 0 || k0 < 0 || i0 + ni > s->n[0] || j0 + nj > s->n[1] || k0 + nk > s->n[2]) {
        set_error("bad cell box");
        return -1;
    }
    for (int i = 0; i < ni; i++)
        for (int j = 0; j < nj; j++)
            for (int k = 0; k < nk; k++)
                if (mask[((size_t)i * nj + j) * nk + k]) s->cell_pec[IDX(s, i0 + i, j0 + j, k0 + k)] = 1;
    return 0;
}

// ---- ports and probes (sim.c:244-340) ----
static int edge_column_ok(const cem_sim *s, int axis, int n1, int n2, int m0, int m1) {
    if (axis < 0 || axis > 2) return 0;
    int ua = (axis + 1) % 3, va = (axis + 2) % 3;
    if (n1 < 0 || n1 > s->n[ua] || n2 < 0 || n2 > s->n[va]) return 0;
    if (m0 < 0 || m1 > s->n[axis] || m1 <= m0) return 0;
    return 1;
}

static size_t column_index(const cem_sim *s, int axis, int n1, int n2, int m) {
    return at3(s, axis, m, (axis + 1) % 3, n1, (axis + 2) % 3, n2);
}

extern "C" int cem_add_lumped_port(cem_sim *s, int axis, int n1, int n2, int m0, int m1, int dir, double R) {
    REQUIRE_SETUP(s);
    if ((int)s->port.size() >= kMaxPorts) {
        set_error("too many ports");
        return -1;
    }
    if (!edge_column_ok(s, axis, n1, n2, m0, m1) || !(R > 0) || (dir != 1 && dir != -1)) {
        set_error("bad port definition");
        return -1;
    }
    LumpedPort p;
    p.axis = axis;
    p.dir = dir;
    p.R = R;
    for (int e = 0; e < m1 - m0; e++) {
        p.edge.push_back(column_index(s, axis, n1, n2, m0 + e));
        p.dl.push_back(s->d[axis][m0 + e]);
    }
    p.cs.assign(p.edge.size(), 0.0);
    s->port.push_back(std::move(p));
    return (int)s->port.size() - 1;
}

extern "C" int cem_set_gaussian(cem_sim *s, int port, double amp, double t0, double tau) {
    if (port < 0 || port >= (int)s->port.size()) {
        set_error("bad port");
        return -1;
    }
    s->src_port = port;
    s->src_amp = amp;
    s->src_t0 = t0;
    s->src_tau = tau;
    dev_sources_changed(s);
    return 0;
}

extern "C" int cem_set_frequencies(cem_sim *s, int nf, const double *f) {
    REQUIRE_SETUP(s);
    if (nf < 1) {
        set_error("need at least one frequency");
        return -1;
    }
    s->nf = nf;
    s->freq.assign(f, f + nf);
    return 0;
}

extern "C" int cem_add_vprobe(cem_sim *s, int axis, int n1, int n2, int m0, int m1, int dir) {
    REQUIRE_SETUP(s);
    if ((int)s->vp.size() >= kMaxProbes) {
        set_error("too many probes");
        return -1;
    }
    if (!edge_column_ok(s, axis, n1, n2, m0, m1)) {
        set_error("bad probe");
        return -1;
    }
    VoltageProbe p;
    p.axis = axis;
    p.dir = dir;
    for (int e = 0; e < m1 - m0; e++) {
        p.edge.push_back(column_index(s, axis, n1, n2, m0 + e));
        p.dl.push_back(s->d[axis][m0 + e] * dir);
    }
    s->vp.push_back(std::move(p));
    return (int)s->vp.size() - 1;
}

extern "C" int cem_add_iprobe(cem_sim *s, int axis, int m, int u0, int u1, int v0, int v1) {
    REQUIRE_SETUP(s);
    if ((int)s->ip.size() >= kMaxProbes) {
        set_error("too many probes");
        return -1;
    }
    int ua = (axis + 1) % 3, va = (axis + 2) % 3;
    if (axis < 0 || axis > 2 || m < 0 || m >= s->n[axis] || u0 < 0 || u1 >= s->n[ua] || u1 <= u0 || v0 < 0 ||
        v1 >= s->n[va] || v1 <= v0) {
        set_error("bad current probe");
        return -1;
    }
    CurrentProbe p;
    auto add = [&](int comp, size_t idx, double w) {
        p.comp.push_back(comp);
        p.idx.push_back(idx);
        p.w.push_back(w);
    };
    // bottom (v0+1/2, +u) and top (v1+1/2, -u) sides: H along ua at nodes u
    for (int u = u0 + 1; u <= u1; u++) {
        add(3 + ua, at3(s, axis, m, ua, u, va, v0), s->dd[ua][u]);
        add(3 + ua, at3(s, axis, m, ua, u, va, v1), -s->dd[ua][u]);
    }
    // right (u1+1/2, +v) and left (u0+1/2, -v) sides: H along va at nodes v
    for (int v = v0 + 1; v <= v1; v++) {
        add(3 + va, at3(s, axis, m, ua, u1, va, v), s->dd[va][v]);
        add(3 + va, at3(s, axis, m, ua, u0, va, v), -s->dd[va][v]);
    }
    s->ip.push_back(std::move(p));
    return (int)s->ip.size() - 1;
}

// ---- finalize (sim.c:344-567) ----
void cem::edge_material(const cem_sim *s, int c, int i, int j, int k, double *eps, double *sig, int *pec) {
    int ua = (c + 1) % 3, va = (c + 2) % 3;
    int ijk[3] = {i, j, k};
    double wsum = 0, es = 0, ss = 0;
    *pec = 0;
    for (int du = -1; du <= 0; du++)
        for (int dv = -1; dv <= 0; dv++) {
            int cc[3] = {ijk[0], ijk[1], ijk[2]};
            cc[ua] += du;
            cc[va] += dv;
            if (cc[ua] < 0 || cc[ua] >= s->n[ua] || cc[va] < 0 || cc[va] >= s->n[va]) continue;
            if (cc[c] >= s->n[c]) continue;
            size_t q = IDX(s, cc[0], cc[1], cc[2]);
            double w = s->d[ua][cc[ua]] * s->d[va][cc[va]];
            if (s->cell_pec[q]) *pec = 1;
            es += w * s->eps_r[q];
            ss += w * s->sigma[q];
            wsum += w;
        }
    *eps = wsum > 0 ? es / wsum : 1.0;
    *sig = wsum > 0 ? ss / wsum : 0.0;
}

// CPML grading (sim.c:369-425): polynomial m = 3, kappa_max = 5, alpha_max = 0.05
static void setup_pml_axis(cem_sim *s, int a) {
    int n = s->n[a];
    s->ide[a].assign(n + 1, 0.0);
    s->idh[a].assign(n, 0.0);
    s->be[a].assign(n + 1, 0.0);
    s->ce[a].assign(n + 1, 0.0);
    s->bh[a].assign(n, 0.0);
    s->ch[a].assign(n, 0.0);
    s->pml_sn[a].assign(n + 1, 0.0);
    s->pml_an[a].assign(n + 1, 0.0);
    s->pml_kn[a].assign(n + 1, 1.0);
    s->pml_sc[a].assign(n, 0.0);
    s->pml_ac[a].assign(n, 0.0);
    s->pml_kc[a].assign(n, 1.0);
    for (int i = 0; i <= n; i++) s->ide[a][i] = s->dd[a][i] > 0 ? 1.0 / s->dd[a][i] : 0.0;
    for (int i = 0; i < n; i++) s->idh[a][i] = 1.0 / s->d[a][i];
    const double m = 3.0, kmax = 5.0, amax = 0.05;
    const cem::dvec &X = s->node[a];
    for (int side = 0; side < 2; side++) {
        int np = s->npml[2 * a + side];
        if (!np) continue;
        double xi = side == 0 ? X[np] : X[n - np];
        double L = side == 0 ? X[np] - X[0] : X[n] - X[n - np];
        double dmean = L / np;
        double smax = 0.8 * (m + 1) / (376.730313668 * dmean);
        auto grade = [&](double pos, double &sg, double &kp, double &al) {
            double depth = side == 0 ? (xi - pos) / L : (pos - xi) / L;
            if (depth <= 0) return false;
            if (depth > 1) depth = 1;
            sg = smax * std::pow(depth, m);
            kp = 1 + (kmax - 1) * std::pow(depth, m);
            al = amax * (1 - depth);
            return true;
        };
        for (int i = 0; i <= n; i++) {
            double sg, kp, al;
            if (!grade(X[i], sg, kp, al)) continue;
            double b = std::exp(-(sg / kp + al) * s->dt / CEM_EPS0);
            s->be[a][i] = b;
            s->ce[a][i] = sg > 0 ? sg / (sg * kp + kp * kp * al) * (b - 1) : 0;
            s->ide[a][i] = 1.0 / (kp * s->dd[a][i]);
            s->pml_sn[a][i] = sg;
            s->pml_kn[a][i] = kp;
            s->pml_an[a][i] = al;
        }
        for (int i = 0; i < n; i++) {
            double sg, kp, al;
            if (!grade(0.5 * (X[i] + X[i + 1]), sg, kp, al)) continue;
            double b = std::exp(-(sg / kp + al) * s->dt / CEM_EPS0);
            s->bh[a][i] = b;
            s->ch[a][i] = sg > 0 ? sg / (sg * kp + kp * kp * al) * (b - 1) : 0;
            s->idh[a][i] = 1.0 / (kp * s->d[a][i]);
            s->pml_sc[a][i] = sg;
            s->pml_kc[a][i] = kp;
            s->pml_ac[a][i] = al;
        }
    }
}

// Compact CPML auxiliary layout (sim.c:641-678): psiX[c][w] belongs to the
// derivative axis (c + w + 1) % 3 and exists only on that axis's slab planes,
// low slab then high slab; the in-plane index runs over the other two axes at
// full node extent.
void cem::pml_ranges(const cem_sim *s, int a, int is_e, int r[2][2]) {
    int n = s->n[a];
    int lo = s->npml[2 * a], hi = s->npml[2 * a + 1];
    if (is_e) {   // E nodes 1..lo and n-hi..n-1
        r[0][0] = 1;
        r[0][1] = lo;
        r[1][0] = n - hi;
        r[1][1] = n - 1;
        if (lo == 0) { r[0][0] = 1; r[0][1] = 0; }
        if (hi == 0) { r[1][0] = 1; r[1][1] = 0; }
    } else {      // H cells 0..lo-1 and n-hi..n-1
        r[0][0] = 0;
        r[0][1] = lo - 1;
        r[1][0] = n - hi;
        r[1][1] = n - 1;
    }
}

size_t cem::pml_per(const cem_sim *s, int a) {
    const int *n = s->n;
    return a == 0 ? (size_t)(n[1] + 1) * (n[2] + 1)
                  : a == 1 ? (size_t)(n[0] + 1) * (n[2] + 1) : (size_t)(n[0] + 1) * (n[1] + 1);
}

static size_t pml_planes(const cem_sim *s, int a, int is_e) {
    int r[2][2];
    pml_ranges(s, a, is_e, r);
    size_t n = 0;
    for (int sd = 0; sd < 2; sd++)
        if (r[sd][1] >= r[sd][0]) n += (size_t)(r[sd][1] - r[sd][0] + 1);
    return n;
}

size_t cem::psi_len(const cem_sim *s, int is_e, int comp, int w) {
    if (!s->has_pml) return 0;
    int a = (comp + w + 1) % 3;
    return pml_planes(s, a, is_e) * pml_per(s, a);
}

extern "C" int cem_finalize(cem_sim *s, double cfl) {
    REQUIRE_SETUP(s);
    if (!s->nf) {
        double f = 1e9;
        cem_set_frequencies(s, 1, &f);
    }
    try {
        double dmin[3];
        for (int a = 0; a < 3; a++) {
            dmin[a] = 1e300;
            for (int i = 0; i < s->n[a]; i++)
                if (s->d[a][i] < dmin[a]) dmin[a] = s->d[a][i];
        }
        if (!(cfl > 0 && cfl <= 1)) cfl = 0.99;
        s->dt = cfl / (CEM_C0 * std::sqrt(1 / (dmin[0] * dmin[0]) + 1 / (dmin[1] * dmin[1]) +
                                          1 / (dmin[2] * dmin[2])));
        for (int a = 0; a < 3; a++) setup_pml_axis(s, a);
        s->has_pml = 0;
        for (int f = 0; f < 6; f++)
            if (s->npml[f]) s->has_pml = 1;

        // E-edge coefficients (sim.c:449-465), one (component, x plane) per task
        const int nx = s->n[0], ny = s->n[1], nz = s->n[2];
        parallel_ranges(3L * (nx + 1), [&](long r0, long r1) {
            for (long r = r0; r < r1; r++) {
                const int c = (int)(r / (nx + 1)), i = (int)(r % (nx + 1));
                for (int j = 0; j <= ny; j++)
                    for (int k = 0; k <= nz; k++) {
                        int ijk[3] = {i, j, k};
                        if (ijk[c] >= s->n[c]) continue;
                        size_t q = IDX(s, i, j, k);
                        double er, sg;
                        int pec;
                        edge_material(s, c, i, j, k, &er, &sg, &pec);
                        if (pec || s->edge_pec[c][q]) {
                            s->ca[c][q] = 0;
                            s->cb[c][q] = 0;
                            continue;
                        }
                        double eps = er * CEM_EPS0, h = sg * s->dt / (2 * eps);
                        s->ca[c][q] = (1 - h) / (1 + h);
                        s->cb[c][q] = (s->dt / eps) / (1 + h);
                    }
            }
        });

        // lumped ports: resistive voltage source per edge (sim.c:467-492)
        for (auto &pt : s->port) {
            double L = 0;
            for (double v : pt.dl) L += v;
            for (size_t e = 0; e < pt.edge.size(); e++) {
                size_t q = pt.edge[e];
                int i = (int)(q / s->si), j = (int)((q % s->si) / s->sj), k = (int)(q % s->sj);
                int ijk[3] = {i, j, k};
                int ua = (pt.axis + 1) % 3, va = (pt.axis + 2) % 3;
                double A = s->dd[ua][ijk[ua]] * s->dd[va][ijk[va]];
                double er, sg;
                int pec;
                edge_material(s, pt.axis, i, j, k, &er, &sg, &pec);
                if (pec || s->edge_pec[pt.axis][q]) {
                    set_error("port edge is PEC");
                    return -1;
                }
                double eps = er * CEM_EPS0;
                double Re = pt.R * pt.dl[e] / L;
                double beta = s->dt * pt.dl[e] / (2 * Re * A * eps) + sg * s->dt / (2 * eps);
                s->ca[pt.axis][q] = (1 - beta) / (1 + beta);
                s->cb[pt.axis][q] = (s->dt / eps) / (1 + beta);
                pt.cs[e] = s->dt / (Re * A * eps * (1 + beta)) * (pt.dl[e] / L) * pt.dir;
            }
        }

        std::vector<unsigned char> is_port;
        if (!s->port.empty() && (s->npole || s->sibc_on)) {
            is_port.assign(s->N * 3, 0);
            for (auto &pt : s->port)
                for (size_t q : pt.edge) is_port[pt.axis * s->N + q] = 1;
        }
        const unsigned char *isp = is_port.empty() ? nullptr : is_port.data();

        // Debye edges: ca/cb with eps_inf + sum beta/2 and the ADE tables (sim.c:501-549)
        if (s->npole) {
            // dispersive edges in (c, i, j, k) order: each task scans a contiguous run
            // of (component, x plane) pairs, the runs are joined in order
            const long nplanes = 3L * (nx + 1);
            const long nt = std::max<long>(1, std::min<long>(64, nplanes));
            std::vector<std::vector<size_t>> lists(nt);
            std::vector<std::vector<int>> comps(nt);
            // cells with any nonzero pole strength: an edge whose four cells have none
            // is not dispersive (edge_debye would return 0), so it is skipped unread
            std::vector<unsigned char> dcell(s->N, 0);
            parallel_ranges((long)s->N, [&](long q0, long q1) {
                for (long q = q0; q < q1; q++)
                    for (int p = 0; p < s->npole; p++)
                        if (s->cell_deps[p][q] != 0) {
                            dcell[q] = 1;
                            break;
                        }
            });
            auto near_dispersive = [&](int c, int i, int j, int k) {
                const int ua = (c + 1) % 3, va = (c + 2) % 3;
                int ijk[3] = {i, j, k};
                for (int du = -1; du <= 0; du++)
                    for (int dv = -1; dv <= 0; dv++) {
                        int cc[3] = {ijk[0], ijk[1], ijk[2]};
                        cc[ua] += du;
                        cc[va] += dv;
                        if (cc[ua] < 0 || cc[ua] >= s->n[ua] || cc[va] < 0 || cc[va] >= s->n[va]) continue;
                        if (dcell[IDX(s, cc[0], cc[1], cc[2])]) return true;
                    }
                return false;
            };
            parallel_ranges(nt, [&](long t0, long t1) {
                double dep[kMaxPoles];
                for (long t = t0; t < t1; t++)
                    for (long r = nplanes * t / nt; r < nplanes * (t + 1) / nt; r++) {
                        const int c = (int)(r / (nx + 1)), i = (int)(r % (nx + 1));
                        for (int j = 0; j <= ny; j++)
                            for (int k = 0; k <= nz; k++) {
                                int ijk[3] = {i, j, k};
                                if (ijk[c] >= s->n[c]) continue;
                                size_t q = IDX(s, i, j, k);
                                if (s->cb[c][q] == 0 || (isp && isp[c * s->N + q])) continue;
                                if (!near_dispersive(c, i, j, k) || !edge_debye(s, c, i, j, k, dep)) continue;
                                comps[t].push_back(c);
                                lists[t].push_back(q);
                            }
                    }
            });
            std::vector<size_t> list;
            std::vector<int> comp;
            for (long t = 0; t < nt; t++) {
                list.insert(list.end(), lists[t].begin(), lists[t].end());
                comp.insert(comp.end(), comps[t].begin(), comps[t].end());
            }
            s->ndisp = (int)list.size();
            s->d_comp = comp;
            s->d_idx = list;
            s->d_cb.assign(s->ndisp, 0.0);
            s->d_beta.assign((size_t)s->ndisp * s->npole, 0.0);
            parallel_ranges(s->ndisp, [&](long e0, long e1) {
                double dep[kMaxPoles];
                for (long e = e0; e < e1; e++) {
                    int c = s->d_comp[e];
                    size_t q = s->d_idx[e];
                    int i = (int)(q / s->si), j = (int)((q % s->si) / s->sj), k = (int)(q % s->sj);
                    edge_debye(s, c, i, j, k, dep);
                    double er, sg;
                    int pec;
                    edge_material(s, c, i, j, k, &er, &sg, &pec);
                    double einst = er * CEM_EPS0;
                    for (int p = 0; p < s->npole; p++) {
                        double tau = 1.0 / (2 * M_PI * s->pole_f[p]);
                        double bt = CEM_EPS0 * dep[p] * s->dt / (tau * (1 + s->dt / (2 * tau)));
                        s->d_beta[(size_t)p * s->ndisp + e] = bt;
                        einst += 0.5 * bt;
                    }
                    double h = sg * s->dt / (2 * einst);
                    s->ca[c][q] = (1 - h) / (1 + h);
                    s->cb[c][q] = (s->dt / einst) / (1 + h);
                    s->d_cb[e] = s->cb[c][q];
                }
            });
            for (int p = 0; p < s->npole; p++) {
                double tau = 1.0 / (2 * M_PI * s->pole_f[p]);
                s->d_k[p] = (1 - s->dt / (2 * tau)) / (1 + s->dt / (2 * tau));
            }
        }
        if (sibc_finalize(s, isp)) return -1;

        s->rotE.assign(2 * s->nf, 0.0);
        for (int f = 0; f < s->nf; f++) {
            double w = 2 * M_PI * s->freq[f];
            s->rotE[2 * f] = std::cos(w * s->dt);
            s->rotE[2 * f + 1] = -std::sin(w * s->dt);
        }
        s->step = 0;
        s->finalized = 1;
        static std::atomic<long> plans{0};
        s->plan_id = ++plans;
        if (modeports_finalize(s)) {
            s->finalized = 0;
            return -1;
        }
    } catch (const std::bad_alloc &) {
        set_error("out of memory");
        return -1;
    }
    return 0;
}

// ---- excitation clone (engine extension; no counterpart in the reference) ----
// A second simulation with a's plan for another excitation of the same board: every
// set-up array is copied (no set-up step runs again, mode-port cross-sections are
// copied with their solved state), the time-dependent state starts from zero at step 0
// on its own device buffers, and the excited port moves to `port`: lumped port `port`
// with a's Gaussian, or, when a drives a mode port, mode port `port` with a's waveform.
extern "C" cem_sim *cem_amd_clone_excitation(const cem_sim *a, int port) {
    if (!a || !a->finalized) {
        set_error("simulation not finalized");
        return nullptr;
    }
    const bool modal = a->msrc >= 0;
    if (port < 0 || port >= (int)(modal ? a->mport.size() : a->port.size())) {
        set_error(modal ? "bad mode port" : "bad port");
        return nullptr;
    }
    cem_sim *b = nullptr;
    try {
        b = new cem_sim();
        for (int x = 0; x < 3; x++) {
            b->n[x] = a->n[x];
            b->node[x] = a->node[x];
            b->d[x] = a->d[x];
            b->dd[x] = a->dd[x];
            b->ca[x] = a->ca[x];
            b->cb[x] = a->cb[x];
            b->edge_pec[x] = a->edge_pec[x];
            b->ide[x] = a->ide[x];
            b->idh[x] = a->idh[x];
            b->be[x] = a->be[x];
            b->ce[x] = a->ce[x];
            b->bh[x] = a->bh[x];
            b->ch[x] = a->ch[x];
            b->pml_sn[x] = a->pml_sn[x];
            b->pml_kn[x] = a->pml_kn[x];
            b->pml_an[x] = a->pml_an[x];
            b->pml_sc[x] = a->pml_sc[x];
            b->pml_kc[x] = a->pml_kc[x];
            b->pml_ac[x] = a->pml_ac[x];
        }
        b->si = a->si;
        b->sj = a->sj;
        b->N = a->N;
        b->eps_r = a->eps_r;
        b->sigma = a->sigma;
        b->cell_pec = a->cell_pec;
        for (int f = 0; f < 6; f++) {
            b->bc[f] = a->bc[f];
            b->npml[f] = a->npml[f];
        }
        b->has_pml = a->has_pml;
        b->dt = a->dt;
        b->step = 0;
        b->finalized = 1;
        b->plan_id = a->plan_id;
        b->port = a->port;
        b->src_port = a->src_port;
        b->src_amp = a->src_amp;
        b->src_t0 = a->src_t0;
        b->src_tau = a->src_tau;
        b->nf = a->nf;
        b->freq = a->freq;
        b->rotE = a->rotE;
        b->vp = a->vp;
        b->ip = a->ip;
        b->mport = a->mport;
        for (auto &mp : b->mport) mp.xs = nullptr;
        for (size_t k = 0; k < a->mport.size(); k++)
            if (a->mport[k].xs && !(b->mport[k].xs = mode2d_clone(a->mport[k].xs))) throw std::bad_alloc();
        b->msrc = a->msrc;
        b->msrc_t0 = a->msrc_t0;
        b->msrc_tau = a->msrc_tau;
        b->npole = a->npole;
        std::copy(a->pole_f, a->pole_f + kMaxPoles, b->pole_f);
        b->cell_deps = a->cell_deps;
        b->ndisp = a->ndisp;
        b->d_comp = a->d_comp;
        b->d_idx = a->d_idx;
        b->d_beta = a->d_beta;
        b->d_cb = a->d_cb;
        std::copy(a->d_k, a->d_k + kMaxPoles, b->d_k);
        b->sibc_on = a->sibc_on;
        b->sibc_rough = a->sibc_rough;
        b->sibc_npole = a->sibc_npole;
        b->sibc_sigma = a->sibc_sigma;
        b->sibc_p1 = a->sibc_p1;
        b->sibc_p2 = a->sibc_p2;
        b->sibc_fit_err = a->sibc_fit_err;
        std::copy(a->sibc_a, a->sibc_a + kSibcTerms, b->sibc_a);
        std::copy(a->sibc_c, a->sibc_c + kSibcTerms, b->sibc_c);
        std::copy(a->sibc_e, a->sibc_e + kSibcTerms, b->sibc_e);
        b->nsb = a->nsb;
        b->sb_comp = a->sb_comp;
        b->sb_nf = a->sb_nf;
        b->sb_hc = a->sb_hc;
        b->sb_idx = a->sb_idx;
        b->sb_hi = a->sb_hi;
        b->sb_sg = a->sb_sg;
        b->sb_mask = a->sb_mask;
    } catch (const std::bad_alloc &) {
        if (b) cem_destroy(b);
        set_error("out of memory");
        return nullptr;
    }
    if (modal)
        b->msrc = port;
    else
        b->src_port = port;
    return b;
}

// ---- results (sim.c:1016-1022): host copies of the device FP64 accumulators ----
static const double *spectrum(const cem_sim *cs, cem::dvec *v) {
    cem_sim *s = const_cast<cem_sim *>(cs);
    if (!s->finalized) return nullptr;
    s->spectra_exposed = true;
    dev_spectra_download(s);
    return v->data();
}

extern "C" const double *cem_port_v(const cem_sim *s, int p) {
    if (p < 0 || p >= (int)s->port.size()) return nullptr;
    return spectrum(s, const_cast<cem::dvec *>(&s->spec_port_v[p]));
}
extern "C" const double *cem_port_i(const cem_sim *s, int p) {
    if (p < 0 || p >= (int)s->port.size()) return nullptr;
    return spectrum(s, const_cast<cem::dvec *>(&s->spec_port_i[p]));
}
extern "C" const double *cem_source_spectrum(const cem_sim *s) {
    return spectrum(s, const_cast<cem::dvec *>(&s->spec_src));
}
extern "C" const double *cem_vprobe(const cem_sim *s, int p) {
    if (p < 0 || p >= (int)s->vp.size()) return nullptr;
    return spectrum(s, const_cast<cem::dvec *>(&s->spec_vp[p]));
}
extern "C" const double *cem_iprobe(const cem_sim *s, int p) {
    if (p < 0 || p >= (int)s->ip.size()) return nullptr;
    return spectrum(s, const_cast<cem::dvec *>(&s->spec_ip[p]));
}
extern "C" double cem_port_impedance(const cem_sim *s, int p) {
    return p >= 0 && p < (int)s->port.size() ? s->port[p].R : 0;
}
extern "C" int cem_excited_port(const cem_sim *s) { return s->src_port; }

// ---- raw fields: host FP64 mirrors of the FP32 device fields ----
extern "C" double *cem_field(cem_sim *s, int comp) {
    if (comp < 0 || comp > 5) return nullptr;
    if (!s->field_mirror[comp]) {
        double *m = new (std::nothrow) double[s->N];
        if (!m) {
            set_error("out of memory");
            return nullptr;
        }
        std::memset(m, 0, s->N * sizeof(double));
        if (s->dev) dev_field_download(s, comp, m);
        s->field_mirror[comp].reset(m);
    }
    return s->field_mirror[comp].get();
}

// ---- stepping entry points (sim.c:867-1012) ----
extern "C" void cem_step(cem_sim *s, long n) {
    if (!s->finalized || n <= 0) return;
    if (dev_ensure(s)) return;
    dev_sync_in(s);
    dev_step(s, n);
    dev_sync_out(s);
}

extern "C" void cem_mt_pre(cem_sim *s) {
    if (!s->finalized || dev_ensure(s)) return;
    dev_sync_in(s);
}
extern "C" void cem_mt_H(cem_sim *s, int i0, int i1) {
    if (!s->finalized || dev_ensure(s)) return;
    dev_mt_H(s, i0, i1);
}
extern "C" void cem_mt_mid(cem_sim *s) { (void)s; }
extern "C" void cem_mt_E(cem_sim *s, int i0, int i1) {
    if (!s->finalized || dev_ensure(s)) return;
    dev_mt_E(s, i0, i1);
}
extern "C" void cem_mt_post(cem_sim *s) {
    if (!s->finalized || dev_ensure(s)) return;
    dev_mt_post(s);
    dev_sync_out(s);
}

extern "C" void cem_set_step(cem_sim *s, long step) {
    s->step = step;
    if (s->dev) dev_set_step(s, step);
}

extern "C" double cem_energy(const cem_sim *cs) {
    cem_sim *s = const_cast<cem_sim *>(cs);
    if (!s->finalized) return 0;
    if (dev_ensure(s)) return NAN;
    dev_sync_in(s);
    return dev_energy(s);
}

extern "C" long cem_run(cem_sim *s, long max_steps, double decay, int check_every) {
    if (!s->finalized) return 0;
    if (check_every < 1) check_every = 200;
    long start = s->step;
    double t_src_end = s->src_port >= 0 ? s->src_t0 + 6 * s->src_tau : 0;
    if (s->msrc >= 0 && s->msrc_t0 + 6 * s->msrc_tau > t_src_end) t_src_end = s->msrc_t0 + 6 * s->msrc_tau;
    double peak = 0;
    while (s->step - start < max_steps) {
        long chunk = check_every;
        if (s->step - start + chunk > max_steps) chunk = max_steps - (s->step - start);
        cem_step(s, chunk);
        double W = cem_energy(s);
        if (W > peak) peak = W;
        if (s->step * s->dt > t_src_end && peak > 0 && W < decay * peak) break;
    }
    return s->step - start;
}

// ---- plan export (export.c:16-178) ----
extern "C" const double *cem_plan_array(const cem_sim *s, int which, int comp) {
    if (!s->finalized || comp < 0 || comp > 2) return nullptr;
    const cem::dvec *v = nullptr;
    switch (which) {
    case 0: v = &s->ca[comp]; break;
    case 1: v = &s->cb[comp]; break;
    case 2: v = &s->ide[comp]; break;
    case 3: v = &s->idh[comp]; break;
    case 4: v = &s->be[comp]; break;
    case 5: v = &s->ce[comp]; break;
    case 6: v = &s->bh[comp]; break;
    case 7: v = &s->ch[comp]; break;
    case 8: v = &s->d[comp]; break;
    case 9: v = &s->dd[comp]; break;
    default: return nullptr;
    }
    return v->data();
}

extern "C" int cem_plan_npml(const cem_sim *s, int face) { return face >= 0 && face < 6 ? s->npml[face] : 0; }
extern "C" int cem_plan_has_pml(const cem_sim *s) { return s->has_pml; }
extern "C" int cem_plan_nport(const cem_sim *s) { return (int)s->port.size(); }
extern "C" int cem_plan_nvprobe(const cem_sim *s) { return (int)s->vp.size(); }
extern "C" int cem_plan_niprobe(const cem_sim *s) { return (int)s->ip.size(); }
extern "C" int cem_plan_ndisp(const cem_sim *s) { return s->ndisp; }
extern "C" int cem_plan_nsibc(const cem_sim *s) { return s->nsb; }

extern "C" int cem_plan_port(const cem_sim *s, int p, double *idx, double *vw, double *cs) {
    if (p < 0 || p >= (int)s->port.size()) return -1;
    const LumpedPort &pt = s->port[p];
    for (size_t e = 0; e < pt.edge.size(); e++) {
        if (idx) idx[e] = (double)(pt.axis * s->N + pt.edge[e]);
        if (vw) vw[e] = pt.dl[e] * pt.dir;
        if (cs) cs[e] = pt.cs[e];
    }
    return (int)pt.edge.size();
}

extern "C" int cem_plan_vprobe(const cem_sim *s, int p, double *idx, double *w) {
    if (p < 0 || p >= (int)s->vp.size()) return -1;
    const VoltageProbe &v = s->vp[p];
    for (size_t e = 0; e < v.edge.size(); e++) {
        if (idx) idx[e] = (double)(v.axis * s->N + v.edge[e]);
        if (w) w[e] = v.dl[e];
    }
    return (int)v.edge.size();
}

extern "C" int cem_plan_iprobe(const cem_sim *s, int p, double *idx, double *w) {
    if (p < 0 || p >= (int)s->ip.size()) return -1;
    const CurrentProbe &c = s->ip[p];
    for (size_t e = 0; e < c.idx.size(); e++) {
        if (idx) idx[e] = (double)((c.comp[e] - 3) * s->N + c.idx[e]);
        if (w) w[e] = c.w[e];
    }
    return (int)c.idx.size();
}

extern "C" int cem_plan_mode_source(const cem_sim *s, double *idx, double *w) {
    if (s->msrc < 0) return 0;
    const ModePort &p = s->mport[s->msrc];
    for (size_t i = 0; i < p.src_idx.size(); i++) {
        if (idx) idx[i] = (double)(p.src_comp[i] * s->N + p.src_idx[i]);
        if (w) w[i] = p.src_w[i];
    }
    return (int)p.src_idx.size();
}

double cem::lumped_waveform(const cem_sim *s, double t) {
    if (s->src_port < 0) return 0.0;
    double x = (t - s->src_t0) / s->src_tau;
    return s->src_amp * std::exp(-x * x);
}

extern "C" void cem_plan_waveforms(const cem_sim *s, long n0, int n, double *lumped, double *mode) {
    for (int q = 0; q < n; q++) {
        double t = (n0 + q + 0.5) * s->dt;
        if (lumped) lumped[q] = lumped_waveform(s, t);
        if (mode) mode[q] = modeport_waveform(s, t);
    }
}

extern "C" int cem_plan_excited_lumped(const cem_sim *s) { return s->src_port; }
extern "C" int cem_plan_npole(const cem_sim *s) { return s->npole; }

extern "C" int cem_plan_disp(const cem_sim *s, double *idx, double *cb, double *beta, double *k) {
    for (int e = 0; e < s->ndisp; e++) {
        if (idx) idx[e] = (double)(s->d_comp[e] * s->N + s->d_idx[e]);
        if (cb) cb[e] = s->d_cb[e];
    }
    if (beta && s->ndisp) std::memcpy(beta, s->d_beta.data(), (size_t)s->npole * s->ndisp * sizeof(double));
    if (k) std::memcpy(k, s->d_k, s->npole * sizeof(double));
    return s->ndisp;
}

extern "C" int cem_plan_sibc(const cem_sim *s, double *idx, double *hidx, double *hw, double *c, double *e,
                             int *nterm) {
    for (int q = 0; q < s->nsb; q++) {
        if (idx) idx[q] = (double)(s->sb_comp[q] * s->N + s->sb_idx[q]);
        for (int f = 0; f < 4; f++) {
            int has = f < s->sb_nf[q];
            if (hidx) hidx[4 * q + f] = has ? (double)(s->sb_hc[4 * q + f] * s->N + s->sb_hi[4 * q + f]) : 0;
            if (hw) hw[4 * q + f] = has ? s->sb_sg[4 * q + f] : 0;
        }
    }
    if (c) std::memcpy(c, s->sibc_c, s->sibc_npole * sizeof(double));
    if (e) std::memcpy(e, s->sibc_e, s->sibc_npole * sizeof(double));
    if (nterm) *nterm = s->sibc_npole;
    return s->nsb;
}

extern "C" void cem_push_samples(cem_sim *s, int n, const double *ip, const double *vold, const double *vnew,
                                 const double *vp) {
    if (!s->finalized || n <= 0 || dev_ensure(s)) return;
    dev_push_samples(s, n, ip, vold, vnew, vp);
}

// ---- checkpoints (export.c:185-242) ----
static const double kCkptMagic = 20261017.0;

static size_t ckpt_body(const cem_sim *s) {
    size_t o = 6 * s->N;
    if (s->has_pml)
        for (int c = 0; c < 3; c++)
            for (int w = 0; w < 2; w++) o += psi_len(s, 1, c, w) + psi_len(s, 0, c, w);
    if (s->ndisp) o += (size_t)s->npole * s->ndisp + s->ndisp;
    if (s->nsb) o += (size_t)s->nsb * s->sibc_npole;
    o += (size_t)(2 * s->port.size() + s->vp.size() + s->ip.size() + 1 + 2) * 2 * s->nf;
    return o;
}

static double ckpt_signature(const cem_sim *s) {
    return (double)(s->port.size() + 100 * s->vp.size() + 10000 * s->ip.size());
}

extern "C" size_t cem_checkpoint_size(const cem_sim *s) {
    if (!s->finalized) return 0;
    return 8 + ckpt_body(s);
}

// Walks the checkpoint body in the reference's block order; `save` selects the direction.
static void ckpt_walk(cem_sim *s, double *buf, int save) {
    size_t o = 0;
    for (int c = 0; c < 6; c++) {
        if (save)
            dev_field_download(s, c, buf + o);
        else
            dev_state_upload(s, 5, c, 0, buf + o);
        o += s->N;
    }
    if (s->has_pml)
        for (int c = 0; c < 3; c++)
            for (int w = 0; w < 2; w++)
                for (int which = 0; which < 2; which++) {   // psiE then psiH (export.c:201)
                    size_t len = psi_len(s, which == 0, c, w);
                    if (save)
                        dev_state_download(s, which, c, w, buf + o);
                    else
                        dev_state_upload(s, which, c, w, buf + o);
                    o += len;
                }
    if (s->ndisp) {
        for (int which = 2; which <= 3; which++) {
            size_t len = which == 2 ? (size_t)s->npole * s->ndisp : (size_t)s->ndisp;
            if (save)
                dev_state_download(s, which, 0, 0, buf + o);
            else
                dev_state_upload(s, which, 0, 0, buf + o);
            o += len;
        }
    }
    if (s->nsb) {
        if (save)
            dev_state_download(s, 4, 0, 0, buf + o);
        else
            dev_state_upload(s, 4, 0, 0, buf + o);
        o += (size_t)s->nsb * s->sibc_npole;
    }
    dev_acc_io(s, buf + o, save);   // accumulators then phasors
}

extern "C" int cem_checkpoint_save(const cem_sim *cs, double *buf) {
    cem_sim *s = const_cast<cem_sim *>(cs);
    if (!s->finalized) {
        set_error("not finalized");
        return -1;
    }
    if (dev_ensure(s)) return -1;
    dev_sync_in(s);
    double hdr[8] = {kCkptMagic, 1, (double)s->N, (double)s->nf, (double)s->step, (double)ckpt_body(s),
                     ckpt_signature(s), (double)(s->ndisp + s->nsb)};
    std::memcpy(buf, hdr, sizeof hdr);
    ckpt_walk(s, buf + 8, 1);
    return 0;
}

extern "C" int cem_checkpoint_load(cem_sim *s, const double *buf) {
    if (!s->finalized) {
        set_error("not finalized");
        return -1;
    }
    if (buf[0] != kCkptMagic || buf[1] != 1 || buf[2] != (double)s->N || buf[3] != (double)s->nf ||
        buf[5] != (double)ckpt_body(s) || buf[6] != ckpt_signature(s) || buf[7] != (double)(s->ndisp + s->nsb)) {
        set_error("checkpoint does not match this simulation");
        return -1;
    }
    if (dev_ensure(s)) return -1;
    ckpt_walk(s, const_cast<double *>(buf + 8), 0);
    s->step = (long)buf[4];
    dev_set_step(s, s->step);
    dev_sync_out(s);
    return 0;
}

// ---- state arrays (export.c:247-268) ----
extern "C" size_t cem_state_len(const cem_sim *s, int which, int comp, int w) {
    switch (which) {
    case 0: return comp >= 0 && comp < 3 && w >= 0 && w < 2 ? psi_len(s, 1, comp, w) : 0;
    case 1: return comp >= 0 && comp < 3 && w >= 0 && w < 2 ? psi_len(s, 0, comp, w) : 0;
    case 2: return s->finalized && s->npole && s->ndisp ? (size_t)s->npole * s->ndisp : 0;
    case 3: return s->finalized && s->npole && s->ndisp ? (size_t)s->ndisp : 0;
    case 4: return s->finalized && s->sibc_on ? (size_t)s->nsb * s->sibc_npole : 0;
    default: return 0;
    }
}

extern "C" double *cem_state_array(cem_sim *s, int which, int comp, int w) {
    if (!s->finalized) return nullptr;
    cem::dvec *v = nullptr;
    switch (which) {
    case 0:
        if (!s->has_pml || comp < 0 || comp > 2 || w < 0 || w > 1) return nullptr;
        v = &s->psiE_mirror[comp][w];
        break;
    case 1:
        if (!s->has_pml || comp < 0 || comp > 2 || w < 0 || w > 1) return nullptr;
        v = &s->psiH_mirror[comp][w];
        break;
    case 2:
        if (!s->ndisp) return nullptr;
        v = &s->dJ_mirror;
        break;
    case 3:
        if (!s->ndisp) return nullptr;
        v = &s->dEold_mirror;
        break;
    case 4:
        if (!s->sibc_on) return nullptr;
        v = &s->sby_mirror;
        break;
    default: return nullptr;
    }
    size_t len = cem_state_len(s, which, comp, w);
    if (v->size() != len + 1) {
        v->assign(len + 1, 0.0);
        if (s->dev) dev_state_download(s, which, comp, w, v->data());
    }
    s->state_exposed = true;
    return v->data();
}
