// Complex 2D eigenmode solver on a Yee cross-section (port profiles, Z0, beta).
//
// Formulation of solver/src/mode2d.c: propagation along +x, transverse (y, z)
// unknowns E_u = Ey (j+1/2, k), E_v = Ez (j, k+1/2), E_w = Ex (j, k);
// beta e = P h, beta h = Q e, so beta^2 e = P Q e, with stretched-coordinate PML
// and lossy / Debye media.  A = P Q is applied matrix-free, assembled in band
// storage by coloured probing, factorised by a banded LU with partial pivoting
// and the mode nearest (k0 n_guess)^2 found by inverse iteration.
//
// The band update runs column-outer / row-inner (contiguous in the column-major
// band), which performs exactly the same floating-point operations per entry as
// the reference's row-outer loop, so results are unchanged.
#include "sim.hpp"
#include "mode_gpu.hpp"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

using namespace cem;

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

struct cem_mode2d {
    int ny = 0, nz = 0;
    dvec y, z, dy, dz, ddy, ddz;
    dvec er_u, er_v, er_w, sg_u, sg_v, sg_w;
    std::vector<unsigned char> pec_u, pec_v, pec_w;
    int npole = 0;
    double pole_f[16] = {};
    dvec dp_u, dp_v, dp_w;
    dvec ps[4], pk[4], pa[4];   // PML sigma/kappa/alpha: y cells, y nodes, z cells, z nodes
    double freq = 0;
    cplx beta = 0;
    std::vector<cplx> eu, ev, hu, hv;
    cem::ModeGpu *gpu = nullptr;   // device context of the GPU solve path, kept across frequencies
};

namespace {

struct Coef {
    std::vector<cplx> iyc, iyn, izc, izn;   // 1/(s d) at y cells, y nodes, z cells, z nodes
    std::vector<cplx> eu, ev, iw;           // absolute complex permittivity; 1/eps at E_w (0 on PEC)
};

inline int IU(int j, int k, int nz) { return j * (nz + 1) + k; }
inline int IV(int j, int k, int nz) { return j * nz + k; }
inline int IN(int j, int k, int nz) { return j * (nz + 1) + k; }
inline int IC(int j, int k, int nz) { return j * nz + k; }

void make_coef(const cem_mode2d *m, double w, Coef &c) {
    const int ny = m->ny, nz = m->nz;
    const int nu = ny * (nz + 1), nv = (ny + 1) * nz, nn = (ny + 1) * (nz + 1);
    const cplx jw(0.0, w * CEM_EPS0);
    auto stretch = [&](int a, int i) { return m->pk[a][i] + m->ps[a][i] / (m->pa[a][i] + jw); };
    c.iyc.resize(ny);
    c.iyn.resize(ny + 1);
    c.izc.resize(nz);
    c.izn.resize(nz + 1);
    for (int j = 0; j < ny; j++) c.iyc[j] = 1.0 / (stretch(0, j) * m->dy[j]);
    for (int j = 0; j <= ny; j++) c.iyn[j] = 1.0 / (stretch(1, j) * m->ddy[j]);
    for (int k = 0; k < nz; k++) c.izc[k] = 1.0 / (stretch(2, k) * m->dz[k]);
    for (int k = 0; k <= nz; k++) c.izn[k] = 1.0 / (stretch(3, k) * m->ddz[k]);
    cplx dfac[16];
    for (int p = 0; p < m->npole; p++) dfac[p] = 1.0 / cplx(1.0, w / (2 * M_PI * m->pole_f[p]));
    auto eps = [&](double er, const dvec &dp, int stride, int q, double sg) {
        cplx e = er;
        for (int p = 0; p < m->npole; p++) e += dp[(size_t)p * stride + q] * dfac[p];
        cplx ee = e * CEM_EPS0;
        return cplx(ee.real() - 0.0, ee.imag() - sg / w);
    };
    c.eu.resize(nu);
    c.ev.resize(nv);
    c.iw.resize(nn);
    for (int q = 0; q < nu; q++) c.eu[q] = eps(m->er_u[q], m->dp_u, nu, q, m->sg_u[q]);
    for (int q = 0; q < nv; q++) c.ev[q] = eps(m->er_v[q], m->dp_v, nv, q, m->sg_v[q]);
    for (int q = 0; q < nn; q++)
        c.iw[q] = m->pec_w[q] ? cplx(0.0) : 1.0 / eps(m->er_w[q], m->dp_w, nn, q, m->sg_w[q]);
}

// h = Q e (hw: cell work array)
void apply_Q(const cem_mode2d *m, const Coef &c, double w, const cplx *eu, const cplx *ev, cplx *hu, cplx *hv,
             cplx *hw) {
    const int ny = m->ny, nz = m->nz;
    for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz; k++)
            hw[IC(j, k, nz)] = (ev[IV(j + 1, k, nz)] - ev[IV(j, k, nz)]) * c.iyc[j] -
                               (eu[IU(j, k + 1, nz)] - eu[IU(j, k, nz)]) * c.izc[k];
    const double cc = 1.0 / (w * CEM_MU0);
    const cplx zero(0.0);
    for (int j = 0; j <= ny; j++)
        for (int k = 0; k < nz; k++) {
            cplx r = (j < ny ? hw[IC(j, k, nz)] : zero) - (j > 0 ? hw[IC(j - 1, k, nz)] : zero);
            hu[IV(j, k, nz)] = -w * c.ev[IV(j, k, nz)] * ev[IV(j, k, nz)] - cc * r * c.iyn[j];
        }
    for (int j = 0; j < ny; j++)
        for (int k = 0; k <= nz; k++) {
            cplx r = (k < nz ? hw[IC(j, k, nz)] : zero) - (k > 0 ? hw[IC(j, k - 1, nz)] : zero);
            hv[IU(j, k, nz)] = w * c.eu[IU(j, k, nz)] * eu[IU(j, k, nz)] - cc * r * c.izn[k];
        }
}

// e = P h (ew: node work array)
void apply_P(const cem_mode2d *m, const Coef &c, double w, const cplx *hu, const cplx *hv, cplx *eu, cplx *ev,
             cplx *ew) {
    const int ny = m->ny, nz = m->nz;
    const cplx zero(0.0);
    for (int j = 0; j <= ny; j++)
        for (int k = 0; k <= nz; k++) {
            cplx du = ((j < ny ? hv[IU(j, k, nz)] : zero) - (j > 0 ? hv[IU(j - 1, k, nz)] : zero)) * c.iyn[j];
            cplx dv = ((k < nz ? hu[IV(j, k, nz)] : zero) - (k > 0 ? hu[IV(j, k - 1, nz)] : zero)) * c.izn[k];
            ew[IN(j, k, nz)] = (du - dv) * c.iw[IN(j, k, nz)];
        }
    const double cc = 1.0 / w;
    for (int j = 0; j < ny; j++)
        for (int k = 0; k <= nz; k++)
            eu[IU(j, k, nz)] = w * CEM_MU0 * hv[IU(j, k, nz)] +
                               cc * (ew[IN(j + 1, k, nz)] - ew[IN(j, k, nz)]) * c.iyc[j];
    for (int j = 0; j <= ny; j++)
        for (int k = 0; k < nz; k++)
            ev[IV(j, k, nz)] = -w * CEM_MU0 * hu[IV(j, k, nz)] +
                               cc * (ew[IN(j, k + 1, nz)] - ew[IN(j, k, nz)]) * c.izc[k];
}

// unknown ordering: j outer, k inner, [E_u, E_v] interleaved (bandwidth ~ 2 (nz + 2))
struct Order {
    std::vector<int> kind, idx;
    int n = 0;
};

void make_order(const cem_mode2d *m, Order &o) {
    const int ny = m->ny, nz = m->nz;
    for (int j = 0; j <= ny; j++)
        for (int k = 0; k <= nz; k++) {
            if (j < ny) { o.kind.push_back(0); o.idx.push_back(IU(j, k, nz)); }
            if (k < nz) { o.kind.push_back(1); o.idx.push_back(IV(j, k, nz)); }
        }
    o.n = (int)o.kind.size();
}

inline bool order_pec(const cem_mode2d *m, const Order &o, int g) {
    return o.kind[g] == 0 ? m->pec_u[o.idx[g]] : m->pec_v[o.idx[g]];
}

// complex band matrix in LAPACK gbtrf layout (kl sub-, ku super-diagonals + kl fill)
struct Band {
    int n, kl, ku, ldab;
    std::vector<cplx> ab;
    std::vector<int> ipiv;
    cplx &at(int i, int j) { return ab[(size_t)j * ldab + (kl + ku + i - j)]; }
};

int band_lu(Band &b) {
    const int n = b.n, kl = b.kl, kv = b.ku + b.kl;
    std::vector<cplx> l(kl + 1);
    for (int j = 0; j < n; j++) {
        int last = j + kl < n - 1 ? j + kl : n - 1;
        int p = j;
        double best = std::abs(b.at(j, j));
        for (int i = j + 1; i <= last; i++) {
            double v = std::abs(b.at(i, j));
            if (v > best) { best = v; p = i; }
        }
        b.ipiv[j] = p;
        if (best == 0) return -1;
        int jmax = j + kv < n - 1 ? j + kv : n - 1;
        if (p != j)
            for (int c = j; c <= jmax; c++) std::swap(b.at(j, c), b.at(p, c));
        const cplx piv = b.at(j, j);
        int nl = 0;
        for (int i = j + 1; i <= last; i++) {
            cplx &lij = b.at(i, j);
            lij /= piv;
            l[i - j - 1] = lij;
            nl = i - j;
        }
        for (int c = j + 1; c <= jmax; c++) {
            const cplx x = b.at(j, c);
            const double xr = x.real(), xi = x.imag();
            cplx *col = &b.at(j + 1, c);   // rows j+1.. of column c are contiguous
            for (int r = 0; r < nl; r++) {
                const double lr = l[r].real(), li = l[r].imag();
                if (lr == 0 && li == 0) continue;
                double pr = lr * xr - li * xi, pi = lr * xi + li * xr;
                col[r] = cplx(col[r].real() - pr, col[r].imag() - pi);
            }
        }
    }
    return 0;
}

void band_solve(Band &b, cplx *x) {
    const int n = b.n, kl = b.kl, kv = b.ku + b.kl;
    for (int j = 0; j < n; j++) {
        int p = b.ipiv[j];
        if (p != j) std::swap(x[j], x[p]);
        int last = j + kl < n - 1 ? j + kl : n - 1;
        for (int i = j + 1; i <= last; i++) x[i] -= b.at(i, j) * x[j];
    }
    for (int j = n - 1; j >= 0; j--) {
        x[j] /= b.at(j, j);
        int first = j - kv > 0 ? j - kv : 0;
        for (int i = first; i < j; i++) x[i] -= b.at(i, j) * x[j];
    }
}

struct Work {
    std::vector<cplx> eu, ev, hu, hv, wk;
};

void apply_A(const cem_mode2d *m, const Coef &c, const Order &o, double w, const cplx *x, cplx *y, Work &W) {
    std::fill(W.eu.begin(), W.eu.end(), cplx(0.0));
    std::fill(W.ev.begin(), W.ev.end(), cplx(0.0));
    for (int g = 0; g < o.n; g++) {
        if (order_pec(m, o, g)) continue;
        if (o.kind[g] == 0) W.eu[o.idx[g]] = x[g];
        else W.ev[o.idx[g]] = x[g];
    }
    apply_Q(m, c, w, W.eu.data(), W.ev.data(), W.hu.data(), W.hv.data(), W.wk.data());
    apply_P(m, c, w, W.hu.data(), W.hv.data(), W.eu.data(), W.ev.data(), W.wk.data());
    for (int g = 0; g < o.n; g++)
        y[g] = order_pec(m, o, g) ? cplx(0.0) : (o.kind[g] == 0 ? W.eu[o.idx[g]] : W.ev[o.idx[g]]);
}

}  // namespace

// ---- API (mode2d.c:57-136, 307-498) ----
extern "C" cem_mode2d *cem_mode2d_create(int ny, int nz, const double *y, const double *z) {
    cem_mode2d *m = new cem_mode2d();
    m->ny = ny;
    m->nz = nz;
    m->y.assign(y, y + ny + 1);
    m->z.assign(z, z + nz + 1);
    m->dy.assign(ny, 0.0);
    m->dz.assign(nz, 0.0);
    m->ddy.assign(ny + 1, 0.0);
    m->ddz.assign(nz + 1, 0.0);
    for (int j = 0; j < ny; j++) {
        m->dy[j] = y[j + 1] - y[j];
        m->ddy[j] += m->dy[j] / 2;
        m->ddy[j + 1] += m->dy[j] / 2;
    }
    for (int k = 0; k < nz; k++) {
        m->dz[k] = z[k + 1] - z[k];
        m->ddz[k] += m->dz[k] / 2;
        m->ddz[k + 1] += m->dz[k] / 2;
    }
    const int nu = ny * (nz + 1), nv = (ny + 1) * nz, nn = (ny + 1) * (nz + 1);
    m->er_u.assign(nu, 1.0);
    m->er_v.assign(nv, 1.0);
    m->er_w.assign(nn, 1.0);
    m->sg_u.assign(nu, 0.0);
    m->sg_v.assign(nv, 0.0);
    m->sg_w.assign(nn, 0.0);
    m->pec_u.assign(nu, 0);
    m->pec_v.assign(nv, 0);
    m->pec_w.assign(nn, 0);
    const int len[4] = {ny, ny + 1, nz, nz + 1};
    for (int a = 0; a < 4; a++) {
        m->ps[a].assign(len[a], 0.0);
        m->pk[a].assign(len[a], 1.0);
        m->pa[a].assign(len[a], 0.0);
    }
    // the outer boundary is PEC
    for (int j = 0; j < ny; j++) m->pec_u[IU(j, 0, nz)] = m->pec_u[IU(j, nz, nz)] = 1;
    for (int k = 0; k < nz; k++) m->pec_v[IV(0, k, nz)] = m->pec_v[IV(ny, k, nz)] = 1;
    for (int j = 0; j <= ny; j++) m->pec_w[IN(j, 0, nz)] = m->pec_w[IN(j, nz, nz)] = 1;
    for (int k = 0; k <= nz; k++) m->pec_w[IN(0, k, nz)] = m->pec_w[IN(ny, k, nz)] = 1;
    m->eu.assign(nu, 0.0);
    m->ev.assign(nv, 0.0);
    m->hu.assign(nv, 0.0);
    m->hv.assign(nu, 0.0);
    return m;
}

extern "C" void cem_mode2d_destroy(cem_mode2d *m) {
    if (!m) return;
    cem::mode_gpu_release(m->gpu);
    delete m;
}

extern "C" void cem_mode2d_set_debye(cem_mode2d *m, int npole, const double *f_p, const double *deps_u,
                                     const double *deps_v, const double *deps_w) {
    const int ny = m->ny, nz = m->nz;
    const size_t nu = (size_t)ny * (nz + 1), nv = (size_t)(ny + 1) * nz, nn = (size_t)(ny + 1) * (nz + 1);
    if (npole > 16) npole = 16;
    m->npole = npole;
    std::memcpy(m->pole_f, f_p, npole * sizeof(double));
    m->dp_u.assign(deps_u, deps_u + npole * nu);
    m->dp_v.assign(deps_v, deps_v + npole * nv);
    m->dp_w.assign(deps_w, deps_w + npole * nn);
}

extern "C" void cem_mode2d_set_media(cem_mode2d *m, const double *epsr_u, const double *epsr_v, const double *epsr_w,
                                     const unsigned char *pec_u, const unsigned char *pec_v,
                                     const unsigned char *pec_w) {
    const int ny = m->ny, nz = m->nz;
    const int nu = ny * (nz + 1), nv = (ny + 1) * nz, nn = (ny + 1) * (nz + 1);
    for (int q = 0; q < nu; q++) { m->er_u[q] = epsr_u[q]; m->pec_u[q] |= pec_u[q]; }
    for (int q = 0; q < nv; q++) { m->er_v[q] = epsr_v[q]; m->pec_v[q] |= pec_v[q]; }
    for (int q = 0; q < nn; q++) { m->er_w[q] = epsr_w[q]; m->pec_w[q] |= pec_w[q]; }
}

extern "C" void cem_mode2d_set_loss(cem_mode2d *m, const double *sig_u, const double *sig_v, const double *sig_w) {
    std::copy(sig_u, sig_u + m->sg_u.size(), m->sg_u.begin());
    std::copy(sig_v, sig_v + m->sg_v.size(), m->sg_v.begin());
    std::copy(sig_w, sig_w + m->sg_w.size(), m->sg_w.begin());
}

extern "C" void cem_mode2d_set_pml(cem_mode2d *m, int which, const double *sigma, const double *kappa,
                                   const double *alpha) {
    if (which < 0 || which > 3) return;
    size_t len = m->ps[which].size();
    std::copy(sigma, sigma + len, m->ps[which].begin());
    std::copy(kappa, kappa + len, m->pk[which].begin());
    std::copy(alpha, alpha + len, m->pa[which].begin());
}

// The shared tail of a solve (mode2d.c:376-399): normalise, fix the global phase so the
// largest component is real, scatter into E_u / E_v, beta = sqrt(lambda) with a
// positive real part, H = Q E / beta.
static void finish_solve(cem_mode2d *m, const Coef &c, const Order &o, double w, double f, const std::vector<cplx> &x,
                         cplx lam, Work &W) {
    const int n = o.n;
    // normalise; global phase so the largest component is real
    double nx2 = 0, amax = 0;
    cplx ph = 1.0;
    for (int g = 0; g < n; g++) {
        nx2 += (x[g] * std::conj(x[g])).real();
        if (std::abs(x[g]) > amax) {
            amax = std::abs(x[g]);
            ph = std::conj(x[g]) / amax;
        }
    }
    double inv = 1 / std::sqrt(nx2);
    std::fill(m->eu.begin(), m->eu.end(), cplx(0.0));
    std::fill(m->ev.begin(), m->ev.end(), cplx(0.0));
    for (int g = 0; g < n; g++) {
        if (order_pec(m, o, g)) continue;
        if (o.kind[g] == 0) m->eu[o.idx[g]] = x[g] * inv * ph;
        else m->ev[o.idx[g]] = x[g] * inv * ph;
    }
    cplx bt = std::sqrt(lam);
    if (bt.real() < 0) bt = -bt;
    m->beta = bt;
    m->freq = f;
    apply_Q(m, c, w, m->eu.data(), m->ev.data(), m->hu.data(), m->hv.data(), W.wk.data());
    for (auto &v : m->hu) v /= bt;
    for (auto &v : m->hv) v /= bt;
}

extern "C" int cem_mode2d_solve(cem_mode2d *m, double f, double neff_guess, const double *start, int maxit,
                                double tol) {
    const int ny = m->ny, nz = m->nz;
    const double w = 2 * M_PI * f, k0 = w / CEM_C0;
    const cplx sigma = k0 * k0 * neff_guess * neff_guess;
    Coef c;
    make_coef(m, w, c);
    Order o;
    make_order(m, o);
    const int n = o.n, bw = 2 * (nz + 2) + 2;
    std::vector<cplx> x;
    Work W;
    try {
        x.assign(n, cplx(0.0));
        W.eu.assign((size_t)ny * (nz + 1), cplx(0.0));
        W.ev.assign((size_t)(ny + 1) * nz, cplx(0.0));
        W.hu.assign((size_t)(ny + 1) * nz, cplx(0.0));
        W.hv.assign((size_t)ny * (nz + 1), cplx(0.0));
        W.wk.assign((size_t)(ny + 1) * (nz + 1), cplx(0.0));
    } catch (const std::bad_alloc &) {
        set_error("out of memory");
        return -1;
    }
    if (start) {
        const cplx *st = reinterpret_cast<const cplx *>(start);
        for (int g = 0; g < n; g++) x[g] = o.kind[g] == 0 ? st[o.idx[g]] : st[ny * (nz + 1) + o.idx[g]];
    } else {
        unsigned sd = 12345u;   // fixed LCG start vector (mode2d.c:345-346)
        for (int g = 0; g < n; g++) {
            sd = sd * 1103515245u + 12345u;
            x[g] = ((sd >> 8) & 0xffff) / 65536.0 - 0.5;
        }
    }
    // PEC unknowns are pushed far from the shift (then A - sigma I)
    const double big = -1e6 * (std::abs(sigma) + 1.0);
    cplx lam = sigma;
    int rc = 1;
    if (cem::mode_gpu_enabled()) {
        // assembly, band LU and the inverse iteration on the GPU (FP64)
        cem::ModeGpuProblem P;
        P.ny = ny;
        P.nz = nz;
        P.w = w;
        P.iyc = c.iyc.data();
        P.iyn = c.iyn.data();
        P.izc = c.izc.data();
        P.izn = c.izn.data();
        P.eu = c.eu.data();
        P.ev = c.ev.data();
        P.iw = c.iw.data();
        P.pec_u = m->pec_u.data();
        P.pec_v = m->pec_v.data();
        P.sigma = sigma;
        P.big = big;
        P.x0 = x.data();
        P.maxit = maxit;
        P.tol = tol;
        std::vector<cplx> xg(n);
        rc = cem::mode_gpu_solve(&m->gpu, P, xg.data(), &lam);
        if (rc < 0) return -1;
        x.swap(xg);
    } else {
        Band B{n, bw, bw, 3 * bw + 1, {}, {}};
        std::vector<cplx> y;
        try {
            B.ab.assign((size_t)B.ldab * n, cplx(0.0));
            B.ipiv.assign(n, 0);
            y.assign(n, cplx(0.0));
        } catch (const std::bad_alloc &) {
            set_error("out of memory");
            return -1;
        }
        // assemble A into band storage by probing with coloured unit vectors
        const int Wc = 2 * bw + 1;
        std::vector<cplx> e(n, cplx(0.0));
        for (int col = 0; col < Wc; col++) {
            std::fill(e.begin(), e.end(), cplx(0.0));
            for (int g = col; g < n; g += Wc) e[g] = 1.0;
            apply_A(m, c, o, w, e.data(), y.data(), W);
            for (int g = col; g < n; g += Wc) {
                int lo = g - bw < 0 ? 0 : g - bw, hi = g + bw > n - 1 ? n - 1 : g + bw;
                for (int r = lo; r <= hi; r++) B.at(r, g) = y[r];
            }
        }
        for (int g = 0; g < n; g++) {
            if (order_pec(m, o, g)) B.at(g, g) = big;
            B.at(g, g) -= sigma;
        }
        if (band_lu(B)) {
            set_error("mode solver: singular pivot in the band LU");
            return -1;
        }
        cplx lam_old = 0.0;
        std::vector<cplx> xo(n, cplx(0.0));
        for (int it = 0; it < maxit; it++) {
            double nx2 = 0;
            for (int g = 0; g < n; g++) nx2 += (x[g] * std::conj(x[g])).real();
            double inv = 1 / std::sqrt(nx2);
            for (int g = 0; g < n; g++) x[g] *= inv;
            // convergence of the phase-aligned vector as well as of lambda
            cplx ph = 0.0;
            for (int g = 0; g < n; g++) ph += std::conj(xo[g]) * x[g];
            double dv = 0;
            if (std::abs(ph) > 0) {
                ph = std::conj(ph) / std::abs(ph);
                for (int g = 0; g < n; g++) {
                    cplx d = x[g] * ph - xo[g];
                    dv += (d * std::conj(d)).real();
                }
            } else {
                dv = 1;
            }
            xo = x;
            y = x;
            band_solve(B, y.data());
            cplx mu = 0.0;
            for (int g = 0; g < n; g++) mu += std::conj(x[g]) * y[g];
            lam = sigma + 1.0 / mu;
            x = y;
            if (it > 2 && std::abs(lam - lam_old) < tol * std::abs(lam) && std::sqrt(dv) < 1e-10) {
                rc = 0;
                break;
            }
            lam_old = lam;
        }
    }
    finish_solve(m, c, o, w, f, x, lam, W);
    return rc;
}

// A batch of solves of one cross-section geometry on the GPU in one launch sequence:
// ms[b] (distinct objects of the same geometry, e.g. clones) receives the mode at f[b],
// with shift shift[b] (complex) and start vector start[b] (nullptr: the fixed LCG
// vector).  The
// device context of ms[0] runs the batch.  rcs[b]: 0 converged, 1 iteration limit;
// its[b] (if not null): iterations run.
// Returns 0, or -1 with the error set.  Each result is the one cem_mode2d_solve gives
// for the same arguments, up to the rounding of the device's reductions.
int cem::mode2d_solve_batch(cem_mode2d *const *ms, int nb, const double *f, const std::complex<double> *shift,
                            const double *const *start, int maxit, double tol, int *rcs, int *its) {
    if (nb < 1) return 0;
    cem_mode2d *m0 = ms[0];
    const int ny = m0->ny, nz = m0->nz;
    Order o;
    make_order(m0, o);
    const int n = o.n;
    std::vector<Coef> c(nb);
    std::vector<std::vector<cplx>> x(nb);
    std::vector<cem::ModeGpuProblem> P(nb);
    std::vector<cplx> lam(nb);
    std::vector<cplx *> xo(nb);
    try {
        for (int b = 0; b < nb; b++) {
            const double w = 2 * M_PI * f[b];
            const cplx sigma = shift[b];
            make_coef(ms[b], w, c[b]);
            x[b].assign(n, cplx(0.0));
            if (start[b]) {
                const cplx *st = reinterpret_cast<const cplx *>(start[b]);
                for (int g = 0; g < n; g++) x[b][g] = o.kind[g] == 0 ? st[o.idx[g]] : st[ny * (nz + 1) + o.idx[g]];
            } else {
                unsigned sd = 12345u;   // fixed LCG start vector (mode2d.c:345-346)
                for (int g = 0; g < n; g++) {
                    sd = sd * 1103515245u + 12345u;
                    x[b][g] = ((sd >> 8) & 0xffff) / 65536.0 - 0.5;
                }
            }
            cem::ModeGpuProblem &q = P[b];
            q.ny = ny;
            q.nz = nz;
            q.w = w;
            q.iyc = c[b].iyc.data();
            q.iyn = c[b].iyn.data();
            q.izc = c[b].izc.data();
            q.izn = c[b].izn.data();
            q.eu = c[b].eu.data();
            q.ev = c[b].ev.data();
            q.iw = c[b].iw.data();
            q.pec_u = m0->pec_u.data();
            q.pec_v = m0->pec_v.data();
            q.sigma = sigma;
            q.big = -1e6 * (std::abs(sigma) + 1.0);
            q.x0 = x[b].data();
            q.maxit = maxit;
            q.tol = tol;
            xo[b] = x[b].data();
        }
    } catch (const std::bad_alloc &) {
        set_error("out of memory");
        return -1;
    }
    if (cem::mode_gpu_solve_batch(&m0->gpu, P.data(), nb, xo.data(), lam.data(), rcs, its)) return -1;
    Work W;
    W.wk.assign((size_t)(ny + 1) * (nz + 1), cplx(0.0));
    for (int b = 0; b < nb; b++) finish_solve(ms[b], c[b], o, 2 * M_PI * f[b], f[b], x[b], lam[b], W);
    return 0;
}

cem_mode2d *cem::mode2d_clone(const cem_mode2d *m) {
    cem_mode2d *c = new cem_mode2d(*m);
    c->gpu = nullptr;
    return c;
}

void cem::mode2d_copy_state(cem_mode2d *dst, const cem_mode2d *src) {
    dst->beta = src->beta;
    dst->freq = src->freq;
    dst->eu = src->eu;
    dst->ev = src->ev;
    dst->hu = src->hu;
    dst->hv = src->hv;
}

bool cem::mode2d_on_gpu() { return cem::mode_gpu_enabled(); }

extern "C" double cem_mode2d_beta(const cem_mode2d *m) { return m->beta.real(); }
extern "C" double cem_mode2d_alpha(const cem_mode2d *m) { return -m->beta.imag(); }
double cem::mode_freq_hint(const cem_mode2d *m) { return m->freq; }
extern "C" const double *cem_mode2d_eu(const cem_mode2d *m) { return reinterpret_cast<const double *>(m->eu.data()); }
extern "C" const double *cem_mode2d_ev(const cem_mode2d *m) { return reinterpret_cast<const double *>(m->ev.data()); }
extern "C" const double *cem_mode2d_hu(const cem_mode2d *m) { return reinterpret_cast<const double *>(m->hu.data()); }
extern "C" const double *cem_mode2d_hv(const cem_mode2d *m) { return reinterpret_cast<const double *>(m->hv.data()); }

// complex power 1/2 sum (Eu Hv* - Ev Hu*) dA, real part (mode2d.c:417-425)
extern "C" double cem_mode2d_power(const cem_mode2d *m) {
    const int ny = m->ny, nz = m->nz;
    cplx P = 0.0;
    for (int j = 0; j < ny; j++)
        for (int k = 0; k <= nz; k++)
            P += m->eu[IU(j, k, nz)] * std::conj(m->hv[IU(j, k, nz)]) * m->dy[j] * m->ddz[k];
    for (int j = 0; j <= ny; j++)
        for (int k = 0; k < nz; k++)
            P -= m->ev[IV(j, k, nz)] * std::conj(m->hu[IV(j, k, nz)]) * m->ddy[j] * m->dz[k];
    return 0.5 * P.real();
}

// V = -sum Ez dz on node column j0, k0..k1; I = circulation around the dual
// rectangle, current in +x (mode2d.c:429-438)
extern "C" void cem_mode2d_vi(const cem_mode2d *m, int j0, int k0, int k1, int ja, int jb, int ka, int kb, double *V,
                              double *Iout) {
    const int nz = m->nz;
    cplx v = 0.0, c = 0.0;
    for (int k = k0; k < k1; k++) v -= m->ev[IV(j0, k, nz)] * m->dz[k];
    for (int j = ja + 1; j <= jb; j++) c += (m->hu[IV(j, ka, nz)] - m->hu[IV(j, kb, nz)]) * m->ddy[j];
    for (int k = ka + 1; k <= kb; k++) c += (m->hv[IU(jb, k, nz)] - m->hv[IU(ja, k, nz)]) * m->ddz[k];
    V[0] = v.real();
    V[1] = v.imag();
    Iout[0] = c.real();
    Iout[1] = c.imag();
}

// conductor attenuation by perturbation over metal surfaces facing field
// cells (mode2d.c:444-498)
extern "C" double cem_mode2d_conductor_alpha(const cem_mode2d *m, double Rs) {
    const int ny = m->ny, nz = m->nz;
    auto solid = [&](int j, int k) {
        return j >= 0 && j < ny && k >= 0 && k < nz && m->pec_u[IU(j, k, nz)] && m->pec_u[IU(j, k + 1, nz)] &&
               m->pec_v[IV(j, k, nz)] && m->pec_v[IV(j + 1, k, nz)];
    };
    const double w = 2 * M_PI * m->freq;
    double loss = 0;
    for (int j = 0; j < ny; j++)           // horizontal surfaces (E_u segments)
        for (int k = 0; k <= nz; k++) {
            if (!m->pec_u[IU(j, k, nz)]) continue;
            if (m->ps[0][j] > 0 || m->ps[3][k] > 0) continue;
            for (int side = 0; side < 2; side++) {
                int kc = side == 0 ? k : k - 1;
                if (kc < 0 || kc >= nz || solid(j, kc)) continue;
                cplx cw = (m->ev[IV(j + 1, kc, nz)] - m->ev[IV(j, kc, nz)]) / m->dy[j] -
                          (m->eu[IU(j, kc + 1, nz)] - m->eu[IU(j, kc, nz)]) / m->dz[kc];
                double hw = std::abs(cw) / (w * CEM_MU0);
                loss += 0.5 * Rs * hw * hw * m->dy[j];
                for (int e = 0; e < 2; e++) {
                    int jn = j + e;
                    if (m->pec_v[IV(jn, kc, nz)]) continue;
                    double hu = std::abs(m->hu[IV(jn, kc, nz)]);
                    loss += 0.5 * Rs * hu * hu * 0.5 * m->dy[j];
                }
            }
        }
    for (int j = 0; j <= ny; j++)          // vertical surfaces (E_v segments)
        for (int k = 0; k < nz; k++) {
            if (!m->pec_v[IV(j, k, nz)]) continue;
            if (m->ps[1][j] > 0 || m->ps[2][k] > 0) continue;
            for (int side = 0; side < 2; side++) {
                int jc = side == 0 ? j : j - 1;
                if (jc < 0 || jc >= ny || solid(jc, k)) continue;
                cplx cw = (m->ev[IV(jc + 1, k, nz)] - m->ev[IV(jc, k, nz)]) / m->dy[jc] -
                          (m->eu[IU(jc, k + 1, nz)] - m->eu[IU(jc, k, nz)]) / m->dz[k];
                double hw = std::abs(cw) / (w * CEM_MU0);
                loss += 0.5 * Rs * hw * hw * m->dz[k];
                for (int e = 0; e < 2; e++) {
                    int kn = k + e;
                    if (m->pec_u[IU(jc, kn, nz)]) continue;
                    double hv = std::abs(m->hv[IU(jc, kn, nz)]);
                    loss += 0.5 * Rs * hv * hv * 0.5 * m->dz[k];
                }
            }
        }
    double P = cem_mode2d_power(m);
    return P > 0 ? loss / (2 * P) : 0;
}

// ---- cross-section of a finalized 3D sim on the x node plane i (sim.c:1026-1079) ----
extern "C" cem_mode2d *cem_cross_section_x(const cem_sim *s, int i, int j0, int j1, int k0, int k1) {
    if (!s->finalized || i < 0 || i >= s->n[0] || j0 < 0 || j1 > s->n[1] || k0 < 0 || k1 > s->n[2] ||
        j1 - j0 < 2 || k1 - k0 < 2) {
        set_error("bad cross-section window");
        return nullptr;
    }
    const int ny = j1 - j0, nz = k1 - k0, np = s->npole;
    cem_mode2d *m = cem_mode2d_create(ny, nz, s->node[1].data() + j0, s->node[2].data() + k0);
    const int nu = ny * (nz + 1), nv = (ny + 1) * nz, nn = (ny + 1) * (nz + 1);
    dvec eu(nu), ev(nv), ew(nn), su(nu, 0.0), sv(nv, 0.0), sw(nn, 0.0);
    std::vector<unsigned char> pu(nu, 0), pv(nv, 0), pw(nn, 0);
    dvec du(np ? (size_t)np * nu : 0, 0.0), dv(np ? (size_t)np * nv : 0, 0.0), dw(np ? (size_t)np * nn : 0, 0.0);
    double dep[kMaxPoles];
    int pec;
    // materials from cell averages; PEC from the update coefficients (sheets included)
    for (int j = 0; j < ny; j++)
        for (int k = 0; k <= nz; k++) {
            size_t q = IDX(s, i, j0 + j, k0 + k);
            int o = j * (nz + 1) + k;
            if (s->cb[1][q] == 0) { pu[o] = 1; eu[o] = 1; continue; }
            edge_material(s, 1, i, j0 + j, k0 + k, &eu[o], &su[o], &pec);
            if (np && edge_debye(s, 1, i, j0 + j, k0 + k, dep))
                for (int p = 0; p < np; p++) du[(size_t)p * nu + o] = dep[p];
        }
    for (int j = 0; j <= ny; j++)
        for (int k = 0; k < nz; k++) {
            size_t q = IDX(s, i, j0 + j, k0 + k);
            int o = j * nz + k;
            if (s->cb[2][q] == 0) { pv[o] = 1; ev[o] = 1; continue; }
            edge_material(s, 2, i, j0 + j, k0 + k, &ev[o], &sv[o], &pec);
            if (np && edge_debye(s, 2, i, j0 + j, k0 + k, dep))
                for (int p = 0; p < np; p++) dv[(size_t)p * nv + o] = dep[p];
        }
    for (int j = 0; j <= ny; j++)
        for (int k = 0; k <= nz; k++) {
            size_t q = IDX(s, i, j0 + j, k0 + k);
            int o = j * (nz + 1) + k;
            if (s->cb[0][q] == 0) { pw[o] = 1; ew[o] = 1; continue; }
            edge_material(s, 0, i, j0 + j, k0 + k, &ew[o], &sw[o], &pec);
            if (np && edge_debye(s, 0, i, j0 + j, k0 + k, dep))
                for (int p = 0; p < np; p++) dw[(size_t)p * nn + o] = dep[p];
        }
    if (np) cem_mode2d_set_debye(m, np, s->pole_f, du.data(), dv.data(), dw.data());
    cem_mode2d_set_media(m, eu.data(), ev.data(), ew.data(), pu.data(), pv.data(), pw.data());
    cem_mode2d_set_loss(m, su.data(), sv.data(), sw.data());
    cem_mode2d_set_pml(m, 0, s->pml_sc[1].data() + j0, s->pml_kc[1].data() + j0, s->pml_ac[1].data() + j0);
    cem_mode2d_set_pml(m, 1, s->pml_sn[1].data() + j0, s->pml_kn[1].data() + j0, s->pml_an[1].data() + j0);
    cem_mode2d_set_pml(m, 2, s->pml_sc[2].data() + k0, s->pml_kc[2].data() + k0, s->pml_ac[2].data() + k0);
    cem_mode2d_set_pml(m, 3, s->pml_sn[2].data() + k0, s->pml_kn[2].data() + k0, s->pml_an[2].data() + k0);
    return m;
}
