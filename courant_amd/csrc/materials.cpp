// Dispersive and lossy-conductor set-up: Debye pole sets and per-edge weights,
// the Djordjevic-Sarkar fit, the surface-impedance (SIBC) Foster fit and the
// surface-edge / face lists.
//
// Follows solver/src/dispersive.c and solver/src/sibc.c of the reference.
// The least-squares fits are Lawson-Hanson NNLS on the normal equations, the
// method the reference uses (sibc.c:44-111), written here so that the fitted
// coefficients agree bit for bit.
#include "sim.hpp"

#include <cmath>
#include <cstring>
#include <vector>

using namespace cem;

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

// ---- Debye (dispersive.c:23-87) ----
extern "C" int cem_set_debye_poles(cem_sim *s, int npole, const double *f_p) {
    if (s->finalized) {
        set_error("already finalized");
        return -1;
    }
    if (s->npole) {
        set_error("Debye poles already set");
        return -1;
    }
    if (npole < 1 || npole > kMaxPoles) {
        set_error("bad pole count");
        return -1;
    }
    for (int p = 0; p < npole; p++)
        if (!(f_p[p] > 0)) {
            set_error("pole frequencies must be positive");
            return -1;
        }
    s->npole = npole;
    s->cell_deps.assign(npole, dvec());
    for (int p = 0; p < npole; p++) {
        s->pole_f[p] = f_p[p];
        s->cell_deps[p].assign(s->N, 0.0);
    }
    return 0;
}

extern "C" int cem_material_box_debye(cem_sim *s, double x0, double x1, double y0, double y1, double z0, double z1,
                                      double eps_inf, double sigma, const double *deps) {
    if (!s->npole) {
        set_error("set Debye poles first");
        return -1;
    }
    if (cem_material_box(s, x0, x1, y0, y1, z0, z1, eps_inf, sigma)) return -1;
    double lo[3] = {x0, y0, z0}, hi[3] = {x1, y1, z1};
    int r[3][2];
    for (int a = 0; a < 3; a++) {
        r[a][0] = s->n[a];
        r[a][1] = -1;
        for (int i = 0; i < s->n[a]; i++) {
            double c = 0.5 * (s->node[a][i] + s->node[a][i + 1]);
            if (c >= lo[a] && c <= hi[a]) {
                if (i < r[a][0]) r[a][0] = i;
                if (i > r[a][1]) r[a][1] = i;
            }
        }
        if (r[a][1] < r[a][0]) return 0;
    }
    for (int i = r[0][0]; i <= r[0][1]; i++)
        for (int j = r[1][0]; j <= r[1][1]; j++)
            for (int k = r[2][0]; k <= r[2][1]; k++) {
                size_t q = IDX(s, i, j, k);
                for (int p = 0; p < s->npole; p++) s->cell_deps[p][q] = deps[p];
            }
    return 0;
}

// area-weighted pole weights of the cells around an edge; 1 if any is non-zero
int cem::edge_debye(const cem_sim *s, int c, int i, int j, int k, double *deps) {
    if (!s->npole) return 0;
    int ua = (c + 1) % 3, va = (c + 2) % 3;
    int ijk[3] = {i, j, k};
    double wsum = 0;
    for (int p = 0; p < s->npole; p++) deps[p] = 0;
    for (int du = -1; du <= 0; du++)
        for (int dv = -1; dv <= 0; dv++) {
            int cc[3] = {ijk[0], ijk[1], ijk[2]};
            cc[ua] += du;
            cc[va] += dv;
            if (cc[ua] < 0 || cc[ua] >= s->n[ua] || cc[va] < 0 || cc[va] >= s->n[va]) continue;
            size_t q = IDX(s, cc[0], cc[1], cc[2]);
            double w = s->d[ua][cc[ua]] * s->d[va][cc[va]];
            for (int p = 0; p < s->npole; p++) deps[p] += w * s->cell_deps[p][q];
            wsum += w;
        }
    int any = 0;
    for (int p = 0; p < s->npole; p++) {
        deps[p] = wsum > 0 ? deps[p] / wsum : 0;
        if (deps[p] != 0) any = 1;
    }
    return any;
}

// ---- Lawson-Hanson NNLS, column-major A (m x n) (sibc.c:44-111) ----
void cem::nnls(int m, int n, const double *A, const double *b, double *x) {
    std::vector<int> passive(n, 0), idx(n);
    std::vector<double> w(n), z(n, 0.0), r(m), M((size_t)n * n), g(n);
    auto col = [&](int j) { return A + (size_t)j * m; };
    for (int j = 0; j < n; j++) x[j] = 0;
    for (int outer = 0; outer < 3 * n; outer++) {
        for (int i = 0; i < m; i++) {
            r[i] = b[i];
            for (int j = 0; j < n; j++) r[i] -= col(j)[i] * x[j];
        }
        int jmax = -1;
        double wmax = 1e-14;
        for (int j = 0; j < n; j++) {
            w[j] = 0;
            for (int i = 0; i < m; i++) w[j] += col(j)[i] * r[i];
            if (!passive[j] && w[j] > wmax) {
                wmax = w[j];
                jmax = j;
            }
        }
        if (jmax < 0) break;
        passive[jmax] = 1;
        for (int inner = 0; inner < 3 * n; inner++) {
            // unconstrained least squares on the passive set: normal equations,
            // Gaussian elimination with partial pivoting
            int np = 0;
            for (int j = 0; j < n; j++)
                if (passive[j]) idx[np++] = j;
            for (int a = 0; a < np; a++) {
                const double *Ca = col(idx[a]);
                g[a] = 0;
                for (int i = 0; i < m; i++) g[a] += Ca[i] * b[i];
                for (int c = 0; c < np; c++) {
                    const double *Cc = col(idx[c]);
                    double acc = 0;
                    for (int i = 0; i < m; i++) acc += Ca[i] * Cc[i];
                    M[a * np + c] = acc;
                }
            }
            for (int a = 0; a < np; a++) {
                int p = a;
                for (int q = a + 1; q < np; q++)
                    if (std::fabs(M[q * np + a]) > std::fabs(M[p * np + a])) p = q;
                if (p != a) {
                    for (int c = 0; c < np; c++) std::swap(M[a * np + c], M[p * np + c]);
                    std::swap(g[a], g[p]);
                }
                for (int q = a + 1; q < np; q++) {
                    double l = M[q * np + a] / M[a * np + a];
                    for (int c = a; c < np; c++) M[q * np + c] -= l * M[a * np + c];
                    g[q] -= l * g[a];
                }
            }
            for (int a = np - 1; a >= 0; a--) {
                double acc = g[a];
                for (int c = a + 1; c < np; c++) acc -= M[a * np + c] * z[idx[c]];
                z[idx[a]] = acc / M[a * np + a];
            }
            bool feasible = true;
            for (int a = 0; a < np; a++)
                if (z[idx[a]] <= 0) feasible = false;
            if (feasible) {
                for (int a = 0; a < np; a++) x[idx[a]] = z[idx[a]];
                break;
            }
            double alpha = 1;
            for (int a = 0; a < np; a++) {
                int j = idx[a];
                if (z[j] <= 0) {
                    double t = x[j] / (x[j] - z[j]);
                    if (t < alpha) alpha = t;
                }
            }
            for (int a = 0; a < np; a++) {
                int j = idx[a];
                x[j] += alpha * (z[j] - x[j]);
                if (x[j] <= 1e-300) {
                    x[j] = 0;
                    passive[j] = 0;
                }
            }
        }
    }
}

// ---- Djordjevic-Sarkar fit (dispersive.c:123-171) ----
// wideband log model between 10^m1 and 10^m2 rad/s, npole Debye poles
// log-spaced over [fmin/10, 10 fmax], NNLS with relative weights
static void ds_logterm(double f, double w1, double w2, double m1, double m2, double *re, double *im) {
    double w = 2 * M_PI * f;
    double nr = w2 * w2 + w * w, dr = w1 * w1 + w * w;
    double ar = std::atan2(w, w2) - std::atan2(w, w1);
    *re = 0.5 * std::log(nr / dr) / std::log(10.0) / (m2 - m1);
    *im = ar / std::log(10.0) / (m2 - m1);
}

extern "C" double cem_fit_djordjevic_sarkar(double dk, double df, double fref, double fmin, double fmax, int npole,
                                            double *eps_inf, double *f_p, double *deps) {
    const double m1 = 3, m2 = 12;
    double w1 = 2 * M_PI * std::pow(10, m1), w2 = 2 * M_PI * std::pow(10, m2);
    double Lr, Li;
    ds_logterm(fref, w1, w2, m1, m2, &Lr, &Li);
    double d_eps = -dk * df / Li;
    double e_inf_ds = dk - d_eps * Lr;
    const int M = 200, K = npole + 1;
    std::vector<double> A((size_t)2 * M * K), b(2 * M), x(K);
    for (int p = 0; p < npole; p++)
        f_p[p] = fmin / 10 * std::pow(100 * fmax / fmin, npole > 1 ? (double)p / (npole - 1) : 0.5);
    double wr = 1 / dk, wi = 1 / (dk * df > 1e-6 ? dk * df : 1e-6);
    for (int i = 0; i < M; i++) {
        double f = fmin * std::pow(fmax / fmin, (double)i / (M - 1)), re, im;
        ds_logterm(f, w1, w2, m1, m2, &re, &im);
        double er = e_inf_ds + d_eps * re, ei = d_eps * im;
        b[i] = er * wr;
        b[M + i] = ei * wi;
        for (int p = 0; p < npole; p++) {
            double r = f / f_p[p], d = 1 + r * r;
            A[(size_t)p * 2 * M + i] = (1 / d) * wr;
            A[(size_t)p * 2 * M + M + i] = (-r / d) * wi;
        }
        A[(size_t)npole * 2 * M + i] = wr;
        A[(size_t)npole * 2 * M + M + i] = 0;
    }
    nnls(2 * M, K, A.data(), b.data(), x.data());
    *eps_inf = x[npole];
    for (int p = 0; p < npole; p++) deps[p] = x[p];
    double worst = 0;
    for (int i = 0; i < M; i++) {
        double f = fmin * std::pow(fmax / fmin, (double)i / (M - 1)), re, im;
        ds_logterm(f, w1, w2, m1, m2, &re, &im);
        double er = e_inf_ds + d_eps * re, ei = d_eps * im;
        double fr = x[npole], fi = 0;
        for (int p = 0; p < npole; p++) {
            double r = f / f_p[p], d = 1 + r * r;
            fr += x[p] / d;
            fi += -x[p] * r / d;
        }
        double t0 = -ei / er, t1 = -fi / fr, e = std::fabs(t1 - t0) / t0;
        if (e > worst) worst = e;
    }
    return worst;
}

// ---- surface impedance of rough copper (sibc.c:26-42) ----
static void zs_target(double f, double sigma, int rough, double p1, double p2, double *re, double *im) {
    double w = 2 * M_PI * f;
    double rs = std::sqrt(w * CEM_MU0 / (2 * sigma));
    double K = 1;
    double delta = std::sqrt(2 / (w * CEM_MU0 * sigma));
    if (rough == 1) {          // Groisse, p1 = Rq
        K = 1 + std::exp(-std::pow(delta / (2 * p1), 1.6));
    } else if (rough == 2) {   // Huray single sphere size, p1 = radius, p2 = spheres per m^2
        double a = p1;
        K = 1 + p2 * 4 * M_PI * a * a / (1 + delta / a + delta * delta / (2 * a * a));
    }
    *re = rs * K;
    *im = rs * K;
}

// Foster form Z(s) = sum_k c_k s/(s + a_k), K = 12 poles one decade beyond the
// band on each side, fitted by NNLS in relative error (sibc.c:115-156)
extern "C" int cem_set_conductor_loss(cem_sim *s, double sigma, int rough, double p1, double p2, double fmin,
                                      double fmax) {
    if (s->finalized) {
        set_error("already finalized");
        return -1;
    }
    if (!(sigma > 0) || !(fmax > fmin) || !(fmin > 0)) {
        set_error("bad conductor-loss parameters");
        return -1;
    }
    s->sibc_on = 1;
    s->sibc_sigma = sigma;
    s->sibc_rough = rough;
    s->sibc_p1 = p1;
    s->sibc_p2 = p2;
    const int K = kSibcTerms, M = 200;
    std::vector<double> A((size_t)2 * M * K), b(2 * M);
    for (int k = 0; k < K; k++) s->sibc_a[k] = 2 * M_PI * fmin / 10 * std::pow(100 * fmax / fmin, (double)k / (K - 1));
    for (int i = 0; i < M; i++) {
        double f = fmin * std::pow(fmax / fmin, (double)i / (M - 1)), w = 2 * M_PI * f, re, im;
        zs_target(f, sigma, rough, p1, p2, &re, &im);
        double scale = 1 / std::hypot(re, im);
        b[i] = re * scale;
        b[M + i] = im * scale;
        for (int k = 0; k < K; k++) {
            double a = s->sibc_a[k];
            double d = w * w + a * a;
            A[(size_t)k * 2 * M + i] = w * w / d * scale;
            A[(size_t)k * 2 * M + M + i] = w * a / d * scale;
        }
    }
    nnls(2 * M, K, A.data(), b.data(), s->sibc_c);
    s->sibc_npole = K;
    double worst = 0;
    for (int i = 0; i < M; i++) {
        double f = fmin * std::pow(fmax / fmin, (double)i / (M - 1)), w = 2 * M_PI * f, re, im, fr = 0, fi = 0;
        zs_target(f, sigma, rough, p1, p2, &re, &im);
        for (int k = 0; k < K; k++) {
            double a = s->sibc_a[k], d = w * w + a * a;
            fr += s->sibc_c[k] * w * w / d;
            fi += s->sibc_c[k] * w * a / d;
        }
        double e = std::hypot(fr - re, fi - im) / std::hypot(re, im);
        if (e > worst) worst = e;
    }
    s->sibc_fit_err = worst;
    return 0;
}

extern "C" double cem_conductor_fit_error(const cem_sim *s) { return s->sibc_fit_err; }

extern "C" void cem_conductor_zs(const cem_sim *s, double f, double *re, double *im) {
    double w = 2 * M_PI * f;
    *re = *im = 0;
    for (int k = 0; k < s->sibc_npole; k++) {
        double a = s->sibc_a[k], d = w * w + a * a;
        *re += s->sibc_c[k] * w * w / d;
        *im += s->sibc_c[k] * w * a / d;
    }
}

// E edge along c at ijk is metal: out of range, a tangential edge on a PEC outer
// face, or a zeroed update coefficient (sibc.c:173-184)
static int edge_is_metal(const cem_sim *s, int c, const int *ijk) {
    for (int a = 0; a < 3; a++) {
        int lim = a == c ? s->n[a] - 1 : s->n[a];
        if (ijk[a] < 0 || ijk[a] > lim) return 1;
    }
    for (int a = 0; a < 3; a++) {
        if (a == c) continue;
        if (ijk[a] == 0 && s->bc[2 * a] == CEM_BC_PEC) return 1;
        if (ijk[a] == s->n[a] && s->bc[2 * a + 1] == CEM_BC_PEC) return 1;
    }
    return s->cb[c][IDX(s, ijk[0], ijk[1], ijk[2])] == 0;
}

// Surface edges and their faces (sibc.c:186-253).  Face order per edge:
// +v (H_u at q, -1), -v (H_u at q - s_v, +1), +u (H_v at q, +1), -u (H_v at q - s_u, -1),
// each weighted by 1/(number of faces).
int cem::sibc_finalize(cem_sim *s, const unsigned char *is_port) {
    if (!s->sibc_on) return 0;
    const int nx = s->n[0], ny = s->n[1], nz = s->n[2];
    const size_t stride[3] = {s->si, s->sj, 1};
    s->sb_comp.clear();
    s->sb_idx.clear();
    s->sb_nf.clear();
    s->sb_hc.clear();
    s->sb_hi.clear();
    s->sb_sg.clear();
    s->sb_mask.clear();
    // surface edges in (c, i, j, k) order: each task scans a contiguous run of
    // (component, x plane) pairs into its own lists, joined in order afterwards
    struct Part {
        std::vector<int> comp, nf, hc;
        std::vector<size_t> idx, hi;
        std::vector<double> sg;
        std::vector<uint8_t> mask;
    };
    const long nplanes = 3L * (nx + 1);
    const long nt = std::max<long>(1, std::min<long>(64, nplanes));
    std::vector<Part> parts(nt);
    parallel_ranges(nt, [&](long t0, long t1) {
        for (long t = t0; t < t1; t++) {
            Part &P = parts[t];
            for (long r = nplanes * t / nt; r < nplanes * (t + 1) / nt; r++) {
                const int c = (int)(r / (nx + 1)), i = (int)(r % (nx + 1));
                const int ua = (c + 1) % 3, va = (c + 2) % 3;
                for (int j = 0; j <= ny; j++)
                    for (int k = 0; k <= nz; k++) {
                        int ijk[3] = {i, j, k};
                        if (ijk[c] >= s->n[c]) continue;
                        size_t q = IDX(s, i, j, k);
                        if (!edge_is_metal(s, c, ijk)) continue;
                        if (is_port && is_port[c * s->N + q]) continue;
                        bool in_pml = false;   // edges in a CPML slab stay PEC
                        for (int a = 0; a < 3; a++) {
                            if (a == c) {
                                if (ijk[a] < s->npml[2 * a] || ijk[a] >= s->n[a] - s->npml[2 * a + 1]) in_pml = true;
                            } else if (ijk[a] < s->npml[2 * a] || ijk[a] > s->n[a] - s->npml[2 * a + 1]) {
                                in_pml = true;
                            }
                        }
                        if (in_pml) continue;
                        int hc[4], nf = 0, mask = 0;
                        size_t hi[4];
                        double sg[4];
                        int nb[3];
                        auto probe = [&](int axis, int delta) {
                            std::memcpy(nb, ijk, sizeof nb);
                            nb[axis] += delta;
                            return !edge_is_metal(s, c, nb);
                        };
                        if (probe(va, +1)) { hc[nf] = ua; hi[nf] = q; sg[nf] = -1; nf++; mask |= 1; }
                        if (probe(va, -1)) { hc[nf] = ua; hi[nf] = q - stride[va]; sg[nf] = 1; nf++; mask |= 2; }
                        if (probe(ua, +1)) { hc[nf] = va; hi[nf] = q; sg[nf] = 1; nf++; mask |= 4; }
                        if (probe(ua, -1)) { hc[nf] = va; hi[nf] = q - stride[ua]; sg[nf] = -1; nf++; mask |= 8; }
                        if (!nf) continue;
                        P.comp.push_back(c);
                        P.idx.push_back(q);
                        P.nf.push_back(nf);
                        P.mask.push_back((uint8_t)mask);
                        for (int f = 0; f < 4; f++) {
                            P.hc.push_back(f < nf ? hc[f] : 0);
                            P.hi.push_back(f < nf ? hi[f] : 0);
                            P.sg.push_back(f < nf ? sg[f] / nf : 0.0);
                        }
                    }
            }
        }
    });
    for (const Part &P : parts) {
        s->sb_comp.insert(s->sb_comp.end(), P.comp.begin(), P.comp.end());
        s->sb_idx.insert(s->sb_idx.end(), P.idx.begin(), P.idx.end());
        s->sb_nf.insert(s->sb_nf.end(), P.nf.begin(), P.nf.end());
        s->sb_mask.insert(s->sb_mask.end(), P.mask.begin(), P.mask.end());
        s->sb_hc.insert(s->sb_hc.end(), P.hc.begin(), P.hc.end());
        s->sb_hi.insert(s->sb_hi.end(), P.hi.begin(), P.hi.end());
        s->sb_sg.insert(s->sb_sg.end(), P.sg.begin(), P.sg.end());
    }
    s->nsb = (int)s->sb_idx.size();
    for (int k = 0; k < s->sibc_npole; k++) s->sibc_e[k] = std::exp(-s->sibc_a[k] * s->dt);
    return 0;
}
