// Mode ports on x-directed lines: definition, source profile from the 2D
// eigenmode at f0, per-frequency wave separation at the monitor plane and the
// 2x2 renormalisation.  Follows solver/src/modeport.c of the reference.
//
// Discrete consistency (modeport.c:12-16): the 2D solve runs at the pre-warped
// frequency (2/dt) sin(w dt/2); its complex eigenvalue kappa maps to the 3D
// propagation constant by kappa = (2/dx) sin(beta dx/2); the voltage is moved to
// the current's half plane with V = (V_m + V_m+1) / (2 cosh(gamma dx/2)).
#include "sim.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

using namespace cem;

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

extern "C" int cem_add_mode_port(cem_sim *s, const cem_mode_port_def *d) {
    if (s->finalized) {
        set_error("already finalized");
        return -1;
    }
    if ((int)s->mport.size() >= kMaxModePorts) {
        set_error("too many mode ports");
        return -1;
    }
    if (d->i_mon < 1 || d->i_mon + 1 > s->n[0] || d->i_src < 1 || d->i_src >= s->n[0] || d->j0 < 0 ||
        d->j1 > s->n[1] || d->k0 < 0 || d->k1 > s->n[2] || (d->dir != 1 && d->dir != -1)) {
        set_error("bad mode port definition");
        return -1;
    }
    if ((d->dir == 1 && d->i_src >= d->i_mon) || (d->dir == -1 && d->i_src <= d->i_mon + 1)) {
        set_error("mode port source must lie behind the monitor (away from the device)");
        return -1;
    }
    ModePort p;
    p.def = *d;
    if (!(p.def.neff0 > 0)) p.def.neff0 = 1.0;
    p.vp0 = cem_add_vprobe(s, CEM_Z, d->i_mon, d->vj, d->vk0, d->vk1, d->vsign);
    p.vp1 = cem_add_vprobe(s, CEM_Z, d->i_mon + 1, d->vj, d->vk0, d->vk1, d->vsign);
    p.ip = cem_add_iprobe(s, CEM_X, d->i_mon, d->ja, d->jb, d->ka, d->kb);
    if (p.vp0 < 0 || p.vp1 < 0 || p.ip < 0) return -1;
    s->mport.push_back(std::move(p));
    return (int)s->mport.size() - 1;
}

extern "C" int cem_add_mode_port_args(cem_sim *s, int dir, int i_src, int i_mon, int j0, int j1, int k0, int k1,
                                      int vj, int vk0, int vk1, int vsign, int ja, int jb, int ka, int kb,
                                      double x_ref, double f0, double neff0) {
    cem_mode_port_def d = {dir, i_src, i_mon, j0, j1, k0, k1, vj, vk0, vk1, vsign, ja, jb, ka, kb, x_ref, f0, neff0};
    return cem_add_mode_port(s, &d);
}

extern "C" int cem_set_mode_excitation(cem_sim *s, int port, double t0, double tau) {
    if (port < 0 || port >= (int)s->mport.size()) {
        set_error("bad mode port");
        return -1;
    }
    s->msrc = port;
    s->msrc_t0 = t0;
    s->msrc_tau = tau;
    dev_sources_changed(s);
    return 0;
}

// frequency at which the continuous-time 2D problem matches the leapfrog (modeport.c:64-67)
static double warp(const cem_sim *s, double f) { return 2.0 / s->dt * std::sin(M_PI * f * s->dt) / (2 * M_PI); }

// differentiated Gaussian, tau d/dt exp(-x^2): zero DC (modeport.c:124-128)
double cem::modeport_waveform(const cem_sim *s, double t) {
    if (s->msrc < 0) return 0;
    double x = (t - s->msrc_t0) / s->msrc_tau;
    return -2 * x * std::exp(-x * x);
}

// Source profile per mode port: real part of the eigenvector at warp(f0),
// normalised by its largest entry, dropped on PEC edges, oriented so that the
// port voltage is positive (modeport.c:70-122).
int cem::modeports_finalize(cem_sim *s) {
    static const double tries[] = {1.0, 1.08, 0.93, 1.16, 0.87};
    for (auto &p : s->mport) {
        const cem_mode_port_def &d = p.def;
        p.xs = cem_cross_section_x(s, d.i_mon, d.j0, d.j1, d.k0, d.k1);
        if (!p.xs) return -1;
    }
    // the eigen-solve at warp(f0), retried at scaled effective indices; the ports are
    // independent, so with the GPU solver they solve at once, one host thread and one
    // device context each, as the frequency sweeps do
    // Returns 1 when solved, 0 when no try converged, -1 on a solver error (device or
    // memory, or a singular band), whose message is kept in err: the error buffer is
    // per thread, so a solving thread hands its message to the caller this way.
    auto solve_f0 = [s](ModePort &p, std::string &err) {
        for (int t = 0; t < 5; t++) {
            const int rc = cem_mode2d_solve(p.xs, warp(s, p.def.f0), p.def.neff0 * tries[t], nullptr, 300, 1e-13);
            if (rc == 0) {
                p.def.neff0 *= tries[t];
                p.xs_f0 = warp(s, p.def.f0);
                return 1;
            }
            if (rc < 0) {
                err = cem_last_error();
                return -1;
            }
        }
        return 0;
    };
    std::vector<int> solved(s->mport.size(), 0);
    std::vector<std::string> errs(s->mport.size());
    if (mode2d_on_gpu() && s->mport.size() > 1) {
        std::vector<std::thread> th;
        for (size_t k = 0; k < s->mport.size(); k++)
            th.emplace_back([&solve_f0, &solved, &errs, s, k]() { solved[k] = solve_f0(s->mport[k], errs[k]); });
        for (auto &t : th) t.join();
    } else {
        for (size_t k = 0; k < s->mport.size(); k++) solved[k] = solve_f0(s->mport[k], errs[k]);
    }
    for (size_t k = 0; k < solved.size(); k++) {
        if (solved[k] < 0) {
            const std::string e = "mode port: " + (errs[k].empty() ? std::string("eigen-solver failure") : errs[k]);
            set_error(e.c_str());
            return -1;
        }
        if (solved[k] == 0) {
            set_error("mode port: eigenmode did not converge");
            return -1;
        }
    }
    for (auto &p : s->mport) {
        const cem_mode_port_def &d = p.def;
        const int ny = d.j1 - d.j0, nz = d.k1 - d.k0;
        const double *eu = cem_mode2d_eu(p.xs), *ev = cem_mode2d_ev(p.xs);
        double vmax = 0;
        for (int i = 0; i < ny * (nz + 1); i++)
            if (std::fabs(eu[2 * i]) > vmax) vmax = std::fabs(eu[2 * i]);
        for (int i = 0; i < (ny + 1) * nz; i++)
            if (std::fabs(ev[2 * i]) > vmax) vmax = std::fabs(ev[2 * i]);
        p.src_comp.clear();
        p.src_idx.clear();
        p.src_w.clear();
        for (int j = 0; j < ny; j++)
            for (int k = 0; k <= nz; k++) {
                double w = eu[2 * (j * (nz + 1) + k)] / vmax;
                if (w == 0) continue;
                size_t idx = IDX(s, d.i_src, d.j0 + j, d.k0 + k);
                if (s->cb[1][idx] == 0) continue;
                p.src_comp.push_back(1);
                p.src_idx.push_back(idx);
                p.src_w.push_back(w);
            }
        for (int j = 0; j <= ny; j++)
            for (int k = 0; k < nz; k++) {
                double w = ev[2 * (j * nz + k)] / vmax;
                if (w == 0) continue;
                size_t idx = IDX(s, d.i_src, d.j0 + j, d.k0 + k);
                if (s->cb[2][idx] == 0) continue;
                p.src_comp.push_back(2);
                p.src_idx.push_back(idx);
                p.src_w.push_back(w);
            }
        double V[2], Iv[2];
        cem_mode2d_vi(p.xs, d.vj - d.j0, d.vk0 - d.k0, d.vk1 - d.k0, d.ja - d.j0, d.jb - d.j0, d.ka - d.k0,
                      d.kb - d.k0, V, Iv);
        if (-V[0] * d.vsign < 0)
            for (auto &w : p.src_w) w = -w;
    }
    return 0;
}

void cem::modeports_free(cem_sim *s) {
    for (auto &p : s->mport) {
        cem_mode2d_destroy(p.xs);
        p.xs = nullptr;
    }
}

// Port quantities at output frequency q from the mode solved on xs
// (modeport.c:170-195): the discrete propagation constant of the 3D grid from the
// 2D one, the conductor attenuation by perturbation when SIBC is on, and Z0 = V / I.
static void record_freq(cem_sim *s, ModePort &p, cem_mode2d *xs, int q) {
    const cem_mode_port_def &d = p.def;
    const double dx = s->d[0][d.i_mon];
    cplx kc(cem_mode2d_beta(xs), -cem_mode2d_alpha(xs));
    cplx b3 = 2 / dx * std::asin(kc * dx / 2.0);
    p.kappa[2 * q] = kc.real();
    p.kappa[2 * q + 1] = kc.imag();
    p.beta[q] = b3.real();
    p.alpha[q] = -b3.imag();
    if (s->sibc_on) {
        double zr, zi;
        cem_conductor_zs(s, s->freq[q], &zr, &zi);
        p.alpha_c[q] = cem_mode2d_conductor_alpha(xs, zr);
        p.alpha[q] += p.alpha_c[q];
    }
    double V[2], Iv[2];
    cem_mode2d_vi(xs, d.vj - d.j0, d.vk0 - d.k0, d.vk1 - d.k0, d.ja - d.j0, d.jb - d.j0, d.ka - d.k0, d.kb - d.k0, V,
                  Iv);
    // cem_mode2d_vi returns -sum Ez dz; the port voltage is vsign * sum Ez dz
    cplx Z = (double)(-d.vsign) * cplx(V[0], V[1]) / cplx(Iv[0], Iv[1]);
    p.z0[2 * q] = Z.real();
    p.z0[2 * q + 1] = Z.imag();
}

// Per output frequency: solve the port mode, walking outward from the frequency
// nearest f0 and seeding each solve with the previous mode (modeport.c:146-201).
// One sweep of the output frequencies outward from q0 (step +1 or -1), each solve
// seeded with the previous mode (modeport.c:160-168), on the cross-section xs.
static int sweep_port(cem_sim *s, ModePort &p, cem_mode2d *xs, int q0, int step, bool write_q0) {
    const int nf = s->nf;
    const cem_mode_port_def &d = p.def;
    const int ny = d.j1 - d.j0, nz = d.k1 - d.k0;
    const int nst = ny * (nz + 1) + (ny + 1) * nz;
    std::vector<double> start(2 * nst);
    cem_mode2d_solve(xs, warp(s, s->freq[q0]), d.neff0, nullptr, 300, 1e-13);
    for (int q = q0; q >= 0 && q < nf; q += step) {
        double fw = warp(s, s->freq[q]);
        double ng = cem_mode2d_beta(xs) / (2 * M_PI * mode_freq_hint(xs) / CEM_C0);
        std::memcpy(start.data(), cem_mode2d_eu(xs), 2 * ny * (nz + 1) * sizeof(double));
        std::memcpy(start.data() + 2 * ny * (nz + 1), cem_mode2d_ev(xs), 2 * (ny + 1) * nz * sizeof(double));
        const int rc = cem_mode2d_solve(xs, fw, ng, start.data(), 300, 1e-13);
        if (std::getenv("COURANT_MODE_DEBUG"))
            std::fprintf(stderr, "mode sweep: port i_mon %d seq q %d rc %d neff %.9f -> %.9f\n", d.i_mon, q, rc, ng,
                         cem_mode2d_beta(xs) / (2 * M_PI * mode_freq_hint(xs) / CEM_C0));
        if (rc > 0) {
            set_error("mode port: eigenmode did not converge at an output frequency");
            return -1;
        }
        if (rc < 0) {   // the solver's own error (device or memory), kept for the caller
            const std::string e = std::string("mode port: ") + cem_last_error();
            set_error(e.c_str());
            return -1;
        }
        if (q == q0 && !write_q0) continue;
        record_freq(s, p, xs, q);
    }
    return 0;
}

static int nearest_freq(const cem_sim *s, double f0) {
    int q0 = 0;
    for (int q = 0; q < s->nf; q++)
        if (std::fabs(s->freq[q] - f0) < std::fabs(s->freq[q0] - f0)) q0 = q;
    return q0;
}

static void reset_port(cem_sim *s, ModePort &p) {
    const int nf = s->nf;
    p.z0.assign(2 * nf, 0.0);
    p.beta.assign(nf, 0.0);
    p.alpha.assign(nf, 0.0);
    p.kappa.assign(2 * nf, 0.0);
    p.alpha_c.assign(nf, 0.0);
}

// Port quantities at every output frequency: two sweeps from the frequency nearest
// f0, upwards then downwards (modeport.c:146-201).
static int solve_port_modes(cem_sim *s, ModePort &p) {
    reset_port(s, p);
    p.xs_f0 = 0;   // the sweeps move xs
    const int q0 = nearest_freq(s, p.def.f0);
    for (int pass = 0; pass < 2; pass++)
        if (sweep_port(s, p, p.xs, q0, pass == 0 ? 1 : -1, true)) return -1;
    p.solved = true;
    return 0;
}

// Frequencies per direction in one batch of the GPU sweep, bounded so that the batch's
// band storage stays within COURANT_MODE_BATCH_GB (default 24 GiB) per port.
static int sweep_batch(const ModePort &p) {
    int k = 32;
    if (const char *v = std::getenv("COURANT_MODE_BATCH")) k = std::max(1, std::atoi(v));
    double gb = 24;
    if (const char *v = std::getenv("COURANT_MODE_BATCH_GB")) gb = std::max(0.1, std::atof(v));
    const int ny = p.def.j1 - p.def.j0, nz = p.def.k1 - p.def.k0;
    const double n = (double)ny * (nz + 1) + (double)(ny + 1) * nz, bw = 2.0 * (nz + 2) + 2;
    const double band = n * (3 * bw + 1) * 16.0;
    const int fit = (int)std::max(1.0, std::floor(gb * (1ull << 30) / band / 2));
    return std::min(k, fit);
}

// The port's sweep on the GPU in batches.  q0 (the output frequency nearest f0) is
// solved as the sequential sweep solves it (a cold solve, then one seeded by it; the
// set-up's solve stands for the cold one when q0 is f0 and xs has not moved).  Then
// each batch takes the next frequencies of each direction (upwards, downwards; at most
// K, within kSpan of the seed's frequency) and solves them at once, each seeded with
// the mode of the last frequency solved in its direction: its start vector, and as the
// shift the square of k0 times its complex effective index, extrapolated linearly in
// frequency from the last two solved (the sequential sweep seeds each frequency with
// its neighbour's vector and real index).  Members that have not converged after
// kPass1 iterations run again from their current vector with their current eigenvalue
// estimate as the shift (a restarted shift: the distance to the wanted eigenvalue, and
// so the convergence ratio of inverse iteration, drops by orders of magnitude).  A
// member whose mode does not continue its seed (overlap below kOverlap, or an index
// more than kIndexJump away from the guess: another eigenpair) is solved again from its
// solved neighbour.  Inverse iteration converges to the same eigenpair from any of
// these seeds, to its tolerance (|d lambda| < 1e-13 |lambda|, phase-aligned change
// < 1e-10).  K per direction grows while the first passes converge and shrinks when
// they do not (up to the memory bound of sweep_batch).
static int sweep_port_batched(cem_sim *s, ModePort &p) {
    constexpr int kPass1 = 24;
    constexpr double kOverlap = 0.98, kIndexJump = 0.02, kSpan = 0.5;
    const int nf = s->nf;
    const cem_mode_port_def &d = p.def;
    const int ny = d.j1 - d.j0, nz = d.k1 - d.k0;
    const size_t nu2 = 2 * (size_t)ny * (nz + 1), nv2 = 2 * (size_t)(ny + 1) * nz;
    const int q0 = nearest_freq(s, d.f0);
    const bool debug = std::getenv("COURANT_MODE_DEBUG") != nullptr;
    cem_mode2d *xs = p.xs;
    auto fail = [](int rc) {
        if (rc > 0) {
            set_error("mode port: eigenmode did not converge at an output frequency");
        } else {
            const std::string e = std::string("mode port: ") + cem_last_error();
            set_error(e.c_str());
        }
        return -1;
    };
    auto k0 = [](double fw) { return 2 * M_PI * fw / CEM_C0; };
    // complex effective index of a solved mode: (beta - j alpha) / k0
    auto index_of = [&](const cem_mode2d *m) {
        return cplx(cem_mode2d_beta(m), -cem_mode2d_alpha(m)) / k0(mode_freq_hint(m));
    };
    auto vec_of = [&](const cem_mode2d *m, std::vector<double> &v) {
        v.resize(nu2 + nv2);
        std::memcpy(v.data(), cem_mode2d_eu(m), nu2 * sizeof(double));
        std::memcpy(v.data() + nu2, cem_mode2d_ev(m), nv2 * sizeof(double));
    };
    struct Dir {
        int step, next, K;
        std::vector<double> st;   // start vector: the last solved mode of this direction
        cplx n = 0, n_prev = 0;   // complex effective indices of the last two solved
        double fa = 0, f_prev = 0;
        bool has_prev = false;
        cplx guess(double fw) const {
            return has_prev && fa != f_prev ? n + (n - n_prev) / (fa - f_prev) * (fw - fa) : n;
        }
    };
    auto take = [&](Dir &dr, const cem_mode2d *m, double fw) {
        if (dr.fa != 0) {
            dr.n_prev = dr.n;
            dr.f_prev = dr.fa;
            dr.has_prev = true;
        }
        dr.n = index_of(m);
        dr.fa = fw;
        vec_of(m, dr.st);
    };
    // |<u, v>| / (|u| |v|) over the transverse E of two modes stored as (eu, ev)
    auto overlap = [&](const double *u, const cem_mode2d *m) {
        const double *eu = cem_mode2d_eu(m), *ev = cem_mode2d_ev(m);
        double pr = 0, pi = 0, nu = 0, nv = 0;
        for (size_t t = 0; t < nu2 + nv2; t += 2) {
            const double *v = t < nu2 ? eu + t : ev + (t - nu2);
            pr += u[t] * v[0] + u[t + 1] * v[1];
            pi += u[t] * v[1] - u[t + 1] * v[0];
            nu += u[t] * u[t] + u[t + 1] * u[t + 1];
            nv += v[0] * v[0] + v[1] * v[1];
        }
        return nu > 0 && nv > 0 ? std::hypot(pr, pi) / std::sqrt(nu * nv) : 0.0;
    };
    const double f_q0 = warp(s, s->freq[q0]);
    // the cold solve at q0 is the set-up's last solve when q0 is f0 itself (same cross-
    // section, frequency, shift and fixed start vector; the solver is deterministic): xs
    // then still holds its result
    if (p.xs_f0 != f_q0) cem_mode2d_solve(xs, f_q0, d.neff0, nullptr, 300, 1e-13);
    p.xs_f0 = 0;
    {
        std::vector<double> st;
        vec_of(xs, st);
        const double ng = cem_mode2d_beta(xs) / k0(mode_freq_hint(xs));
        if (const int rc = cem_mode2d_solve(xs, f_q0, ng, st.data(), 300, 1e-13)) return fail(rc);
    }
    record_freq(s, p, xs, q0);
    const int kmax = sweep_batch(p);
    Dir dirs[2] = {{1, q0 + 1, std::min(8, kmax)}, {-1, q0 - 1, std::min(8, kmax)}};
    for (auto &dr : dirs) take(dr, xs, f_q0);
    std::vector<cem_mode2d *> pool;   // result objects of a batch (pool[0] holds the batch's device context)
    struct Guard {
        std::vector<cem_mode2d *> &v;
        ~Guard() {
            for (auto *x : v) cem_mode2d_destroy(x);
        }
    } guard{pool};
    auto left = [&](const Dir &dr) { return dr.step > 0 ? nf - dr.next : dr.next + 1; };
    auto count = [&](const Dir &dr) {
        int n = 0;
        while (n < dr.K && n < left(dr)) {
            const double fw = warp(s, s->freq[dr.next + n * dr.step]);
            if (n > 0 && std::fabs(fw - dr.fa) > kSpan * dr.fa) break;
            n++;
        }
        return n;
    };
    // solve `idx` (members of pool) at frequencies f with shifts sh from starts st
    auto solve = [&](const std::vector<int> &idx, const std::vector<double> &f, const std::vector<cplx> &sh,
                     const std::vector<const double *> &st, int maxit, std::vector<int> &rc, std::vector<int> &it) {
        const int m = (int)idx.size();
        std::vector<cem_mode2d *> ms(m);
        for (int b = 0; b < m; b++) ms[b] = pool[idx[b]];
        rc.assign(m, 1);
        it.assign(m, 0);
        return mode2d_solve_batch(ms.data(), m, f.data(), sh.data(), st.data(), maxit, 1e-13, rc.data(), it.data());
    };
    while (left(dirs[0]) > 0 || left(dirs[1]) > 0) {
        std::vector<int> qs, di;
        const int take_n[2] = {count(dirs[0]), count(dirs[1])};
        for (int k = 0; k < 2; k++)
            for (int u = 0; u < take_n[k]; u++) {
                qs.push_back(dirs[k].next + u * dirs[k].step);
                di.push_back(k);
            }
        const int B = (int)qs.size();
        while ((int)pool.size() < B) pool.push_back(mode2d_clone(xs));
        std::vector<double> f(B);
        std::vector<cplx> sh(B), guess(B);
        std::vector<const double *> start(B);
        const std::vector<double> seed[2] = {dirs[0].st, dirs[1].st};   // the anchors move while results are taken
        std::vector<int> all(B);
        for (int b = 0; b < B; b++) {
            all[b] = b;
            f[b] = warp(s, s->freq[qs[b]]);
            guess[b] = dirs[di[b]].guess(f[b]);
            sh[b] = k0(f[b]) * k0(f[b]) * guess[b] * guess[b];
            start[b] = seed[di[b]].data();
        }
        std::vector<int> rc1, it1;
        if (solve(all, f, sh, start, kPass1, rc1, it1)) return fail(-1);
        // second pass: restarted shifts for the members still converging
        std::vector<int> rest;
        for (int b = 0; b < B; b++)
            if (rc1[b]) rest.push_back(b);
        std::vector<int> rc(B, 0);
        int npass2[2] = {0, 0};
        if (!rest.empty()) {
            std::vector<double> f2;
            std::vector<cplx> sh2;
            std::vector<std::vector<double>> v2(rest.size());
            std::vector<const double *> st2;
            for (size_t r = 0; r < rest.size(); r++) {
                const int b = rest[r];
                const cplx bt(cem_mode2d_beta(pool[b]), -cem_mode2d_alpha(pool[b]));
                f2.push_back(f[b]);
                sh2.push_back(bt * bt);
                vec_of(pool[b], v2[r]);
                st2.push_back(v2[r].data());
                npass2[di[b]]++;
            }
            std::vector<int> rc2, it2;
            if (solve(rest, f2, sh2, st2, 300, rc2, it2)) return fail(-1);
            for (size_t r = 0; r < rest.size(); r++) rc[rest[r]] = rc2[r];
        }
        int nre[2] = {0, 0};
        for (int b = 0; b < B; b++) {
            Dir &dr = dirs[di[b]];
            const cplx nb = index_of(pool[b]);
            const bool other = std::abs(nb - guess[b]) > kIndexJump * std::abs(guess[b]) ||
                               overlap(start[b], pool[b]) < kOverlap;
            if (rc[b] || other) {   // solved again from its solved neighbour
                nre[di[b]]++;
                const std::vector<int> one = {b};
                const cplx g = dr.guess(f[b]);
                std::vector<int> r1, i1;
                if (solve(one, {f[b]}, {k0(f[b]) * k0(f[b]) * g * g}, {dr.st.data()}, 300, r1, i1)) return fail(-1);
                if (debug)
                    std::fprintf(stderr, "mode sweep: port i_mon %d re-solve q %d (rc %d, other %d) -> rc %d, %d its\n",
                                 d.i_mon, qs[b], rc[b], (int)other, r1[0], i1[0]);
                if (r1[0]) return fail(r1[0]);
            }
            record_freq(s, p, pool[b], qs[b]);
            take(dr, pool[b], f[b]);   // batch members of a direction come outward in order
            if (qs[b] == 0) mode2d_copy_state(xs, pool[b]);   // the cross-section ends where the sequential solve ends
        }
        for (int k = 0; k < 2; k++) {
            Dir &dr = dirs[k];
            const int n = take_n[k];
            if (n <= 0) continue;
            int worst = 0;
            for (int b = 0; b < B; b++)
                if (di[b] == k) worst = std::max(worst, it1[b]);
            if (debug)
                std::fprintf(stderr, "mode sweep: port i_mon %d dir %+d: %d frequencies from %d, first pass <= %d "
                             "iterations, %d restarted, %d re-solved\n", d.i_mon, dr.step, n, dr.next, worst,
                             npass2[k], nre[k]);
            dr.next += n * dr.step;
            if (nre[k] || 4 * npass2[k] > n) dr.K = std::max(1, dr.K / 2);
            else if (npass2[k] == 0 && worst <= 16) dr.K = std::min(kmax, 2 * dr.K);
        }
    }
    return 0;
}

// With the GPU solver the ports sweep at once, one host thread and one device context
// each, in frequency batches (sweep_port_batched); COURANT_MODE_SWEEP=seq runs the
// sequential sweeps instead, each port's upward and downward sweep on a thread of its
// own (a clone of the cross-section for the downward one).
static int solve_all_ports(cem_sim *s) {
    const char *mode = std::getenv("COURANT_MODE_SWEEP");
    if (!(mode && std::string(mode) == "seq")) {
        std::vector<ModePort *> ports;
        for (auto &p : s->mport)
            if (!p.solved && p.xs) {
                reset_port(s, p);
                ports.push_back(&p);
            }
        std::vector<int> rc(ports.size(), 0);
        std::vector<std::string> err(ports.size());
        std::vector<std::thread> th;
        for (size_t k = 0; k < ports.size(); k++)
            th.emplace_back([s, &ports, &rc, &err, k]() {
                rc[k] = sweep_port_batched(s, *ports[k]);
                if (rc[k]) err[k] = cem_last_error();
            });
        for (auto &t : th) t.join();
        int r = 0;
        for (size_t k = 0; k < ports.size(); k++) {
            if (rc[k]) {
                set_error(err[k].c_str());
                r = -1;
            } else {
                ports[k]->solved = true;
            }
        }
        return r;
    }
    struct Job {
        ModePort *p;
        cem_mode2d *xs;
        int step;
        int rc = 0;
        std::string err;
    };
    std::vector<Job> jobs;
    for (auto &p : s->mport)
        if (!p.solved && p.xs) {
            reset_port(s, p);
            p.xs_f0 = 0;   // the sweeps move xs
            jobs.push_back({&p, p.xs, 1, 0, {}});
            jobs.push_back({&p, mode2d_clone(p.xs), -1, 0, {}});
        }
    std::vector<std::thread> th;
    for (auto &jb : jobs)
        th.emplace_back([s, &jb]() {
            const int q0 = nearest_freq(s, jb.p->def.f0);
            jb.rc = sweep_port(s, *jb.p, jb.xs, q0, jb.step, jb.step == 1);
            if (jb.rc) jb.err = cem_last_error();
        });
    for (auto &t : th) t.join();
    int rc = 0;
    for (size_t k = 0; k < jobs.size(); k += 2) {
        Job &up = jobs[k], &down = jobs[k + 1];
        mode2d_copy_state(up.p->xs, down.xs);   // the cross-section ends where the sequential solve ends
        cem_mode2d_destroy(down.xs);
        if (up.rc || down.rc) {
            set_error((up.rc ? up.err : down.err).c_str());
            rc = -1;
        } else {
            up.p->solved = true;
        }
    }
    return rc;
}

// Incident and outgoing waves at the reference plane (modeport.c:203-227)
// COURANT_WAVES_HOST=1 runs the same FP64 formula on the host (the check path of
// the device wave separation); by default it runs on the GPU.
static bool waves_on_host() {
    const char *v = std::getenv("COURANT_WAVES_HOST");
    return v && std::atoi(v) != 0;
}

extern "C" int cem_mode_port_waves(cem_sim *s, int port, double *a, double *b, double *z0, double *beta) {
    if (port < 0 || port >= (int)s->mport.size() || !s->finalized) {
        set_error("bad mode port");
        return -1;
    }
    ModePort &p = s->mport[port];
    if (!p.solved) {
        const int rc = mode2d_on_gpu() ? solve_all_ports(s) : solve_port_modes(s, p);
        if (!p.solved) return -1;
        (void)rc;
    }
    const cem_mode_port_def &d = p.def;
    const double dx = s->d[0][d.i_mon];
    const double xmon = 0.5 * (s->node[0][d.i_mon] + s->node[0][d.i_mon + 1]);
    const double dist = d.dir * (d.x_ref - xmon);
    if (!waves_on_host()) {
        // FP64 on the GPU, next to the DFT accumulators; only a and b come back
        std::vector<double> mode(4 * (size_t)s->nf);
        for (int q = 0; q < s->nf; q++) {
            mode[4 * q] = p.z0[2 * q];
            mode[4 * q + 1] = p.z0[2 * q + 1];
            mode[4 * q + 2] = p.alpha[q];
            mode[4 * q + 3] = p.beta[q];
        }
        if (dev_mode_waves(s, p.vp0, p.vp1, p.ip, mode.data(), dx, dist, d.dir, a, b)) return -1;
        for (int q = 0; q < s->nf; q++) {
            if (z0) z0[q] = p.z0[2 * q];
            if (beta) beta[q] = p.beta[q];
        }
        return 0;
    }
    const double *V0 = cem_vprobe(s, p.vp0), *V1 = cem_vprobe(s, p.vp1), *Ix = cem_iprobe(s, p.ip);
    if (!V0 || !V1 || !Ix) return -1;
    for (int q = 0; q < s->nf; q++) {
        cplx Z(p.z0[2 * q], p.z0[2 * q + 1]);
        cplx g(p.alpha[q], p.beta[q]);
        cplx c = std::cosh(g * dx / 2.0);
        cplx V = cplx(V0[2 * q] + V1[2 * q], V0[2 * q + 1] + V1[2 * q + 1]) / (2.0 * c);
        cplx Iin = (double)d.dir * cplx(Ix[2 * q], Ix[2 * q + 1]);
        cplx sq = 2.0 * std::sqrt(Z);
        cplx A = (V + Z * Iin) / sq, Bw = (V - Z * Iin) / sq;
        A *= std::exp(-g * dist);
        Bw *= std::exp(g * dist);
        a[2 * q] = A.real();
        a[2 * q + 1] = A.imag();
        b[2 * q] = Bw.real();
        b[2 * q + 1] = Bw.imag();
        if (z0) z0[q] = Z.real();
        if (beta) beta[q] = p.beta[q];
    }
    return 0;
}

extern "C" int cem_mode_port_alpha(const cem_sim *s, int port, double *alpha) {
    if (port < 0 || port >= (int)s->mport.size() || !s->mport[port].solved) {
        set_error("mode port not solved");
        return -1;
    }
    std::memcpy(alpha, s->mport[port].alpha.data(), s->nf * sizeof(double));
    return 0;
}

extern "C" int cem_mode_port_alpha_c(const cem_sim *s, int port, double *alpha_c) {
    if (port < 0 || port >= (int)s->mport.size() || !s->mport[port].solved) {
        set_error("mode port not solved");
        return -1;
    }
    std::memcpy(alpha_c, s->mport[port].alpha_c.data(), s->nf * sizeof(double));
    return 0;
}

// Z = sqrt(Zd) (I + S)(I - S)^-1 sqrt(Zd);  S' = (Z - zref)(Z + zref)^-1  (modeport.c:241-264)
extern "C" void cem_s2_renormalize(double *Sv, double z1, double z2, double zref) {
    cplx S[2][2] = {{cplx(Sv[0], Sv[1]), cplx(Sv[2], Sv[3])}, {cplx(Sv[4], Sv[5]), cplx(Sv[6], Sv[7])}};
    cplx A[2][2] = {{1.0 - S[0][0], -S[0][1]}, {-S[1][0], 1.0 - S[1][1]}};
    cplx det = A[0][0] * A[1][1] - A[0][1] * A[1][0];
    cplx Ai[2][2] = {{A[1][1] / det, -A[0][1] / det}, {-A[1][0] / det, A[0][0] / det}};
    cplx B[2][2] = {{1.0 + S[0][0], S[0][1]}, {S[1][0], 1.0 + S[1][1]}};
    cplx Zn[2][2];
    for (int r = 0; r < 2; r++)
        for (int c = 0; c < 2; c++) Zn[r][c] = B[r][0] * Ai[0][c] + B[r][1] * Ai[1][c];
    const double sq[2] = {std::sqrt(z1), std::sqrt(z2)};
    cplx Z[2][2];
    for (int r = 0; r < 2; r++)
        for (int c = 0; c < 2; c++) Z[r][c] = sq[r] * Zn[r][c] * sq[c];
    cplx P[2][2] = {{Z[0][0] + zref, Z[0][1]}, {Z[1][0], Z[1][1] + zref}};
    cplx M[2][2] = {{Z[0][0] - zref, Z[0][1]}, {Z[1][0], Z[1][1] - zref}};
    cplx dp = P[0][0] * P[1][1] - P[0][1] * P[1][0];
    cplx Pi[2][2] = {{P[1][1] / dp, -P[0][1] / dp}, {-P[1][0] / dp, P[0][0] / dp}};
    cplx R[2][2];
    for (int r = 0; r < 2; r++)
        for (int c = 0; c < 2; c++) R[r][c] = M[r][0] * Pi[0][c] + M[r][1] * Pi[1][c];
    Sv[0] = R[0][0].real(); Sv[1] = R[0][0].imag();
    Sv[2] = R[0][1].real(); Sv[3] = R[0][1].imag();
    Sv[4] = R[1][0].real(); Sv[5] = R[1][0].imag();
    Sv[6] = R[1][1].real(); Sv[7] = R[1][1].imag();
}
