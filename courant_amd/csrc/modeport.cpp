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
    const double dx = s->d[0][d.i_mon];
    cem_mode2d_solve(xs, warp(s, s->freq[q0]), d.neff0, nullptr, 300, 1e-13);
    for (int q = q0; q >= 0 && q < nf; q += step) {
        double fw = warp(s, s->freq[q]);
        double ng = cem_mode2d_beta(xs) / (2 * M_PI * mode_freq_hint(xs) / CEM_C0);
        std::memcpy(start.data(), cem_mode2d_eu(xs), 2 * ny * (nz + 1) * sizeof(double));
        std::memcpy(start.data() + 2 * ny * (nz + 1), cem_mode2d_ev(xs), 2 * (ny + 1) * nz * sizeof(double));
        const int rc = cem_mode2d_solve(xs, fw, ng, start.data(), 300, 1e-13);
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
        cem_mode2d_vi(xs, d.vj - d.j0, d.vk0 - d.k0, d.vk1 - d.k0, d.ja - d.j0, d.jb - d.j0, d.ka - d.k0,
                      d.kb - d.k0, V, Iv);
        // cem_mode2d_vi returns -sum Ez dz; the port voltage is vsign * sum Ez dz
        cplx Z = (double)(-d.vsign) * cplx(V[0], V[1]) / cplx(Iv[0], Iv[1]);
        p.z0[2 * q] = Z.real();
        p.z0[2 * q + 1] = Z.imag();
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
    const int q0 = nearest_freq(s, p.def.f0);
    for (int pass = 0; pass < 2; pass++)
        if (sweep_port(s, p, p.xs, q0, pass == 0 ? 1 : -1, true)) return -1;
    p.solved = true;
    return 0;
}

// With the GPU solver, the sweeps of every unsolved port run at once, one host
// thread and one device context each (a clone of the port's cross-section for the
// downward sweep).  Each sweep is the sequence the sequential solve runs, and the
// device solve is deterministic, so the results are the same; the downward sweep
// leaves q0 to the upward one, which computes the same value.
static int solve_all_ports(cem_sim *s) {
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
