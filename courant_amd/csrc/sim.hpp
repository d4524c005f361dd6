// Internal state of one simulation (host side).  The layout of the set-up
// data mirrors the reference's state model (solver/src/sim_internal.h:63-130)
// because the plan export (cem_plan_*), the checkpoint format and the bit-exact
// integer plan depend on it; the time-dependent state lives on the GPU
// (engine.hip) and is only mirrored here on request.
#pragma once

#include "courant.h"

#include <algorithm>
#include <complex>
#include <cstdint>
#include <memory>
#include <thread>
#include <vector>

namespace cem {

constexpr int kMaxPorts = 16;      // sim_internal.h:14
constexpr int kMaxProbes = 256;    // sim_internal.h:15
constexpr int kMaxModePorts = 8;   // sim_internal.h:42
constexpr int kMaxPoles = 16;      // sim_internal.h:43
constexpr int kSibcTerms = 12;     // sibc.c:122

using cplx = std::complex<double>;
using dvec = std::vector<double>;

void set_error(const char *msg);

struct LumpedPort {                // sim.c:258-276
    int axis = 0, dir = 1;
    std::vector<size_t> edge;      // field index of each edge
    dvec dl;                       // edge lengths
    dvec cs;                       // source coefficient per unit Vs (finalize)
    double R = 50;
};

struct VoltageProbe {              // sim.c:294-307
    int axis = 0, dir = 1;
    std::vector<size_t> edge;
    dvec dl;                       // signed lengths (times dir)
};

struct CurrentProbe {              // sim.c:309-340
    std::vector<int> comp;         // 3..5
    std::vector<size_t> idx;
    dvec w;                        // signed dual lengths
};

struct ModePort {                  // sim_internal.h:45-61
    cem_mode_port_def def{};
    int vp0 = -1, vp1 = -1, ip = -1;
    std::vector<int> src_comp;
    std::vector<size_t> src_idx;
    dvec src_w;
    cem_mode2d *xs = nullptr;
    double xs_f0 = 0;   // warped frequency of the set-up's cold solve while xs still holds it, else 0
    dvec z0, beta, kappa, alpha, alpha_c;   // per output frequency
    bool solved = false;
};

struct Device;                     // engine.hip

}  // namespace cem

struct cem_sim {
    int n[3] = {0, 0, 0};
    cem::dvec node[3], d[3], dd[3];
    size_t si = 0, sj = 0, N = 0;

    // update coefficients (f64 plan, exported by cem_plan_array)
    cem::dvec ca[3], cb[3];
    std::vector<uint8_t> edge_pec[3];

    // cell materials (index i*si + j*sj + k, i<nx, j<ny, k<nz used)
    cem::dvec eps_r, sigma;
    std::vector<uint8_t> cell_pec;

    int bc[6] = {0, 0, 0, 0, 0, 0}, npml[6] = {0, 0, 0, 0, 0, 0};
    cem::dvec ide[3], idh[3], be[3], ce[3], bh[3], ch[3];
    cem::dvec pml_sn[3], pml_kn[3], pml_an[3], pml_sc[3], pml_kc[3], pml_ac[3];
    int has_pml = 0;

    double dt = 0;
    long step = 0;
    int finalized = 0;
    long plan_id = 0;              // set by finalize; an excitation clone keeps its source's

    std::vector<cem::LumpedPort> port;
    int src_port = -1;
    double src_amp = 0, src_t0 = 0, src_tau = 1;

    int nf = 0;
    cem::dvec freq;
    cem::dvec rotE;                // e^{-j w dt}, 2*nf

    std::vector<cem::VoltageProbe> vp;
    std::vector<cem::CurrentProbe> ip;

    std::vector<cem::ModePort> mport;
    int msrc = -1;
    double msrc_t0 = 0, msrc_tau = 1;

    // Debye dispersion (dispersive.c, sim.c:501-549)
    int npole = 0;
    double pole_f[cem::kMaxPoles] = {};
    std::vector<cem::dvec> cell_deps;
    int ndisp = 0;
    std::vector<int> d_comp;
    std::vector<size_t> d_idx;
    cem::dvec d_beta, d_cb;        // beta pole-major [p*ndisp + e]
    double d_k[cem::kMaxPoles] = {};

    // surface-impedance copper (sibc.c)
    int sibc_on = 0, sibc_rough = 0, sibc_npole = 0;
    double sibc_sigma = 0, sibc_p1 = 0, sibc_p2 = 0, sibc_fit_err = 0;
    double sibc_a[cem::kSibcTerms] = {}, sibc_c[cem::kSibcTerms] = {}, sibc_e[cem::kSibcTerms] = {};
    int nsb = 0;
    std::vector<int> sb_comp, sb_nf, sb_hc;
    std::vector<size_t> sb_idx, sb_hi;
    cem::dvec sb_sg;
    std::vector<uint8_t> sb_mask;  // faces present: bit0 +v, bit1 -v, bit2 +u, bit3 -u

    // host FP64 mirrors handed out through cem_field / cem_state_array / spectra
    std::unique_ptr<double[]> field_mirror[6];
    cem::dvec psiE_mirror[3][2], psiH_mirror[3][2];
    cem::dvec dJ_mirror, dEold_mirror, sby_mirror;
    bool state_exposed = false;
    cem::dvec spec_port_v[cem::kMaxPorts], spec_port_i[cem::kMaxPorts];
    std::vector<cem::dvec> spec_vp, spec_ip;
    cem::dvec spec_src;
    bool spectra_exposed = false;

    cem::Device *dev = nullptr;
};

namespace cem {

// fn(lo, hi) over [0, n) in contiguous ranges on up to 16 host threads; set-up
// loops use it where every result is computed by the same code whichever
// thread runs it (the plan stays bit-identical)
template <class F>
void parallel_ranges(long n, F fn) {
    long nt = std::min<long>(16, std::max(1u, std::thread::hardware_concurrency()));
    nt = std::min(nt, n);
    if (nt <= 1) {
        fn(0L, n);
        return;
    }
    std::vector<std::thread> th;
    for (long t = 0; t < nt; t++) th.emplace_back(fn, n * t / nt, n * (t + 1) / nt);
    for (auto &x : th) x.join();
}

inline size_t IDX(const cem_sim *s, int i, int j, int k) {
    return (size_t)i * s->si + (size_t)j * s->sj + (size_t)k;
}

// set-up helpers shared between translation units
void edge_material(const cem_sim *s, int c, int i, int j, int k, double *eps, double *sig, int *pec);
int edge_debye(const cem_sim *s, int c, int i, int j, int k, double *deps);
size_t psi_len(const cem_sim *s, int is_e, int comp, int w);
void pml_ranges(const cem_sim *s, int a, int is_e, int r[2][2]);
size_t pml_per(const cem_sim *s, int a);
int sibc_finalize(cem_sim *s, const unsigned char *is_port);
int modeports_finalize(cem_sim *s);
void modeports_free(cem_sim *s);
double modeport_waveform(const cem_sim *s, double t);
double lumped_waveform(const cem_sim *s, double t);
double mode_freq_hint(const cem_mode2d *m);
// independent copy of a cross-section (own solve state, no device context)
cem_mode2d *mode2d_clone(const cem_mode2d *m);
// solved mode of src (beta, frequency, E, H) into dst
void mode2d_copy_state(cem_mode2d *dst, const cem_mode2d *src);
// true when cem_mode2d_solve runs on the GPU
bool mode2d_on_gpu();
// solves of one cross-section geometry at several frequencies in one GPU batch
int mode2d_solve_batch(cem_mode2d *const *ms, int nb, const double *f, const std::complex<double> *shift,
                       const double *const *start, int maxit, double tol, int *rcs, int *its);
void nnls(int m, int n, const double *A, const double *b, double *x);

// device engine (engine.hip)
void dev_destroy(cem_sim *s);
int dev_ensure(cem_sim *s);                       // allocate + upload the plan on first use
void dev_sources_changed(cem_sim *s);
void dev_step(cem_sim *s, long nsteps);
void dev_mt_H(cem_sim *s, int i0, int i1);
void dev_mt_E(cem_sim *s, int i0, int i1);
void dev_mt_post(cem_sim *s);
double dev_energy(cem_sim *s);
void dev_sync_in(cem_sim *s);                     // host mirrors -> device (before stepping)
void dev_sync_out(cem_sim *s);                    // device -> host mirrors (after stepping)
void dev_field_download(cem_sim *s, int comp, double *dst);
void dev_spectra_download(cem_sim *s);
void dev_push_samples(cem_sim *s, int n, const double *ip, const double *vold, const double *vnew,
                      const double *vp);
void dev_state_download(cem_sim *s, int which, int comp, int w, double *dst);
void dev_state_upload(cem_sim *s, int which, int comp, int w, const double *src);
void dev_phasors(cem_sim *s, double *phE, double *phH, int download);
void dev_acc_io(cem_sim *s, double *buf, int download);   // all DFT accumulators, checkpoint order
void dev_set_step(cem_sim *s, long step);
int dev_mode_waves(cem_sim *s, int vp0, int vp1, int ip, const double *mode, double dx, double dist, int dir,
                   double *a, double *b);   // wave separation of one mode port on the device

}  // namespace cem
