/*
 * oracle.c - CPU restatement of the reference FDTD step (TEST INFRASTRUCTURE ONLY).
 *
 * Drives a finalized plan (cem_plan_* export, solver/src/export.c:16-178)
 * through the reference's per-step order contract (solver/src/sim.c:867-933):
 *   pre   phasor resync every 1024 steps                     sim.c:867-870
 *   H     bulk curl-E update, then CPML psi_H corrections     sim.c:579-606, 680-754
 *   mid   current-probe DFT, port V before, Debye save        sim.c:872-892
 *   E     bulk E update, then CPML psi_E corrections          sim.c:608-637, 756-832
 *   post  Debye correct, SIBC, lumped and modal sources,
 *         Debye update, port and voltage-probe DFTs, rotate   sim.c:894-933
 * With ORC_REAL = double every update uses the reference's expression order,
 * so results agree with the reference core to the last bit (checked in tests).
 * ORC_REAL = float gives an FP32 emulation of the GPU step; with ORC_HALF_STORE as
 * well, every stored state is rounded to binary16 (the accuracy check of a
 * half-precision storage variant), and with ORC_HALF_J_ONLY too, only the Debye
 * currents are (the check of a variant that halves just the largest auxiliary
 * array).
 */
#include "oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef ORC_REAL
#define ORC_REAL double
#endif
typedef ORC_REAL real;

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#ifdef ORC_HALF_STORE
/* FP16-storage emulation (with ORC_REAL = float): every field, psi, Debye current
 * and SIBC state is rounded to IEEE binary16 where a storage variant would write
 * it, while arithmetic stays FP32.  Round to nearest even on the float's bits:
 * 10 significand bits kept for normal halves, the 2^-24 grid below 2^-14
 * (subnormal halves), overflow to infinity from 65520 up. */
#include <stdint.h>
static float half_round(float x) {
    union { float f; uint32_t u; } v = {x};
    const uint32_t sign = v.u & 0x80000000u;
    uint32_t a = v.u & 0x7fffffffu;
    if (a >= 0x7f800000u) return x;
    if (a >= 0x477ff000u) { v.u = sign | 0x7f800000u; return v.f; }
    if (a < 0x38800000u) {
        v.f = nearbyintf(fabsf(x) * 16777216.0f) / 16777216.0f;
        v.u |= sign;
        return v.f;
    }
    a += 0x0fffu + ((a >> 13) & 1u);
    v.u = sign | (a & ~0x1fffu);
    return v.f;
}
/* Stored values sit on a binary16 grid scaled by a power of two per array kind (a
 * storage variant folds the scale into its coefficients): the scale changes the
 * representable range, not the relative precision.  CPML psi carries a 1/d factor
 * (~1e7 for 1e3 V/m over 40 um) and Debye J a 1/dt factor, far beyond binary16's
 * 65504 unscaled. */
#define HALF_SCALE_E 1.0f
#define HALF_SCALE_H 64.0f
#define HALF_SCALE_PSI (1.0f / 65536.0f)
#define HALF_SCALE_J (1.0f / 4294967296.0f)
#define HALF_SCALE_Y 64.0f
static void half_store(real *p, size_t n, float scale) {
    for (size_t i = 0; i < n; i++) p[i] = half_round(p[i] * scale) / scale;
}
/* the rounding itself, exported for its check against an IEEE binary16 conversion */
double orc_half_round(double x) { return (double)half_round((float)x); }
#endif
#define ORC_EPS0 8.8541878128e-12
#define ORC_MU0 1.25663706212e-6

typedef struct {
    int n;
    size_t *idx;       /* global index comp*N + q */
    double *w;
} termlist;

struct orc_sim {
    int n[3];
    size_t si, sj, N;
    double dt;
    long step;
    real *F;                       /* E (3N) then H (3N) */
    real *ca, *cb;                 /* 3N each */
    real *ide[3], *idh[3], *be[3], *ce[3], *bh[3], *ch[3];
    double *d[3], *dd[3];
    int npml[6], has_pml;
    real *psiE[3][2], *psiH[3][2];
    int nport, src_port;
    termlist *pv;                  /* port V terms (dl*dir) */
    termlist *pcs;                 /* port source coefficients */
    double *R, *vold;
    double src_amp, src_t0, src_tau;
    int nvp, nip;
    termlist *vp, *ip;
    int msrc_on;
    termlist msrc;
    double msrc_t0, msrc_tau;
    int ndisp, npole;
    size_t *d_idx;
    real *d_cb, *d_beta, *d_J, *d_Eold;
    double d_k[16];
    int nsibc, nterm;
    size_t *sb_idx, *sb_hidx;
    real *sb_hw, *sb_y;
    double sb_c[16], sb_e[16];
    int nf;
    double *freq, *phE, *phH, *rot;
    double *accV, *accI, *accVp, *accIp, *accSrc;
};

static real *rdup(const double *src, size_t n) {
    real *r = malloc((n ? n : 1) * sizeof(real));
    for (size_t i = 0; i < n; i++) r[i] = (real)src[i];
    return r;
}

static double *ddup(const double *src, size_t n) {
    double *r = malloc((n ? n : 1) * sizeof(double));
    if (n) memcpy(r, src, n * sizeof(double));
    return r;
}

static void terms(termlist *t, int n, const double *idx, const double *w) {
    t->n = n;
    t->idx = malloc((n ? n : 1) * sizeof(size_t));
    t->w = malloc((n ? n : 1) * sizeof(double));
    for (int i = 0; i < n; i++) {
        t->idx[i] = (size_t)idx[i];
        t->w[i] = w[i];
    }
}

/* CPML slab ranges and compact psi layout (sim.c:641-678) */
static void pml_ranges(const orc_sim *s, int a, int is_e, int r[2][2]) {
    int n = s->n[a], lo = s->npml[2 * a], hi = s->npml[2 * a + 1];
    if (is_e) {
        r[0][0] = 1; r[0][1] = lo;
        r[1][0] = n - hi; r[1][1] = n - 1;
        if (lo == 0) { r[0][0] = 1; r[0][1] = 0; }
        if (hi == 0) { r[1][0] = 1; r[1][1] = 0; }
    } else {
        r[0][0] = 0; r[0][1] = lo - 1;
        r[1][0] = n - hi; r[1][1] = n - 1;
    }
}

static size_t pml_per(const orc_sim *s, int a) {
    const int *n = s->n;
    return a == 0 ? (size_t)(n[1] + 1) * (n[2] + 1) : a == 1 ? (size_t)(n[0] + 1) * (n[2] + 1)
                                                             : (size_t)(n[0] + 1) * (n[1] + 1);
}

static size_t pml_len(const orc_sim *s, int a, int is_e) {
    int r[2][2];
    pml_ranges(s, a, is_e, r);
    size_t planes = 0;
    for (int sd = 0; sd < 2; sd++)
        if (r[sd][1] >= r[sd][0]) planes += (size_t)(r[sd][1] - r[sd][0] + 1);
    return planes * pml_per(s, a);
}

static size_t pml_base(const orc_sim *s, int a, int sd, int idx, int r[2][2]) {
    size_t lo = r[0][1] >= r[0][0] ? (size_t)(r[0][1] - r[0][0] + 1) : 0;
    size_t plane = sd == 0 ? (size_t)(idx - r[0][0]) : lo + (size_t)(idx - r[1][0]);
    return plane * pml_per(s, a);
}

orc_sim *orc_create(const orc_plan *p) {
    orc_sim *s = calloc(1, sizeof *s);
    for (int a = 0; a < 3; a++) s->n[a] = p->n[a];
    s->sj = (size_t)p->n[2] + 1;
    s->si = s->sj * ((size_t)p->n[1] + 1);
    s->N = s->si * ((size_t)p->n[0] + 1);
    s->dt = p->dt;
    s->F = calloc(6 * s->N, sizeof(real));
    s->ca = malloc(3 * s->N * sizeof(real));
    s->cb = malloc(3 * s->N * sizeof(real));
    for (int c = 0; c < 3; c++)
        for (size_t q = 0; q < s->N; q++) {
            s->ca[c * s->N + q] = (real)p->ca[c][q];
            s->cb[c * s->N + q] = (real)p->cb[c][q];
        }
    for (int a = 0; a < 3; a++) {
        int n = p->n[a];
        s->ide[a] = rdup(p->ide[a], n + 1);
        s->idh[a] = rdup(p->idh[a], n);
        s->be[a] = rdup(p->be[a], n + 1);
        s->ce[a] = rdup(p->ce[a], n + 1);
        s->bh[a] = rdup(p->bh[a], n);
        s->ch[a] = rdup(p->ch[a], n);
        s->d[a] = ddup(p->d[a], n);
        s->dd[a] = ddup(p->dd[a], n + 1);
    }
    for (int f = 0; f < 6; f++) s->npml[f] = p->npml[f];
    s->has_pml = p->has_pml;
    if (s->has_pml)
        for (int c = 0; c < 3; c++)
            for (int w = 0; w < 2; w++) {
                int a = (c + w + 1) % 3;
                s->psiE[c][w] = calloc(pml_len(s, a, 1) + 1, sizeof(real));
                s->psiH[c][w] = calloc(pml_len(s, a, 0) + 1, sizeof(real));
            }
    s->nport = p->nport;
    s->src_port = p->src_port;
    s->src_amp = p->src_amp;
    s->src_t0 = p->src_t0;
    s->src_tau = p->src_tau;
    s->pv = calloc(p->nport + 1, sizeof(termlist));
    s->pcs = calloc(p->nport + 1, sizeof(termlist));
    s->R = ddup(p->port_R, p->nport);
    s->vold = calloc(p->nport + 1, sizeof(double));
    for (int q = 0, o = 0; q < p->nport; o += p->port_n[q], q++) {
        terms(&s->pv[q], p->port_n[q], p->port_idx + o, p->port_vw + o);
        terms(&s->pcs[q], p->port_n[q], p->port_idx + o, p->port_cs + o);
    }
    s->nvp = p->nvp;
    s->vp = calloc(p->nvp + 1, sizeof(termlist));
    for (int q = 0, o = 0; q < p->nvp; o += p->vp_n[q], q++) terms(&s->vp[q], p->vp_n[q], p->vp_idx + o, p->vp_w + o);
    s->nip = p->nip;
    s->ip = calloc(p->nip + 1, sizeof(termlist));
    for (int q = 0, o = 0; q < p->nip; o += p->ip_n[q], q++) {
        terms(&s->ip[q], p->ip_n[q], p->ip_idx + o, p->ip_w + o);
        for (int e = 0; e < s->ip[q].n; e++) s->ip[q].idx[e] += 3 * s->N;   /* plan H indices are H-relative */
    }
    s->msrc_on = p->msrc_on;
    terms(&s->msrc, p->msrc_on ? p->nmsrc : 0, p->msrc_idx, p->msrc_w);
    s->msrc_t0 = p->msrc_t0;
    s->msrc_tau = p->msrc_tau;
    s->ndisp = p->ndisp;
    s->npole = p->npole;
    s->d_idx = malloc((p->ndisp + 1) * sizeof(size_t));
    for (int e = 0; e < p->ndisp; e++) s->d_idx[e] = (size_t)p->disp_idx[e];
    s->d_cb = rdup(p->disp_cb, p->ndisp);
    s->d_beta = rdup(p->disp_beta, (size_t)p->ndisp * p->npole);
    s->d_J = calloc((size_t)p->ndisp * p->npole + 1, sizeof(real));
    s->d_Eold = calloc(p->ndisp + 1, sizeof(real));
    for (int q = 0; q < p->npole; q++) s->d_k[q] = p->disp_k[q];
    s->nsibc = p->nsibc;
    s->nterm = p->nterm;
    s->sb_idx = malloc((p->nsibc + 1) * sizeof(size_t));
    s->sb_hidx = malloc((4 * (size_t)p->nsibc + 1) * sizeof(size_t));
    s->sb_hw = rdup(p->sibc_hw, 4 * (size_t)p->nsibc);
    s->sb_y = calloc((size_t)p->nsibc * p->nterm + 1, sizeof(real));
    for (int e = 0; e < p->nsibc; e++) {
        s->sb_idx[e] = (size_t)p->sibc_idx[e];
        for (int f = 0; f < 4; f++) s->sb_hidx[4 * e + f] = (size_t)p->sibc_hidx[4 * e + f];
    }
    for (int t = 0; t < p->nterm; t++) {
        s->sb_c[t] = p->sibc_c[t];
        s->sb_e[t] = p->sibc_e[t];
    }
    s->nf = p->nf;
    s->freq = ddup(p->freq, p->nf);
    s->phE = calloc(2 * p->nf, sizeof(double));
    s->phH = calloc(2 * p->nf, sizeof(double));
    s->rot = malloc(2 * p->nf * sizeof(double));
    for (int f = 0; f < p->nf; f++) {
        double w = 2 * M_PI * s->freq[f];
        s->rot[2 * f] = cos(w * s->dt);
        s->rot[2 * f + 1] = -sin(w * s->dt);
    }
    size_t blk = 2 * (size_t)p->nf;
    s->accV = calloc(blk * (p->nport + 1), sizeof(double));
    s->accI = calloc(blk * (p->nport + 1), sizeof(double));
    s->accVp = calloc(blk * (p->nvp + 1), sizeof(double));
    s->accIp = calloc(blk * (p->nip + 1), sizeof(double));
    s->accSrc = calloc(blk, sizeof(double));
    return s;
}

static void tl_free(termlist *t) { free(t->idx); free(t->w); }

void orc_destroy(orc_sim *s) {
    if (!s) return;
    free(s->F); free(s->ca); free(s->cb);
    for (int a = 0; a < 3; a++) {
        free(s->ide[a]); free(s->idh[a]); free(s->be[a]); free(s->ce[a]); free(s->bh[a]); free(s->ch[a]);
        free(s->d[a]); free(s->dd[a]);
        for (int w = 0; w < 2; w++) { free(s->psiE[a][w]); free(s->psiH[a][w]); }
    }
    for (int q = 0; q < s->nport; q++) { tl_free(&s->pv[q]); tl_free(&s->pcs[q]); }
    for (int q = 0; q < s->nvp; q++) tl_free(&s->vp[q]);
    for (int q = 0; q < s->nip; q++) tl_free(&s->ip[q]);
    tl_free(&s->msrc);
    free(s->pv); free(s->pcs); free(s->R); free(s->vold); free(s->vp); free(s->ip);
    free(s->d_idx); free(s->d_cb); free(s->d_beta); free(s->d_J); free(s->d_Eold);
    free(s->sb_idx); free(s->sb_hidx); free(s->sb_hw); free(s->sb_y);
    free(s->freq); free(s->phE); free(s->phH); free(s->rot);
    free(s->accV); free(s->accI); free(s->accVp); free(s->accIp); free(s->accSrc);
    free(s);
}

static inline int imax(int a, int b) { return a > b ? a : b; }

/* ---- H: bulk (sim.c:579-606) then CPML (sim.c:680-754) ---- */
static void step_H(orc_sim *s) {
    const int nx = s->n[0], ny = s->n[1], nz = s->n[2];
    const size_t si = s->si, sj = s->sj, N = s->N;
    const real dth = (real)(s->dt / ORC_MU0);
    real *Ex = s->F, *Ey = s->F + N, *Ez = s->F + 2 * N, *Hx = s->F + 3 * N, *Hy = s->F + 4 * N, *Hz = s->F + 5 * N;
    const real *ix = s->idh[0], *iy = s->idh[1], *iz = s->idh[2];
    for (int i = 0; i <= nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                size_t q = i * si + j * sj + k;
                Hx[q] -= dth * ((Ez[q + sj] - Ez[q]) * iy[j] - (Ey[q + 1] - Ey[q]) * iz[k]);
            }
    for (int i = 0; i < nx; i++)
        for (int j = 0; j <= ny; j++)
            for (int k = 0; k < nz; k++) {
                size_t q = i * si + j * sj + k;
                Hy[q] -= dth * ((Ex[q + 1] - Ex[q]) * iz[k] - (Ez[q + si] - Ez[q]) * ix[i]);
            }
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k <= nz; k++) {
                size_t q = i * si + j * sj + k;
                Hz[q] -= dth * ((Ey[q + si] - Ey[q]) * ix[i] - (Ex[q + sj] - Ex[q]) * iy[j]);
            }
    if (!s->has_pml) return;
    int r[2][2];
    pml_ranges(s, 0, 0, r);     /* d/dx: Hy (+dEz/dx), Hz (-dEy/dx) */
    for (int sd = 0; sd < 2; sd++)
        for (int i = r[sd][0]; i <= r[sd][1]; i++) {
            real b = s->bh[0][i], c = s->ch[0][i], id = (real)(1.0 / s->d[0][i]);
            size_t base = pml_base(s, 0, sd, i, r);
            real *pHy = s->psiH[1][1], *pHz = s->psiH[2][0];
            for (int j = 0; j <= ny; j++)
                for (int k = 0; k < nz; k++) {
                    size_t q = i * si + j * sj + k, p = base + (size_t)j * (nz + 1) + k;
                    pHy[p] = b * pHy[p] + c * (Ez[q + si] - Ez[q]) * id;
                    Hy[q] += dth * pHy[p];
                }
            for (int j = 0; j < ny; j++)
                for (int k = 0; k <= nz; k++) {
                    size_t q = i * si + j * sj + k, p = base + (size_t)j * (nz + 1) + k;
                    pHz[p] = b * pHz[p] + c * (Ey[q + si] - Ey[q]) * id;
                    Hz[q] -= dth * pHz[p];
                }
        }
    pml_ranges(s, 1, 0, r);     /* d/dy: Hz (+dEx/dy), Hx (-dEz/dy) */
    for (int sd = 0; sd < 2; sd++)
        for (int j = r[sd][0]; j <= r[sd][1]; j++) {
            real b = s->bh[1][j], c = s->ch[1][j], id = (real)(1.0 / s->d[1][j]);
            size_t base = pml_base(s, 1, sd, j, r);
            real *pHz = s->psiH[2][1], *pHx = s->psiH[0][0];
            for (int i = 0; i < nx; i++)
                for (int k = 0; k <= nz; k++) {
                    size_t q = i * si + j * sj + k, p = base + (size_t)i * (nz + 1) + k;
                    pHz[p] = b * pHz[p] + c * (Ex[q + sj] - Ex[q]) * id;
                    Hz[q] += dth * pHz[p];
                }
            for (int i = 0; i <= nx; i++)
                for (int k = 0; k < nz; k++) {
                    size_t q = i * si + j * sj + k, p = base + (size_t)i * (nz + 1) + k;
                    pHx[p] = b * pHx[p] + c * (Ez[q + sj] - Ez[q]) * id;
                    Hx[q] -= dth * pHx[p];
                }
        }
    pml_ranges(s, 2, 0, r);     /* d/dz: Hx (+dEy/dz), Hy (-dEx/dz) */
    for (int sd = 0; sd < 2; sd++)
        for (int k = r[sd][0]; k <= r[sd][1]; k++) {
            real b = s->bh[2][k], c = s->ch[2][k], id = (real)(1.0 / s->d[2][k]);
            size_t base = pml_base(s, 2, sd, k, r);
            real *pHx = s->psiH[0][1], *pHy = s->psiH[1][0];
            for (int i = 0; i <= nx; i++)
                for (int j = 0; j < ny; j++) {
                    size_t q = i * si + j * sj + k, p = base + (size_t)i * (ny + 1) + j;
                    pHx[p] = b * pHx[p] + c * (Ey[q + 1] - Ey[q]) * id;
                    Hx[q] += dth * pHx[p];
                }
            for (int i = 0; i < nx; i++)
                for (int j = 0; j <= ny; j++) {
                    size_t q = i * si + j * sj + k, p = base + (size_t)i * (ny + 1) + j;
                    pHy[p] = b * pHy[p] + c * (Ex[q + 1] - Ex[q]) * id;
                    Hy[q] -= dth * pHy[p];
                }
        }
}

/* ---- E: bulk (sim.c:608-637) then CPML (sim.c:756-832) ---- */
static void step_E(orc_sim *s) {
    const int nx = s->n[0], ny = s->n[1], nz = s->n[2];
    const size_t si = s->si, sj = s->sj, N = s->N;
    real *Ex = s->F, *Ey = s->F + N, *Ez = s->F + 2 * N, *Hx = s->F + 3 * N, *Hy = s->F + 4 * N, *Hz = s->F + 5 * N;
    const real *ix = s->ide[0], *iy = s->ide[1], *iz = s->ide[2];
    const real *cax = s->ca, *cay = s->ca + N, *caz = s->ca + 2 * N;
    const real *cbx = s->cb, *cby = s->cb + N, *cbz = s->cb + 2 * N;
    for (int i = 0; i < nx; i++)
        for (int j = 1; j < ny; j++)
            for (int k = 1; k < nz; k++) {
                size_t q = i * si + j * sj + k;
                Ex[q] = cax[q] * Ex[q] + cbx[q] * ((Hz[q] - Hz[q - sj]) * iy[j] - (Hy[q] - Hy[q - 1]) * iz[k]);
            }
    for (int i = 1; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 1; k < nz; k++) {
                size_t q = i * si + j * sj + k;
                Ey[q] = cay[q] * Ey[q] + cby[q] * ((Hx[q] - Hx[q - 1]) * iz[k] - (Hz[q] - Hz[q - si]) * ix[i]);
            }
    for (int i = 1; i < nx; i++)
        for (int j = 1; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                size_t q = i * si + j * sj + k;
                Ez[q] = caz[q] * Ez[q] + cbz[q] * ((Hy[q] - Hy[q - si]) * ix[i] - (Hx[q] - Hx[q - sj]) * iy[j]);
            }
    if (!s->has_pml) return;
    int r[2][2];
    pml_ranges(s, 0, 1, r);     /* d/dx: Ey (-dHz/dx), Ez (+dHy/dx) */
    for (int sd = 0; sd < 2; sd++)
        for (int i = r[sd][0]; i <= r[sd][1]; i++) {
            real b = s->be[0][i], c = s->ce[0][i], id = (real)(1.0 / s->dd[0][i]);
            size_t base = pml_base(s, 0, sd, i, r);
            real *pEy = s->psiE[1][1], *pEz = s->psiE[2][0];
            for (int j = 0; j < ny; j++)
                for (int k = 1; k < nz; k++) {
                    size_t q = i * si + j * sj + k, p = base + (size_t)j * (nz + 1) + k;
                    pEy[p] = b * pEy[p] + c * (Hz[q] - Hz[q - si]) * id;
                    Ey[q] -= cby[q] * pEy[p];
                }
            for (int j = 1; j < ny; j++)
                for (int k = 0; k < nz; k++) {
                    size_t q = i * si + j * sj + k, p = base + (size_t)j * (nz + 1) + k;
                    pEz[p] = b * pEz[p] + c * (Hy[q] - Hy[q - si]) * id;
                    Ez[q] += cbz[q] * pEz[p];
                }
        }
    pml_ranges(s, 1, 1, r);     /* d/dy: Ez (-dHx/dy), Ex (+dHz/dy) */
    for (int sd = 0; sd < 2; sd++)
        for (int j = r[sd][0]; j <= r[sd][1]; j++) {
            real b = s->be[1][j], c = s->ce[1][j], id = (real)(1.0 / s->dd[1][j]);
            size_t base = pml_base(s, 1, sd, j, r);
            real *pEz = s->psiE[2][1], *pEx = s->psiE[0][0];
            for (int i = imax(1, 0); i < nx; i++)
                for (int k = 0; k < nz; k++) {
                    size_t q = i * si + j * sj + k, p = base + (size_t)i * (nz + 1) + k;
                    pEz[p] = b * pEz[p] + c * (Hx[q] - Hx[q - sj]) * id;
                    Ez[q] -= cbz[q] * pEz[p];
                }
            for (int i = 0; i < nx; i++)
                for (int k = 1; k < nz; k++) {
                    size_t q = i * si + j * sj + k, p = base + (size_t)i * (nz + 1) + k;
                    pEx[p] = b * pEx[p] + c * (Hz[q] - Hz[q - sj]) * id;
                    Ex[q] += cbx[q] * pEx[p];
                }
        }
    pml_ranges(s, 2, 1, r);     /* d/dz: Ex (-dHy/dz), Ey (+dHx/dz) */
    for (int sd = 0; sd < 2; sd++)
        for (int k = r[sd][0]; k <= r[sd][1]; k++) {
            real b = s->be[2][k], c = s->ce[2][k], id = (real)(1.0 / s->dd[2][k]);
            size_t base = pml_base(s, 2, sd, k, r);
            real *pEx = s->psiE[0][1], *pEy = s->psiE[1][0];
            for (int i = 0; i < nx; i++)
                for (int j = 1; j < ny; j++) {
                    size_t q = i * si + j * sj + k, p = base + (size_t)i * (ny + 1) + j;
                    pEx[p] = b * pEx[p] + c * (Hy[q] - Hy[q - 1]) * id;
                    Ex[q] -= cbx[q] * pEx[p];
                }
            for (int i = 1; i < nx; i++)
                for (int j = 0; j < ny; j++) {
                    size_t q = i * si + j * sj + k, p = base + (size_t)i * (ny + 1) + j;
                    pEy[p] = b * pEy[p] + c * (Hx[q] - Hx[q - 1]) * id;
                    Ey[q] += cby[q] * pEy[p];
                }
        }
}

static void dft_add(double *acc, const double *ph, int nf, double v) {
    for (int f = 0; f < nf; f++) {
        acc[2 * f] += v * ph[2 * f];
        acc[2 * f + 1] += v * ph[2 * f + 1];
    }
}

static void rotate(double *ph, const double *rot, int nf) {
    for (int f = 0; f < nf; f++) {
        double a = ph[2 * f], b = ph[2 * f + 1], c = rot[2 * f], d = rot[2 * f + 1];
        ph[2 * f] = a * c - b * d;
        ph[2 * f + 1] = a * d + b * c;
    }
}

static double sum_terms(const orc_sim *s, const termlist *t) {
    double v = 0;
    for (int e = 0; e < t->n; e++) v += (double)s->F[t->idx[e]] * t->w[e];
    return v;
}

static double lumped_wave(const orc_sim *s, double t) {
    if (s->src_port < 0) return 0.0;
    double x = (t - s->src_t0) / s->src_tau;
    return s->src_amp * exp(-x * x);
}

static double mode_wave(const orc_sim *s, double t) {
    if (!s->msrc_on) return 0;
    double x = (t - s->msrc_t0) / s->msrc_tau;
    return -2 * x * exp(-x * x);
}

#if defined(ORC_HALF_STORE) && !defined(ORC_HALF_J_ONLY)
static void half_store_psi(orc_sim *s, int is_e) {
    if (!s->has_pml) return;
    for (int c = 0; c < 3; c++)
        for (int w = 0; w < 2; w++) {
            const int a = (c + w + 1) % 3;
            half_store(is_e ? s->psiE[c][w] : s->psiH[c][w], pml_len(s, a, is_e), HALF_SCALE_PSI);
        }
}
#endif

void orc_step(orc_sim *s, long nsteps) {
    const int nf = s->nf;
    const size_t blk = 2 * (size_t)nf;
    for (long it = 0; it < nsteps; it++) {
        long n = s->step;
        if ((n & 1023) == 0) {   /* pre (sim.c:867-870, 847-853) */
            double tE = (n + 1) * s->dt, tH = (n + 0.5) * s->dt;
            for (int f = 0; f < nf; f++) {
                double w = 2 * M_PI * s->freq[f];
                s->phE[2 * f] = cos(w * tE); s->phE[2 * f + 1] = -sin(w * tE);
                s->phH[2 * f] = cos(w * tH); s->phH[2 * f + 1] = -sin(w * tH);
            }
        }
        step_H(s);
#if defined(ORC_HALF_STORE) && !defined(ORC_HALF_J_ONLY)
        half_store(s->F + 3 * s->N, 3 * s->N, HALF_SCALE_H);
        half_store_psi(s, 0);
#endif
        /* mid (sim.c:872-892): current probes, port V before, Debye save */
        for (int p = 0; p < s->nip; p++) dft_add(s->accIp + p * blk, s->phH, nf, sum_terms(s, &s->ip[p]) * s->dt);
        for (int p = 0; p < s->nport; p++) s->vold[p] = sum_terms(s, &s->pv[p]);
        for (int e = 0; e < s->ndisp; e++) s->d_Eold[e] = s->F[s->d_idx[e]];
        step_E(s);
#if defined(ORC_HALF_STORE) && !defined(ORC_HALF_J_ONLY)
        half_store(s->F, 3 * s->N, HALF_SCALE_E);
        half_store_psi(s, 1);
#endif
        /* post (sim.c:894-933) */
        double tH = (s->step + 0.5) * s->dt;
        for (int e = 0; e < s->ndisp; e++) {     /* Debye correct (dispersive.c:93-100) */
            real acc = 0;
            for (int p = 0; p < s->npole; p++) acc += (real)(0.5 * (1 + s->d_k[p])) * s->d_J[(size_t)p * s->ndisp + e];
            s->F[s->d_idx[e]] -= s->d_cb[e] * acc;
        }
        for (int e = 0; e < s->nsibc; e++) {     /* SIBC (sibc.c:256-269) */
            real J = 0;
            for (int f = 0; f < 4; f++)
                if (s->sb_hw[4 * e + f] != 0) J += s->sb_hw[4 * e + f] * s->F[3 * s->N + s->sb_hidx[4 * e + f]];
            real E = 0, *y = s->sb_y + (size_t)e * s->nterm;
            for (int t = 0; t < s->nterm; t++) {
                y[t] = (real)s->sb_e[t] * y[t] + (real)(1 - s->sb_e[t]) * J;
                E += (real)s->sb_c[t] * (J - y[t]);
            }
            s->F[s->sb_idx[e]] = E;
        }
        double Vs = lumped_wave(s, tH);
        if (s->src_port >= 0) {
            termlist *cs = &s->pcs[s->src_port];
            for (int e = 0; e < cs->n; e++) s->F[cs->idx[e]] += (real)(cs->w[e] * Vs);
            dft_add(s->accSrc, s->phH, nf, Vs * s->dt);
        }
        if (s->msrc_on) {
            double g = mode_wave(s, tH);
            for (int e = 0; e < s->msrc.n; e++) s->F[s->msrc.idx[e]] += (real)(s->msrc.w[e] * g);
            dft_add(s->accSrc, s->phH, nf, g * s->dt);
        }
        if (s->ndisp) {                           /* Debye update (dispersive.c:102-112) */
            double idt = 1.0 / s->dt;
            for (int e = 0; e < s->ndisp; e++) {
                real dE = (s->F[s->d_idx[e]] - s->d_Eold[e]) * (real)idt;
                for (int p = 0; p < s->npole; p++) {
                    size_t q = (size_t)p * s->ndisp + e;
                    s->d_J[q] = (real)s->d_k[p] * s->d_J[q] + s->d_beta[q] * dE;
                }
            }
        }
#if defined(ORC_HALF_J_ONLY)
        half_store(s->d_J, (size_t)s->npole * s->ndisp, HALF_SCALE_J);
#elif defined(ORC_HALF_STORE)
        half_store(s->F, 3 * s->N, HALF_SCALE_E);
        half_store(s->d_J, (size_t)s->npole * s->ndisp, HALF_SCALE_J);
        half_store(s->sb_y, (size_t)s->nsibc * s->nterm, HALF_SCALE_Y);
#endif
        for (int p = 0; p < s->nport; p++) {
            double V = 0.5 * (sum_terms(s, &s->pv[p]) + s->vold[p]);
            double vs = p == s->src_port ? Vs : 0.0;
            double cur = (vs - V) / s->R[p];
            dft_add(s->accV + p * blk, s->phH, nf, V * s->dt);
            dft_add(s->accI + p * blk, s->phH, nf, cur * s->dt);
        }
        for (int p = 0; p < s->nvp; p++) dft_add(s->accVp + p * blk, s->phE, nf, sum_terms(s, &s->vp[p]) * s->dt);
        rotate(s->phE, s->rot, nf);
        rotate(s->phH, s->rot, nf);
        s->step++;
    }
}

long orc_step_count(const orc_sim *s) { return s->step; }

void orc_field(const orc_sim *s, int comp, double *out) {
    const real *f = s->F + (size_t)comp * s->N;
    for (size_t q = 0; q < s->N; q++) out[q] = (double)f[q];
}

void orc_field_set(orc_sim *s, int comp, const double *in) {
    real *f = s->F + (size_t)comp * s->N;
    for (size_t q = 0; q < s->N; q++) f[q] = (real)in[q];
}

double orc_field_at(const orc_sim *s, int comp, size_t q) { return (double)s->F[(size_t)comp * s->N + q]; }

const double *orc_spectrum(const orc_sim *s, int kind, int idx) {
    size_t blk = 2 * (size_t)s->nf;
    switch (kind) {
    case 0: return idx >= 0 && idx < s->nport ? s->accV + idx * blk : NULL;
    case 1: return idx >= 0 && idx < s->nport ? s->accI + idx * blk : NULL;
    case 2: return idx >= 0 && idx < s->nvp ? s->accVp + idx * blk : NULL;
    case 3: return idx >= 0 && idx < s->nip ? s->accIp + idx * blk : NULL;
    case 4: return s->accSrc;
    default: return NULL;
    }
}

/* W = sum 1/2 eps E^2 V_dual + 1/2 mu0 H^2 V_primal, eps = dt/cb (sim.c:970-994) */
double orc_energy(const orc_sim *s) {
    double we = 0, wm = 0;
    const int *n = s->n;
    for (int c = 0; c < 3; c++) {
        int ua = (c + 1) % 3, va = (c + 2) % 3;
        for (int i = 0; i <= n[0]; i++)
            for (int j = 0; j <= n[1]; j++)
                for (int k = 0; k <= n[2]; k++) {
                    int ijk[3] = {i, j, k};
                    size_t q = i * s->si + j * s->sj + k;
                    double cbv = s->cb[c * s->N + q];
                    if (ijk[c] < n[c] && cbv != 0) {
                        double vol = s->d[c][ijk[c]] * s->dd[ua][ijk[ua]] * s->dd[va][ijk[va]];
                        double eps = s->dt / cbv, e = s->F[c * s->N + q];
                        we += 0.5 * eps * e * e * vol;
                    }
                    if (ijk[ua] < n[ua] && ijk[va] < n[va]) {
                        double vol = s->dd[c][ijk[c]] * s->d[ua][ijk[ua]] * s->d[va][ijk[va]];
                        double h = s->F[(3 + c) * s->N + q];
                        wm += 0.5 * ORC_MU0 * h * h * vol;
                    }
                }
    }
    return we + wm;
}

/* ---- FP64 port post-processing ---- */
void orc_mode_waves(int nf, const double *V0, const double *V1, const double *Ix, const double *z0,
                    const double *alpha, const double *beta, double dx, double d, int dir, double *a, double *b) {
    for (int q = 0; q < nf; q++) {
        double complex Z = z0[2 * q] + I * z0[2 * q + 1];
        double complex g = alpha[q] + I * beta[q];
        double complex c = ccosh(g * dx / 2);
        double complex V = ((V0[2 * q] + V1[2 * q]) + I * (V0[2 * q + 1] + V1[2 * q + 1])) / (2 * c);
        double complex Iin = dir * (Ix[2 * q] + I * Ix[2 * q + 1]);
        double complex sq = 2 * csqrt(Z);
        double complex A = (V + Z * Iin) / sq, B = (V - Z * Iin) / sq;
        A *= cexp(-g * d);
        B *= cexp(g * d);
        a[2 * q] = creal(A); a[2 * q + 1] = cimag(A);
        b[2 * q] = creal(B); b[2 * q + 1] = cimag(B);
    }
}

void orc_s2_renormalize(double *Sv, double z1, double z2, double zref) {
    double complex S11 = Sv[0] + I * Sv[1], S12 = Sv[2] + I * Sv[3], S21 = Sv[4] + I * Sv[5], S22 = Sv[6] + I * Sv[7];
    /* Z = sqrt(Zd) (1 + S)(1 - S)^-1 sqrt(Zd) */
    double complex a11 = 1 - S11, a12 = -S12, a21 = -S21, a22 = 1 - S22;
    double complex det = a11 * a22 - a12 * a21;
    double complex i11 = a22 / det, i12 = -a12 / det, i21 = -a21 / det, i22 = a11 / det;
    double complex b11 = 1 + S11, b12 = S12, b21 = S21, b22 = 1 + S22;
    double complex n11 = b11 * i11 + b12 * i21, n12 = b11 * i12 + b12 * i22;
    double complex n21 = b21 * i11 + b22 * i21, n22 = b21 * i12 + b22 * i22;
    double s1 = sqrt(z1), s2 = sqrt(z2);
    double complex Z11 = s1 * n11 * s1, Z12 = s1 * n12 * s2, Z21 = s2 * n21 * s1, Z22 = s2 * n22 * s2;
    /* S' = (Z - zref)(Z + zref)^-1 */
    double complex p11 = Z11 + zref, p12 = Z12, p21 = Z21, p22 = Z22 + zref;
    double complex m11 = Z11 - zref, m12 = Z12, m21 = Z21, m22 = Z22 - zref;
    double complex dp = p11 * p22 - p12 * p21;
    double complex q11 = p22 / dp, q12 = -p12 / dp, q21 = -p21 / dp, q22 = p11 / dp;
    double complex r11 = m11 * q11 + m12 * q21, r12 = m11 * q12 + m12 * q22;
    double complex r21 = m21 * q11 + m22 * q21, r22 = m21 * q12 + m22 * q22;
    Sv[0] = creal(r11); Sv[1] = cimag(r11); Sv[2] = creal(r12); Sv[3] = cimag(r12);
    Sv[4] = creal(r21); Sv[5] = cimag(r21); Sv[6] = creal(r22); Sv[7] = cimag(r22);
}
