/*
 * oracle.h - CPU restatement of the reference time stepper (TEST INFRASTRUCTURE).
 *
 * This library is a checker: only tests/, the smoke test and bench.py's CPU
 * baseline load it.  It restates, in plain C, the per-step algorithm of the
 * reference core (solver/src/sim.c:579-933, dispersive.c:89-112,
 * sibc.c:256-269, modeport.c:124-135) and its FP64 port post-processing
 * (modeport.c:203-264), driven by a finalized update plan as exported through
 * the cem_plan_* functions (export.c:16-178).
 *
 * ORC_REAL selects the field arithmetic: double (default, the f64 oracle) or
 * float (an FP32 CPU emulation used to separate FP32 rounding from GPU defects).
 * DFT accumulators are always double.
 */
#ifndef COURANT_ORACLE_H
#define COURANT_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int n[3];
    double dt;
    int npml[6];
    int has_pml;
    const double *ca[3], *cb[3];              /* N each */
    const double *ide[3], *idh[3];            /* n+1 / n, include 1/kappa */
    const double *be[3], *ce[3], *bh[3], *ch[3];
    const double *d[3], *dd[3];
    /* lumped ports: concatenated edge tables (plan order), counts per port */
    int nport;
    const int *port_n;
    const double *port_idx, *port_vw, *port_cs, *port_R;
    int src_port;
    double src_amp, src_t0, src_tau;
    /* voltage and current probes */
    int nvp;
    const int *vp_n;
    const double *vp_idx, *vp_w;
    int nip;
    const int *ip_n;
    const double *ip_idx, *ip_w;
    /* excited mode port source */
    int nmsrc, msrc_on;
    const double *msrc_idx, *msrc_w;
    double msrc_t0, msrc_tau;
    /* Debye ADE */
    int ndisp, npole;
    const double *disp_idx, *disp_cb, *disp_beta, *disp_k;
    /* surface impedance */
    int nsibc, nterm;
    const double *sibc_idx, *sibc_hidx, *sibc_hw, *sibc_c, *sibc_e;
    /* output frequencies */
    int nf;
    const double *freq;
} orc_plan;

typedef struct orc_sim orc_sim;

orc_sim *orc_create(const orc_plan *p);
void orc_destroy(orc_sim *s);
void orc_step(orc_sim *s, long n);
long orc_step_count(const orc_sim *s);
/* field component c (0..2 E, 3..5 H) converted to double into out (N values) */
void orc_field(const orc_sim *s, int comp, double *out);
void orc_field_set(orc_sim *s, int comp, const double *in);
double orc_field_at(const orc_sim *s, int comp, size_t q);
/* DFT accumulator, 2*nf interleaved: kind 0 port V, 1 port I, 2 vprobe, 3 iprobe, 4 source */
const double *orc_spectrum(const orc_sim *s, int kind, int idx);
double orc_energy(const orc_sim *s);

/* Wave separation at a mode-port monitor (modeport.c:203-227): inputs are the
 * two voltage-probe spectra, the current-probe spectrum, the port's complex
 * Z0 (2*nf), alpha and beta (nf), the monitor cell width dx, the signed
 * distance d = dir (x_ref - x_mon) and dir; outputs a and b (2*nf). */
void orc_mode_waves(int nf, const double *V0, const double *V1, const double *Ix, const double *z0,
                    const double *alpha, const double *beta, double dx, double d, int dir, double *a, double *b);
/* liboracle_f16s.so only (ORC_HALF_STORE): x rounded to IEEE binary16 and back */
double orc_half_round(double x);
/* 2x2 renormalisation (modeport.c:242-264) */
void orc_s2_renormalize(double *S, double z1, double z2, double zref);

#ifdef __cplusplus
}
#endif
#endif
