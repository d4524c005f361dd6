// MI355X (gfx950) time-stepping engine behind the courant.h ABI.
//
// Per step (order contract of solver/src/sim.c:867-933):
//   k_update_h : bulk H curl-E update fused with the CPML psi_H recursions
//                (sim.c:579-606, 680-754), one thread per (j,k) column of an
//                i-segment, FP32;
//   k_update_e : bulk E update fused with CPML psi_E (sim.c:608-637, 756-832);
//                (ca, cb) come from a per-component table addressed by 8-bit
//                indices, since a grid has only a handful of distinct pairs;
//   k_post     : the post-phase per-edge physics over the compact edge lists,
//                in reference order: Debye correction (dispersive.c:93-100),
//                SIBC overwrite (sibc.c:256-269), lumped and modal sources
//                (sim.c:900-909, modeport.c:130-135) and the Debye current
//                update (dispersive.c:102-112); one thread per listed edge, the
//                lists in the reference's own (comp, i, j, k) order;
//   k_gather   : port / probe line and loop integrals, one wavefront per
//                sample in one workgroup, FP64 sums reduced with fixed-order
//                shuffles, written to a per-step sample series; it then
//                advances the device step and series-row counters;
//   k_dft      : every chunk of steps, the FP64 running DFT of all samples
//                (sim.c:840-861, export.c:121-149), one thread per
//                (accumulator, frequency), phasors resynchronised every 1024
//                steps and rotated in between exactly as the reference does.
// Fields are component-major FP32 arrays with the reference index
// i*si + j*sj + k, so every index list of the plan applies unchanged.  With
// the step and series row held on the device, every launch argument of a step
// is constant, and runs of steps are replayed from a captured HIP graph.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <chrono>
#include <array>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "courant_amd.h"
#include "devmem.hpp"
#include "sim.hpp"

using namespace cem;

namespace cem {

constexpr int kBlock = 256;        // 4 wavefronts of 64
constexpr int kIseg = 2;           // default i-planes per thread in the stencil kernels
constexpr int kSeriesCap = 512;    // steps recorded between DFT flushes
constexpr int kGraphSteps = 64;
constexpr int kGatherBlock = 1024;    // steps per captured chunk (divides kSeriesCap)

#define HIP_OK(expr)                                                                  \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess) {                                                       \
            char m_[200];                                                             \
            std::snprintf(m_, sizeof m_, "HIP error %s at %s:%d", hipGetErrorString(e_), \
                          __FILE__, __LINE__);                                        \
            set_error(m_);                                                            \
            return -1;                                                                \
        }                                                                             \
    } while (0)

#define HIP_CHECK_VOID(expr)                                                          \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess) {                                                       \
            std::fprintf(stderr, "courant: HIP error %s at %s:%d\n", hipGetErrorString(e_), \
                         __FILE__, __LINE__);                                         \
            std::abort();                                                             \
        }                                                                             \
    } while (0)

// ---------------------------------------------------------------------------
// kernel argument blocks
// ---------------------------------------------------------------------------
struct Grid {
    int nx, ny, nz, S;      // S = (ny+1)(nz+1) points per x plane
    long long si, sj, N;
    int nchunk;             // ceil(S / 64) wavefront chunks per plane
};

struct Pml {                // CPML per derivative axis a: slot per index (-1 outside a slab)
    const int *slot[3];
    const float *b[3], *c[3];
    long long per[3];       // points per slab plane
    int zpl;                // slab planes of the z-derivative arrays (both faces)
};

// psi of the x and y derivatives keeps the reference's slab-major order
// (slot * per + in-plane index), which is k-fastest and so coalesced for the
// column threads.  The z-derivative arrays are stored slot-fastest on the
// device, ((i (ny+1) + j) zpl + slot): the lanes of a wavefront that lie in a
// z slab then touch consecutive words instead of one slab plane each
// (measured: z-face CPML cost 92 us of H and 105 us of E on 600x400x160 in
// slab-major order).  State export transposes to the reference order.
__device__ __forceinline__ long long psi_z(const Grid &g, const Pml &m, int i, int j, int sz) {
    return ((long long)i * (g.ny + 1) + j) * m.zpl + sz;
}

struct HArgs {
    Grid g;
    const float *E;         // 3N
    float *H;               // 3N
    const float *hs[3];     // dt/mu0 * idh per axis
    Pml pml;
    float *psi[3][2];       // psiH[c][w]
    float dth;
    int has_pml;
};

// Per-edge physics parameters, resident on the device so the E kernel's
// argument block stays small.
struct PostParams {
    int ndisp, npole, nsb, nterm;
    float kc[kMaxPoles];    // (1 + k)/2
    float km1[kMaxPoles];   // k - 1 (computed without cancellation)
    float zc[kSibcTerms];   // c_k
    float om[kSibcTerms];   // 1 - e_k, from -expm1(-a_k dt)
};

struct EArgs {
    Grid g;
    float *E;               // 3N
    const float *H;         // 3N
    const float *ca, *cb;   // 3N each (uncompressed coefficients)
    const unsigned *cidx;   // N: packed 8-bit coefficient indices of Ex, Ey, Ez (compressed)
    const float2 *ctab;     // 3 x 256 (ca, cb) pairs (compressed)
    const float *es[3];     // ide per axis
    Pml pml;
    float *psi[3][2];       // psiE[c][w]
    int has_pml;
};

// Device-side step and series-row counters: every per-step kernel argument is
// then constant, so a chunk of steps can be captured once and replayed.
struct Ctl {
    long long step;          // step about to complete (sets the source waveform time)
    int row;                 // series row of this step's samples
    unsigned pad;
};

struct Waveform {            // excitation parameters (sim.c:834-838, modeport.c:124-128)
    int has_lumped, has_mode;
    double amp, t0, tau, mt0, mtau, dt;
};

__device__ __forceinline__ void waveforms(const Waveform &w, long long n, float *vs, float *gm) {
    const double tH = ((double)n + 0.5) * w.dt;
    double a = 0, b = 0;
    if (w.has_lumped) {
        const double x = (tH - w.t0) / w.tau;
        a = w.amp * exp(-x * x);
    }
    if (w.has_mode) {
        const double x = (tH - w.mt0) / w.mtau;
        b = -2 * x * exp(-x * x);
    }
    *vs = (float)a;
    *gm = (float)b;
}

// Post-phase edge lists.  One thread per edge runs every step that applies to
// its edge, in the reference order; an edge can be in several lists (a
// dielectric edge tangential to a PEC outer face is both dispersive and a
// surface edge), so the lists are linked: threads [0, ndisp) take the
// dispersive edges (with their SIBC slot and source slot, if any), then nsbo
// surface edges that are not dispersive, then nxs remaining source edges.
// Indices are global E indices comp*N + q.
struct PostArgs {
    float *E;               // 3N
    const float *H;         // 3N
    long long N, si, sj;
    const PostParams *pp;
    int ndisp, nsb, nsbo, nxs;
    const unsigned *d_idx;
    const unsigned *d_mat;  // per dispersive edge: row of dtab (bits 0-29), has an SIBC slot
                            // (bit 31), has a source slot (bit 30)
    const float *dtab;      // per material: cb, then beta/dt per pole
    int dstride;            // 1 + npole
    // the few dispersive edges that are also SIBC or source edges: ascending edge
    // numbers with their slots (found by binary search from the flag bits)
    const int *dsb_t, *dsb_u, *dsrc_t, *dsrc_x;
    int ndsb, ndsrc;
    float *J;               // npole * ndisp (pole-major, reference order)
    float *Eold;            // E of each dispersive edge before this step's E update
    const unsigned *sb_idx; // per SIBC slot (reference order)
    const unsigned char *face;
    float *Y;               // nterm * nsb (term-major)
    const unsigned *sbo;    // SIBC slots of the surface edges that are not dispersive
    const int *sb_src;      // source slot per SIBC slot, -1 if none (nullptr: none at all)
    const unsigned *x_idx;  // source edges neither dispersive nor surface
    const int *x_src;       // their source slots
    const float *wl, *wm;   // lumped cs and modal weight per source slot
    Waveform wave;
    Ctl *ctl;
};

struct GatherArgs {
    const float *F;         // 6N: E then H
    const long long *tidx;  // term field index into F
    const double *tw;       // term weight
    const int *toff;        // nsamp + 1
    const double *post;     // per-sample factor applied after the sum
    int nsamp, nip, nport, nvp;
    double *vlast;          // port V after the previous E update
    double *ser_ip, *ser_vold, *ser_vnew, *ser_vp;   // sample series, row-major
    int mid_only;           // 1: only refresh vlast (V before the E update)
    Ctl *ctl;
};

struct DftArgs {
    int nf, nport, nvp, nip, nacc, rows;
    long long step0;
    double dt;
    const double *omega;    // 2 pi f
    const double *rot;      // e^{-j w dt}, 2 nf
    const double *ph_in;    // phE (2nf) | phH (2nf) at the first row
    double *ph_out;         // same, advanced past the last row
    const double *ser_ip, *ser_vold, *ser_vnew, *ser_vp;
    const double *R;        // port resistances
    int src_port, has_lumped, has_mode, direct;
    double amp, t0, tau;    // lumped Gaussian (sim.c:834-838)
    double mt0, mtau;       // modal differentiated Gaussian (modeport.c:124-128)
    double *acc;            // nacc * 2nf
    Ctl *ctl;               // series row counter, reset after the flush
};

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
// Workgroups are dealt round-robin to the 8 XCDs (each with its own L2).  Remap
// the dispatch id so that XCD x runs one contiguous band of logical tiles: tiles
// that share halo rows (neighbouring p blocks, consecutive i segments) then
// share one L2.  A bijection on [0, G).
__device__ __forceinline__ int xcd_tile(int wg, int G) {
    const int q = G >> 3, r = G & 7, x = wg & 7;
    return x * q + min(x, r) + (wg >> 3);
}

// One thread per (j,k) column of an i-segment; the E values of plane i+1 are
// carried in registers into the next iteration.
__global__ void __launch_bounds__(kBlock) k_update_h(HArgs a, int i0, int i1, int iseg, int npblk) {
    const Grid g = a.g;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int pblk = tile % npblk, seg = tile / npblk;
    const int p = pblk * kBlock + threadIdx.x;
    if (p >= g.S) return;
    const int j = p / (g.nz + 1), k = p - j * (g.nz + 1);
    const int ib = i0 + seg * iseg;
    const int ie = min(i1, ib + iseg);
    const float *__restrict__ Ex = a.E;
    const float *__restrict__ Ey = a.E + g.N;
    const float *__restrict__ Ez = a.E + 2 * g.N;
    float *__restrict__ Hx = a.H;
    float *__restrict__ Hy = a.H + g.N;
    float *__restrict__ Hz = a.H + 2 * g.N;
    const bool jin = j < g.ny, kin = k < g.nz;
    const float hy = jin ? a.hs[1][j] : 0.f, hz = kin ? a.hs[2][k] : 0.f;
    const int sy = a.has_pml && jin ? a.pml.slot[1][j] : -1;
    const int sz = a.has_pml && kin ? a.pml.slot[2][k] : -1;
    const float by = sy >= 0 ? a.pml.b[1][j] : 0.f, cy = sy >= 0 ? a.pml.c[1][j] : 0.f;
    const float bz = sz >= 0 ? a.pml.b[2][k] : 0.f, cz = sz >= 0 ? a.pml.c[2][k] : 0.f;
    long long q = (long long)ib * g.si + p;
    float ey = Ey[q], ez = Ez[q];
    for (int i = ib; i < ie; i++, q += g.si) {
        const bool iin = i < g.nx;
        const float ex = Ex[q];
        const float ey_k = kin ? Ey[q + 1] : 0.f, ex_k = kin ? Ex[q + 1] : 0.f;
        const float ez_j = jin ? Ez[q + g.sj] : 0.f, ex_j = jin ? Ex[q + g.sj] : 0.f;
        const float ey_i = iin ? Ey[q + g.si] : 0.f, ez_i = iin ? Ez[q + g.si] : 0.f;
        const float dz_ey = ey_k - ey, dy_ez = ez_j - ez;
        const float dz_ex = ex_k - ex, dy_ex = ex_j - ex;
        const float dx_ez = ez_i - ez, dx_ey = ey_i - ey;
        const float hx = iin ? a.hs[0][i] : 0.f;
        const int sx = a.has_pml && iin ? a.pml.slot[0][i] : -1;
        // psi of the y and z slabs, loaded for every lane through clamped offsets
        // (offset 0 lies inside every psi array) so the loads are not behind
        // lane-divergent branches; stores stay conditional
        const long long oy = sy >= 0 ? sy * a.pml.per[1] + (long long)i * (g.nz + 1) + k : 0;
        const long long oz = sz >= 0 ? psi_z(g, a.pml, i, j, sz) : 0;
        const float v0y = a.has_pml ? a.psi[0][0][oy] : 0.f, v0z = a.has_pml ? a.psi[0][1][oz] : 0.f;
        const float v1z = a.has_pml ? a.psi[1][0][oz] : 0.f, v2y = a.has_pml ? a.psi[2][1][oy] : 0.f;
        if (jin && kin) {   // Hx: all i
            float h = Hx[q] - (dy_ez * hy - dz_ey * hz);
            if (sy >= 0) {
                const float v = by * v0y + cy * dy_ez;
                a.psi[0][0][oy] = v;
                h -= a.dth * v;
            }
            if (sz >= 0) {
                const float v = bz * v0z + cz * dz_ey;
                a.psi[0][1][oz] = v;
                h += a.dth * v;
            }
            Hx[q] = h;
        }
        if (iin && kin) {   // Hy: j <= ny
            float h = Hy[q] - (dz_ex * hz - dx_ez * hx);
            if (sx >= 0) {
                float *ps = a.psi[1][1] + sx * a.pml.per[0] + p;
                const float v = a.pml.b[0][i] * *ps + a.pml.c[0][i] * dx_ez;
                *ps = v;
                h += a.dth * v;
            }
            if (sz >= 0) {
                const float v = bz * v1z + cz * dz_ex;
                a.psi[1][0][oz] = v;
                h -= a.dth * v;
            }
            Hy[q] = h;
        }
        if (iin && jin) {   // Hz: k <= nz
            float h = Hz[q] - (dx_ey * hx - dy_ex * hy);
            if (sx >= 0) {
                float *ps = a.psi[2][0] + sx * a.pml.per[0] + p;
                const float v = a.pml.b[0][i] * *ps + a.pml.c[0][i] * dx_ey;
                *ps = v;
                h -= a.dth * v;
            }
            if (sy >= 0) {
                const float v = by * v2y + cy * dy_ex;
                a.psi[2][1][oy] = v;
                h += a.dth * v;
            }
            Hz[q] = h;
        }
        ey = ey_i;
        ez = ez_i;
    }
}

// One thread per (j,k) column of an i-segment; the H values of plane i-1 are
// carried in registers from the previous iteration.
template <bool kCmp>
__global__ void __launch_bounds__(kBlock) k_update_e(EArgs a, int i0, int i1, int iseg, int npblk) {
    const Grid g = a.g;
    // (ca, cb) table read through the caches: a handful of pairs per component,
    // so lookups hit L1; staging it in LDS cost a workgroup barrier before the
    // first field load (measured 6 % slower on the 4.6 M-point board)
    const float2 *__restrict__ tab = a.ctab;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int pblk = tile % npblk, seg = tile / npblk;
    const int p = pblk * kBlock + threadIdx.x;
    if (p >= g.S) return;
    const int j = p / (g.nz + 1), k = p - j * (g.nz + 1);
    const int ib = i0 + seg * iseg;
    const int ie = min(i1, ib + iseg);
    float *__restrict__ Ex = a.E;
    float *__restrict__ Ey = a.E + g.N;
    float *__restrict__ Ez = a.E + 2 * g.N;
    const float *__restrict__ Hx = a.H;
    const float *__restrict__ Hy = a.H + g.N;
    const float *__restrict__ Hz = a.H + 2 * g.N;
    const float *__restrict__ ca = a.ca;
    const float *__restrict__ cb = a.cb;
    const unsigned *__restrict__ cidx = a.cidx;
    const bool jb = j >= 1 && j < g.ny, kb = k >= 1 && k < g.nz;   // interior node ranges
    const float ey_s = a.es[1][j], ez_s = a.es[2][k];
    const int sy = a.has_pml ? a.pml.slot[1][j] : -1;
    const int sz = a.has_pml ? a.pml.slot[2][k] : -1;
    const float by = sy >= 0 ? a.pml.b[1][j] : 0.f, cy = sy >= 0 ? a.pml.c[1][j] : 0.f;
    const float bz = sz >= 0 ? a.pml.b[2][k] : 0.f, cz = sz >= 0 ? a.pml.c[2][k] : 0.f;
    long long q = (long long)ib * g.si + p;
    float hy_m = ib >= 1 ? Hy[q - g.si] : 0.f, hz_m = ib >= 1 ? Hz[q - g.si] : 0.f;
    for (int i = ib; i < ie; i++, q += g.si) {
        const bool ib_ = i >= 1 && i < g.nx;
        const float hx = Hx[q], hy = Hy[q], hz = Hz[q];
        const float hx_k = k >= 1 ? Hx[q - 1] : 0.f, hy_k = k >= 1 ? Hy[q - 1] : 0.f;
        const float hx_j = j >= 1 ? Hx[q - g.sj] : 0.f, hz_j = j >= 1 ? Hz[q - g.sj] : 0.f;
        const float ex_s = a.es[0][i];
        const int sx = a.has_pml ? a.pml.slot[0][i] : -1;
        float2 cx, cy2, cz2;   // (ca, cb) of Ex, Ey, Ez at q
        if (kCmp) {
            const unsigned ci = cidx[q];
            cx = tab[ci & 255u];
            cy2 = tab[256 + ((ci >> 8) & 255u)];
            cz2 = tab[512 + ((ci >> 16) & 255u)];
        } else {
            cx = make_float2(i < g.nx ? ca[q] : 0.f, i < g.nx ? cb[q] : 0.f);
            cy2 = make_float2(j < g.ny ? ca[g.N + q] : 0.f, j < g.ny ? cb[g.N + q] : 0.f);
            cz2 = make_float2(k < g.nz ? ca[2 * g.N + q] : 0.f, k < g.nz ? cb[2 * g.N + q] : 0.f);
        }
        // own E of the three components, loaded for every lane ahead of the
        // component branches (q indexes every component array; values of the
        // components a point does not update go unused): 439 -> 397 us on the
        // 38.8 M-point board, where the k < nz branch splits every row's last wave
        const float ex0 = Ex[q], ey0 = Ey[q], ez0 = Ez[q];
        // psi of the y and z slabs, loaded for every lane through clamped offsets
        // (offset 0 lies inside every psi array) so the loads are not behind
        // lane-divergent branches; stores stay conditional
        const long long oy = sy >= 0 ? sy * a.pml.per[1] + (long long)i * (g.nz + 1) + k : 0;
        const long long oz = sz >= 0 ? psi_z(g, a.pml, i, j, sz) : 0;
        const float v0y = a.has_pml ? a.psi[0][0][oy] : 0.f, v0z = a.has_pml ? a.psi[0][1][oz] : 0.f;
        const float v1z = a.has_pml ? a.psi[1][0][oz] : 0.f, v2y = a.has_pml ? a.psi[2][1][oy] : 0.f;
        if (i < g.nx) {     // Ex
            const float e0 = ex0, cbv = cx.y;
            float e = e0;
            if (jb && kb) {
                const float dy_hz = hz - hz_j, dz_hy = hy - hy_k;
                e = cx.x * e0 + cbv * (dy_hz * ey_s - dz_hy * ez_s);
                if (sy >= 0) {
                    const float v = by * v0y + cy * dy_hz;
                    a.psi[0][0][oy] = v;
                    e += cbv * v;
                }
                if (sz >= 0) {
                    const float v = bz * v0z + cz * dz_hy;
                    a.psi[0][1][oz] = v;
                    e -= cbv * v;
                }
            }
            Ex[q] = e;
        }
        if (j < g.ny) {     // Ey
            const float e0 = ey0, cbv = cy2.y;
            float e = e0;
            if (ib_ && kb) {
                const float dz_hx = hx - hx_k, dx_hz = hz - hz_m;
                e = cy2.x * e0 + cbv * (dz_hx * ez_s - dx_hz * ex_s);
                if (sx >= 0) {
                    float *ps = a.psi[1][1] + sx * a.pml.per[0] + p;
                    const float v = a.pml.b[0][i] * *ps + a.pml.c[0][i] * dx_hz;
                    *ps = v;
                    e -= cbv * v;
                }
                if (sz >= 0) {
                    const float v = bz * v1z + cz * dz_hx;
                    a.psi[1][0][oz] = v;
                    e += cbv * v;
                }
            }
            Ey[q] = e;
        }
        if (k < g.nz) {     // Ez
            const float e0 = ez0, cbv = cz2.y;
            float e = e0;
            if (ib_ && jb) {
                const float dx_hy = hy - hy_m, dy_hx = hx - hx_j;
                e = cz2.x * e0 + cbv * (dx_hy * ex_s - dy_hx * ey_s);
                if (sx >= 0) {
                    float *ps = a.psi[2][0] + sx * a.pml.per[0] + p;
                    const float v = a.pml.b[0][i] * *ps + a.pml.c[0][i] * dx_hy;
                    *ps = v;
                    e += cbv * v;
                }
                if (sy >= 0) {
                    const float v = by * v2y + cy * dy_hx;
                    a.psi[2][1][oy] = v;
                    e -= cbv * v;
                }
            }
            Ez[q] = e;
        }
        hy_m = hy;
        hz_m = hz;
    }
}

// surface-impedance overwrite of the E edge of SIBC slot u (sibc.c:256-269):
// J from up to four H faces, then the recursive-convolution terms
__device__ __forceinline__ float sibc_edge(const PostArgs &a, const PostParams &pp, int u) {
    const unsigned q = a.sb_idx[u];
    const int c = (int)(q / a.N);
    const long long qq = (long long)q - c * a.N;
    const int cu = (c + 1) % 3, cv = (c + 2) % 3;
    const long long su = cu == 0 ? a.si : cu == 1 ? a.sj : 1;
    const long long sv = cv == 0 ? a.si : cv == 1 ? a.sj : 1;
    const float *Hu = a.H + cu * a.N, *Hv = a.H + cv * a.N;
    const unsigned m = a.face[u];
    const float inv = 1.0f / (float)__builtin_popcount(m);
    float Js = 0.f;
    if (m & 1u) Js += -inv * Hu[qq];
    if (m & 2u) Js += inv * Hu[qq - sv];
    if (m & 4u) Js += inv * Hv[qq];
    if (m & 8u) Js += -inv * Hv[qq - su];
    float es = 0.f;
    for (int k = 0; k < pp.nterm; k++) {
        float *y = a.Y + (long long)k * a.nsb + u;
        float yv = *y;
        yv += pp.om[k] * (Js - yv);
        *y = yv;
        es += pp.zc[k] * (Js - yv);
    }
    return es;
}

// Probe line and loop integrals of one step (one wavefront per sample, FP64 sums
// reduced with fixed-order shuffles) written to the sample series row.
__device__ void gather_samples(const GatherArgs &a, int row, int wave0, int nwave) {
    const int lane = threadIdx.x & 63;
    for (int w = wave0; w < a.nsamp; w += nwave) {
        const bool is_port = w >= a.nip && w < a.nip + a.nport;
        if (a.mid_only && !is_port) continue;
        double acc = 0.0;
        for (int t = a.toff[w] + lane; t < a.toff[w + 1]; t += 64) {
            acc += (double)a.F[a.tidx[t]] * a.tw[t];
        }
        for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off);
        if (lane != 0) continue;
        const double v = acc * a.post[w];
        if (w < a.nip) {
            a.ser_ip[(long long)row * a.nip + w] = v;
        } else if (is_port) {
            const int pp = w - a.nip;
            if (!a.mid_only) {
                a.ser_vold[(long long)row * a.nport + pp] = a.vlast[pp];
                a.ser_vnew[(long long)row * a.nport + pp] = v;
            }
            a.vlast[pp] = v;
        } else {
            a.ser_vp[(long long)row * a.nvp + (w - a.nip - a.nport)] = v;
        }
    }
}

// Post phase of one step: Debye, SIBC and source edges.
constexpr unsigned kDispSibc = 0x80000000u, kDispSrc = 0x40000000u, kDispMat = 0x3fffffffu;

// position of edge t in an ascending list that holds it
__device__ __forceinline__ int find_edge(const int *list, int n, int t) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (list[mid] < t) lo = mid + 1; else hi = mid;
    }
    return lo;
}

__global__ void __launch_bounds__(kBlock) k_post(PostArgs a) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    const PostParams &pp = *a.pp;
    const long long n = a.ctl->step;
    float vs, gm;
    waveforms(a.wave, n, &vs, &gm);
    if (t < a.ndisp) {                       // dispersive edge
        const unsigned q = a.d_idx[t];
        const unsigned m = a.d_mat[t];
        const float *mat = a.dtab + (m & kDispMat) * a.dstride;
        float e = a.E[q];
        // every pole's J loaded once, together (unrolled, guarded), and kept in
        // registers for the update: 396 -> 342 us on the 19 M-edge board
        float jv[kMaxPoles];
#pragma unroll
        for (int p = 0; p < kMaxPoles; p++) jv[p] = p < pp.npole ? a.J[(long long)p * a.ndisp + t] : 0.f;
        float acc = 0.f;
#pragma unroll
        for (int p = 0; p < kMaxPoles; p++)
            if (p < pp.npole) acc += pp.kc[p] * jv[p];
        e -= mat[0] * acc;
        if (m & kDispSibc) e = sibc_edge(a, pp, a.dsb_u[find_edge(a.dsb_t, a.ndsb, t)]);
        if (m & kDispSrc) {
            const int x = a.dsrc_x[find_edge(a.dsrc_t, a.ndsrc, t)];
            e += a.wl[x] * vs + a.wm[x] * gm;
        }
        const float de = e - a.Eold[t];
#pragma unroll
        for (int p = 0; p < kMaxPoles; p++)
            if (p < pp.npole) a.J[(long long)p * a.ndisp + t] = jv[p] + pp.km1[p] * jv[p] + mat[1 + p] * de;
        a.E[q] = e;
        a.Eold[t] = e;                       // E of this edge before the next E update
    } else if (t - a.ndisp < a.nsbo) {       // surface edge without dispersion
        const int slot = (int)a.sbo[t - a.ndisp];
        float e = sibc_edge(a, pp, slot);
        if (a.sb_src) {
            const int x = a.sb_src[slot];
            if (x >= 0) e += a.wl[x] * vs + a.wm[x] * gm;
        }
        a.E[a.sb_idx[slot]] = e;
    } else if (t - a.ndisp - a.nsbo < a.nxs) {   // source edge only
        const int u = t - a.ndisp - a.nsbo;
        const unsigned q = a.x_idx[u];
        const int x = a.x_src[u];
        a.E[q] += a.wl[x] * vs + a.wm[x] * gm;
    }
}

// The step's samples (one workgroup), then the advance of the device step and
// series-row counters, once every wave has read them.
__global__ void __launch_bounds__(kGatherBlock) k_gather(GatherArgs a) {
    const int row = a.ctl->row;
    gather_samples(a, row, threadIdx.x >> 6, kGatherBlock / 64);
    if (a.mid_only) return;
    __syncthreads();
    if (threadIdx.x == 0) {
        a.ctl->row = row + 1;
        a.ctl->step += 1;
    }
}

__global__ void k_set_ctl(Ctl *c, long long step, int row) {
    c->step = step;
    c->row = row;
}

// E of every dispersive edge, taken as E_old of the next E update
__global__ void __launch_bounds__(kBlock) k_prime_eold(const float *E, const unsigned *d_idx, float *Eold, int n) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t < n) Eold[t] = E[d_idx[t]];
}

__device__ __forceinline__ void dft_acc(double *acc, double v, double pr, double pi) {
    acc[0] += v * pr;
    acc[1] += v * pi;
}

__device__ __forceinline__ void phasor_at(double w, double t, double *pr, double *pi) {
    *pr = cos(w * t);
    *pi = -sin(w * t);
}

// One thread per (accumulator, frequency); accumulators: port V, port I,
// voltage probes, current probes, source spectrum.  Threads with acc_id == nacc
// only advance the phasor pair of their frequency (ping-pong buffers, so no
// thread reads a phasor another block has already advanced).
__global__ void k_dft(DftArgs a) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nf * (a.nacc + 1)) return;
    const int f = t % a.nf, acc_id = t / a.nf;
    int kind, idx;
    if (acc_id < a.nport) { kind = 0; idx = acc_id; }
    else if (acc_id < 2 * a.nport) { kind = 1; idx = acc_id - a.nport; }
    else if (acc_id < 2 * a.nport + a.nvp) { kind = 2; idx = acc_id - 2 * a.nport; }
    else if (acc_id < 2 * a.nport + a.nvp + a.nip) { kind = 3; idx = acc_id - 2 * a.nport - a.nvp; }
    else if (acc_id < a.nacc) { kind = 4; idx = 0; }
    else { kind = 5; idx = 0; }
    const double w = a.omega[f], rr = a.rot[2 * f], ri = a.rot[2 * f + 1];
    double er = a.ph_in[2 * f], ei = a.ph_in[2 * f + 1];
    double hr = a.ph_in[2 * a.nf + 2 * f], hi = a.ph_in[2 * a.nf + 2 * f + 1];
    double acc[2] = {0, 0};
    double *dst = kind < 5 ? a.acc + (long long)acc_id * 2 * a.nf + 2 * f : nullptr;
    if (dst) { acc[0] = dst[0]; acc[1] = dst[1]; }
    for (int r = 0; r < a.rows; r++) {
        const long long n = a.step0 + r;
        if (a.direct || (n & 1023) == 0) {
            phasor_at(w, (double)(n + 1) * a.dt, &er, &ei);
            phasor_at(w, ((double)n + 0.5) * a.dt, &hr, &hi);
        }
        const double tH = ((double)n + 0.5) * a.dt;
        double vs = 0, gm = 0;
        if ((kind == 1 || kind == 4) && a.has_lumped) {
            const double x = (tH - a.t0) / a.tau;
            vs = a.amp * exp(-x * x);
        }
        if (kind == 4 && a.has_mode) {
            const double x = (tH - a.mt0) / a.mtau;
            gm = -2 * x * exp(-x * x);
        }
        switch (kind) {
        case 0: {
            const double V = 0.5 * (a.ser_vnew[(long long)r * a.nport + idx] + a.ser_vold[(long long)r * a.nport + idx]);
            dft_acc(acc, V * a.dt, hr, hi);
            break;
        }
        case 1: {
            const double V = 0.5 * (a.ser_vnew[(long long)r * a.nport + idx] + a.ser_vold[(long long)r * a.nport + idx]);
            const double I = ((idx == a.src_port ? vs : 0.0) - V) / a.R[idx];
            dft_acc(acc, I * a.dt, hr, hi);
            break;
        }
        case 2: dft_acc(acc, a.ser_vp[(long long)r * a.nvp + idx] * a.dt, er, ei); break;
        case 3: dft_acc(acc, a.ser_ip[(long long)r * a.nip + idx] * a.dt, hr, hi); break;
        case 4:
            if (a.has_lumped) dft_acc(acc, vs * a.dt, hr, hi);
            if (a.has_mode) dft_acc(acc, gm * a.dt, hr, hi);
            break;
        default: break;
        }
        double nr = er * rr - ei * ri, ni = er * ri + ei * rr;
        er = nr;
        ei = ni;
        nr = hr * rr - hi * ri;
        ni = hr * ri + hi * rr;
        hr = nr;
        hi = ni;
    }
    if (dst) { dst[0] = acc[0]; dst[1] = acc[1]; }
    if (t == 0) a.ctl->row = 0;   // the series has been consumed
    if (kind == 5) {
        if (a.direct) {   // leave the phasors where the next step expects them (export.c:146-148)
            const long long n = a.step0 + a.rows;
            phasor_at(w, (double)(n + 1) * a.dt, &er, &ei);
            phasor_at(w, ((double)n + 0.5) * a.dt, &hr, &hi);
        }
        a.ph_out[2 * f] = er;
        a.ph_out[2 * f + 1] = ei;
        a.ph_out[2 * a.nf + 2 * f] = hr;
        a.ph_out[2 * a.nf + 2 * f + 1] = hi;
    }
}

struct EnergyArgs {
    Grid g;
    const float *E, *H, *cb;
    const unsigned *cidx;
    const float2 *ctab;
    const double *d[3], *dd[3];
    double dt;
};

// W = sum 1/2 eps E^2 V_dual + 1/2 mu0 H^2 V_primal, eps = dt/cb (sim.c:970-994)
__global__ void __launch_bounds__(kBlock) k_energy(EnergyArgs a, double *partial) {
    __shared__ double red[kBlock];
    const Grid &g = a.g;
    double acc = 0.0;
    const long long stride = (long long)gridDim.x * kBlock;
    for (long long q = (long long)blockIdx.x * kBlock + threadIdx.x; q < g.N; q += stride) {
        const int i = (int)(q / g.si);
        const int rem = (int)(q - (long long)i * g.si);
        const int j = rem / (g.nz + 1), k = rem - j * (g.nz + 1);
        const int ijk[3] = {i, j, k};
        const int n[3] = {g.nx, g.ny, g.nz};
        for (int c = 0; c < 3; c++) {
            const int u = (c + 1) % 3, v = (c + 2) % 3;
            const float cbv = a.cidx ? a.ctab[c * 256 + ((a.cidx[q] >> (8 * c)) & 255u)].y : a.cb[c * g.N + q];
            if (ijk[c] < n[c] && cbv != 0.f) {
                const double vol = a.d[c][ijk[c]] * a.dd[u][ijk[u]] * a.dd[v][ijk[v]];
                const double eps = a.dt / (double)cbv;
                const double e = a.E[c * g.N + q];
                acc += 0.5 * eps * e * e * vol;
            }
            if (ijk[u] < n[u] && ijk[v] < n[v]) {
                const double vol = a.dd[c][ijk[c]] * a.d[u][ijk[u]] * a.d[v][ijk[v]];
                const double h = a.H[c * g.N + q];
                acc += 0.5 * CEM_MU0 * h * h * vol;
            }
        }
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void k_sum(const double *partial, int n, double *out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0;
        for (int i = 0; i < n; i++) s += partial[i];
        *out = s;
    }
}

// plain streaming copy, 16 B per lane per access: the measured-bandwidth reference
// Stream copy (calibration): four independent 16-byte non-temporal loads per thread,
// then the four stores, one pass over the data.
typedef float cf4 __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(kBlock) k_copy(const cf4 *__restrict__ src, cf4 *__restrict__ dst, long long n) {
    const long long base = (long long)blockIdx.x * kBlock * 4 + threadIdx.x;
    cf4 v[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const long long i = base + u * kBlock;
        if (i < n) v[u] = __builtin_nontemporal_load(src + i);
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const long long i = base + u * kBlock;
        if (i < n) __builtin_nontemporal_store(v[u], dst + i);
    }
}

__global__ void k_f32_to_f64(const float *src, double *dst, long long n) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; t < n; t += stride) dst[t] = src[t];
}

__global__ void k_f64_to_f32(const double *src, float *dst, long long n) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; t < n; t += stride) dst[t] = (float)src[t];
}

// ---------------------------------------------------------------------------
// FP64 port post-processing on the device
// ---------------------------------------------------------------------------
struct zd {
    double x, y;
};
__device__ __forceinline__ zd zadd(zd a, zd b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ zd zsub(zd a, zd b) { return {a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ zd zscale(zd a, double s) { return {a.x * s, a.y * s}; }
__device__ __forceinline__ zd zmul(zd a, zd b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
// Smith's quotient: no intermediate overflow for the magnitudes on this path
__device__ __forceinline__ zd zdiv(zd a, zd b) {
    if (fabs(b.x) >= fabs(b.y)) {
        const double r = b.y / b.x, den = b.x + b.y * r;
        return {(a.x + a.y * r) / den, (a.y - a.x * r) / den};
    }
    const double r = b.x / b.y, den = b.x * r + b.y;
    return {(a.x * r + a.y) / den, (a.y * r - a.x) / den};
}
__device__ __forceinline__ zd zexp(zd z) {
    double sn, cs;
    sincos(z.y, &sn, &cs);
    const double e = exp(z.x);
    return {e * cs, e * sn};
}
__device__ __forceinline__ zd zcosh(zd z) {
    double sn, cs;
    sincos(z.y, &sn, &cs);
    return {cosh(z.x) * cs, sinh(z.x) * sn};
}
// principal square root (branch cut on the negative real axis)
__device__ __forceinline__ zd zsqrt(zd z) {
    if (z.x == 0.0 && z.y == 0.0) return {0.0, z.y};
    const double t = sqrt(0.5 * (fabs(z.x) + hypot(z.x, z.y)));
    if (z.x >= 0.0) return {t, z.y / (2.0 * t)};
    return {fabs(z.y) / (2.0 * t), copysign(t, z.y)};
}

struct ModeWaveArgs {
    int nf, dir;
    double dx, dist;               // monitor cell length; signed distance monitor -> reference plane
    const double *V0, *V1, *Ix;    // DFT accumulators of the two monitor V probes and the I probe
    const double *mode;            // per frequency: Re Z0, Im Z0, alpha, beta of the port mode
    double *out;                   // a (2 nf) then b (2 nf)
};

// Wave separation of one mode port (modeport.c:203-227), one thread per
// frequency: V at the monitor plane from the two staggered V probes with the
// discrete-dispersion correction 1/cosh(gamma dx / 2), I from the loop probe,
// a, b = (V +- Z0 I) / 2 sqrt(Z0), each moved to the reference plane by
// exp(-+gamma d).
__global__ void k_mode_waves(ModeWaveArgs a) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= a.nf) return;
    const zd Z = {a.mode[4 * q], a.mode[4 * q + 1]};
    const zd g = {a.mode[4 * q + 2], a.mode[4 * q + 3]};
    const zd c = zcosh({g.x * a.dx / 2.0, g.y * a.dx / 2.0});
    const zd V = zdiv({a.V0[2 * q] + a.V1[2 * q], a.V0[2 * q + 1] + a.V1[2 * q + 1]}, zscale(c, 2.0));
    const zd I = {(double)a.dir * a.Ix[2 * q], (double)a.dir * a.Ix[2 * q + 1]};
    const zd sq = zscale(zsqrt(Z), 2.0);
    const zd ZI = zmul(Z, I);
    const zd A = zmul(zdiv(zadd(V, ZI), sq), zexp({-g.x * a.dist, -g.y * a.dist}));
    const zd B = zmul(zdiv(zsub(V, ZI), sq), zexp({g.x * a.dist, g.y * a.dist}));
    a.out[2 * q] = A.x;
    a.out[2 * q + 1] = A.y;
    a.out[2 * a.nf + 2 * q] = B.x;
    a.out[2 * a.nf + 2 * q + 1] = B.y;
}

// 2x2 renormalisation (modeport.c:242-264) of nf frequencies in place, one
// thread per frequency: Z = sqrt(Zd)(I + S)(I - S)^-1 sqrt(Zd),
// S' = (Z - zref)(Z + zref)^-1 with Zd = diag(z1, z2).
__global__ void k_s2_renormalize(double *S, int nf, const double *z1, const double *z2, double zref) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nf) return;
    double *v = S + 8 * q;
    const zd s00 = {v[0], v[1]}, s01 = {v[2], v[3]}, s10 = {v[4], v[5]}, s11 = {v[6], v[7]};
    const zd one = {1.0, 0.0};
    const zd a00 = zsub(one, s00), a01 = {-s01.x, -s01.y}, a10 = {-s10.x, -s10.y}, a11 = zsub(one, s11);
    const zd det = zsub(zmul(a00, a11), zmul(a01, a10));
    const zd i00 = zdiv(a11, det), i01 = zdiv({-a01.x, -a01.y}, det);
    const zd i10 = zdiv({-a10.x, -a10.y}, det), i11 = zdiv(a00, det);
    const zd b00 = zadd(one, s00), b01 = s01, b10 = s10, b11 = zadd(one, s11);
    const zd n00 = zadd(zmul(b00, i00), zmul(b01, i10)), n01 = zadd(zmul(b00, i01), zmul(b01, i11));
    const zd n10 = zadd(zmul(b10, i00), zmul(b11, i10)), n11 = zadd(zmul(b10, i01), zmul(b11, i11));
    const double r1 = sqrt(z1[q]), r2 = sqrt(z2[q]);
    const zd z00 = zscale(zscale(n00, r1), r1), z01 = zscale(zscale(n01, r1), r2);
    const zd z10 = zscale(zscale(n10, r2), r1), z11 = zscale(zscale(n11, r2), r2);
    const zd p00 = {z00.x + zref, z00.y}, p11 = {z11.x + zref, z11.y};
    const zd m00 = {z00.x - zref, z00.y}, m11 = {z11.x - zref, z11.y};
    const zd dp = zsub(zmul(p00, p11), zmul(z01, z10));
    const zd q00 = zdiv(p11, dp), q01 = zdiv({-z01.x, -z01.y}, dp);
    const zd q10 = zdiv({-z10.x, -z10.y}, dp), q11 = zdiv(p00, dp);
    const zd r00 = zadd(zmul(m00, q00), zmul(z01, q10)), r01 = zadd(zmul(m00, q01), zmul(z01, q11));
    const zd r10 = zadd(zmul(z10, q00), zmul(m11, q10)), r11 = zadd(zmul(z10, q01), zmul(m11, q11));
    v[0] = r00.x; v[1] = r00.y;
    v[2] = r01.x; v[3] = r01.y;
    v[4] = r10.x; v[5] = r10.y;
    v[6] = r11.x; v[7] = r11.y;
}

// ---------------------------------------------------------------------------
// device state of one simulation
// ---------------------------------------------------------------------------
template <class T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    hipStream_t owner = nullptr;   // stream the block was allocated on and is freed on
    int alloc(size_t count, hipStream_t st) {
        n = count;
        owner = st;
        if (!count) return 0;
        if (dev_alloc((void **)&p, count * sizeof(T), st) != hipSuccess) {
            p = nullptr;
            set_error("device out of memory");
            return -1;
        }
        return 0;
    }
    int upload(const T *src, size_t count, hipStream_t st) {   // synchronous: src may be a temporary
        if (!count) return 0;
        if (hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) {
            set_error("device upload failed");
            return -1;
        }
        return 0;
    }
    int zero(hipStream_t st) {
        if (!n) return 0;
        return hipMemsetAsync(p, 0, n * sizeof(T), st) == hipSuccess ? 0 : -1;
    }
    void release() {
        dev_free(p, owner);
        p = nullptr;
        n = 0;
    }
};

// Per-kernel GPU time from event pairs recorded around each launch on the
// engine's stream; pairs are only read back (and the host synchronised) when
// the pool is full or the statistics are requested.
struct KernelTimer {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pool;
    size_t used = 0;
    double total_ms = 0;
    long count = 0;
    void fold() {
        if (!used) return;
        HIP_CHECK_VOID(hipEventSynchronize(pool[used - 1].second));
        for (size_t k = 0; k < used; k++) {
            float ms = 0;
            HIP_CHECK_VOID(hipEventElapsedTime(&ms, pool[k].first, pool[k].second));
            total_ms += ms;
        }
        count += (long)used;
        used = 0;
    }
    void destroy() {
        for (auto &e : pool) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
        pool.clear();
        used = 0;
    }
};

struct Device {
    hipStream_t st = nullptr;
    Grid g{};
    DBuf<float> F;               // E (3N) then H (3N)
    DBuf<float> ca, cb;          // 3N each, only when more than 256 distinct pairs
    DBuf<unsigned> cidx;         // N packed coefficient indices
    DBuf<float2> ctab;           // 3 x 256 (ca, cb)
    bool compressed = false;
    DBuf<float> axis1d;          // hs[3], es[3], pml b/c for H and E per axis
    DBuf<int> slots;             // CPML slot tables
    DBuf<float> psi;             // all psi arrays
    float *psiE[3][2] = {}, *psiH[3][2] = {};
    size_t psiE_len[3][2] = {}, psiH_len[3][2] = {};
    const float *hs[3] = {}, *es[3] = {};
    Pml pmlH{}, pmlE{};
    // post-phase edge lists
    DBuf<unsigned> d_idx, sb_idx, sbo, x_idx;
    DBuf<int> sb_src, x_src;
    DBuf<int> dsb_t, dsb_u, dsrc_t, dsrc_x;   // dispersive edges with SIBC / source slots
    std::vector<unsigned> d_mat_host;          // material rows of the dispersive edges, no flags
    std::vector<int> dsb_t_host;               // ascending dispersive edges with an SIBC slot
    int nsbo = 0;
    bool has_sbsrc = false;
    DBuf<float> J, dtab, Eold, Y, wl, wm;
    DBuf<unsigned> d_mat;
    int dstride = 1;
    DBuf<unsigned char> face;
    DBuf<PostParams> pp;
    DBuf<Ctl> ctl;
    long long ctl_step = -1;     // what ctl holds once the queued work has run
    int ctl_row = -1;
    int nxs = 0;
    int ndsb = 0, ndsrc = 0;
    bool eold_stale = false;
    bool sources_dirty = true;
    // probes
    DBuf<long long> tidx;
    DBuf<double> tw, post;
    DBuf<int> toff;
    int nsamp = 0;
    DBuf<double> vlast;
    bool vlast_stale = false;
    // series and DFT
    DBuf<double> ser;            // ip | vold | vnew | vp, kSeriesCap rows each
    double *ser_ip = nullptr, *ser_vold = nullptr, *ser_vnew = nullptr, *ser_vp = nullptr;
    int rows = 0;
    long long row0 = 0;
    DBuf<double> omega, rot, acc, R;
    DBuf<double> ph[2];          // ping-pong phasors: phE (2nf) | phH (2nf)
    int ph_cur = 0;
    int nacc = 0;
    // energy
    DBuf<double> d1d, partial;
    int energy_blocks = 0;
    // launch shape: x planes per thread in the stencil kernels, set by the autotuner
    // on the first steps unless fixed by COURANT_ISEG_H / COURANT_ISEG_E
    int iseg_h = kIseg, iseg_e = kIseg;
    bool tuned = false;
    // timing
    bool profiling = false;
    KernelTimer tH, tE, tG, tD, tP;
    // captured chunk of kGraphSteps steps, valid while its launch arguments are unchanged
    bool use_graph = true;
    hipGraphExec_t gexec = nullptr;
    std::vector<unsigned char> gkey;
};

}  // namespace cem

// ---------------------------------------------------------------------------
// plan upload
// ---------------------------------------------------------------------------
namespace {

// Slot of global E index g in the dispersive / surface edge lists (both sorted by
// global index, the reference's (comp, i, j, k) order), or -1.
long dispersive_slot(const cem_sim *s, unsigned long long g) {
    size_t lo = 0, hi = s->ndisp;
    while (lo < hi) {
        const size_t mid = (lo + hi) / 2;
        if ((unsigned long long)s->d_comp[mid] * s->N + s->d_idx[mid] < g) lo = mid + 1; else hi = mid;
    }
    return lo < (size_t)s->ndisp && (unsigned long long)s->d_comp[lo] * s->N + s->d_idx[lo] == g ? (long)lo : -1;
}

long surface_slot(const cem_sim *s, unsigned long long g) {
    size_t lo = 0, hi = s->nsb;
    while (lo < hi) {
        const size_t mid = (lo + hi) / 2;
        if ((unsigned long long)s->sb_comp[mid] * s->N + s->sb_idx[mid] < g) lo = mid + 1; else hi = mid;
    }
    return lo < (size_t)s->nsb && (unsigned long long)s->sb_comp[lo] * s->N + s->sb_idx[lo] == g ? (long)lo : -1;
}

// Excited sources as (global E index, lumped cs, modal weight) slots; edges that
// are also dispersive are flagged in d_mat and listed in dsrc_t / dsrc_x, the
// others go to x_idx.
int upload_sources(cem_sim *s, Device *D) {
    std::vector<std::pair<unsigned long long, std::pair<double, double>>> src;
    if (s->src_port >= 0) {
        const LumpedPort &pt = s->port[s->src_port];
        for (size_t e = 0; e < pt.edge.size(); e++)
            src.push_back({(unsigned long long)pt.axis * s->N + pt.edge[e], {pt.cs[e], 0.0}});
    }
    if (s->msrc >= 0) {
        const ModePort &mp = s->mport[s->msrc];
        for (size_t e = 0; e < mp.src_idx.size(); e++)
            src.push_back({(unsigned long long)mp.src_comp[e] * s->N + mp.src_idx[e], {0.0, mp.src_w[e]}});
    }
    std::stable_sort(src.begin(), src.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
    std::vector<unsigned long long> where;
    std::vector<float> wl, wm;
    for (auto &e : src) {
        if (!where.empty() && where.back() == e.first) {
            wl.back() += (float)e.second.first;
            wm.back() += (float)e.second.second;
            continue;
        }
        where.push_back(e.first);
        wl.push_back((float)e.second.first);
        wm.push_back((float)e.second.second);
    }
    std::vector<int> dsrc_t, dsrc_x, sb_src(s->nsb, -1), x_src;
    std::vector<unsigned> x_idx;
    D->has_sbsrc = false;
    for (size_t k = 0; k < where.size(); k++) {   // ascending global index, so t ascends too
        const unsigned long long g = where[k];
        const long t = dispersive_slot(s, g);
        const long u = t < 0 ? surface_slot(s, g) : -1;
        if (t >= 0) {
            dsrc_t.push_back((int)t);
            dsrc_x.push_back((int)k);
        } else if (u >= 0) {
            sb_src[u] = (int)k;
            D->has_sbsrc = true;
        } else {
            x_idx.push_back((unsigned)g);
            x_src.push_back((int)k);
        }
    }
    D->nxs = (int)x_idx.size();
    for (auto *b : {&D->wl, &D->wm}) b->release();
    D->x_idx.release();
    D->x_src.release();
    D->dsrc_t.release();
    D->dsrc_x.release();
    D->sb_src.release();
    if (D->wl.alloc(wl.size(), D->st) || D->wm.alloc(wm.size(), D->st) || D->x_idx.alloc(x_idx.size(), D->st) ||
        D->x_src.alloc(x_src.size(), D->st))
        return -1;
    if (D->wl.upload(wl.data(), wl.size(), D->st) || D->wm.upload(wm.data(), wm.size(), D->st) ||
        D->x_idx.upload(x_idx.data(), x_idx.size(), D->st) || D->x_src.upload(x_src.data(), x_src.size(), D->st))
        return -1;
    D->ndsrc = (int)dsrc_t.size();
    if (D->dsrc_t.alloc(dsrc_t.size(), D->st) || D->dsrc_x.alloc(dsrc_x.size(), D->st) ||
        D->dsrc_t.upload(dsrc_t.data(), dsrc_t.size(), D->st) || D->dsrc_x.upload(dsrc_x.data(), dsrc_x.size(), D->st))
        return -1;
    // material rows of the dispersive edges with the SIBC / source flag bits
    if (s->ndisp) {
        std::vector<unsigned> m(D->d_mat_host);
        for (int t : D->dsb_t_host) m[t] |= kDispSibc;
        for (int t : dsrc_t) m[t] |= kDispSrc;
        if (D->d_mat.upload(m.data(), m.size(), D->st)) return -1;
    }
    if (D->has_sbsrc && (D->sb_src.alloc(sb_src.size(), D->st) || D->sb_src.upload(sb_src.data(), sb_src.size(), D->st)))
        return -1;
    D->sources_dirty = false;
    return 0;
}

int upload_plan(cem_sim *s, Device *D) {
    const int nx = s->n[0], ny = s->n[1], nz = s->n[2];
    Grid &g = D->g;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.S = (ny + 1) * (nz + 1);
    g.si = (long long)s->si;
    g.sj = (long long)s->sj;
    g.N = (long long)s->N;
    g.nchunk = (g.S + 63) / 64;
    const size_t N = s->N;
    hipStream_t st = D->st;

    // fields (zero) and coefficients
    if (D->F.alloc(6 * N, D->st) || D->F.zero(st)) return -1;
    // Coefficient compression: the (ca, cb) pairs of each component take few
    // distinct values (one per material mix), so each point stores three 8-bit
    // indices into a 256-entry table; more than 256 pairs falls back to arrays.
    {
        std::vector<unsigned> idx(N, 0);
        std::vector<float2> tab(3 * 256, make_float2(0.f, 0.f));
        bool ok = std::getenv("COURANT_COEF_ARRAYS") == nullptr;   // switch to force the array path
        for (int c = 0; c < 3 && ok; c++) {
            std::vector<std::pair<unsigned long long, unsigned>> seen;   // (ca bits << 32 | cb bits) -> slot
            unsigned long long last_key = ~0ull;
            unsigned last_slot = 0;
            for (size_t q = 0; q < N && ok; q++) {
                const float fa = (float)s->ca[c][q], fb = (float)s->cb[c][q];
                unsigned ua, ub;
                std::memcpy(&ua, &fa, 4);
                std::memcpy(&ub, &fb, 4);
                const unsigned long long key = ((unsigned long long)ua << 32) | ub;
                unsigned slot = last_slot;
                if (key != last_key) {
                    auto it = std::find_if(seen.begin(), seen.end(), [&](const auto &e) { return e.first == key; });
                    if (it == seen.end()) {
                        if (seen.size() == 256) { ok = false; break; }
                        slot = (unsigned)seen.size();
                        seen.push_back({key, slot});
                        tab[c * 256 + slot] = make_float2(fa, fb);
                    } else {
                        slot = it->second;
                    }
                    last_key = key;
                    last_slot = slot;
                }
                idx[q] |= slot << (8 * c);
            }
        }
        D->compressed = ok;
        if (ok) {
            if (D->cidx.alloc(N, D->st) || D->cidx.upload(idx.data(), N, st)) return -1;
            if (D->ctab.alloc(3 * 256, D->st) || D->ctab.upload(tab.data(), 3 * 256, st)) return -1;
        }
    }
    if (!D->compressed && (D->ca.alloc(3 * N, D->st) || D->cb.alloc(3 * N, D->st))) return -1;
    if (!D->compressed) {
        std::vector<float> tmp(N);
        for (int c = 0; c < 3; c++) {
            for (size_t q = 0; q < N; q++) tmp[q] = (float)s->ca[c][q];
            if (hipMemcpyAsync(D->ca.p + c * N, tmp.data(), N * sizeof(float), hipMemcpyHostToDevice, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess) return -1;
            for (size_t q = 0; q < N; q++) tmp[q] = (float)s->cb[c][q];
            if (hipMemcpyAsync(D->cb.p + c * N, tmp.data(), N * sizeof(float), hipMemcpyHostToDevice, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess) return -1;
        }
    }

    // 1D axis arrays: hs (n), es (n+1), bH, cH (n), bE, cE (n+1) per axis
    const double dth = s->dt / CEM_MU0;
    std::vector<float> ax;
    size_t off_hs[3], off_es[3], off_bh[3], off_ch[3], off_be[3], off_ce[3];
    for (int a = 0; a < 3; a++) {
        const int n = s->n[a];
        off_hs[a] = ax.size();
        for (int i = 0; i < n; i++) ax.push_back((float)(dth * s->idh[a][i]));
        off_es[a] = ax.size();
        for (int i = 0; i <= n; i++) ax.push_back((float)s->ide[a][i]);
        off_bh[a] = ax.size();
        for (int i = 0; i < n; i++) ax.push_back((float)s->bh[a][i]);
        off_ch[a] = ax.size();
        for (int i = 0; i < n; i++) ax.push_back((float)(s->ch[a][i] * (1.0 / s->d[a][i])));
        off_be[a] = ax.size();
        for (int i = 0; i <= n; i++) ax.push_back((float)s->be[a][i]);
        off_ce[a] = ax.size();
        for (int i = 0; i <= n; i++) ax.push_back((float)(s->dd[a][i] > 0 ? s->ce[a][i] * (1.0 / s->dd[a][i]) : 0.0));
    }
    if (D->axis1d.alloc(ax.size(), D->st) || D->axis1d.upload(ax.data(), ax.size(), st)) return -1;
    for (int a = 0; a < 3; a++) {
        D->hs[a] = D->axis1d.p + off_hs[a];
        D->es[a] = D->axis1d.p + off_es[a];
        D->pmlH.b[a] = D->axis1d.p + off_bh[a];
        D->pmlH.c[a] = D->axis1d.p + off_ch[a];
        D->pmlE.b[a] = D->axis1d.p + off_be[a];
        D->pmlE.c[a] = D->axis1d.p + off_ce[a];
    }

    // CPML slot tables and psi storage (compact layout of sim.c:655-678)
    std::vector<int> slots;
    size_t off_sh[3], off_se[3];
    for (int a = 0; a < 3; a++) {
        for (int is_e = 0; is_e < 2; is_e++) {
            const int len = s->n[a] + (is_e ? 1 : 0);
            (is_e ? off_se : off_sh)[a] = slots.size();
            std::vector<int> sl(len + 1, -1);
            if (s->has_pml) {
                int r[2][2];
                pml_ranges(s, a, is_e, r);
                size_t lo = r[0][1] >= r[0][0] ? (size_t)(r[0][1] - r[0][0] + 1) : 0;
                for (int sd = 0; sd < 2; sd++)
                    for (int x = r[sd][0]; x <= r[sd][1]; x++)
                        sl[x] = (int)(sd == 0 ? x - r[0][0] : lo + (x - r[1][0]));
            }
            slots.insert(slots.end(), sl.begin(), sl.end());
        }
        D->pmlH.per[a] = D->pmlE.per[a] = (long long)pml_per(s, a);
    }
    {
        const size_t per = pml_per(s, 2);
        D->pmlH.zpl = per ? (int)(psi_len(s, 0, 0, 1) / per) : 0;
        D->pmlE.zpl = per ? (int)(psi_len(s, 1, 0, 1) / per) : 0;
    }
    if (D->slots.alloc(slots.size(), D->st) || D->slots.upload(slots.data(), slots.size(), st)) return -1;
    for (int a = 0; a < 3; a++) {
        D->pmlH.slot[a] = D->slots.p + off_sh[a];
        D->pmlE.slot[a] = D->slots.p + off_se[a];
    }
    size_t psi_total = 0;
    for (int c = 0; c < 3; c++)
        for (int w = 0; w < 2; w++) {
            D->psiE_len[c][w] = psi_len(s, 1, c, w);
            D->psiH_len[c][w] = psi_len(s, 0, c, w);
            psi_total += D->psiE_len[c][w] + D->psiH_len[c][w];
        }
    if (D->psi.alloc(psi_total + 1, D->st) || D->psi.zero(st)) return -1;
    {
        size_t o = 0;
        for (int c = 0; c < 3; c++)
            for (int w = 0; w < 2; w++) {
                D->psiE[c][w] = D->psi.p + o;
                o += D->psiE_len[c][w];
                D->psiH[c][w] = D->psi.p + o;
                o += D->psiH_len[c][w];
            }
    }

    // Debye ADE tables
    {
        if (3 * N > 0xffffffffull) {
            set_error("grid too large for 32-bit edge indices");
            return -1;
        }
        // Dispersive edges share a handful of materials: (cb, beta_p / dt) rows of
        // a table, found by exact FP32 value, replace 4 + 4P bytes per edge.
        const int np = s->npole;
        std::vector<unsigned> idx(s->ndisp), mat(s->ndisp);
        std::vector<float> tab;
        std::map<std::vector<float>, unsigned> rows;
        std::vector<float> key(1 + np);
        for (int e = 0; e < s->ndisp; e++) {
            idx[e] = (unsigned)(s->d_comp[e] * N + s->d_idx[e]);
            key[0] = (float)s->d_cb[e];
            for (int p = 0; p < np; p++) key[1 + p] = (float)(s->d_beta[(size_t)p * s->ndisp + e] * (1.0 / s->dt));
            auto it = rows.find(key);
            if (it == rows.end()) {
                it = rows.emplace(key, (unsigned)rows.size()).first;
                tab.insert(tab.end(), key.begin(), key.end());
            }
            mat[e] = it->second;
        }
        D->dstride = 1 + np;
        if (D->d_idx.alloc(s->ndisp, D->st) || D->d_idx.upload(idx.data(), s->ndisp, st)) return -1;
        if (rows.size() > kDispMat) {
            set_error("too many Debye materials");
            return -1;
        }
        D->d_mat_host = mat;   // uploaded with its flag bits by upload_sources
        if (D->d_mat.alloc(s->ndisp, D->st)) return -1;
        if (D->dtab.alloc(tab.size() + 1, D->st) || D->dtab.upload(tab.data(), tab.size(), st)) return -1;
        if (D->Eold.alloc(s->ndisp, D->st) || D->Eold.zero(st)) return -1;
        const size_t nj = (size_t)np * s->ndisp;
        if (D->J.alloc(nj, D->st) || D->J.zero(st)) return -1;
        PostParams pp{};
        pp.ndisp = s->ndisp;
        pp.npole = s->npole;
        pp.nsb = s->nsb;
        pp.nterm = s->sibc_npole;
        for (int p = 0; p < s->npole; p++) {
            const double tau = 1.0 / (2 * M_PI * s->pole_f[p]);
            const double x = s->dt / (2 * tau);
            pp.kc[p] = (float)(1.0 / (1.0 + x));          // (1 + k)/2
            pp.km1[p] = (float)(-2.0 * x / (1.0 + x));     // k - 1
        }
        for (int t = 0; t < s->sibc_npole; t++) {
            pp.zc[t] = (float)s->sibc_c[t];
            pp.om[t] = (float)(-std::expm1(-s->sibc_a[t] * s->dt));
        }
        if (D->pp.alloc(1, D->st) || D->pp.upload(&pp, 1, st)) return -1;
        if (D->ctl.alloc(1, D->st) || D->ctl.zero(st)) return -1;
        D->ctl_step = 0;
        D->ctl_row = 0;
    }
    // SIBC tables
    {
        std::vector<unsigned> idx(s->nsb), sbo;
        std::vector<int> dsb_u;
        D->dsb_t_host.clear();
        for (int e = 0; e < s->nsb; e++) {   // ascending global index, so t ascends too
            idx[e] = (unsigned)(s->sb_comp[e] * N + s->sb_idx[e]);
            const long t = dispersive_slot(s, idx[e]);
            if (t >= 0) {
                D->dsb_t_host.push_back((int)t);
                dsb_u.push_back(e);
            } else {
                sbo.push_back((unsigned)e);
            }
        }
        D->nsbo = (int)sbo.size();
        if (D->sb_idx.alloc(s->nsb, D->st) || D->sb_idx.upload(idx.data(), s->nsb, st)) return -1;
        if (D->sbo.alloc(sbo.size(), D->st) || D->sbo.upload(sbo.data(), sbo.size(), st)) return -1;
        D->ndsb = (int)dsb_u.size();
        if (D->dsb_t.alloc(dsb_u.size(), D->st) || D->dsb_u.alloc(dsb_u.size(), D->st) ||
            D->dsb_t.upload(D->dsb_t_host.data(), dsb_u.size(), st) || D->dsb_u.upload(dsb_u.data(), dsb_u.size(), st))
            return -1;
        const size_t ny_ = (size_t)s->nsb * s->sibc_npole;
        if (D->Y.alloc(ny_, D->st) || D->Y.zero(st) || D->face.alloc(s->nsb, D->st)) return -1;
        if (D->face.upload(s->sb_mask.data(), s->nsb, st)) return -1;
    }
    if (upload_sources(s, D)) return -1;

    // probes: samples ip | ports | vp
    {
        std::vector<long long> tidx;
        std::vector<double> tw, post;
        std::vector<int> toff;
        for (auto &p : s->ip) {
            toff.push_back((int)tidx.size());
            for (size_t e = 0; e < p.idx.size(); e++) {
                tidx.push_back((long long)p.comp[e] * (long long)N + (long long)p.idx[e]);
                tw.push_back(p.w[e]);
            }
            post.push_back(1.0);
        }
        for (auto &p : s->port) {
            toff.push_back((int)tidx.size());
            for (size_t e = 0; e < p.edge.size(); e++) {
                tidx.push_back((long long)p.axis * (long long)N + (long long)p.edge[e]);
                tw.push_back(p.dl[e]);
            }
            post.push_back((double)p.dir);
        }
        for (auto &p : s->vp) {
            toff.push_back((int)tidx.size());
            for (size_t e = 0; e < p.edge.size(); e++) {
                tidx.push_back((long long)p.axis * (long long)N + (long long)p.edge[e]);
                tw.push_back(p.dl[e]);
            }
            post.push_back(1.0);
        }
        toff.push_back((int)tidx.size());
        D->nsamp = (int)post.size();
        if (D->tidx.alloc(tidx.size() + 1, D->st) || D->tw.alloc(tw.size() + 1, D->st) || D->toff.alloc(toff.size(), D->st) ||
            D->post.alloc(post.size() + 1, D->st))
            return -1;
        if (D->tidx.upload(tidx.data(), tidx.size(), st) || D->tw.upload(tw.data(), tw.size(), st) ||
            D->toff.upload(toff.data(), toff.size(), st) || D->post.upload(post.data(), post.size(), st))
            return -1;
        if (D->vlast.alloc(s->port.size() + 1, D->st) || D->vlast.zero(st)) return -1;
    }
    // series and DFT state
    {
        const size_t nip = s->ip.size(), np = s->port.size(), nvp = s->vp.size();
        const size_t per = nip + 2 * np + nvp;
        if (D->ser.alloc(per * kSeriesCap + 1, D->st)) return -1;
        double *b = D->ser.p;
        D->ser_ip = b;
        b += nip * kSeriesCap;
        D->ser_vold = b;
        b += np * kSeriesCap;
        D->ser_vnew = b;
        b += np * kSeriesCap;
        D->ser_vp = b;
        const int nf = s->nf;
        D->nacc = (int)(2 * np + nvp + nip + 1);
        std::vector<double> om(nf), R(np + 1, 1.0);
        for (int f = 0; f < nf; f++) om[f] = 2 * M_PI * s->freq[f];
        for (size_t p = 0; p < np; p++) R[p] = s->port[p].R;
        if (D->omega.alloc(nf, D->st) || D->omega.upload(om.data(), nf, st)) return -1;
        if (D->rot.alloc(2 * nf, D->st) || D->rot.upload(s->rotE.data(), 2 * nf, st)) return -1;
        for (int b2 = 0; b2 < 2; b2++)
            if (D->ph[b2].alloc(4 * nf, D->st) || D->ph[b2].zero(st)) return -1;
        D->ph_cur = 0;
        if (D->acc.alloc((size_t)D->nacc * 2 * nf, D->st) || D->acc.zero(st)) return -1;
        if (D->R.alloc(np + 1, D->st) || D->R.upload(R.data(), np + 1, st)) return -1;
        D->rows = 0;
        D->row0 = 0;
    }
    // energy
    {
        std::vector<double> d1;
        for (int a = 0; a < 3; a++) d1.insert(d1.end(), s->d[a].begin(), s->d[a].end());
        for (int a = 0; a < 3; a++) d1.insert(d1.end(), s->dd[a].begin(), s->dd[a].end());
        if (D->d1d.alloc(d1.size(), D->st) || D->d1d.upload(d1.data(), d1.size(), st)) return -1;
        D->energy_blocks = 1024;
        if (D->partial.alloc(D->energy_blocks + 1, D->st)) return -1;
    }
    HIP_OK(hipStreamSynchronize(st));
    return 0;
}

HArgs h_args(cem_sim *s, Device *D) {
    HArgs a;
    std::memset(&a, 0, sizeof a);   // padding too: the bytes key the captured graph
    a.g = D->g;
    a.E = D->F.p;
    a.H = D->F.p + 3 * D->g.N;
    for (int i = 0; i < 3; i++) a.hs[i] = D->hs[i];
    a.pml = D->pmlH;
    for (int c = 0; c < 3; c++)
        for (int w = 0; w < 2; w++) a.psi[c][w] = D->psiH[c][w];
    a.dth = (float)(s->dt / CEM_MU0);
    a.has_pml = s->has_pml;
    return a;
}

EArgs e_args(cem_sim *s, Device *D) {
    EArgs a;
    std::memset(&a, 0, sizeof a);
    a.g = D->g;
    a.E = D->F.p;
    a.H = D->F.p + 3 * D->g.N;
    a.ca = D->ca.p;
    a.cb = D->cb.p;
    a.cidx = D->compressed ? D->cidx.p : nullptr;
    a.ctab = D->compressed ? D->ctab.p : nullptr;
    for (int i = 0; i < 3; i++) a.es[i] = D->es[i];
    a.pml = D->pmlE;
    for (int c = 0; c < 3; c++)
        for (int w = 0; w < 2; w++) a.psi[c][w] = D->psiE[c][w];
    a.has_pml = s->has_pml;
    return a;
}

// Kernel launch on the engine stream; when profiling, the launch carries a start /
// stop event pair that the runtime stamps with the dispatch's own begin and end.
template <typename F, typename... A>
void launch(Device *D, KernelTimer &t, F kernel, dim3 grid, dim3 block, A... args) {
    if (!D->profiling) {
        hipLaunchKernelGGL(kernel, grid, block, 0, D->st, args...);
        return;
    }
    if (t.used == 4096) t.fold();
    if (t.used == t.pool.size()) {
        hipEvent_t a, b;
        HIP_CHECK_VOID(hipEventCreate(&a));
        HIP_CHECK_VOID(hipEventCreate(&b));
        t.pool.push_back({a, b});
    }
    hipExtLaunchKernelGGL(kernel, grid, block, 0, D->st, t.pool[t.used].first, t.pool[t.used].second, 0, args...);
    t.used++;
}

void launch_h(cem_sim *s, Device *D, int i0, int i1) {
    if (i1 <= i0) return;
    HArgs a = h_args(s, D);
    const int npblk = (D->g.S + kBlock - 1) / kBlock, nseg = (i1 - i0 + D->iseg_h - 1) / D->iseg_h;
    launch(D, D->tH, k_update_h, dim3(npblk * nseg), dim3(kBlock), a, i0, i1, D->iseg_h, npblk);
}

void launch_e(cem_sim *s, Device *D, int i0, int i1) {
    if (i1 <= i0) return;
    EArgs a = e_args(s, D);
    const int npblk = (D->g.S + kBlock - 1) / kBlock, nseg = (i1 - i0 + D->iseg_e - 1) / D->iseg_e;
    if (D->compressed)
        launch(D, D->tE, k_update_e<true>, dim3(npblk * nseg), dim3(kBlock), a, i0, i1, D->iseg_e, npblk);
    else
        launch(D, D->tE, k_update_e<false>, dim3(npblk * nseg), dim3(kBlock), a, i0, i1, D->iseg_e, npblk);
}

void launch_dft(cem_sim *s, Device *D, int direct);

PostArgs post_args(cem_sim *s, Device *D) {
    PostArgs a;
    std::memset(&a, 0, sizeof a);
    a.E = D->F.p;
    a.H = D->F.p + 3 * D->g.N;
    a.N = D->g.N;
    a.si = D->g.si;
    a.sj = D->g.sj;
    a.pp = D->pp.p;
    a.ndisp = s->ndisp;
    a.nsb = s->nsb;
    a.nsbo = D->nsbo;
    a.nxs = D->nxs;
    a.d_idx = D->d_idx.p;
    a.d_mat = D->d_mat.p;
    a.dtab = D->dtab.p;
    a.dstride = D->dstride;
    a.dsb_t = D->dsb_t.p;
    a.dsb_u = D->dsb_u.p;
    a.ndsb = D->ndsb;
    a.dsrc_t = D->dsrc_t.p;
    a.dsrc_x = D->dsrc_x.p;
    a.ndsrc = D->ndsrc;
    a.sbo = D->sbo.p;
    a.sb_src = D->has_sbsrc ? D->sb_src.p : nullptr;
    a.J = D->J.p;
    a.Eold = D->Eold.p;
    a.sb_idx = D->sb_idx.p;
    a.face = D->face.p;
    a.Y = D->Y.p;
    a.x_idx = D->x_idx.p;
    a.x_src = D->x_src.p;
    a.wl = D->wl.p;
    a.wm = D->wm.p;
    a.wave.has_lumped = s->src_port >= 0;
    a.wave.has_mode = s->msrc >= 0;
    a.wave.amp = s->src_amp;
    a.wave.t0 = s->src_t0;
    a.wave.tau = s->src_tau;
    a.wave.mt0 = s->msrc_t0;
    a.wave.mtau = s->msrc_tau;
    a.wave.dt = s->dt;
    a.ctl = D->ctl.p;
    return a;
}

GatherArgs gather_args(cem_sim *s, Device *D, int mid_only) {
    GatherArgs a;
    std::memset(&a, 0, sizeof a);
    a.F = D->F.p;
    a.tidx = D->tidx.p;
    a.tw = D->tw.p;
    a.toff = D->toff.p;
    a.post = D->post.p;
    a.nsamp = D->nsamp;
    a.nip = (int)s->ip.size();
    a.nport = (int)s->port.size();
    a.nvp = (int)s->vp.size();
    a.vlast = D->vlast.p;
    a.ser_ip = D->ser_ip;
    a.ser_vold = D->ser_vold;
    a.ser_vnew = D->ser_vnew;
    a.ser_vp = D->ser_vp;
    a.mid_only = mid_only;
    a.ctl = D->ctl.p;
    return a;
}

int post_blocks(cem_sim *s, Device *D) {
    return (s->ndisp + D->nsbo + D->nxs + kBlock - 1) / kBlock;
}

// The device counters follow the host's step and row; after a host-side change
// of either (resume, sample push, set-up) they are written before the next step.
void sync_ctl(cem_sim *s, Device *D) {
    if (D->ctl_step == s->step && D->ctl_row == D->rows) return;
    hipLaunchKernelGGL(k_set_ctl, dim3(1), dim3(1), 0, D->st, D->ctl.p, (long long)s->step, D->rows);
    D->ctl_step = s->step;
    D->ctl_row = D->rows;
}

// Post phase and probe samples of one step, then the host's step / row
// bookkeeping and the DFT flush when the series is full.
void launch_post(cem_sim *s, Device *D) {
    sync_ctl(s, D);
    if (const int nb = post_blocks(s, D)) {
        PostArgs a = post_args(s, D);
        launch(D, D->tP, k_post, dim3(nb), dim3(kBlock), a);
    }
    GatherArgs g = gather_args(s, D, 0);
    launch(D, D->tG, k_gather, dim3(1), dim3(kGatherBlock), g);
    if (D->rows == 0) D->row0 = s->step;
    D->rows++;
    s->step++;
    D->ctl_step = s->step;
    D->ctl_row = D->rows;
    if (D->rows == kSeriesCap) launch_dft(s, D, 0);
}

void launch_dft(cem_sim *s, Device *D, int direct) {
    if (!D->rows) return;
    DftArgs a{};
    a.nf = s->nf;
    a.nport = (int)s->port.size();
    a.nvp = (int)s->vp.size();
    a.nip = (int)s->ip.size();
    a.nacc = D->nacc;
    a.rows = D->rows;
    a.step0 = D->row0;
    a.dt = s->dt;
    a.omega = D->omega.p;
    a.rot = D->rot.p;
    a.ph_in = D->ph[D->ph_cur].p;
    a.ph_out = D->ph[1 - D->ph_cur].p;
    a.ser_ip = D->ser_ip;
    a.ser_vold = D->ser_vold;
    a.ser_vnew = D->ser_vnew;
    a.ser_vp = D->ser_vp;
    a.R = D->R.p;
    a.src_port = s->src_port;
    a.has_lumped = s->src_port >= 0;
    a.has_mode = s->msrc >= 0;
    a.direct = direct;
    a.amp = s->src_amp;
    a.t0 = s->src_t0;
    a.tau = s->src_tau;
    a.mt0 = s->msrc_t0;
    a.mtau = s->msrc_tau;
    a.acc = D->acc.p;
    a.ctl = D->ctl.p;
    const int nt = a.nf * (a.nacc + 1);
    launch(D, D->tD, k_dft, dim3((nt + 127) / 128), dim3(128), a);
    D->ph_cur = 1 - D->ph_cur;
    D->row0 += D->rows;
    D->rows = 0;
    D->ctl_row = 0;
}

// V of every port before the E update (after a host write of E or a resume)
void launch_gather_mid(cem_sim *s, Device *D) {
    if (!D->nsamp) return;
    GatherArgs a = gather_args(s, D, 1);
    hipLaunchKernelGGL(k_gather, dim3(1), dim3(kGatherBlock), 0, D->st, a);
}

int ensure_vlast(cem_sim *s, Device *D) {
    if (D->vlast_stale) {
        launch_gather_mid(s, D);
        D->vlast_stale = false;
    }
    if (D->eold_stale && s->ndisp) {
        hipLaunchKernelGGL(k_prime_eold, dim3((s->ndisp + kBlock - 1) / kBlock), dim3(kBlock), 0, D->st, D->F.p,
                           D->d_idx.p, D->Eold.p, s->ndisp);
    }
    D->eold_stale = false;
    return 0;
}

// host <-> device copies ordered after all work queued on the engine stream
void copy_sync(Device *D, void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
    if (!bytes) return;
    HIP_CHECK_VOID(hipMemcpyAsync(dst, src, bytes, kind, D->st));
    HIP_CHECK_VOID(hipStreamSynchronize(D->st));
}

void download_f32(Device *D, const float *src, double *dst, size_t n) {
    std::vector<float> tmp(n);
    HIP_CHECK_VOID(hipMemcpyAsync(tmp.data(), src, n * sizeof(float), hipMemcpyDeviceToHost, D->st));
    HIP_CHECK_VOID(hipStreamSynchronize(D->st));
    for (size_t q = 0; q < n; q++) dst[q] = tmp[q];
}

void upload_f32(Device *D, float *dst, const double *src, size_t n) {
    std::vector<float> tmp(n);
    for (size_t q = 0; q < n; q++) tmp[q] = (float)src[q];
    HIP_CHECK_VOID(hipMemcpyAsync(dst, tmp.data(), n * sizeof(float), hipMemcpyHostToDevice, D->st));
    HIP_CHECK_VOID(hipStreamSynchronize(D->st));
}

void drop_graph(Device *D) {
    if (D->gexec) (void)hipGraphExecDestroy(D->gexec);
    D->gexec = nullptr;
    D->gkey.clear();
}

// kGraphSteps full steps (H, E, post + samples) replayed from one captured graph.
// Every argument of these launches is fixed for a given plan and source set (the
// step and series row live on the device), so the graph is captured once and
// captured again only when an argument changes.  The DFT flush is launched by
// the host after the chunk that fills the series.
int run_chunk(cem_sim *s, Device *D) {
    sync_ctl(s, D);
    const int nx = s->n[0];
    const HArgs ha = h_args(s, D);
    const EArgs ea = e_args(s, D);
    const PostArgs pa = post_args(s, D);
    const GatherArgs ga = gather_args(s, D, 0);
    const int geo[4] = {D->iseg_h, D->iseg_e, post_blocks(s, D), (int)D->compressed};
    std::vector<unsigned char> key;
    for (auto pr : {std::make_pair((const void *)&ha, sizeof ha), std::make_pair((const void *)&ea, sizeof ea),
                    std::make_pair((const void *)&pa, sizeof pa), std::make_pair((const void *)&ga, sizeof ga),
                    std::make_pair((const void *)geo, sizeof geo)}) {
        const unsigned char *b = (const unsigned char *)pr.first;
        key.insert(key.end(), b, b + pr.second);
    }
    if (!D->gexec || key != D->gkey) {
        drop_graph(D);
        hipGraph_t graph = nullptr;
        if (hipStreamBeginCapture(D->st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            D->use_graph = false;
            return -1;
        }
        const int npblk = (D->g.S + kBlock - 1) / kBlock;
        const int gh = npblk * ((nx + 1 + D->iseg_h - 1) / D->iseg_h);
        const int ge = npblk * ((nx + D->iseg_e - 1) / D->iseg_e);
        for (int k = 0; k < kGraphSteps; k++) {
            hipLaunchKernelGGL(k_update_h, dim3(gh), dim3(kBlock), 0, D->st, ha, 0, nx + 1, D->iseg_h, npblk);
            if (D->compressed)
                hipLaunchKernelGGL(k_update_e<true>, dim3(ge), dim3(kBlock), 0, D->st, ea, 0, nx, D->iseg_e, npblk);
            else
                hipLaunchKernelGGL(k_update_e<false>, dim3(ge), dim3(kBlock), 0, D->st, ea, 0, nx, D->iseg_e, npblk);
            if (geo[2]) hipLaunchKernelGGL(k_post, dim3(geo[2]), dim3(kBlock), 0, D->st, pa);
            hipLaunchKernelGGL(k_gather, dim3(1), dim3(kGatherBlock), 0, D->st, ga);
        }
        const hipError_t ce = hipStreamEndCapture(D->st, &graph);
        if (ce != hipSuccess || !graph || hipGraphInstantiate(&D->gexec, graph, nullptr, nullptr, 0) != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            D->gexec = nullptr;
            D->use_graph = false;   // fall back to direct launches for this simulation
            return -1;
        }
        (void)hipGraphDestroy(graph);
        D->gkey.swap(key);
    }
    if (hipGraphLaunch(D->gexec, D->st) != hipSuccess) {
        set_error("graph launch failed");
        drop_graph(D);
        D->use_graph = false;
        return -1;
    }
    if (D->rows == 0) D->row0 = s->step;
    D->rows += kGraphSteps;
    s->step += kGraphSteps;
    D->ctl_step = s->step;
    D->ctl_row = D->rows;
    if (D->rows == kSeriesCap) launch_dft(s, D, 0);
    return 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// engine entry points (called from setup.cpp)
// ---------------------------------------------------------------------------
int cem::dev_ensure(cem_sim *s) {
    if (s->dev) {
        if (s->dev->sources_dirty && upload_sources(s, s->dev)) return -1;
        return 0;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        set_error("no GPU device available for the courant engine");
        return -1;
    }
    Device *D = new Device();
    if (hipStreamCreateWithFlags(&D->st, hipStreamNonBlocking) != hipSuccess) {
        set_error("cannot create HIP stream");
        delete D;
        return -1;
    }
    s->dev = D;
    if (upload_plan(s, D)) {
        dev_destroy(s);
        return -1;
    }
    D->profiling = std::getenv("COURANT_PROFILE") != nullptr;
    if (const char *v = std::getenv("COURANT_GRAPH")) D->use_graph = std::atoi(v) != 0;
    if (const char *v = std::getenv("COURANT_ISEG_H")) D->iseg_h = std::max(1, std::atoi(v)), D->tuned = true;
    if (const char *v = std::getenv("COURANT_ISEG_E")) D->iseg_e = std::max(1, std::atoi(v)), D->tuned = true;
    if (const char *v = std::getenv("COURANT_AUTOTUNE"))
        if (std::atoi(v) == 0) D->tuned = true;
    s->spec_src.assign(2 * s->nf, 0.0);
    for (size_t p = 0; p < s->port.size(); p++) {
        s->spec_port_v[p].assign(2 * s->nf, 0.0);
        s->spec_port_i[p].assign(2 * s->nf, 0.0);
    }
    s->spec_vp.assign(s->vp.size(), dvec(2 * s->nf, 0.0));
    s->spec_ip.assign(s->ip.size(), dvec(2 * s->nf, 0.0));
    // phasors of step 0 are set by the DFT kernel's resync at n % 1024 == 0
    for (int c = 0; c < 6; c++)
        if (s->field_mirror[c]) D->vlast_stale = D->eold_stale = true;
    if (s->step != 0) D->vlast_stale = D->eold_stale = true;
    return 0;
}

void cem::dev_destroy(cem_sim *s) {
    Device *D = s->dev;
    if (!D) return;
    (void)hipStreamSynchronize(D->st);
    D->F.release();
    D->ca.release();
    D->cb.release();
    D->cidx.release();
    D->ctab.release();
    D->axis1d.release();
    D->slots.release();
    D->psi.release();
    drop_graph(D);
    D->J.release();
    D->dtab.release();
    D->d_mat.release();
    D->ctl.release();
    D->Y.release();
    D->wl.release();
    D->wm.release();
    D->face.release();
    D->pp.release();
    for (auto *b : {&D->d_idx, &D->sb_idx, &D->sbo, &D->x_idx}) b->release();
    for (auto *b : {&D->dsb_t, &D->dsb_u, &D->dsrc_t, &D->dsrc_x}) b->release();
    D->sb_src.release();
    D->x_src.release();
    D->Eold.release();
    D->tidx.release();
    D->tw.release();
    D->post.release();
    D->toff.release();
    D->vlast.release();
    D->ser.release();
    D->omega.release();
    D->rot.release();
    D->ph[0].release();
    D->ph[1].release();
    D->acc.release();
    D->R.release();
    D->d1d.release();
    D->partial.release();
    for (KernelTimer *t : {&D->tH, &D->tE, &D->tG, &D->tD, &D->tP}) t->destroy();
    (void)hipStreamDestroy(D->st);
    delete D;
    s->dev = nullptr;
}

void cem::dev_sources_changed(cem_sim *s) {
    if (s->dev) s->dev->sources_dirty = true;
}

// Autotuner of the stencil launch shape (i-planes walked per thread by the H and
// E updates).  Which thread updates a point does not change any value, so the
// candidates run on the simulation's own first steps.  The two shapes are
// searched jointly: a kernel's time depends on the other's tile order through
// what it leaves in the caches (on the 38.8 M-point board H at 8 planes takes
// 341 us after E at 4 and 371 us after E at 2).  Each (H, E) pair runs one
// settling step and kTuneSteps steps with event-stamped launches, in two rounds;
// the pair with the lowest best-of-two H + E time wins.  The choice is kept for
// later simulations of the same grid shape in this process (a batch of equal
// boards tunes once).  COURANT_TUNE_DEBUG prints the measured times.
constexpr int kTuneCand[] = {1, 2, 4, 8};
constexpr int kTuneNc = sizeof kTuneCand / sizeof kTuneCand[0];
constexpr int kTuneSteps = 2;
constexpr int kTuneRounds = 2;
constexpr long kTuneCost = 1 + (long)kTuneRounds * kTuneNc * kTuneNc * (1 + kTuneSteps);

std::mutex g_tune_mu;
std::map<std::array<long long, 5>, std::pair<int, int>> g_tune;

long autotune(cem_sim *s, Device *D, long budget) {
    const std::array<long long, 5> key = {s->n[0], s->n[1], s->n[2], s->has_pml, (long long)D->compressed};
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune.find(key);
        if (it != g_tune.end()) {
            D->iseg_h = it->second.first;
            D->iseg_e = it->second.second;
            D->tuned = true;
            return 0;
        }
    }
    if (budget < kTuneCost) return 0;   // too few steps in this call: tune on a later one
    const int nx = s->n[0];
    auto one_step = [&] {
        launch_h(s, D, 0, nx + 1);
        launch_e(s, D, 0, nx);
        launch_post(s, D);
    };
    one_step();                         // warm-up step (first launches, caches)
    double t[kTuneNc][kTuneNc];
    for (int r = 0; r < kTuneRounds; r++)
        for (int ch = 0; ch < kTuneNc; ch++)
            for (int ce = 0; ce < kTuneNc; ce++) {
                D->iseg_h = kTuneCand[ch];
                D->iseg_e = kTuneCand[ce];
                one_step();             // settling step, untimed
                D->profiling = true;
                for (KernelTimer *k : {&D->tH, &D->tE}) {
                    k->fold();
                    k->total_ms = 0;
                    k->count = 0;
                }
                for (int st = 0; st < kTuneSteps; st++) one_step();
                D->tH.fold();
                D->tE.fold();
                D->profiling = false;
                const double v = D->tH.total_ms + D->tE.total_ms;
                t[ch][ce] = r ? std::min(t[ch][ce], v) : v;
            }
    int bh = 0, be = 0;
    for (int ch = 0; ch < kTuneNc; ch++)
        for (int ce = 0; ce < kTuneNc; ce++)
            if (t[ch][ce] < t[bh][be]) bh = ch, be = ce;
    for (KernelTimer *k : {&D->tH, &D->tE, &D->tG, &D->tD, &D->tP}) {
        k->fold();
        k->total_ms = 0;
        k->count = 0;
    }
    D->iseg_h = kTuneCand[bh];
    D->iseg_e = kTuneCand[be];
    D->tuned = true;
    if (std::getenv("COURANT_TUNE_DEBUG")) {
        std::fprintf(stderr, "courant autotune H+E us per step (rows H, columns E = 1 2 4 8):\n");
        for (int ch = 0; ch < kTuneNc; ch++) {
            std::fprintf(stderr, "  H %d:", kTuneCand[ch]);
            for (int ce = 0; ce < kTuneNc; ce++) std::fprintf(stderr, " %7.1f", t[ch][ce] / kTuneSteps * 1e3);
            std::fprintf(stderr, "\n");
        }
        std::fprintf(stderr, "  -> H %d, E %d\n", D->iseg_h, D->iseg_e);
    }
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[key] = {D->iseg_h, D->iseg_e};
    return kTuneCost;
}

void cem::dev_step(cem_sim *s, long nsteps) {
    Device *D = s->dev;
    ensure_vlast(s, D);
    const int nx = s->n[0];
    long it = 0;
    if (!D->tuned && !D->profiling) it += autotune(s, D, nsteps);
    while (it < nsteps) {
        // whole chunks from a captured graph once the series row is chunk-aligned
        if (D->use_graph && !D->profiling && D->rows % kGraphSteps == 0 && nsteps - it >= kGraphSteps &&
            run_chunk(s, D) == 0) {
            it += kGraphSteps;
            continue;
        }
        launch_h(s, D, 0, nx + 1);
        launch_e(s, D, 0, nx);
        launch_post(s, D);
        it++;
    }
}

void cem::dev_mt_H(cem_sim *s, int i0, int i1) {
    Device *D = s->dev;
    ensure_vlast(s, D);
    launch_h(s, D, std::max(0, i0), std::min(s->n[0] + 1, i1));
}

void cem::dev_mt_E(cem_sim *s, int i0, int i1) {
    launch_e(s, s->dev, std::max(0, i0), std::min(s->n[0], i1));
}

void cem::dev_mt_post(cem_sim *s) {
    launch_post(s, s->dev);
}

double cem::dev_energy(cem_sim *s) {
    Device *D = s->dev;
    EnergyArgs a{};
    a.g = D->g;
    a.E = D->F.p;
    a.H = D->F.p + 3 * D->g.N;
    a.cb = D->cb.p;
    a.cidx = D->compressed ? D->cidx.p : nullptr;
    a.ctab = D->compressed ? D->ctab.p : nullptr;
    size_t o = 0;
    for (int c = 0; c < 3; c++) {
        a.d[c] = D->d1d.p + o;
        o += s->n[c];
    }
    for (int c = 0; c < 3; c++) {
        a.dd[c] = D->d1d.p + o;
        o += s->n[c] + 1;
    }
    a.dt = s->dt;
    hipLaunchKernelGGL(k_energy, dim3(D->energy_blocks), dim3(kBlock), 0, D->st, a, D->partial.p);
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(64), 0, D->st, D->partial.p, D->energy_blocks,
                       D->partial.p + D->energy_blocks);
    double W = 0;
    HIP_CHECK_VOID(hipMemcpyAsync(&W, D->partial.p + D->energy_blocks, sizeof(double), hipMemcpyDeviceToHost, D->st));
    HIP_CHECK_VOID(hipStreamSynchronize(D->st));
    return W;
}

void cem::dev_field_download(cem_sim *s, int comp, double *dst) {
    Device *D = s->dev;
    if (!D) {
        std::memset(dst, 0, s->N * sizeof(double));
        return;
    }
    download_f32(D, D->F.p + (size_t)comp * s->N, dst, s->N);
}

// slab planes of a z-derivative psi array (stored slot-fastest), 0 for every other state
static int psi_z_planes(const cem_sim *s, const Device *D, int which, int comp, int w) {
    if ((which != 0 && which != 1) || (comp + w + 1) % 3 != 2 || !s->has_pml) return 0;
    return which == 0 ? D->pmlE.zpl : D->pmlH.zpl;
}

static float *state_ptr(cem_sim *s, int which, int comp, int w, size_t *len) {
    Device *D = s->dev;
    switch (which) {
    case 0: *len = D->psiE_len[comp][w]; return D->psiE[comp][w];
    case 1: *len = D->psiH_len[comp][w]; return D->psiH[comp][w];
    case 2: *len = (size_t)s->npole * s->ndisp; return D->J.p;
    case 5: *len = s->N; return D->F.p + (size_t)comp * s->N;
    default: *len = 0; return nullptr;
    }
}

void cem::dev_state_download(cem_sim *s, int which, int comp, int w, double *dst) {
    Device *D = s->dev;
    if (which == 3) {       // Debye E_old: E of each dispersive edge before the next E update
        ensure_vlast(s, D);
        download_f32(D, D->Eold.p, dst, s->ndisp);
        return;
    }
    if (which == 4) {       // SIBC states: device term-major -> reference edge-major
        const size_t K = s->sibc_npole, n = s->nsb;
        std::vector<double> t(K * n);
        download_f32(D, D->Y.p, t.data(), K * n);
        for (size_t e = 0; e < n; e++)
            for (size_t k = 0; k < K; k++) dst[e * K + k] = t[k * n + e];
        return;
    }
    size_t len = 0;
    float *p = state_ptr(s, which, comp, w, &len);
    if (!p) return;
    const int zpl = psi_z_planes(s, D, which, comp, w);
    if (!zpl) {
        download_f32(D, p, dst, len);
        return;
    }
    std::vector<double> t(len);     // device slot-fastest -> reference slab-major
    download_f32(D, p, t.data(), len);
    const size_t per = len / zpl;
    for (size_t r = 0; r < per; r++)
        for (int z = 0; z < zpl; z++) dst[z * per + r] = t[r * zpl + z];
}

void cem::dev_state_upload(cem_sim *s, int which, int comp, int w, const double *src) {
    Device *D = s->dev;
    if (which == 3) {        // E_old is taken from E again before the next E update
        D->eold_stale = true;
        return;
    }
    if (which == 4) {
        const size_t K = s->sibc_npole, n = s->nsb;
        std::vector<double> t(K * n);
        for (size_t e = 0; e < n; e++)
            for (size_t k = 0; k < K; k++) t[k * n + e] = src[e * K + k];
        upload_f32(D, D->Y.p, t.data(), K * n);
        return;
    }
    size_t len = 0;
    float *p = state_ptr(s, which, comp, w, &len);
    if (p) {
        const int zpl = psi_z_planes(s, D, which, comp, w);
        if (!zpl) {
            upload_f32(D, p, src, len);
        } else {
            std::vector<double> t(len);
            const size_t per = len / zpl;
            for (size_t r = 0; r < per; r++)
                for (int z = 0; z < zpl; z++) t[r * zpl + z] = src[z * per + r];
            upload_f32(D, p, t.data(), len);
        }
    }
    if (which == 5 && comp < 3) D->vlast_stale = D->eold_stale = true;
}

void cem::dev_sync_in(cem_sim *s) {
    Device *D = s->dev;
    for (int c = 0; c < 6; c++)
        if (s->field_mirror[c]) dev_state_upload(s, 5, c, 0, s->field_mirror[c].get());
    if (s->state_exposed) {
        for (int c = 0; c < 3; c++)
            for (int w = 0; w < 2; w++) {
                if (s->psiE_mirror[c][w].size()) dev_state_upload(s, 0, c, w, s->psiE_mirror[c][w].data());
                if (s->psiH_mirror[c][w].size()) dev_state_upload(s, 1, c, w, s->psiH_mirror[c][w].data());
            }
        if (s->dJ_mirror.size()) dev_state_upload(s, 2, 0, 0, s->dJ_mirror.data());
        if (s->sby_mirror.size()) dev_state_upload(s, 4, 0, 0, s->sby_mirror.data());
    }
    (void)D;
}

void cem::dev_sync_out(cem_sim *s) {
    for (int c = 0; c < 6; c++)
        if (s->field_mirror[c]) dev_field_download(s, c, s->field_mirror[c].get());
    if (s->state_exposed) {
        for (int c = 0; c < 3; c++)
            for (int w = 0; w < 2; w++) {
                if (s->psiE_mirror[c][w].size()) dev_state_download(s, 0, c, w, s->psiE_mirror[c][w].data());
                if (s->psiH_mirror[c][w].size()) dev_state_download(s, 1, c, w, s->psiH_mirror[c][w].data());
            }
        if (s->dJ_mirror.size()) dev_state_download(s, 2, 0, 0, s->dJ_mirror.data());
        if (s->dEold_mirror.size()) dev_state_download(s, 3, 0, 0, s->dEold_mirror.data());
        if (s->sby_mirror.size()) dev_state_download(s, 4, 0, 0, s->sby_mirror.data());
    }
    if (s->spectra_exposed) dev_spectra_download(s);
}

void cem::dev_spectra_download(cem_sim *s) {
    Device *D = s->dev;
    if (!D) return;   // nothing stepped yet: spectra are zero
    launch_dft(s, D, 0);
    const int nf = s->nf;
    std::vector<double> all((size_t)D->nacc * 2 * nf);
    HIP_CHECK_VOID(hipMemcpyAsync(all.data(), D->acc.p, all.size() * sizeof(double), hipMemcpyDeviceToHost, D->st));
    HIP_CHECK_VOID(hipStreamSynchronize(D->st));
    const size_t np = s->port.size(), nvp = s->vp.size(), nip = s->ip.size();
    auto blk = [&](size_t a) { return all.data() + a * 2 * nf; };
    for (size_t p = 0; p < np; p++) {
        std::memcpy(s->spec_port_v[p].data(), blk(p), 2 * nf * sizeof(double));
        std::memcpy(s->spec_port_i[p].data(), blk(np + p), 2 * nf * sizeof(double));
    }
    for (size_t p = 0; p < nvp; p++) std::memcpy(s->spec_vp[p].data(), blk(2 * np + p), 2 * nf * sizeof(double));
    for (size_t p = 0; p < nip; p++) std::memcpy(s->spec_ip[p].data(), blk(2 * np + nvp + p), 2 * nf * sizeof(double));
    std::memcpy(s->spec_src.data(), blk(2 * np + nvp + nip), 2 * nf * sizeof(double));
}

// Wave separation of one mode port on the device from the DFT accumulators it
// already holds; mode = (Re Z0, Im Z0, alpha, beta) per frequency.
int cem::dev_mode_waves(cem_sim *s, int vp0, int vp1, int ip, const double *mode, double dx, double dist, int dir,
                        double *a, double *b) {
    if (dev_ensure(s)) return -1;
    Device *D = s->dev;
    launch_dft(s, D, 0);
    const int nf = s->nf;
    const size_t np = s->port.size(), nvp = s->vp.size();
    const double *acc = D->acc.p;
    ModeWaveArgs w{};
    w.nf = nf;
    w.dir = dir;
    w.dx = dx;
    w.dist = dist;
    w.V0 = acc + (2 * np + vp0) * 2 * (size_t)nf;
    w.V1 = acc + (2 * np + vp1) * 2 * (size_t)nf;
    w.Ix = acc + (2 * np + nvp + ip) * 2 * (size_t)nf;
    DBuf<double> md, out;
    int rc = -1;
    if (!md.alloc(4 * (size_t)nf, D->st) && !md.upload(mode, 4 * (size_t)nf, D->st) && !out.alloc(4 * (size_t)nf, D->st)) {
        w.mode = md.p;
        w.out = out.p;
        hipLaunchKernelGGL(k_mode_waves, dim3((nf + 127) / 128), dim3(128), 0, D->st, w);
        if (hipGetLastError() == hipSuccess &&
            hipMemcpyAsync(a, out.p, 2 * nf * sizeof(double), hipMemcpyDeviceToHost, D->st) == hipSuccess &&
            hipMemcpyAsync(b, out.p + 2 * nf, 2 * nf * sizeof(double), hipMemcpyDeviceToHost, D->st) == hipSuccess &&
            hipStreamSynchronize(D->st) == hipSuccess)
            rc = 0;
        else
            set_error("mode port: device wave separation failed");
    }
    md.release();
    out.release();
    return rc;
}

// host-recorded samples of an external engine (export.c:121-149): the reference
// recomputes the phasors from cos/sin at every sample, so the DFT runs "direct"
void cem::dev_push_samples(cem_sim *s, int n, const double *ip, const double *vold, const double *vnew,
                           const double *vp) {
    Device *D = s->dev;
    launch_dft(s, D, 0);
    const size_t nip = s->ip.size(), np = s->port.size(), nvp = s->vp.size();
    int done = 0;
    while (done < n) {
        const int m = std::min(n - done, kSeriesCap);
        if (nip) copy_sync(D, D->ser_ip, ip + (size_t)done * nip, m * nip * sizeof(double), hipMemcpyHostToDevice);
        if (np) {
            copy_sync(D, D->ser_vold, vold + (size_t)done * np, m * np * sizeof(double), hipMemcpyHostToDevice);
            copy_sync(D, D->ser_vnew, vnew + (size_t)done * np, m * np * sizeof(double), hipMemcpyHostToDevice);
        }
        if (nvp) copy_sync(D, D->ser_vp, vp + (size_t)done * nvp, m * nvp * sizeof(double), hipMemcpyHostToDevice);
        D->row0 = s->step;
        D->rows = m;
        launch_dft(s, D, 1);
        s->step += m;
        done += m;
    }
    if (s->spectra_exposed) dev_spectra_download(s);
}

void cem::dev_phasors(cem_sim *s, double *phE, double *phH, int download) {
    Device *D = s->dev;
    launch_dft(s, D, 0);
    const size_t nb = 2 * s->nf * sizeof(double);
    double *cur = D->ph[D->ph_cur].p;
    HIP_CHECK_VOID(hipStreamSynchronize(D->st));
    if (download) {
        copy_sync(D, phE, cur, nb, hipMemcpyDeviceToHost);
        copy_sync(D, phH, cur + 2 * s->nf, nb, hipMemcpyDeviceToHost);
    } else {
        copy_sync(D, cur, phE, nb, hipMemcpyHostToDevice);
        copy_sync(D, cur + 2 * s->nf, phH, nb, hipMemcpyHostToDevice);
    }
}

// all DFT accumulators in checkpoint order: port V/I pairs, vprobes, iprobes,
// source, then phE, phH (export.c:204-210)
void cem::dev_acc_io(cem_sim *s, double *buf, int download) {
    Device *D = s->dev;
    launch_dft(s, D, 0);
    const int nf = s->nf;
    const size_t blk = 2 * (size_t)nf;
    const size_t np = s->port.size(), nvp = s->vp.size(), nip = s->ip.size();
    std::vector<double> all((size_t)D->nacc * blk);
    if (download) {
        copy_sync(D, all.data(), D->acc.p, all.size() * sizeof(double), hipMemcpyDeviceToHost);
    }
    // device order: V(np) I(np) vp ip src; checkpoint order: (V, I) per port, vp, ip, src
    size_t o = 0;
    auto io = [&](size_t dev_block) {
        if (download)
            std::memcpy(buf + o, all.data() + dev_block * blk, blk * sizeof(double));
        else
            std::memcpy(all.data() + dev_block * blk, buf + o, blk * sizeof(double));
        o += blk;
    };
    for (size_t p = 0; p < np; p++) {
        io(p);
        io(np + p);
    }
    for (size_t p = 0; p < nvp; p++) io(2 * np + p);
    for (size_t p = 0; p < nip; p++) io(2 * np + nvp + p);
    io(2 * np + nvp + nip);
    if (!download) copy_sync(D, D->acc.p, all.data(), all.size() * sizeof(double), hipMemcpyHostToDevice);
    dev_phasors(s, buf + o, buf + o + blk, download);
}

void cem::dev_set_step(cem_sim *s, long step) {
    Device *D = s->dev;
    if (!D) return;
    launch_dft(s, D, 0);
    D->row0 = step;
}

// ---------------------------------------------------------------------------
// engine extensions (include/courant_amd.h)
// ---------------------------------------------------------------------------
extern "C" int cem_amd_kernel_stats(cem_sim *s, int which, double *total_ms, long *count) {
    if (!s || !s->dev) return -1;
    KernelTimer *t[5] = {&s->dev->tH, &s->dev->tE, &s->dev->tG, &s->dev->tD, &s->dev->tP};
    if (which < 0 || which > 4) return -1;
    t[which]->fold();
    *total_ms = t[which]->total_ms;
    *count = t[which]->count;
    return 0;
}

extern "C" int cem_amd_set_profiling(cem_sim *s, int on) {
    if (!s || !s->finalized || dev_ensure(s)) return -1;
    s->dev->profiling = on != 0;
    for (KernelTimer *t : {&s->dev->tH, &s->dev->tE, &s->dev->tG, &s->dev->tD, &s->dev->tP}) {
        t->fold();
        t->total_ms = 0;
        t->count = 0;
    }
    return 0;
}

extern "C" int cem_amd_synchronize(cem_sim *s) {
    if (!s || !s->dev) return -1;
    return hipStreamSynchronize(s->dev->st) == hipSuccess ? 0 : -1;
}

// Advance n steps without touching host mirrors or spectra (timed loops).
extern "C" int cem_amd_step_device(cem_sim *s, long n) {
    if (!s || !s->finalized || dev_ensure(s)) return -1;
    dev_step(s, n);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// 2x2 renormalisation of nf frequencies on the device, FP64, in place.
extern "C" int cem_amd_s2_renormalize(double *S, int nf, const double *z1, const double *z2, double zref) {
    if (!S || !z1 || !z2 || nf < 0) {
        set_error("s2 renormalize: bad arguments");
        return -1;
    }
    if (nf == 0) return 0;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        set_error("no GPU device available for the courant engine");
        return -1;
    }
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
        set_error("cannot create HIP stream");
        return -1;
    }
    DBuf<double> dS, d1, d2;
    int rc = -1;
    const size_t n = (size_t)nf;
    if (!dS.alloc(8 * n, st) && !d1.alloc(n, st) && !d2.alloc(n, st) && !dS.upload(S, 8 * n, st) &&
        !d1.upload(z1, n, st) && !d2.upload(z2, n, st)) {
        hipLaunchKernelGGL(k_s2_renormalize, dim3((nf + 127) / 128), dim3(128), 0, st, dS.p, nf, d1.p, d2.p, zref);
        if (hipGetLastError() == hipSuccess &&
            hipMemcpyAsync(S, dS.p, 8 * n * sizeof(double), hipMemcpyDeviceToHost, st) == hipSuccess &&
            hipStreamSynchronize(st) == hipSuccess)
            rc = 0;
        else
            set_error("s2 renormalize: device failure");
    }
    dS.release();
    d1.release();
    d2.release();
    (void)hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
    return rc;
}

// Allocation-free renormalisation.  Its device staging (S, then z1, then z2) is
// this file-scope array rather than memory from the pool: kNoAllocMaxNf
// frequencies at most, 10 doubles each (5 MiB, held by every GPU the library
// runs on until the process ends).  Every call on one GPU shares that GPU's
// array, so calls that overlap on one GPU race; each call overwrites all of
// the staging it reads, so nothing has to start at zero.
constexpr int kNoAllocMaxNf = 65536;
static __device__ __align__(256) double scratch[10 * kNoAllocMaxNf];

namespace {
std::mutex g_scratch_mu;
std::vector<double *> g_scratch_dev;   // device address of `scratch`, per device

// Device address of this device's copy of `scratch`, looked up once per device
// (with that device current) by hipGetSymbolAddress; the array's own name is a
// host-side stand-in and never a device address.
double *scratch_on(int dev) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch_dev.size() <= dev) g_scratch_dev.resize(dev + 1, nullptr);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            char m[160];
            std::snprintf(m, sizeof m, "s2 renormalize (no-alloc): hipGetSymbolAddress failed: %s", hipGetErrorString(e));
            set_error(m);
            return nullptr;
        }
        g_scratch_dev[dev] = (double *)p;
    }
    return g_scratch_dev[dev];
}
}  // namespace

extern "C" int cem_amd_s2_renormalize_noalloc(double *S, int nf, const double *z1, const double *z2, double zref) {
    if (!S || !z1 || !z2 || nf < 0) {
        set_error("s2 renormalize (no-alloc): bad arguments");
        return -1;
    }
    if (nf > kNoAllocMaxNf) {
        char m[160];
        std::snprintf(m, sizeof m, "s2 renormalize (no-alloc): %d frequencies exceed the fixed staging of %d", nf,
                      kNoAllocMaxNf);
        set_error(m);
        return -1;
    }
    if (nf == 0) return 0;
    int ndev = 0, dev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1 || hipGetDevice(&dev) != hipSuccess) {
        set_error("no GPU device available for the courant engine");
        return -1;
    }
    double *base = scratch_on(dev);
    if (!base) return -1;
    const size_t n = (size_t)nf;
    double *dS = base, *d1 = base + 8 * n, *d2 = base + 9 * n;
    hipStream_t st = nullptr;   // the device's null stream: no stream is created
    if (hipMemcpyAsync(dS, S, 8 * n * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(d1, z1, n * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(d2, z2, n * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess) {
        set_error("s2 renormalize (no-alloc): upload failed");
        return -1;
    }
    hipLaunchKernelGGL(k_s2_renormalize, dim3((nf + 127) / 128), dim3(128), 0, st, dS, nf, d1, d2, zref);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(S, dS, 8 * n * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        set_error("s2 renormalize (no-alloc): device failure");
        return -1;
    }
    return 0;
}

// Stream-copy bandwidth of this GPU: copies `bytes` (rounded down to 16 B) between
// two device buffers `iters` times and returns the best GB/s (read + write bytes).
extern "C" int cem_amd_launch_shape(cem_sim *s, int *iseg_h, int *iseg_e) {
    if (!s || !s->dev) {
        set_error("simulation has not run on the GPU yet");
        return -1;
    }
    if (iseg_h) *iseg_h = s->dev->iseg_h;
    if (iseg_e) *iseg_e = s->dev->iseg_e;
    return s->dev->tuned ? 1 : 0;
}

extern "C" double cem_amd_copy_bandwidth(size_t bytes, int iters) {
    const long long n = (long long)(bytes / 16);
    if (n <= 0 || iters <= 0) return -1;
    float4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, n * 16) != hipSuccess) return -1;
    if (hipMalloc(&b, n * 16) != hipSuccess) {
        (void)hipFree(a);
        return -1;
    }
    (void)hipMemset(a, 0, n * 16);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const long long grid = (n + 4 * kBlock - 1) / (4 * kBlock);
    double best = 0;
    for (int it = 0; it < iters; it++) {
        (void)hipEventRecord(e0, nullptr);
        hipLaunchKernelGGL(k_copy, dim3((unsigned)grid), dim3(kBlock), 0, nullptr, (const cf4 *)a, (cf4 *)b, n);
        (void)hipEventRecord(e1, nullptr);
        (void)hipEventSynchronize(e1);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (ms > 0) best = std::max(best, 2.0 * n * 16 / (ms * 1e-3) / 1e9);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return best;
}
