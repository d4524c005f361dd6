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
#include <unordered_map>
#include <vector>

#include "courant_amd.h"
#include "devmem.hpp"
#include "graph.hpp"
#include "sim.hpp"

using namespace cem;

namespace cem {

constexpr int kBlock = 256;        // 4 wavefronts of 64
constexpr int kIseg = 2;           // default i-planes per thread in the stencil kernels
constexpr int kSeriesCap = 512;    // steps recorded between DFT flushes
constexpr int kGraphSteps = 64;    // steps per captured chunk (even; divides kSeriesCap)
constexpr int kGatherBlock = 1024;
constexpr int kStepRows = 8;       // most owned rows per tile of the one-pass step
constexpr int kStepLanes = 1024;   // most lanes per workgroup of the one-pass step

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
// Device grid.  Every field array is (nx+1) x (ny+1) rows of kp floats, k
// fastest, where the row pitch kp is nz+1 rounded up to a multiple of 4: each
// row starts on a 16-byte boundary, so a lane of the stencil kernels moves the
// four consecutive k points of one "group" with a single 128-bit access and a
// wavefront moves 1 KiB of an array per instruction.  The kp - nz - 1 padding
// points of a row hold zero.  Plan indices (i si + j sj + k of the reference)
// are translated to i si_dev + j kp + k when they are uploaded.
struct Grid {
    int nx, ny, nz;
    int kp, ng;             // row pitch in floats, float4 groups per row (kp / 4)
    int S4;                 // groups per x plane: (ny + 1) ng
    long long si, sj, N;    // device strides (si = (ny+1) kp, sj = kp), points per component
    long long si4, N4;      // si / 4, N / 4
};

typedef float v4f __attribute__((ext_vector_type(4)));

// CPML of one kernel (H or E): per derivative axis a the slab slot of each index
// along a (-1 outside the slabs) and the recursion coefficients b, c.  The psi
// arrays are laid out for the 128-bit lanes:
//   x derivative: [slot][j][kp]                float4 index  slot S4 + p
//   y derivative: [slot][i][kp]                float4 index  (slot (nx+1) + i) ng + gk
//   z derivative: per (i, j) column, the k groups that hold slab points, one
//                 float4 each in k order       float4 index  (i (ny+1) + j) zg + zo[gk]
// (state export converts to and from the reference's slab-major order).
struct Pml {
    const int *slot[3];
    const float *b[3], *c[3];
    const int *zo;          // per k group: its place in a column's z layout, -1 outside
    int zg;                 // k groups per column in the z layout
};

struct HArgs {
    Grid g;
    const float *E;         // 3N
    float *H;               // 3N
    const float *hs[3];     // dt/mu0 * idh per axis
    Pml pml;
    float *psi[3][2];       // psiH[c][w]
    float dth;
    int has_pml;
    int xcd;                // 1: XCD-banded tile order (xcd_tile), 0: dispatch order in chunks
    int chunk;              // log2 of the tiles per chunk in dispatch order (chunk_tile)
    int rev;                // 1: tiles from the last (the sweep starts where the E sweep ended)
    // pair: a second right-hand side of the same plan (the other excitation of a 2-port),
    // whose fields and psi_H these are; odd tiles of the launch update it
    int pair;
    const float *E2;
    float *H2;
    float *psi2[3][2];
};

// Per-edge physics parameters, resident on the device so the post kernel's
// argument block stays small.
struct PostParams {
    int ndisp, npole, nsb, nterm;
    float kc[kMaxPoles];    // (1 + k)/2
    float km1[kMaxPoles];   // k - 1 (computed without cancellation)
    float zc[kSibcTerms];   // c_k
    float om[kSibcTerms];   // 1 - e_k, from -expm1(-a_k dt)
};

constexpr int kFusedPoles = 8;     // most Debye poles the E kernel updates itself
constexpr int kDebyeRows = 256;    // most (beta_p / dt) material rows it holds in LDS

struct EArgs {
    Grid g;
    float *E;               // 3N
    const float *H;         // 3N
    const float *ca, *cb;   // 3N each (uncompressed coefficients)
    const unsigned *cidx;   // N: 8-bit classes of Ex, Ey, Ez (compressed coefficients)
    const float2 *ctab;     // 3 x 256 (ca, cb) per class
    const int *crow;        // 3 x 256 Debye material row of a class; -1: not updated here
    int ncls;               // classes in use per component (slots 0..ncls-1 of each 256)
    // one byte per point instead of cidx where at most 256 distinct class triples occur:
    // pidx[point] indexes ptab, whose entry is the point's packed cidx word
    const unsigned char *pidx;
    const unsigned *ptab;
    int nptab;
    const float *es[3];     // ide per axis
    Pml pml;
    float *psi[3][2];       // psiE[c][w]
    int has_pml;
    int xcd;                // 1: XCD-banded tile order (xcd_tile), 0: dispatch order in chunks
    int chunk;              // log2 of the tiles per chunk in dispatch order (chunk_tile)
    // Debye ADE of the dispersive edges this kernel updates.  J is held over a box
    // of the grid, [i - bi0][j - bj0][comp][pole][k - 4 bg0]: the lanes of a wave
    // read one pole of consecutive points as one contiguous run, and one (i, j) row
    // of the box is one contiguous block
    float *J;
    const float *btab;      // per material row: beta_p / dt for each pole
    int nrow;
    int bi0, bi1, bj0, bj1, bg0, bg1;
    float kc[kFusedPoles], km1[kFusedPoles];
    // pair: the second right-hand side's fields, psi_E and Debye currents (odd tiles)
    int pair;
    float *E2;
    const float *H2;
    float *psi2[3][2];
    float *J2;
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
// surface edge), so the lists are linked: threads [0, ndisp) take the listed
// dispersive edges (with their SIBC slot and source slot, if any), then nsbo
// surface edges that are not dispersive, then nxs remaining source edges.
// Indices are device E indices comp*N + q.  When the E kernel updates the
// Debye currents itself, only the dispersive edges it cannot take (surface
// edges, mode-port source edges, edges on the x = max face) are listed here;
// their J entries then live in the E kernel's box and d_jb gives each one's
// index there.
struct PostArgs {
    float *E;               // 3N
    const float *H;         // 3N
    long long N, si, sj;    // device strides
    const PostParams *pp;
    int ndisp, nsb, nsbo, nxs;
    const unsigned *d_idx;
    const unsigned *d_mat;  // per listed dispersive edge: row of dtab (bits 0-29), has an SIBC
                            // slot (bit 31), has a source slot (bit 30)
    const float *dtab;      // per material: cb, then beta/dt per pole
    int dstride;            // 1 + npole
    // the few listed dispersive edges that are also SIBC or source edges: ascending
    // edge numbers with their slots (found by binary search from the flag bits)
    const int *dsb_t, *dsb_u, *dsrc_t, *dsrc_x;
    int ndsb, ndsrc;
    float *J;               // pole p of listed edge t at (d_jb ? d_jb[t] : t) + p * jstride
    const long long *d_jb;
    long long jstride;
    float *Eold;            // E of each listed dispersive edge before this step's E update
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

// Dispatch order in chunks of 2^s consecutive tiles: the n-th workgroup dealt to XCD x
// runs tile n mod 2^s of chunk 8 (n >> s) + x, so each XCD sweeps runs of 2^s tiles
// (the halo rows of a run's inner tile boundaries stay in its L2) while the eight XCDs
// stream one window of 8 chunks; tiles past the last whole round of 8 chunks run in
// dispatch order.  s = 0 is plain dispatch order.  A bijection on [0, G).
__device__ __forceinline__ int chunk_tile(int wg, int G, int s) {
    if (s == 0 || wg >= ((G >> (3 + s)) << (3 + s))) return wg;
    const int x = wg & 7, n = wg >> 3;
    return ((((n >> s) << 3) + x) << s) + (n & ((1 << s) - 1));
}

// Lane width W: each thread of the stencil kernels owns W consecutive k points
// of one (i, j) row (a "group"; W divides kp), moved with one W-wide load or
// store per array.
template <int W> struct LaneVec;
template <> struct LaneVec<1> { typedef float T; };
template <> struct LaneVec<2> { typedef float T __attribute__((ext_vector_type(2))); };
template <> struct LaneVec<4> { typedef float T __attribute__((ext_vector_type(4))); };

template <int W>
__device__ __forceinline__ void ldw(const float *__restrict__ p, float (&x)[W]) {
    typedef typename LaneVec<W>::T V;
    const V v = *(const V *)p;
    if constexpr (W == 1) {
        x[0] = v;
    } else {
#pragma unroll
        for (int e = 0; e < W; e++) x[e] = v[e];
    }
}

template <int W>
__device__ __forceinline__ void stw(float *__restrict__ p, const float (&x)[W]) {
    typedef typename LaneVec<W>::T V;
    V v;
    if constexpr (W == 1) {
        v = x[0];
    } else {
#pragma unroll
        for (int e = 0; e < W; e++) v[e] = x[e];
    }
    *(V *)p = v;
}

template <int W>
__device__ __forceinline__ void zerow(float (&x)[W]) {
#pragma unroll
    for (int e = 0; e < W; e++) x[e] = 0.f;
}

// The same at a 32-bit byte offset from a base pointer: the load or store takes the
// base in scalar registers and the offset in one vector register, with no 64-bit
// address arithmetic per access.
template <int W>
__device__ __forceinline__ void ldb(const float *b, unsigned off, float (&x)[W]) {
    ldw<W>((const float *)((const char *)b + off), x);
}

template <int W>
__device__ __forceinline__ void stb(float *b, unsigned off, const float (&x)[W]) {
    stw<W>((float *)((char *)b + off), x);
}

__device__ __forceinline__ float lb(const float *b, unsigned off) {
    return *(const float *)((const char *)b + off);
}

// One H component at one point (sim.c:579-606) with the CPML recursions of the slabs
// that hold it (sim.c:680-754), from the point's E differences, with the reference's
// expression in its order.  Every kernel that updates H calls these, so all of them
// give the same bits.
__device__ __forceinline__ void hx_point(float &hxv, float dz_ey, float dy_ez, float hy, float hz, int sy, float by,
                                         float cy, bool inz, float bz, float cz, float dth, float &py0, float &pz0) {
    float h = hxv - (dy_ez * hy - dz_ey * hz);
    if (sy >= 0) {
        const float v = by * py0 + cy * dy_ez;
        py0 = v;
        h -= dth * v;
    }
    if (inz) {
        const float v = bz * pz0 + cz * dz_ey;
        pz0 = v;
        h += dth * v;
    }
    hxv = h;
}

__device__ __forceinline__ void hy_point(float &hyv, float dz_ex, float dx_ez, float hx, float hz, int sx, float bx,
                                         float cx, bool inz, float bz, float cz, float dth, float &px1, float &pz1) {
    float h = hyv - (dz_ex * hz - dx_ez * hx);
    if (sx >= 0) {
        const float v = bx * px1 + cx * dx_ez;
        px1 = v;
        h += dth * v;
    }
    if (inz) {
        const float v = bz * pz1 + cz * dz_ex;
        pz1 = v;
        h -= dth * v;
    }
    hyv = h;
}

__device__ __forceinline__ void hz_point(float &hzv, float dx_ey, float dy_ex, float hx, float hy, int sx, float bx,
                                         float cx, int sy, float by, float cy, float dth, float &px2, float &py2) {
    float h = hzv - (dx_ey * hx - dy_ex * hy);
    if (sx >= 0) {
        const float v = bx * px2 + cx * dx_ey;
        px2 = v;
        h -= dth * v;
    }
    if (sy >= 0) {
        const float v = by * py2 + cy * dy_ex;
        py2 = v;
        h += dth * v;
    }
    hzv = h;
}

// One H point: each component only where the reference updates it (Hx: j < ny, k < nz;
// Hy: i < nx, k < nz; Hz: i < nx, j < ny).
__device__ __forceinline__ void h_point(float &hxv, float &hyv, float &hzv, float dz_ey, float dy_ez, float dz_ex,
                                        float dy_ex, float dx_ez, float dx_ey, float hx, float hy, float hz, bool jin,
                                        bool kin, bool iin, int sy, float by, float cy, int sx, float bx, float cx,
                                        bool inz, float bz, float cz, float dth, float &py0, float &py2, float &px1,
                                        float &px2, float &pz0, float &pz1) {
    if (jin && kin) hx_point(hxv, dz_ey, dy_ez, hy, hz, sy, by, cy, inz, bz, cz, dth, py0, pz0);
    if (iin && kin) hy_point(hyv, dz_ex, dx_ez, hx, hz, sx, bx, cx, inz, bz, cz, dth, px1, pz1);
    if (iin && jin) hz_point(hzv, dx_ey, dy_ex, hx, hy, sx, bx, cx, sy, by, cy, dth, px2, py2);
}

// H update (sim.c:579-606) fused with the CPML psi_H recursions (sim.c:680-754).
// One thread per group of an i-segment; the k + 1 neighbour of the group's last
// point comes from one 32-bit load, the E of plane i+1 is carried in registers
// into the next iteration.  Each point is updated with the reference's
// expression in the reference's order.
template <int W>
__global__ void __launch_bounds__(kBlock) k_update_h(HArgs a, int i0, int i1, int iseg, int npblk) {
    const Grid g = a.g;
    const int ngw = g.kp / W;
    int tile = a.xcd ? xcd_tile(blockIdx.x, gridDim.x) : chunk_tile(blockIdx.x, gridDim.x, a.chunk);
    // reversed, the H sweep starts on the planes the E sweep before it touched last (still
    // in the MALL) and ends on the planes the E sweep after it starts with
    if (a.rev) tile = (int)gridDim.x - 1 - tile;
    if (a.pair) {   // the two right-hand sides of a tile are neighbouring tiles (one XCD)
        if (tile & 1) {
            a.E = a.E2;
            a.H = a.H2;
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int w = 0; w < 2; w++) a.psi[c][w] = a.psi2[c][w];
        }
        tile >>= 1;
    }
    const int pblk = tile % npblk, seg = tile / npblk;
    const int p = pblk * kBlock + threadIdx.x;
    if (p >= (g.ny + 1) * ngw) return;
    const int j = p / ngw, gk = p - j * ngw, k0 = W * gk;
    const int ib = i0 + seg * iseg;
    const int ie = min(i1, ib + iseg);
    const float *__restrict__ Ex = a.E;
    const float *__restrict__ Ey = a.E + g.N;
    const float *__restrict__ Ez = a.E + 2 * g.N;
    float *__restrict__ Hx = a.H;
    float *__restrict__ Hy = a.H + g.N;
    float *__restrict__ Hz = a.H + 2 * g.N;
    const bool jin = j < g.ny;
    const float hy = jin ? a.hs[1][j] : 0.f;
    bool kin[W], inz[W];
    float hz[W], bz[W], cz[W];
    const int sy = a.has_pml && jin ? a.pml.slot[1][j] : -1;
    const float by = sy >= 0 ? a.pml.b[1][j] : 0.f, cy = sy >= 0 ? a.pml.c[1][j] : 0.f;
    const int zo = a.has_pml ? a.pml.zo[k0 / 4] : -1;
#pragma unroll
    for (int e = 0; e < W; e++) {
        const int k = k0 + e;
        kin[e] = k < g.nz;
        hz[e] = kin[e] ? a.hs[2][k] : 0.f;
        inz[e] = zo >= 0 && kin[e] && a.pml.slot[2][k] >= 0;
        bz[e] = inz[e] ? a.pml.b[2][k] : 0.f;
        cz[e] = inz[e] ? a.pml.c[2][k] : 0.f;
    }
    // byte offsets of the group within a component array (32 bits: 12 N bytes < 4 GiB, set-up
    // checks), added to scalar base pointers by the loads themselves
    const unsigned si4 = 4u * (unsigned)g.si, sj4 = 4u * (unsigned)g.sj;
    const unsigned kn = k0 + W < g.kp ? 4u * W : 0u;   // on to the next group's first point
    unsigned q = 4u * ((unsigned)ib * (unsigned)g.si + (unsigned)j * (unsigned)g.sj + (unsigned)k0);
    float ey[W], ez[W];
    ldb<W>(Ey, q, ey);
    ldb<W>(Ez, q, ez);
    for (int i = ib; i < ie; i++, q += si4) {
        const bool iin = i < g.nx;
        float ex[W], ezj[W], exj[W], eyi[W], ezi[W], hxv[W], hyv[W], hzv[W];
        ldb<W>(Ex, q, ex);
        const float eyk = lb(Ey, q + kn), exk = lb(Ex, q + kn);
        if (jin) {
            ldb<W>(Ez, q + sj4, ezj);
            ldb<W>(Ex, q + sj4, exj);
        } else {
            zerow<W>(ezj);
            zerow<W>(exj);
        }
        if (iin) {
            ldb<W>(Ey, q + si4, eyi);
            ldb<W>(Ez, q + si4, ezi);
        } else {
            zerow<W>(eyi);
            zerow<W>(ezi);
        }
        const float hx = iin ? a.hs[0][i] : 0.f;
        const int sx = a.has_pml && iin ? a.pml.slot[0][i] : -1;
        ldb<W>(Hx, q, hxv);
        ldb<W>(Hy, q, hyv);
        ldb<W>(Hz, q, hzv);
        // psi of the slabs that hold this group (x: the plane, y: the row, z: the lane)
        float py0[W], py2[W], px1[W], px2[W], pz0[W], pz1[W];
        float bx = 0.f, cx = 0.f;
        unsigned oy = 0, ox = 0, oz = 0;   // byte offsets in the psi arrays
        if (sy >= 0) {
            oy = 4u * (((unsigned)sy * (unsigned)(g.nx + 1) + (unsigned)i) * (unsigned)g.kp + (unsigned)k0);
            ldb<W>(a.psi[0][0], oy, py0);
            ldb<W>(a.psi[2][1], oy, py2);
        }
        if (sx >= 0) {
            ox = (unsigned)sx * si4 + (unsigned)j * sj4 + 4u * (unsigned)k0;
            ldb<W>(a.psi[1][1], ox, px1);
            ldb<W>(a.psi[2][0], ox, px2);
            bx = a.pml.b[0][i];
            cx = a.pml.c[0][i];
        }
        if (zo >= 0) {
            oz = 4u * ((((unsigned)i * (unsigned)(g.ny + 1) + (unsigned)j) * (unsigned)a.pml.zg + (unsigned)zo) * 4u +
                       (unsigned)(k0 & 3));
            ldb<W>(a.psi[0][1], oz, pz0);
            ldb<W>(a.psi[1][0], oz, pz1);
        }
#pragma unroll
        for (int e = 0; e < W; e++) {
            const float ey_k = e + 1 < W ? ey[(e + 1) % W] : eyk, ex_k = e + 1 < W ? ex[(e + 1) % W] : exk;
            h_point(hxv[e], hyv[e], hzv[e], ey_k - ey[e], ezj[e] - ez[e], ex_k - ex[e], exj[e] - ex[e],
                    ezi[e] - ez[e], eyi[e] - ey[e], hx, hy, hz[e], jin, kin[e], iin, sy, by, cy, sx, bx, cx, inz[e],
                    bz[e], cz[e], a.dth, py0[e], py2[e], px1[e], px2[e], pz0[e], pz1[e]);
        }
        if (jin) stb<W>(Hx, q, hxv);
        if (iin) stb<W>(Hy, q, hyv);
        if (iin && jin) stb<W>(Hz, q, hzv);
        // psi entries of points not updated are stored back unchanged
        if (sy >= 0) {
            stb<W>(a.psi[0][0], oy, py0);
            stb<W>(a.psi[2][1], oy, py2);
        }
        if (sx >= 0) {
            stb<W>(a.psi[1][1], ox, px1);
            stb<W>(a.psi[2][0], ox, px2);
        }
        if (zo >= 0) {
            stb<W>(a.psi[0][1], oz, pz0);
            stb<W>(a.psi[1][0], oz, pz1);
        }
#pragma unroll
        for (int e = 0; e < W; e++) {
            ey[e] = eyi[e];
            ez[e] = ezi[e];
        }
    }
}

// One E point (sim.c:608-637) with the CPML recursions of the slabs that hold it
// (sim.c:756-832), from the point's six H differences; each component only where
// the reference updates it.  Shared by the E sweep and the one-pass step kernel.
__device__ __forceinline__ void e_point(float &exn, float &eyn, float &ezn, float ex0, float ey0, float ez0, float cax,
                                        float cbx, float cay, float cby, float caz, float cbz, float dy_hz, float dz_hy,
                                        float dz_hx, float dx_hz, float dx_hy, float dy_hx, float ex_s, float ey_s,
                                        float ez_s, bool jb, bool jlt, bool ib_, bool kb, bool kz, int sy, float by,
                                        float cy, int sx, float bx, float cx, bool inz, float bz, float cz, float &py0,
                                        float &py2, float &px1, float &px2, float &pz0, float &pz1) {
    if (jb && kb) {   // Ex
        float v0 = cax * ex0 + cbx * (dy_hz * ey_s - dz_hy * ez_s);
        if (sy >= 0) {
            const float v = by * py0 + cy * dy_hz;
            py0 = v;
            v0 += cbx * v;
        }
        if (inz) {
            const float v = bz * pz0 + cz * dz_hy;
            pz0 = v;
            v0 -= cbx * v;
        }
        exn = v0;
    }
    if (jlt && ib_ && kb) {   // Ey
        float v0 = cay * ey0 + cby * (dz_hx * ez_s - dx_hz * ex_s);
        if (sx >= 0) {
            const float v = bx * px1 + cx * dx_hz;
            px1 = v;
            v0 -= cby * v;
        }
        if (inz) {
            const float v = bz * pz1 + cz * dz_hx;
            pz1 = v;
            v0 += cby * v;
        }
        eyn = v0;
    }
    if (kz && ib_ && jb) {   // Ez
        float v0 = caz * ez0 + cbz * (dx_hy * ex_s - dy_hx * ey_s);
        if (sx >= 0) {
            const float v = bx * px2 + cx * dx_hy;
            px2 = v;
            v0 += cbz * v;
        }
        if (sy >= 0) {
            const float v = by * py2 + cy * dy_hx;
            py2 = v;
            v0 -= cbz * v;
        }
        ezn = v0;
    }
}

// Debye ADE of one dispersive edge of a class with material row bt (beta_p / dt):
// the correction E -= cb sum (1+k_p)/2 J_p (dispersive.c:93-100) after the bulk
// update, then J_p += (k_p - 1) J_p + beta_p/dt (E_new - E_old) (dispersive.c:102-112).
template <int NP>
__device__ __forceinline__ void debye_point(float &en, float e0, float cb, float *jp, const float *bt,
                                            const float *kc, const float *km1) {
    float acc = 0.f;
#pragma unroll
    for (int pp = 0; pp < NP; pp++) acc += kc[pp] * jp[pp];
    const float ev = en - cb * acc;
    en = ev;
    const float de = ev - e0;
#pragma unroll
    for (int pp = 0; pp < NP; pp++) jp[pp] = jp[pp] + km1[pp] * jp[pp] + bt[pp] * de;
}

// E update (sim.c:608-637) fused with CPML psi_E (sim.c:756-832) and, for NP > 0,
// with the Debye ADE of the dispersive edges whose class carries a material row:
// the correction E -= cb sum (1+k_p)/2 J_p (dispersive.c:93-100) after the bulk
// update and the current update J_p += (k_p - 1) J_p + beta_p/dt (E_new - E_old)
// (dispersive.c:102-112), with E_old the value this kernel has just read, in the
// reference's order and with the reference's expressions.  Same lane shape as
// the H update; the H of plane i-1 is carried in registers.  The (ca, cb) class
// table, the classes' Debye rows and the material rows sit in LDS.
// CMP: 0 coefficient arrays, 1 class words (cidx), 2 class bytes (pidx -> ptab -> class word)
template <int NP, int CMP, int W>
__global__ void __launch_bounds__(kBlock) k_update_e(EArgs a, int i0, int i1, int iseg, int npblk) {
    constexpr bool kCmp = CMP > 0;
    __shared__ unsigned s_ptab[CMP == 2 ? 256 : 1];
    __shared__ float2 s_tab[kCmp ? 768 : 1];
    __shared__ int s_row[(kCmp && NP > 0) ? 768 : 1];
    __shared__ float s_bt[(kCmp && NP > 0) ? kDebyeRows * NP : 1];
    if (kCmp) {
        for (int t = threadIdx.x; t < 3 * a.ncls; t += kBlock) {   // the classes in use only
            const int c = t / a.ncls, o = c * 256 + (t - c * a.ncls);
            s_tab[o] = a.ctab[o];
            if (NP > 0) s_row[o] = a.crow[o];
        }
        if (CMP == 2)
            for (int t = threadIdx.x; t < a.nptab; t += kBlock) s_ptab[t] = a.ptab[t];
        if (NP > 0)
            for (int t = threadIdx.x; t < a.nrow * NP; t += kBlock) s_bt[t] = a.btab[t];
        __syncthreads();
    }
    const Grid g = a.g;
    const int ngw = g.kp / W;
    int tile = a.xcd ? xcd_tile(blockIdx.x, gridDim.x) : chunk_tile(blockIdx.x, gridDim.x, a.chunk);
    if (a.pair) {   // the two right-hand sides of a tile are neighbouring tiles (one XCD)
        if (tile & 1) {
            a.E = a.E2;
            a.H = a.H2;
            a.J = a.J2;
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int w = 0; w < 2; w++) a.psi[c][w] = a.psi2[c][w];
        }
        tile >>= 1;
    }
    const int pblk = tile % npblk, seg = tile / npblk;
    const int p = pblk * kBlock + threadIdx.x;
    if (p >= (g.ny + 1) * ngw) return;
    const int j = p / ngw, gk = p - j * ngw, k0 = W * gk;
    const int ib = i0 + seg * iseg;
    const int ie = min(i1, ib + iseg);
    float *__restrict__ Ex = a.E;
    float *__restrict__ Ey = a.E + g.N;
    float *__restrict__ Ez = a.E + 2 * g.N;
    const float *__restrict__ Hx = a.H;
    const float *__restrict__ Hy = a.H + g.N;
    const float *__restrict__ Hz = a.H + 2 * g.N;
    const bool jb = j >= 1 && j < g.ny;
    const float ey_s = a.es[1][j];
    bool kb[W], kz[W], inz[W];
    float ez_s[W], bz[W], cz[W];
    const int sy = a.has_pml ? a.pml.slot[1][j] : -1;
    const float by = sy >= 0 ? a.pml.b[1][j] : 0.f, cy = sy >= 0 ? a.pml.c[1][j] : 0.f;
    const int zo = a.has_pml ? a.pml.zo[k0 / 4] : -1;
#pragma unroll
    for (int e = 0; e < W; e++) {
        const int k = k0 + e;
        kb[e] = k >= 1 && k < g.nz;
        kz[e] = k < g.nz;
        ez_s[e] = k <= g.nz ? a.es[2][k] : 0.f;
        inz[e] = zo >= 0 && kz[e] && a.pml.slot[2][k] >= 0;
        bz[e] = inz[e] ? a.pml.b[2][k] : 0.f;
        cz[e] = inz[e] ? a.pml.c[2][k] : 0.f;
    }
    const bool box_jk = NP > 0 && j >= a.bj0 && j < a.bj1 && k0 >= 4 * a.bg0 && k0 < 4 * a.bg1;
    // byte offsets of the group within a component array (32 bits: 12 N bytes < 4 GiB, set-up
    // checks), added to scalar base pointers by the loads themselves
    const unsigned si4 = 4u * (unsigned)g.si, sj4 = 4u * (unsigned)g.sj;
    const unsigned km = k0 ? 4u : 0u;   // back to the previous group's last point
    unsigned q = 4u * ((unsigned)ib * (unsigned)g.si + (unsigned)j * (unsigned)g.sj + (unsigned)k0);
    float hym[W], hzm[W];
    if (ib >= 1) {
        ldb<W>(Hy, q - si4, hym);
        ldb<W>(Hz, q - si4, hzm);
    } else {
        zerow<W>(hym);
        zerow<W>(hzm);
    }
    for (int i = ib; i < ie; i++, q += si4) {
        const bool ib_ = i >= 1 && i < g.nx;
        // Debye currents of the group first, every component and pole at once, so
        // that their loads are in flight together with the field loads
        const bool in_box = kCmp && NP > 0 && box_jk && i >= a.bi0 && i < a.bi1;
        constexpr int NPE = NP > 0 ? NP : 1;          // poles per point
        float jv[3][W * NPE];                         // [comp][point e * NPE + pole]
        const unsigned bkp = 4u * (unsigned)(a.bg1 - a.bg0);   // pole stride in the box row (floats)
        unsigned jo = 0;                              // byte offset of the group in its (i, j) box row
        if (in_box) {
            jo = 4u * (((unsigned)(i - a.bi0) * (unsigned)(a.bj1 - a.bj0) + (unsigned)(j - a.bj0)) * 3u * NPE * bkp +
                       (unsigned)(k0 - 4 * a.bg0));
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int pp = 0; pp < NPE; pp++) {
                    float t[W];
                    ldb<W>(a.J + (c * NPE + pp) * bkp, jo, t);
#pragma unroll
                    for (int e = 0; e < W; e++) jv[c][e * NPE + pp] = t[e];
                }
        }
        float hx[W], hy[W], hz[W], hxj[W], hzj[W];
        ldb<W>(Hx, q, hx);
        ldb<W>(Hy, q, hy);
        ldb<W>(Hz, q, hz);
        const float hxk = lb(Hx, q - km), hyk = lb(Hy, q - km);
        if (j >= 1) {
            ldb<W>(Hx, q - sj4, hxj);
            ldb<W>(Hz, q - sj4, hzj);
        } else {
            zerow<W>(hxj);
            zerow<W>(hzj);
        }
        const float ex_s = a.es[0][i];
        const int sx = a.has_pml ? a.pml.slot[0][i] : -1;
        unsigned cls[W];
        float ca[3][W], cb[3][W];
        if constexpr (CMP == 2) {
            const unsigned char *pb = a.pidx + (q >> 2);   // q / 4: the point index
            if constexpr (W == 4) {
                const uchar4 b4 = *(const uchar4 *)pb;
                cls[0] = s_ptab[b4.x];
                cls[1 % W] = s_ptab[b4.y];
                cls[2 % W] = s_ptab[b4.z];
                cls[3 % W] = s_ptab[b4.w];
            } else if constexpr (W == 2) {
                const uchar2 b2 = *(const uchar2 *)pb;
                cls[0] = s_ptab[b2.x];
                cls[1 % W] = s_ptab[b2.y];
            } else {
                cls[0] = s_ptab[*pb];
            }
        } else if constexpr (CMP == 1) {
            if constexpr (W == 4) {
                const uint4 c4 = *(const uint4 *)((const char *)a.cidx + q);
                cls[0] = c4.x;
                cls[1 % W] = c4.y;
                cls[2 % W] = c4.z;
                cls[3 % W] = c4.w;
            } else if constexpr (W == 2) {
                const uint2 c2 = *(const uint2 *)((const char *)a.cidx + q);
                cls[0] = c2.x;
                cls[1 % W] = c2.y;
            } else {
                cls[0] = *(const unsigned *)((const char *)a.cidx + q);
            }
        }
        if constexpr (kCmp) {
#pragma unroll
            for (int e = 0; e < W; e++)
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const float2 t = s_tab[c * 256 + ((cls[e] >> (8 * c)) & 255u)];
                    ca[c][e] = t.x;
                    cb[c][e] = t.y;
                }
        } else {
#pragma unroll
            for (int c = 0; c < 3; c++) {
                ldb<W>(a.ca + c * g.N, q, ca[c]);
                ldb<W>(a.cb + c * g.N, q, cb[c]);
            }
        }
        float ex0[W], ey0[W], ez0[W];
        ldb<W>(Ex, q, ex0);
        ldb<W>(Ey, q, ey0);
        ldb<W>(Ez, q, ez0);
        float exn[W], eyn[W], ezn[W];
#pragma unroll
        for (int e = 0; e < W; e++) {
            exn[e] = ex0[e];
            eyn[e] = ey0[e];
            ezn[e] = ez0[e];
        }
        float py0[W], py2[W], px1[W], px2[W], pz0[W], pz1[W];
        float bx = 0.f, cx = 0.f;
        unsigned oy = 0, ox = 0, oz = 0;   // byte offsets in the psi arrays
        if (sy >= 0) {
            oy = 4u * (((unsigned)sy * (unsigned)(g.nx + 1) + (unsigned)i) * (unsigned)g.kp + (unsigned)k0);
            ldb<W>(a.psi[0][0], oy, py0);
            ldb<W>(a.psi[2][1], oy, py2);
        }
        if (sx >= 0) {
            ox = (unsigned)sx * si4 + (unsigned)j * sj4 + 4u * (unsigned)k0;
            ldb<W>(a.psi[1][1], ox, px1);
            ldb<W>(a.psi[2][0], ox, px2);
            bx = a.pml.b[0][i];
            cx = a.pml.c[0][i];
        }
        if (zo >= 0) {
            oz = 4u * ((((unsigned)i * (unsigned)(g.ny + 1) + (unsigned)j) * (unsigned)a.pml.zg + (unsigned)zo) * 4u +
                       (unsigned)(k0 & 3));
            ldb<W>(a.psi[0][1], oz, pz0);
            ldb<W>(a.psi[1][0], oz, pz1);
        }
#pragma unroll
        for (int e = 0; e < W; e++) {
            const float hx_k = e ? hx[(e + W - 1) % W] : hxk, hy_k = e ? hy[(e + W - 1) % W] : hyk;
            e_point(exn[e], eyn[e], ezn[e], ex0[e], ey0[e], ez0[e], ca[0][e], cb[0][e], ca[1][e], cb[1][e], ca[2][e],
                    cb[2][e], hz[e] - hzj[e], hy[e] - hy_k, hx[e] - hx_k, hz[e] - hzm[e], hy[e] - hym[e],
                    hx[e] - hxj[e], ex_s, ey_s, ez_s[e], jb, j < g.ny, ib_, kb[e], kz[e], sy, by, cy, sx, bx, cx,
                    inz[e], bz[e], cz[e], py0[e], py2[e], px1[e], px2[e], pz0[e], pz1[e]);
        }
        if (in_box) {
#pragma unroll
            for (int c = 0; c < 3; c++) {
                float *en = c == 0 ? exn : c == 1 ? eyn : ezn;
                const float *e0 = c == 0 ? ex0 : c == 1 ? ey0 : ez0;
#pragma unroll
                for (int e = 0; e < W; e++) {
                    const int row = s_row[c * 256 + ((cls[e] >> (8 * c)) & 255u)];
                    if (row >= 0) debye_point<NP>(en[e], e0[e], cb[c][e], jv[c] + e * NPE, s_bt + row * NP, a.kc, a.km1);
                }
#pragma unroll
                for (int pp = 0; pp < NPE; pp++) {
                    float t[W];
#pragma unroll
                    for (int e = 0; e < W; e++) t[e] = jv[c][e * NPE + pp];
                    stb<W>(a.J + (c * NPE + pp) * bkp, jo, t);
                }
            }
        }
        stb<W>(Ex, q, exn);
        if (j < g.ny) stb<W>(Ey, q, eyn);
        stb<W>(Ez, q, ezn);
        if (sy >= 0) {
            stb<W>(a.psi[0][0], oy, py0);
            stb<W>(a.psi[2][1], oy, py2);
        }
        if (sx >= 0) {
            stb<W>(a.psi[1][1], ox, px1);
            stb<W>(a.psi[2][0], ox, px2);
        }
        if (zo >= 0) {
            stb<W>(a.psi[0][1], oz, pz0);
            stb<W>(a.psi[1][0], oz, pz1);
        }
#pragma unroll
        for (int e = 0; e < W; e++) {
            hym[e] = hy[e];
            hzm[e] = hz[e];
        }
    }
}

// One-pass step: the H update of an x plane and then the E update of the same plane
// in one sweep over x, so each field component is read once and written once per
// step (the two sweeps above read E and H twice between them).  The old fields and
// psi_H are read from one buffer set and the new ones written to the other; psi_E
// and the Debye currents, which only their own point reads, are updated in place.
// A workgroup owns a tile of tj rows (j0 .. j0+tj-1) by kw points of k (k0 ..
// k0+kw-1) and walks iseg planes of x.  Its lanes hold one point each, on tj + 2
// rows of kw + 2 columns:
//   row 0 / column 0          (j0 - 1, k0 - 1) H only, for the j - 1 and k - 1
//                             neighbours of the E update
//   rows 1 .. tj, columns 1 .. kw   owned: H, E, their stores
//   row tj + 1 / column kw + 1      publishes its E, the j + 1 / k + 1 neighbour
//                             of the H update
// The E of the plane and the new H pass between lanes through LDS; the H of plane
// i - 1 stays in registers, and a workgroup starting inside x computes that plane's
// Hy, Hz at its points first.  Every point is updated by h_point / e_point /
// debye_point, the functions of the two sweeps, so the two paths give the same bits.
// Measured on the 38.8 M-point four-layer board it moves 3.4 GB a step (the sweeps
// 4.3 GB) but takes 1.36 ms against their 0.91 ms: at 113 VGPRs one 1024-lane
// workgroup fills a CU, its two barriers a plane leave the loads of each plane
// unhidden, and it issues 2.4x the sweeps' VALU work (SGPR spills among it).
struct SArgs {
    HArgs h;                // old E, H (h.E, h.H) and old psi_H (h.psi)
    EArgs e;                // coefficients, psi_E, Debye box (e.E / e.H unused)
    float *En, *Hn;         // new E (3N), H (3N)
    float *psiHn[3][2];     // new psi_H
    int tj, ntj;            // owned rows per tile, tiles over j
    int kw, ntk, kl;        // owned k points per tile, tiles over k, lanes per row (kw + 2)
};

template <int NP, bool kCmp>
__global__ void __launch_bounds__(1024) k_step(SArgs f, int iseg) {
    extern __shared__ float lds[];
    const Grid g = f.h.g;
    const int tj = f.tj, kp = g.kp, kl = f.kl;
    float *sE = lds;                                    // [3][tj + 2][kl]
    float *sH = lds + 3 * (tj + 2) * kl;                // [3][tj + 1][kl]
    float2 *s_tab = (float2 *)(sH + 3 * (tj + 1) * kl); // kCmp: 3 x 256 (ca, cb)
    int *s_row = (int *)(s_tab + (kCmp ? 768 : 0));     // kCmp, NP > 0: 3 x 256 Debye rows
    float *s_bt = (float *)(s_row + ((kCmp && NP > 0) ? 768 : 0));   // nrow x NP
    const EArgs &a = f.e;
    if (kCmp) {
        for (int t = threadIdx.x; t < 3 * a.ncls; t += blockDim.x) {   // the classes in use only
            const int c = t / a.ncls, o = c * 256 + (t - c * a.ncls);
            s_tab[o] = a.ctab[o];
            if (NP > 0) s_row[o] = a.crow[o];
        }
        if (NP > 0)
            for (int t = threadIdx.x; t < a.nrow * NP; t += blockDim.x) s_bt[t] = a.btab[t];
    }
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int jt = tile % f.ntj, kt = (tile / f.ntj) % f.ntk, seg = tile / (f.ntj * f.ntk);
    const int r = threadIdx.x / kl, c = threadIdx.x - r * kl;
    const int j = jt * tj - 1 + r, kr = kt * f.kw - 1 + c;
    const bool lane = r < tj + 2 && j >= 0 && j <= g.ny && kr >= 0 && kr < kp;
    const bool hrow = lane && r <= tj && c <= f.kw;
    const bool own = lane && r >= 1 && r <= tj && c >= 1 && c <= f.kw;
    const int k = lane ? kr : 0;
    const int ib = seg * iseg, ie = min(g.nx + 1, ib + iseg);
    const float *__restrict__ Exo = f.h.E;
    const float *__restrict__ Eyo = f.h.E + g.N;
    const float *__restrict__ Ezo = f.h.E + 2 * g.N;
    const float *__restrict__ Hxo = f.h.H;
    const float *__restrict__ Hyo = f.h.H + g.N;
    const float *__restrict__ Hzo = f.h.H + 2 * g.N;
    // H-side constants of the lane (k_update_h at one point per lane)
    const Pml &ph = f.h.pml;
    const bool jin = j < g.ny, kin = k < g.nz;
    const float hy_s = lane && jin ? f.h.hs[1][j] : 0.f;
    const float hz_s = kin ? f.h.hs[2][k] : 0.f;
    const int syh = f.h.has_pml && lane && jin ? ph.slot[1][j] : -1;
    const float byh = syh >= 0 ? ph.b[1][j] : 0.f, cyh = syh >= 0 ? ph.c[1][j] : 0.f;
    const int zoh = f.h.has_pml ? ph.zo[k / 4] : -1;
    const bool inzh = zoh >= 0 && kin && ph.slot[2][k] >= 0;
    const float bzh = inzh ? ph.b[2][k] : 0.f, czh = inzh ? ph.c[2][k] : 0.f;
    const int cn = k + 1 < kp ? c + 1 : c;   // LDS column of the k + 1 neighbour
    // E-side constants (k_update_e at one point per lane; owned rows only)
    const Pml &pe = a.pml;
    const bool jb = j >= 1 && j < g.ny, kb = k >= 1 && k < g.nz, kz = k < g.nz;
    const float ey_s = own ? a.es[1][j] : 0.f;
    const float ez_s = k <= g.nz ? a.es[2][k] : 0.f;
    const int sye = a.has_pml && own ? pe.slot[1][j] : -1;
    const float bye = sye >= 0 ? pe.b[1][j] : 0.f, cye = sye >= 0 ? pe.c[1][j] : 0.f;
    const int zoe = a.has_pml ? pe.zo[k / 4] : -1;
    const bool inze = zoe >= 0 && kz && pe.slot[2][k] >= 0;
    const float bze = inze ? pe.b[2][k] : 0.f, cze = inze ? pe.c[2][k] : 0.f;
    const bool box_jk = NP > 0 && own && j >= a.bj0 && j < a.bj1 && k >= 4 * a.bg0 && k < 4 * a.bg1;
    const int cm = k ? c - 1 : c;            // of the k - 1 neighbour
    // 32-bit element offsets (3N < 2^32: set-up checks) keep addresses in SGPR base + VGPR offset form
    const unsigned si = (unsigned)g.si, sj = (unsigned)g.sj, N = (unsigned)g.N;
    unsigned q = (unsigned)ib * si + (unsigned)(lane ? j : 0) * sj + k;
    // E of plane ib at the lane's point; H of plane ib - 1 (Hy, Hz) at owned points
    float ex = 0.f, ey = 0.f, ez = 0.f, hym = 0.f, hzm = 0.f;
    if (lane) {
        ex = Exo[q];
        ey = Eyo[q];
        ez = Ezo[q];
    }
    if (own && ib >= 1) {
        const unsigned qm = q - si;
        const int im = ib - 1;
        const float exm = Exo[qm], eym = Eyo[qm], ezm = Ezo[qm];
        const float exk = Exo[k + 1 < kp ? qm + 1 : qm];
        const float exj = jin ? Exo[qm + sj] : 0.f;
        const int sx = f.h.has_pml ? ph.slot[0][im] : -1;
        float hxv = 0.f, hyv = Hyo[qm], hzv = Hzo[qm];
        float py0 = 0.f, py2 = 0.f, px1 = 0.f, px2 = 0.f, pz0 = 0.f, pz1 = 0.f, bx = 0.f, cx = 0.f;
        if (syh >= 0) py2 = f.h.psi[2][1][((unsigned)syh * (g.nx + 1) + im) * (unsigned)kp + k];
        if (sx >= 0) {
            const unsigned ox = (unsigned)sx * si + (unsigned)j * sj + k;
            px1 = f.h.psi[1][1][ox];
            px2 = f.h.psi[2][0][ox];
            bx = ph.b[0][im];
            cx = ph.c[0][im];
        }
        if (zoh >= 0) pz1 = f.h.psi[1][0][(((unsigned)im * (g.ny + 1) + j) * ph.zg + zoh) * 4 + (k & 3)];
        // Hx is not needed here: its E differences are passed as zero and its result dropped
        h_point(hxv, hyv, hzv, 0.f, 0.f, exk - exm, exj - exm, ez - ezm, ey - eym, f.h.hs[0][im], hy_s, hz_s,
                jin, kin, true, syh, byh, cyh, sx, bx, cx, inzh, bzh, czh, f.h.dth, py0, py2, px1, px2, pz0, pz1);
        hym = hyv;
        hzm = hzv;
    }
    const int rE = tj + 2, rH = tj + 1;
    for (int i = ib; i < ie; i++, q += si) {
        const bool iin = i < g.nx;
        // loads of this plane: E of plane i + 1, old H and psi_H, and the E side's inputs
        float exi = 0.f, eyi = 0.f, ezi = 0.f;
        if (lane && iin) {
            exi = Exo[q + si];
            eyi = Eyo[q + si];
            ezi = Ezo[q + si];
        }
        float hxv = 0.f, hyv = 0.f, hzv = 0.f;
        float hpy0 = 0.f, hpy2 = 0.f, hpx1 = 0.f, hpx2 = 0.f, hpz0 = 0.f, hpz1 = 0.f, hbx = 0.f, hcx = 0.f;
        const int sxh = f.h.has_pml && iin ? ph.slot[0][i] : -1;
        unsigned hoy = 0, hox = 0, hoz = 0;
        if (hrow) {
            hxv = Hxo[q];
            hyv = Hyo[q];
            hzv = Hzo[q];
            if (syh >= 0) {
                hoy = ((unsigned)syh * (g.nx + 1) + i) * (unsigned)kp + k;
                hpy0 = f.h.psi[0][0][hoy];
                hpy2 = f.h.psi[2][1][hoy];
            }
            if (sxh >= 0) {
                hox = (unsigned)sxh * si + (unsigned)j * sj + k;
                hpx1 = f.h.psi[1][1][hox];
                hpx2 = f.h.psi[2][0][hox];
                hbx = ph.b[0][i];
                hcx = ph.c[0][i];
            }
            if (zoh >= 0) {
                hoz = (((unsigned)i * (g.ny + 1) + j) * ph.zg + zoh) * 4 + (k & 3);
                hpz0 = f.h.psi[0][1][hoz];
                hpz1 = f.h.psi[1][0][hoz];
            }
        }
        // E-side inputs of the plane at owned points: classes or coefficients, psi_E,
        // Debye currents (box layout of EArgs)
        constexpr int NPE = NP > 0 ? NP : 1;
        unsigned cls = 0;
        float ca[3], cb[3], jv[3][NPE];
        float *Jl = nullptr;
        long long bkp = 0;
        bool in_box = false;
        int sx = -1;
        float py0 = 0.f, py2 = 0.f, px1 = 0.f, px2 = 0.f, pz0 = 0.f, pz1 = 0.f, bx = 0.f, cx = 0.f;
        unsigned oy = 0, ox = 0, oz = 0;
        auto load_e = [&]() {
            sx = a.has_pml ? pe.slot[0][i] : -1;
            if (kCmp) {
                cls = a.cidx[q];
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const float2 t = s_tab[c * 256 + ((cls >> (8 * c)) & 255u)];
                    ca[c] = t.x;
                    cb[c] = t.y;
                }
            } else {
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    ca[c] = a.ca[c * N + q];
                    cb[c] = a.cb[c * N + q];
                }
            }
            in_box = kCmp && box_jk && i >= a.bi0 && i < a.bi1;
            if (in_box) {
                bkp = 4LL * (a.bg1 - a.bg0);
                Jl = a.J + ((long long)(i - a.bi0) * (a.bj1 - a.bj0) + (j - a.bj0)) * 3 * NPE * bkp + (k - 4 * a.bg0);
#pragma unroll
                for (int c = 0; c < 3; c++)
#pragma unroll
                    for (int pp = 0; pp < NPE; pp++) jv[c][pp] = Jl[(c * NPE + pp) * bkp];
            }
            if (sye >= 0) {
                oy = ((unsigned)sye * (g.nx + 1) + i) * (unsigned)kp + k;
                py0 = a.psi[0][0][oy];
                py2 = a.psi[2][1][oy];
            }
            if (sx >= 0) {
                ox = (unsigned)sx * si + (unsigned)j * sj + k;
                px1 = a.psi[1][1][ox];
                px2 = a.psi[2][0][ox];
                bx = pe.b[0][i];
                cx = pe.c[0][i];
            }
            if (zoe >= 0) {
                oz = (((unsigned)i * (g.ny + 1) + j) * pe.zg + zoe) * 4 + (k & 3);
                pz0 = a.psi[0][1][oz];
                pz1 = a.psi[1][0][oz];
            }
        };
        if (lane) {
            sE[(0 * rE + r) * kl + c] = ex;
            sE[(1 * rE + r) * kl + c] = ey;
            sE[(2 * rE + r) * kl + c] = ez;
        }
        __syncthreads();   // E of plane i complete in LDS; last plane's H reads done
        if (hrow) {
            const float eyk = sE[(1 * rE + r) * kl + cn], exk = sE[(0 * rE + r) * kl + cn];
            const float ezj = jin ? sE[(2 * rE + r + 1) * kl + c] : 0.f;
            const float exj = jin ? sE[(0 * rE + r + 1) * kl + c] : 0.f;
            h_point(hxv, hyv, hzv, eyk - ey, ezj - ez, exk - ex, exj - ex, ezi - ez, eyi - ey,
                    iin ? f.h.hs[0][i] : 0.f, hy_s, hz_s, jin, kin, iin, syh, byh, cyh, sxh, hbx, hcx, inzh, bzh, czh,
                    f.h.dth, hpy0, hpy2, hpx1, hpx2, hpz0, hpz1);
            sH[(0 * rH + r) * kl + c] = hxv;
            sH[(1 * rH + r) * kl + c] = hyv;
            sH[(2 * rH + r) * kl + c] = hzv;
        }
        if (own) {
            f.Hn[q] = hxv;
            f.Hn[q + N] = hyv;
            f.Hn[q + 2 * N] = hzv;
            // every psi_H entry of the new set is written, updated or not
            if (syh >= 0) {
                f.psiHn[0][0][hoy] = hpy0;
                f.psiHn[2][1][hoy] = hpy2;
            }
            if (sxh >= 0) {
                f.psiHn[1][1][hox] = hpx1;
                f.psiHn[2][0][hox] = hpx2;
            }
            if (zoh >= 0) {
                f.psiHn[0][1][hoz] = hpz0;
                f.psiHn[1][0][hoz] = hpz1;
            }
        }
        __syncthreads();   // H of plane i complete in LDS
        if (own) {
            float exn = ex, eyn = ey, ezn = ez;
            if (iin) {
                const bool ib_ = i >= 1;
                const float hxk = sH[(0 * rH + r) * kl + cm], hyk = sH[(1 * rH + r) * kl + cm];
                const float hxj = j >= 1 ? sH[(0 * rH + r - 1) * kl + c] : 0.f;
                const float hzj = j >= 1 ? sH[(2 * rH + r - 1) * kl + c] : 0.f;
                const float ex_s = a.es[0][i];
                load_e();
                e_point(exn, eyn, ezn, ex, ey, ez, ca[0], cb[0], ca[1], cb[1], ca[2], cb[2], hzv - hzj, hyv - hyk,
                        hxv - hxk, hzv - hzm, hyv - hym, hxv - hxj, ex_s, ey_s, ez_s, jb, jin, ib_, kb, kz, sye, bye,
                        cye, sx, bx, cx, inze, bze, cze, py0, py2, px1, px2, pz0, pz1);
                if (in_box) {
#pragma unroll
                    for (int c = 0; c < 3; c++) {
                        float &en = c == 0 ? exn : c == 1 ? eyn : ezn;
                        const float e0 = c == 0 ? ex : c == 1 ? ey : ez;
                        const int row = s_row[c * 256 + ((cls >> (8 * c)) & 255u)];
                        if (row >= 0) debye_point<NP>(en, e0, cb[c], jv[c], s_bt + row * NP, a.kc, a.km1);
#pragma unroll
                        for (int pp = 0; pp < NPE; pp++) Jl[(c * NPE + pp) * bkp] = jv[c][pp];
                    }
                }
                if (sye >= 0) {
                    a.psi[0][0][oy] = py0;
                    a.psi[2][1][oy] = py2;
                }
                if (sx >= 0) {
                    a.psi[1][1][ox] = px1;
                    a.psi[2][0][ox] = px2;
                }
                if (zoe >= 0) {
                    a.psi[0][1][oz] = pz0;
                    a.psi[1][0][oz] = pz1;
                }
            }
            f.En[q] = exn;
            f.En[q + N] = eyn;
            f.En[q + 2 * N] = ezn;
            hym = hyv;
            hzm = hzv;
        }
        ex = exi;
        ey = eyi;
        ez = ezi;
    }
}

// One-pass step without barriers.  Each lane owns one point (j, k) of the (j, k)
// plane, walks an x segment, and at every plane computes the new H it needs itself:
// H at (i, j, k) in full (stored), and Hx, Hz at (i, j - 1, k) again (the row below
// is another lane's; recomputing its two components from the old fields costs VALU
// work but no memory traffic or synchronisation).  The H of (i, j, k - 1) comes from
// the lane below by a cross-lane shift: a wavefront's lanes hold 63 consecutive
// points of the flattened (j, k) plane after one halo lane, which computes the H of
// the point before them and owns nothing.  The H of plane i - 1 is carried in
// registers; a segment starting inside x computes that plane's Hy, Hz first.  Old E,
// H and psi_H are read from one buffer set and the new values written to the other
// (so no lane sees another's writes); psi_E and the Debye currents are updated in
// place by their owner.  Every value is computed by hx/hy/hz_point, e_point and
// debye_point, the functions of the two sweeps, so all paths give the same bits.
// Measured on the 38.8 M-point four-layer board: 1.02 ms a step against 0.93 ms for
// the two sweeps; it moves the one-pass step's bytes but issues ~1.7x the sweeps'
// instructions per point (the recomputed row, 64-bit address arithmetic, scalar
// register spills of its large argument block) at 127 VGPRs (4 waves per SIMD).
template <int NP, bool kCmp>
__global__ void __launch_bounds__(kBlock) k_step_rows(SArgs f, int iseg, int npblk) {
    __shared__ float2 s_tab[kCmp ? 768 : 1];
    __shared__ int s_row[(kCmp && NP > 0) ? 768 : 1];
    __shared__ float s_bt[(kCmp && NP > 0) ? kDebyeRows * NP : 1];
    const EArgs &a = f.e;
    if (kCmp) {
        for (int t = threadIdx.x; t < 3 * a.ncls; t += kBlock) {   // the classes in use only
            const int c = t / a.ncls, o = c * 256 + (t - c * a.ncls);
            s_tab[o] = a.ctab[o];
            if (NP > 0) s_row[o] = a.crow[o];
        }
        if (NP > 0)
            for (int t = threadIdx.x; t < a.nrow * NP; t += kBlock) s_bt[t] = a.btab[t];
        __syncthreads();
    }
    const Grid g = f.h.g;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int pblk = tile % npblk, seg = tile / npblk;
    const int lane = threadIdx.x & 63;
    const int P = (g.ny + 1) * g.kp;
    const int p0 = (pblk * (kBlock / 64) + (int)(threadIdx.x >> 6)) * 63;   // first owned point of the wave
    if (p0 >= P) return;                                                   // whole wavefronts only
    const int p = p0 + lane - 1;
    const bool valid = p >= 0 && p < P;
    const bool own = valid && lane >= 1;
    const int pc = valid ? p : 0;
    const int j = pc / g.kp, k = pc - j * g.kp;
    const int ib = seg * iseg, ie = min(g.nx + 1, ib + iseg);
    const float *__restrict__ Exo = f.h.E;
    const float *__restrict__ Eyo = f.h.E + g.N;
    const float *__restrict__ Ezo = f.h.E + 2 * g.N;
    const float *__restrict__ Hxo = f.h.H;
    const float *__restrict__ Hyo = f.h.H + g.N;
    const float *__restrict__ Hzo = f.h.H + 2 * g.N;
    const Pml &ph = f.h.pml;
    const Pml &pe = a.pml;
    // rows j (owned) and j - 1 (Hx, Hz recomputed); column k
    const bool jin = j < g.ny, kin = k < g.nz, jm_ok = j >= 1;
    const int jm = jm_ok ? j - 1 : 0;
    const float hy_s = jin ? f.h.hs[1][j] : 0.f, hy_m = jm_ok ? f.h.hs[1][jm] : 0.f;
    const float hz_s = kin ? f.h.hs[2][k] : 0.f;
    const int syh = f.h.has_pml && jin ? ph.slot[1][j] : -1;
    const float byh = syh >= 0 ? ph.b[1][j] : 0.f, cyh = syh >= 0 ? ph.c[1][j] : 0.f;
    const int sym = f.h.has_pml && jm_ok ? ph.slot[1][jm] : -1;
    const float bym = sym >= 0 ? ph.b[1][jm] : 0.f, cym = sym >= 0 ? ph.c[1][jm] : 0.f;
    const int zoh = f.h.has_pml ? ph.zo[k / 4] : -1;
    const bool inzh = zoh >= 0 && kin && ph.slot[2][k] >= 0;
    const float bzh = inzh ? ph.b[2][k] : 0.f, czh = inzh ? ph.c[2][k] : 0.f;
    const bool jb = j >= 1 && j < g.ny, kb = k >= 1 && k < g.nz, kz = k < g.nz;
    const float ey_s = a.es[1][j];
    const float ez_s = k <= g.nz ? a.es[2][k] : 0.f;
    const int sye = a.has_pml ? pe.slot[1][j] : -1;
    const float bye = sye >= 0 ? pe.b[1][j] : 0.f, cye = sye >= 0 ? pe.c[1][j] : 0.f;
    const int zoe = a.has_pml ? pe.zo[k / 4] : -1;
    const bool inze = zoe >= 0 && kz && pe.slot[2][k] >= 0;
    const float bze = inze ? pe.b[2][k] : 0.f, cze = inze ? pe.c[2][k] : 0.f;
    const bool box_jk = NP > 0 && own && j >= a.bj0 && j < a.bj1 && k >= 4 * a.bg0 && k < 4 * a.bg1;
    // 32-bit element offsets (3N < 2^32: set-up checks)
    const unsigned si = (unsigned)g.si, sj = (unsigned)g.sj, N = (unsigned)g.N;
    const unsigned kn = k + 1 < g.kp ? 1u : 0u;          // the k + 1 neighbour (itself at a row's end)
    const unsigned dm = jm_ok ? sj : 0u;                  // row j - 1 (row j itself at j = 0, unused)
    const unsigned dp = jin ? sj : 0u;                    // row j + 1 (row j itself at j = ny, unused)
    unsigned q = (unsigned)ib * si + (unsigned)j * sj + k;
    // H of plane ib - 1 (Hy, Hz) at the owned point, for the x differences of E at ib
    float hyp = 0.f, hzp = 0.f;
    if (own && ib >= 1) {
        const unsigned qm = q - si;
        const int im = ib - 1;
        const float exm = Exo[qm], eym = Eyo[qm], ezm = Ezo[qm];
        const float exk = Exo[qm + kn], exj = Exo[qm + dp];
        const float eyi = Eyo[q], ezi = Ezo[q];
        const int sx = f.h.has_pml ? ph.slot[0][im] : -1;
        const float bx = sx >= 0 ? ph.b[0][im] : 0.f, cx = sx >= 0 ? ph.c[0][im] : 0.f;
        const float hx_s = f.h.hs[0][im];
        hyp = Hyo[qm];
        hzp = Hzo[qm];
        float px1 = 0.f, px2 = 0.f, pz1 = 0.f, py2 = 0.f;
        if (syh >= 0) py2 = f.h.psi[2][1][((unsigned)syh * (g.nx + 1) + im) * (unsigned)g.kp + k];
        if (sx >= 0) {
            const unsigned ox = (unsigned)sx * si + (unsigned)j * sj + k;
            px1 = f.h.psi[1][1][ox];
            px2 = f.h.psi[2][0][ox];
        }
        if (zoh >= 0) pz1 = f.h.psi[1][0][(((unsigned)im * (g.ny + 1) + j) * ph.zg + zoh) * 4 + (k & 3)];
        if (kin) hy_point(hyp, exk - exm, ezi - ezm, hx_s, hz_s, sx, bx, cx, inzh, bzh, czh, f.h.dth, px1, pz1);
        if (jin) hz_point(hzp, eyi - eym, exj - exm, hx_s, hy_s, sx, bx, cx, syh, byh, cyh, f.h.dth, px2, py2);
    }
    // E of plane ib carried into the loop: Ey, Ez of row j, Ey of row j - 1
    float ey = Eyo[q], ez = Ezo[q], eym = Eyo[q - dm];
    for (int i = ib; i < ie; i++, q += si) {
        const bool iin = i < g.nx;
        const unsigned qm = q - dm;
        // Debye currents first, so that their loads are in flight with the field loads
        const bool in_box = kCmp && NP > 0 && box_jk && iin && i >= a.bi0 && i < a.bi1;
        constexpr int NPE = NP > 0 ? NP : 1;
        float jv[3][NPE];
        float *Jl = nullptr;
        long long bkp = 0;
        if (in_box) {
            bkp = 4 * (a.bg1 - a.bg0);
            Jl = a.J + ((long long)(i - a.bi0) * (a.bj1 - a.bj0) + (j - a.bj0)) * 3 * NPE * bkp + (k - 4 * a.bg0);
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int pp = 0; pp < NPE; pp++) jv[c][pp] = Jl[(c * NPE + pp) * bkp];
        }
        // old E of planes i (rows j - 1, j, j + 1; k + 1) and i + 1 (rows j - 1, j)
        const float ex = Exo[q], exk = Exo[q + kn], eyk = Eyo[q + kn];
        const float exm = Exo[qm], ezm = Ezo[qm], eykm = Eyo[qm + kn];
        const float exj = Exo[q + dp], ezj = Ezo[q + dp];
        float eyi = 0.f, ezi = 0.f, eyim = 0.f;
        if (iin) {
            eyi = Eyo[q + si];
            ezi = Ezo[q + si];
            eyim = Eyo[qm + si];
        }
        // old H of rows j and j - 1
        float hxv = Hxo[q], hyv = Hyo[q], hzv = Hzo[q];
        float hxm = Hxo[qm], hzm = Hzo[qm];
        const float hx_s = iin ? f.h.hs[0][i] : 0.f;
        const int sxh = f.h.has_pml && iin ? ph.slot[0][i] : -1;
        const float bxh = sxh >= 0 ? ph.b[0][i] : 0.f, cxh = sxh >= 0 ? ph.c[0][i] : 0.f;
        // old psi_H of rows j and j - 1
        float hpy0 = 0.f, hpy2 = 0.f, hpx1 = 0.f, hpx2 = 0.f, hpz0 = 0.f, hpz1 = 0.f;
        float mpy0 = 0.f, mpy2 = 0.f, mpx2 = 0.f, mpz0 = 0.f;
        unsigned hoy = 0, hox = 0, hoz = 0;
        if (syh >= 0) {
            hoy = ((unsigned)syh * (g.nx + 1) + i) * (unsigned)g.kp + k;
            hpy0 = f.h.psi[0][0][hoy];
            hpy2 = f.h.psi[2][1][hoy];
        }
        if (sym >= 0) {
            const unsigned o = ((unsigned)sym * (g.nx + 1) + i) * (unsigned)g.kp + k;
            mpy0 = f.h.psi[0][0][o];
            mpy2 = f.h.psi[2][1][o];
        }
        if (sxh >= 0) {
            hox = (unsigned)sxh * si + (unsigned)j * sj + k;
            hpx1 = f.h.psi[1][1][hox];
            hpx2 = f.h.psi[2][0][hox];
            mpx2 = f.h.psi[2][0][hox - dm];
        }
        if (zoh >= 0) {
            hoz = (((unsigned)i * (g.ny + 1) + j) * ph.zg + zoh) * 4 + (k & 3);
            hpz0 = f.h.psi[0][1][hoz];
            hpz1 = f.h.psi[1][0][hoz];
            mpz0 = f.h.psi[0][1][hoz - (jm_ok ? (unsigned)ph.zg * 4 : 0u)];
        }
        // E-side inputs: classes or coefficients, psi_E
        const int sx = a.has_pml && iin ? pe.slot[0][i] : -1;
        unsigned cls = 0;
        float ca[3], cb[3];
        if (kCmp) {
            cls = a.cidx[q];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const float2 t = s_tab[c * 256 + ((cls >> (8 * c)) & 255u)];
                ca[c] = t.x;
                cb[c] = t.y;
            }
        } else {
#pragma unroll
            for (int c = 0; c < 3; c++) {
                ca[c] = a.ca[c * N + q];
                cb[c] = a.cb[c * N + q];
            }
        }
        float py0 = 0.f, py2 = 0.f, px1 = 0.f, px2 = 0.f, pz0 = 0.f, pz1 = 0.f, bx = 0.f, cx = 0.f;
        unsigned oy = 0, ox = 0, oz = 0;
        if (iin) {
            if (sye >= 0) {
                oy = ((unsigned)sye * (g.nx + 1) + i) * (unsigned)g.kp + k;
                py0 = a.psi[0][0][oy];
                py2 = a.psi[2][1][oy];
            }
            if (sx >= 0) {
                ox = (unsigned)sx * si + (unsigned)j * sj + k;
                px1 = a.psi[1][1][ox];
                px2 = a.psi[2][0][ox];
                bx = pe.b[0][i];
                cx = pe.c[0][i];
            }
            if (zoe >= 0) {
                oz = (((unsigned)i * (g.ny + 1) + j) * pe.zg + zoe) * 4 + (k & 3);
                pz0 = a.psi[0][1][oz];
                pz1 = a.psi[1][0][oz];
            }
        }
        // new H of row j, and Hx, Hz of row j - 1
        h_point(hxv, hyv, hzv, eyk - ey, ezj - ez, exk - ex, exj - ex, ezi - ez, eyi - ey, hx_s, hy_s, hz_s, jin, kin,
                iin, syh, byh, cyh, sxh, bxh, cxh, inzh, bzh, czh, f.h.dth, hpy0, hpy2, hpx1, hpx2, hpz0, hpz1);
        if (jm_ok) {
            if (kin) hx_point(hxm, eykm - eym, ez - ezm, hy_m, hz_s, sym, bym, cym, inzh, bzh, czh, f.h.dth, mpy0, mpz0);
            if (iin) hz_point(hzm, eyim - eym, ex - exm, hx_s, hy_m, sxh, bxh, cxh, sym, bym, cym, f.h.dth, mpx2, mpy2);
        }
        // H of (j, k - 1) from the lane below (the point itself at k = 0: unused there)
        const float hxk0 = __shfl_up(hxv, 1u, 64), hyk0 = __shfl_up(hyv, 1u, 64);
        const float hxk = k ? hxk0 : hxv, hyk = k ? hyk0 : hyv;
        if (own) {
            f.Hn[q] = hxv;
            f.Hn[q + N] = hyv;
            f.Hn[q + 2 * N] = hzv;
            // every psi_H entry of the new set is written, updated or not
            if (syh >= 0) {
                f.psiHn[0][0][hoy] = hpy0;
                f.psiHn[2][1][hoy] = hpy2;
            }
            if (sxh >= 0) {
                f.psiHn[1][1][hox] = hpx1;
                f.psiHn[2][0][hox] = hpx2;
            }
            if (zoh >= 0) {
                f.psiHn[0][1][hoz] = hpz0;
                f.psiHn[1][0][hoz] = hpz1;
            }
            float exn = ex, eyn = ey, ezn = ez;
            if (iin) {
                const bool ib_ = i >= 1;
                const float hxj = jm_ok ? hxm : 0.f, hzj = jm_ok ? hzm : 0.f;
                e_point(exn, eyn, ezn, ex, ey, ez, ca[0], cb[0], ca[1], cb[1], ca[2], cb[2], hzv - hzj, hyv - hyk,
                        hxv - hxk, hzv - hzp, hyv - hyp, hxv - hxj, a.es[0][i], ey_s, ez_s, jb, jin, ib_, kb, kz, sye,
                        bye, cye, sx, bx, cx, inze, bze, cze, py0, py2, px1, px2, pz0, pz1);
                if (in_box) {
#pragma unroll
                    for (int c = 0; c < 3; c++) {
                        float &en = c == 0 ? exn : c == 1 ? eyn : ezn;
                        const float e0 = c == 0 ? ex : c == 1 ? ey : ez;
                        const int row = s_row[c * 256 + ((cls >> (8 * c)) & 255u)];
                        if (row >= 0) debye_point<NP>(en, e0, cb[c], jv[c], s_bt + row * NP, a.kc, a.km1);
#pragma unroll
                        for (int pp = 0; pp < NPE; pp++) Jl[(c * NPE + pp) * bkp] = jv[c][pp];
                    }
                }
                if (sye >= 0) {
                    a.psi[0][0][oy] = py0;
                    a.psi[2][1][oy] = py2;
                }
                if (sx >= 0) {
                    a.psi[1][1][ox] = px1;
                    a.psi[2][0][ox] = px2;
                }
                if (zoe >= 0) {
                    a.psi[0][1][oz] = pz0;
                    a.psi[1][0][oz] = pz1;
                }
            }
            f.En[q] = exn;
            f.En[q + N] = eyn;
            f.En[q + 2 * N] = ezn;
        }
        hyp = hyv;
        hzp = hzv;
        ey = eyi;
        ez = ezi;
        eym = eyim;
    }
}

// surface-impedance overwrite of the E edge of SIBC slot u (sibc.c:256-269):
// J from up to four H faces, then the recursive-convolution terms
__device__ __forceinline__ float sibc_edge(const PostArgs &a, const PostParams &pp, int u) {
    // every term's state loaded at once, ahead of the face values (a loop of load, update,
    // store would wait for each load in turn: the stores may alias the next term's load)
    float yv[kSibcTerms];
#pragma unroll
    for (int k = 0; k < kSibcTerms; k++) yv[k] = k < pp.nterm ? a.Y[(long long)k * a.nsb + u] : 0.f;
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
#pragma unroll
    for (int k = 0; k < kSibcTerms; k++)
        if (k < pp.nterm) {
            yv[k] += pp.om[k] * (Js - yv[k]);
            es += pp.zc[k] * (Js - yv[k]);
        }
#pragma unroll
    for (int k = 0; k < kSibcTerms; k++)
        if (k < pp.nterm) a.Y[(long long)k * a.nsb + u] = yv[k];
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

// the post-phase physics of listed edge t (dispersive, then surface-only, then
// source-only edges) at source waveform values vs (lumped) and gm (modal)
__device__ __forceinline__ void post_edge(const PostArgs &a, const PostParams &pp, int t, float vs, float gm) {
    if (t < a.ndisp) {                       // listed dispersive edge
        const unsigned q = a.d_idx[t];
        const unsigned m = a.d_mat[t];
        const float *mat = a.dtab + (m & kDispMat) * a.dstride;
        const long long jb = a.d_jb ? a.d_jb[t] : (long long)t;
        float e = a.E[q];
        // every pole's J loaded once, together (unrolled, guarded), and kept in
        // registers for the update
        float jv[kMaxPoles];
#pragma unroll
        for (int p = 0; p < kMaxPoles; p++) jv[p] = p < pp.npole ? a.J[jb + p * a.jstride] : 0.f;
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
            if (p < pp.npole) a.J[jb + p * a.jstride] = jv[p] + pp.km1[p] * jv[p] + mat[1 + p] * de;
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

__global__ void __launch_bounds__(kBlock) k_post(PostArgs a) {
    const PostParams &pp = *a.pp;
    float vs, gm;
    waveforms(a.wave, a.ctl->step, &vs, &gm);
    post_edge(a, pp, blockIdx.x * kBlock + threadIdx.x, vs, gm);
}

// Short post-phase lists (a small board, or lumped sources only): one workgroup runs every
// listed edge and then the step's probe samples and the counter advance (k_gather's work),
// saving a launch a step.  The workgroup's barrier orders its E writes before its sample
// reads (one workgroup: no device-wide fence).
constexpr int kSmallPost = 4 * kGatherBlock;   // most listed edges for the one-workgroup form

__global__ void __launch_bounds__(kGatherBlock) k_post_gather(PostArgs a, GatherArgs g) {
    const PostParams &pp = *a.pp;
    float vs, gm;
    waveforms(a.wave, a.ctl->step, &vs, &gm);
    const int n = a.ndisp + a.nsbo + a.nxs;
    for (int t = threadIdx.x; t < n; t += kGatherBlock) post_edge(a, pp, t, vs, gm);
    __syncthreads();
    const int row = g.ctl->row;
    gather_samples(g, row, threadIdx.x >> 6, kGatherBlock / 64);
    __syncthreads();
    if (threadIdx.x == 0) {
        g.ctl->row = row + 1;
        g.ctl->step += 1;
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

// E of every listed dispersive edge, taken as E_old of the next E update
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
        const int j = rem / g.kp, k = rem - j * g.kp;
        if (k > g.nz) continue;   // row padding
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
    // summed time of the dispatches stamped on pairs [a, b) (their stream synchronised)
    double span(size_t a, size_t b) const {
        double ms = 0;
        for (size_t k = a; k < b; k++) {
            float x = 0;
            if (hipEventElapsedTime(&x, pool[k].first, pool[k].second) == hipSuccess) ms += x;
        }
        return ms;
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
    DBuf<float> F;               // E (3N) then H (3N); two such sets for the one-pass step
    int par = 0;                 // set holding the current fields (and psi_H)
    bool onepass = false;        // steps run k_step (H and E in one sweep, set to set)
    bool steprows = false;       // ... as k_step_rows (no tiles: row j - 1 recomputed per lane)
    Device *partner = nullptr;   // while a pair steps: the second right-hand side (same plan),
    cem_sim *partner_sim = nullptr;   // whose H and E updates share this device's launches
    int tj = 0, ntj = 0;         // k_step: owned rows per tile, tiles over j
    int kw = 0, ntk = 0;         // k_step: owned k points per tile, tiles over k
    int iseg_s = 8;              // k_step: x planes per workgroup
    DBuf<float> ca, cb;          // 3N each, only when more than 256 distinct pairs
    DBuf<unsigned> cidx;         // N packed coefficient classes
    DBuf<float2> ctab;           // 3 x 256 (ca, cb)
    DBuf<int> crow;              // 3 x 256 Debye material row of a class (-1: none)
    int ncls = 256;              // classes in use per component
    DBuf<unsigned char> pidx;    // N class bytes (at most 256 distinct class words), else empty
    DBuf<unsigned> ptab;         // 256 class words indexed by pidx
    int nptab = 0;
    bool compressed = false;
    // Debye currents updated inside the E kernel (fused) or by the post kernel
    bool fused = false;
    int npole = 0;
    int nlisted = 0;             // dispersive edges the post kernel takes
    DBuf<float> btab;            // fused: beta_p / dt per material row
    int nrow = 0;
    int bi0 = 0, bi1 = 0, bj0 = 0, bj1 = 0, bg0 = 0, bg1 = 0;   // J box (k in groups)
    long long bkp = 0;           // k points of a J box row (the pole stride)
    DBuf<long long> d_jb;        // fused: J box index of each listed edge
    std::vector<long long> jbox_of;   // fused: J box index of pole 0 of every dispersive edge (reference order)
    DBuf<float> axis1d;          // hs[3], es[3], pml b/c for H and E per axis
    DBuf<int> slots;             // CPML slot tables, then the z-layout group tables
    std::vector<int> slots_host;
    size_t slot_off[2][3] = {}, zo_off[2] = {};   // [is_e][axis] offsets into slots
    DBuf<float> psi;             // all psi arrays
    float *psiE[3][2] = {}, *psiHs[2][3][2] = {};   // psi_H per field set
    size_t psiE_len[3][2] = {}, psiH_len[3][2] = {};
    const float *hs[3] = {}, *es[3] = {};
    Pml pmlH{}, pmlE{};
    // post-phase edge lists
    DBuf<unsigned> d_idx, sb_idx, sbo, x_idx;
    DBuf<int> sb_src, x_src;
    DBuf<int> dsb_t, dsb_u, dsrc_t, dsrc_x;   // listed dispersive edges with SIBC / source slots
    std::vector<unsigned> d_mat_host;          // material rows of the listed dispersive edges, no flags
    std::vector<int> dsb_t_host;               // ascending listed dispersive edges with an SIBC slot
    std::vector<int> listed_of;                // listed slot of each dispersive edge (reference order), -1
    int nsbo = 0;
    bool has_sbsrc = false;
    DBuf<float> J, dtab, Eold, Y, wl, wm;
    DBuf<unsigned> d_mat;
    int dstride = 1;
    DBuf<unsigned char> face;
    DBuf<PostParams> pp;
    PostParams pp_host{};
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
    // on the first steps unless fixed by COURANT_ISEG_H / COURANT_ISEG_E / COURANT_ISEG_S
    int iseg_h = kIseg, iseg_e = kIseg;
    int wh = 4, we = 1;          // lane widths (k points per thread) of the H and E updates
    bool wh_fixed = false;       // H lane width set by COURANT_LANES_H (else chosen by the tuner)
    int xcd = 1;                 // tile order of the H and E updates (1: XCD bands, 0: dispatch order)
    bool xcd_fixed = false;      // set by COURANT_TILE_ORDER (else chosen by the tuner)
    int chunk = 3;               // dispatch order: log2 of the tiles per chunk (chunk_tile)
    int hrev = 1;                // H sweep over its tiles from the last (zig-zag with the E sweep)
    bool hrev_fixed = false;     // set by COURANT_H_REVERSE (else chosen by the tuner)
    bool tuned = false;
    // timing
    bool profiling = false;
    KernelTimer tH, tE, tG, tD, tP, tS;
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

// Device index of the reference's global index comp N + q (i si + j sj + k).
unsigned long long dev_index(const cem_sim *s, const Grid &g, unsigned long long gidx) {
    const unsigned long long c = gidx / s->N, q = gidx % s->N;
    const unsigned long long i = q / s->si, r = q % s->si, j = r / s->sj, k = r % s->sj;
    return c * (unsigned long long)g.N + i * (unsigned long long)g.si + j * (unsigned long long)g.sj + k;
}

// Excited sources as (device E index, lumped cs, modal weight) slots; edges that
// are also listed dispersive edges are flagged in d_mat and listed in dsrc_t /
// dsrc_x, surface edges get sb_src, the others go to x_idx.  Every dispersive
// edge a source can drive (the mode ports' source planes; lumped-port edges are
// never dispersive) is a listed edge, whichever port is excited.
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
            const int lt = D->listed_of[t];
            if (lt < 0) {
                set_error("source edge missing from the post-phase edge list");
                return -1;
            }
            dsrc_t.push_back(lt);
            dsrc_x.push_back((int)k);
        } else if (u >= 0) {
            sb_src[u] = (int)k;
            D->has_sbsrc = true;
        } else {
            x_idx.push_back((unsigned)dev_index(s, D->g, g));
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
    // material rows of the listed dispersive edges with the SIBC / source flag bits
    if (D->nlisted) {
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

// Floats of the psi array (is_e, comp, w) in the device layout (see Pml).
size_t psi_dev_len(const cem_sim *s, const Device *D, int is_e, int c, int w) {
    if (!s->has_pml) return 0;
    const int a = (c + w + 1) % 3;
    const size_t planes = psi_len(s, is_e, c, w) / pml_per(s, a);
    const Grid &g = D->g;
    if (a == 0) return planes * (size_t)g.S4 * 4;
    if (a == 1) return planes * (size_t)(g.nx + 1) * g.kp;
    return planes ? (size_t)(g.nx + 1) * (g.ny + 1) * (is_e ? D->pmlE : D->pmlH).zg * 4 : 0;
}

// Device float index of every entry of the reference's psi array (is_e, comp, w):
// slab-major, slot * per + in-plane index with the in-plane index running over
// the other two axes at full node extent, the later axis fastest (sim.c:641-678).
std::vector<size_t> psi_dev_map(const cem_sim *s, const Device *D, int is_e, int c, int w) {
    const int a = (c + w + 1) % 3;
    const size_t per = pml_per(s, a), len = psi_len(s, is_e, c, w);
    const Grid &g = D->g;
    const int nx = s->n[0], ny = s->n[1], nz = s->n[2];
    std::vector<size_t> m(len);
    std::vector<int> kof;   // z slab slot -> k
    if (a == 2) {
        const int *sl = D->slots_host.data() + D->slot_off[is_e][2];
        for (int k = 0; k <= nz + is_e; k++)
            if (sl[k] >= 0) {
                if ((int)kof.size() <= sl[k]) kof.resize(sl[k] + 1, 0);
                kof[sl[k]] = k;
            }
    }
    const int *zo = D->slots_host.data() + D->zo_off[is_e];
    const int zg = (is_e ? D->pmlE : D->pmlH).zg;
    for (size_t r = 0; r < len; r++) {
        const size_t slot = r / per, rr = r % per;
        if (a == 0) {
            const size_t j = rr / (nz + 1), k = rr % (nz + 1);
            m[r] = slot * (size_t)g.S4 * 4 + j * g.kp + k;
        } else if (a == 1) {
            const size_t i = rr / (nz + 1), k = rr % (nz + 1);
            m[r] = (slot * (nx + 1) + i) * g.kp + k;
        } else {
            const int k = kof[slot];
            m[r] = (rr * zg + zo[k / 4]) * 4 + (k % 4);
        }
    }
    (void)ny;
    return m;
}

// Coefficient classes: per component the (ca, cb) pairs of a grid take a handful
// of values (one per material mix), so each point stores three 8-bit class
// numbers; a class also carries the Debye material row of the edges the E kernel
// updates (row[e] >= 0 for those, -1 elsewhere).  Class 0 of every component is
// (0, 0, no row): the row padding.  Returns false when a component needs more
// than 256 classes.
bool build_classes(const cem_sim *s, const Grid &g, const std::vector<int> &row_of, std::vector<unsigned> &idx,
                   std::vector<float2> &tab, std::vector<int> &rowtab, int &ncls) {
    const int nx = s->n[0], ny = s->n[1], nz = s->n[2];
    idx.assign((size_t)g.N, 0u);
    tab.assign(3 * 256, make_float2(0.f, 0.f));
    rowtab.assign(3 * 256, -1);
    ncls = 1;
    size_t t = 0;   // dispersive edges are sorted by (comp, q): walk them with q
    for (int c = 0; c < 3; c++) {
        std::vector<std::pair<std::array<unsigned, 3>, unsigned>> seen;
        seen.push_back({{0u, 0u, 0xffffffffu}, 0u});
        std::array<unsigned, 3> last_key = {0u, 0u, 0xffffffffu};
        unsigned last_slot = 0;
        for (int i = 0; i <= nx; i++)
            for (int j = 0; j <= ny; j++) {
                const size_t q0 = (size_t)i * s->si + (size_t)j * s->sj;
                const size_t d0 = (size_t)i * g.si + (size_t)j * g.sj;
                for (int k = 0; k <= nz; k++) {
                    const size_t q = q0 + k;
                    int row = -1;
                    while (t < (size_t)s->ndisp && (s->d_comp[t] < c || (s->d_comp[t] == c && s->d_idx[t] < q))) t++;
                    if (t < (size_t)s->ndisp && s->d_comp[t] == c && s->d_idx[t] == q) row = row_of[t];
                    const float fa = (float)s->ca[c][q], fb = (float)s->cb[c][q];
                    std::array<unsigned, 3> key;
                    std::memcpy(&key[0], &fa, 4);
                    std::memcpy(&key[1], &fb, 4);
                    key[2] = (unsigned)row;
                    unsigned slot = last_slot;
                    if (key != last_key) {
                        auto it = std::find_if(seen.begin(), seen.end(), [&](const auto &e) { return e.first == key; });
                        if (it == seen.end()) {
                            if (seen.size() == 256) return false;
                            slot = (unsigned)seen.size();
                            seen.push_back({key, slot});
                            tab[c * 256 + slot] = make_float2(fa, fb);
                            rowtab[c * 256 + slot] = row;
                        } else {
                            slot = it->second;
                        }
                        last_key = key;
                        last_slot = slot;
                    }
                    idx[d0 + k] |= slot << (8 * c);
                }
            }
        ncls = std::max(ncls, (int)seen.size());
    }
    return true;
}

// Class bytes: when the grid holds at most 256 distinct class words (the three
// components' classes of a point), the E update reads one byte per point and looks
// the word up in a 256-entry table instead of reading the 4-byte word.  Returns -1
// on a device error; leaves pidx empty when there are too many distinct words.
int build_class_bytes(const std::vector<unsigned> &cls, Device *D, hipStream_t st) {
    std::unordered_map<unsigned, unsigned char> slot;
    std::vector<unsigned> tab;
    std::vector<unsigned char> b(cls.size());
    unsigned last = 0;
    unsigned char lastb = 0;
    for (size_t q = 0; q < cls.size(); q++) {
        const unsigned v = cls[q];
        if (q == 0 || v != last) {
            auto it = slot.find(v);
            if (it == slot.end()) {
                if (tab.size() == 256) return 0;
                it = slot.emplace(v, (unsigned char)tab.size()).first;
                tab.push_back(v);
            }
            last = v;
            lastb = it->second;
        }
        b[q] = lastb;
    }
    D->nptab = (int)tab.size();
    tab.resize(256, 0u);
    if (D->ptab.alloc(256, D->st) || D->ptab.upload(tab.data(), 256, st)) return -1;
    if (D->pidx.alloc(b.size(), D->st) || D->pidx.upload(b.data(), b.size(), st)) return -1;
    return 0;
}

int upload_plan(cem_sim *s, Device *D) {
    const int nx = s->n[0], ny = s->n[1], nz = s->n[2];
    Grid &g = D->g;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.kp = (nz + 1 + 3) / 4 * 4;
    g.ng = g.kp / 4;
    g.S4 = (ny + 1) * g.ng;
    g.sj = g.kp;
    g.si = (long long)(ny + 1) * g.kp;
    g.N = (long long)(nx + 1) * g.si;
    g.si4 = g.si / 4;
    g.N4 = g.N / 4;
    const size_t N = s->N;
    const size_t Nd = (size_t)g.N;
    hipStream_t st = D->st;
    // edge indices (3 N) and the stencils' byte offsets within one component array (4 N)
    // are 32-bit
    if (4 * Nd > 0xffffffffull) {
        set_error("grid too large for 32-bit offsets (more than 1.07e9 grid points)");
        return -1;
    }
    // device index of a reference point index (same component)
    auto dq = [&](size_t q) -> size_t {
        const size_t i = q / s->si, r = q % s->si, j = r / s->sj, k = r % s->sj;
        return i * (size_t)g.si + j * (size_t)g.sj + k;
    };

    // fields (zero, row padding included).  Full steps run the one-pass step only on
    // request (COURANT_STEP=onepass at set-up): on the four-layer board it takes 1.36 ms
    // a step against 0.91 ms for the two sweeps.  Its workgroup holds a tile of tj rows
    // by kw points of k plus a halo row and a halo column on each side, at most
    // kStepLanes lanes; whole rows of k (one k tile) are taken when two rows fit,
    // otherwise the fewest k tiles that leave two rows.
    {
        const char *mode = std::getenv("COURANT_STEP");
        int tj = 0, kw = 0, ntk = 0;
        for (int n = 1; n <= 8 && tj < 2; n++) {
            kw = (g.kp + n - 1) / n;
            tj = std::min(kStepRows, kStepLanes / (kw + 2) - 2);
            ntk = n;
        }
        D->onepass = tj >= 2 && mode && std::strcmp(mode, "onepass") == 0;
        D->steprows = mode && std::strcmp(mode, "rows") == 0;
        if (D->steprows) D->onepass = true, tj = 1, kw = g.kp, ntk = 1;
        D->tj = D->onepass ? tj : 0;
        D->ntj = D->onepass ? (ny + 1 + tj - 1) / tj : 0;
        D->kw = D->onepass ? kw : 0;
        D->ntk = D->onepass ? ntk : 0;
        D->par = 0;
    }
    if (D->F.alloc((D->onepass ? 12 : 6) * Nd, D->st) || D->F.zero(st)) return -1;

    // Debye materials: dispersive edges share a handful of (cb, beta_p / dt) rows
    const int np = s->npole, nd = s->ndisp;
    D->npole = np;
    std::vector<unsigned> mat(nd);
    std::vector<float> tab;
    {
        std::map<std::vector<float>, unsigned> rows;
        std::vector<float> key(1 + np);
        for (int e = 0; e < nd; e++) {
            key[0] = (float)s->d_cb[e];
            for (int p = 0; p < np; p++) key[1 + p] = (float)(s->d_beta[(size_t)p * nd + e] * (1.0 / s->dt));
            auto it = rows.find(key);
            if (it == rows.end()) {
                it = rows.emplace(key, (unsigned)rows.size()).first;
                tab.insert(tab.end(), key.begin(), key.end());
            }
            mat[e] = it->second;
        }
        if (rows.size() > kDispMat) {
            set_error("too many Debye materials");
            return -1;
        }
        D->nrow = (int)rows.size();
    }
    // Dispersive edges the post kernel must take in reference order with other
    // physics: surface (SIBC) edges, any mode port's source edges, and the edges
    // on the x = max face, which the E kernel's i range does not visit.
    std::vector<uint8_t> linked(nd, 0);
    for (int e = 0; e < s->nsb; e++) {
        const long t = dispersive_slot(s, (unsigned long long)s->sb_comp[e] * N + s->sb_idx[e]);
        if (t >= 0) linked[t] = 1;
    }
    for (auto &mp : s->mport)
        for (size_t e = 0; e < mp.src_idx.size(); e++) {
            const long t = dispersive_slot(s, (unsigned long long)mp.src_comp[e] * N + mp.src_idx[e]);
            if (t >= 0) linked[t] = 1;
        }
    for (auto &pt : s->port)
        for (size_t e = 0; e < pt.edge.size(); e++) {
            const long t = dispersive_slot(s, (unsigned long long)pt.axis * N + pt.edge[e]);
            if (t >= 0) linked[t] = 1;
        }
    int bi0 = nx + 1, bi1 = 0, bj0 = ny + 1, bj1 = 0, bk0 = nz + 1, bk1 = 0;
    for (int t = 0; t < nd; t++) {
        const size_t q = s->d_idx[t];
        const int i = (int)(q / s->si), j = (int)(q % s->si / s->sj), k = (int)(q % s->sj);
        if (i >= nx) linked[t] = 1;
        bi0 = std::min(bi0, i);
        bi1 = std::max(bi1, i + 1);
        bj0 = std::min(bj0, j);
        bj1 = std::max(bj1, j + 1);
        bk0 = std::min(bk0, k);
        bk1 = std::max(bk1, k + 1);
    }
    // the E kernel takes the Debye currents when the poles fit its registers, the
    // material rows its LDS, and the J box is not much larger than the edge list
    bool fused = np >= 1 && np <= kFusedPoles && D->nrow <= kDebyeRows && nd > 0 &&
                 std::getenv("COURANT_DEBYE_LIST") == nullptr && std::getenv("COURANT_COEF_ARRAYS") == nullptr;
    if (fused) {
        bk0 = bk0 / 4 * 4;
        bk1 = (bk1 + 3) / 4 * 4;
        const size_t box = (size_t)(bi1 - bi0) * (bj1 - bj0) * (bk1 - bk0);
        if (3 * box > 8 * (size_t)nd + ((size_t)1 << 22)) fused = false;
        if (4 * 3 * (size_t)np * box > 0xffffffffull) fused = false;   // 32-bit byte offsets in the box
    }
    // coefficient classes, with the Debye rows of the fused edges
    std::vector<unsigned> cls;
    std::vector<float2> ctab;
    std::vector<int> crow;
    bool compressed = false;
    if (std::getenv("COURANT_COEF_ARRAYS") == nullptr) {
        std::vector<int> row_of(nd, -1);
        if (fused)
            for (int t = 0; t < nd; t++)
                if (!linked[t]) row_of[t] = (int)mat[t];
        compressed = build_classes(s, g, row_of, cls, ctab, crow, D->ncls);
        if (!compressed && fused) {   // retry with every dispersive edge in the post kernel
            fused = false;
            std::fill(row_of.begin(), row_of.end(), -1);
            compressed = build_classes(s, g, row_of, cls, ctab, crow, D->ncls);
        }
    }
    if (!compressed) fused = false;
    D->compressed = compressed;
    D->fused = fused;
    if (compressed) {
        if (D->cidx.alloc(Nd, D->st) || D->cidx.upload(cls.data(), Nd, st)) return -1;
        if (D->ctab.alloc(3 * 256, D->st) || D->ctab.upload(ctab.data(), 3 * 256, st)) return -1;
        if (D->crow.alloc(3 * 256, D->st) || D->crow.upload(crow.data(), 3 * 256, st)) return -1;
        if (!std::getenv("COURANT_CLASS_WORDS") && build_class_bytes(cls, D, st)) return -1;
        std::vector<unsigned>().swap(cls);
    } else {
        if (D->ca.alloc(3 * Nd, D->st) || D->cb.alloc(3 * Nd, D->st)) return -1;
        std::vector<float> tmp(3 * Nd, 0.f);
        for (int pass = 0; pass < 2; pass++) {
            const dvec *arr = pass ? s->cb : s->ca;
            std::fill(tmp.begin(), tmp.end(), 0.f);
            for (int c = 0; c < 3; c++)
                for (size_t q = 0; q < N; q++) tmp[c * Nd + dq(q)] = (float)arr[c][q];
            if ((pass ? D->cb : D->ca).upload(tmp.data(), 3 * Nd, st)) return -1;
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

    // CPML slot tables (compact slab numbering of sim.c:655-678) and, per kernel,
    // the z-layout place of each k group that holds z-slab points
    std::vector<int> slots;
    size_t off_sh[3], off_se[3], off_zh = 0, off_ze = 0;
    for (int is_e = 0; is_e < 2; is_e++) {
        for (int a = 0; a < 3; a++) {
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
        const size_t zs = (is_e ? off_se : off_sh)[2];
        const int len = nz + (is_e ? 1 : 0);
        (is_e ? off_ze : off_zh) = slots.size();
        int zg = 0;
        for (int gk = 0; gk < g.ng; gk++) {
            bool any = false;
            for (int k = 4 * gk; k < 4 * gk + 4 && k < len; k++) any |= slots[zs + k] >= 0;
            slots.push_back(any ? zg++ : -1);
        }
        (is_e ? D->pmlE : D->pmlH).zg = zg;
    }
    if (D->slots.alloc(slots.size(), D->st) || D->slots.upload(slots.data(), slots.size(), st)) return -1;
    D->slots_host = slots;
    for (int a = 0; a < 3; a++) {
        D->pmlH.slot[a] = D->slots.p + off_sh[a];
        D->pmlE.slot[a] = D->slots.p + off_se[a];
        D->slot_off[0][a] = off_sh[a];
        D->slot_off[1][a] = off_se[a];
    }
    D->pmlH.zo = D->slots.p + off_zh;
    D->pmlE.zo = D->slots.p + off_ze;
    D->zo_off[0] = off_zh;
    D->zo_off[1] = off_ze;
    // psi arrays in the device layout (see Pml)
    size_t psi_total = 0;
    for (int c = 0; c < 3; c++)
        for (int w = 0; w < 2; w++) {
            D->psiE_len[c][w] = psi_dev_len(s, D, 1, c, w);
            D->psiH_len[c][w] = psi_dev_len(s, D, 0, c, w);
            psi_total += D->psiE_len[c][w] + (D->onepass ? 2 : 1) * D->psiH_len[c][w];
        }
    if (D->psi.alloc(psi_total + 4, D->st) || D->psi.zero(st)) return -1;
    {
        size_t o = 0;
        for (int c = 0; c < 3; c++)
            for (int w = 0; w < 2; w++) {
                D->psiE[c][w] = D->psi.p + o;
                o += D->psiE_len[c][w];
                for (int set = 0; set < (D->onepass ? 2 : 1); set++) {
                    D->psiHs[set][c][w] = D->psi.p + o;
                    o += D->psiH_len[c][w];
                }
            }
    }

    // Debye tables.  Listed edges (all of them when the E kernel does not take the
    // currents) keep the reference's order; their J is the reference's pole-major
    // array, or, when the E kernel takes the others, the box entry of the edge.
    {
        D->listed_of.assign(nd, -1);
        std::vector<unsigned> lidx, lmat;
        std::vector<long long> ljb;
        D->jbox_of.clear();
        if (fused) {
            D->bi0 = bi0;
            D->bi1 = bi1;
            D->bj0 = bj0;
            D->bj1 = bj1;
            D->bg0 = bk0 / 4;
            D->bg1 = bk1 / 4;
            D->bkp = bk1 - bk0;
            D->jbox_of.resize(nd);
        }
        for (int t = 0; t < nd; t++) {
            const size_t q = s->d_idx[t];
            if (fused) {
                const int i = (int)(q / s->si), j = (int)(q % s->si / s->sj), k = (int)(q % s->sj);
                D->jbox_of[t] = (((long long)(i - bi0) * (bj1 - bj0) + (j - bj0)) * 3 + s->d_comp[t]) * np * D->bkp +
                                (k - bk0);
            }
            if (fused && !linked[t]) continue;
            D->listed_of[t] = (int)lidx.size();
            lidx.push_back((unsigned)((size_t)s->d_comp[t] * Nd + dq(q)));
            lmat.push_back(mat[t]);
            if (fused) ljb.push_back(D->jbox_of[t]);
        }
        D->nlisted = (int)lidx.size();
        D->dstride = 1 + np;
        if (D->d_idx.alloc(lidx.size(), D->st) || D->d_idx.upload(lidx.data(), lidx.size(), st)) return -1;
        D->d_mat_host = lmat;   // uploaded with its flag bits by upload_sources
        if (D->d_mat.alloc(lmat.size(), D->st)) return -1;
        if (fused && (D->d_jb.alloc(ljb.size(), D->st) || D->d_jb.upload(ljb.data(), ljb.size(), st))) return -1;
        if (D->dtab.alloc(tab.size() + 1, D->st) || D->dtab.upload(tab.data(), tab.size(), st)) return -1;
        if (fused) {
            std::vector<float> bt((size_t)D->nrow * np);
            for (int r = 0; r < D->nrow; r++)
                for (int p = 0; p < np; p++) bt[(size_t)r * np + p] = tab[(size_t)r * (1 + np) + 1 + p];
            if (D->btab.alloc(bt.size(), D->st) || D->btab.upload(bt.data(), bt.size(), st)) return -1;
        }
        if (D->Eold.alloc(lidx.size(), D->st) || D->Eold.zero(st)) return -1;
        const size_t nj = fused ? (size_t)(bi1 - bi0) * (bj1 - bj0) * 3 * D->bkp * np : (size_t)np * nd;
        if (D->J.alloc(nj, D->st) || D->J.zero(st)) return -1;
        PostParams pp{};
        pp.ndisp = D->nlisted;
        pp.npole = np;
        pp.nsb = s->nsb;
        pp.nterm = s->sibc_npole;
        for (int p = 0; p < np; p++) {
            const double tau = 1.0 / (2 * M_PI * s->pole_f[p]);
            const double x = s->dt / (2 * tau);
            pp.kc[p] = (float)(1.0 / (1.0 + x));          // (1 + k)/2
            pp.km1[p] = (float)(-2.0 * x / (1.0 + x));     // k - 1
        }
        for (int t = 0; t < s->sibc_npole; t++) {
            pp.zc[t] = (float)s->sibc_c[t];
            pp.om[t] = (float)(-std::expm1(-s->sibc_a[t] * s->dt));
        }
        D->pp_host = pp;
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
            const unsigned long long gi = (unsigned long long)s->sb_comp[e] * N + s->sb_idx[e];
            idx[e] = (unsigned)dev_index(s, g, gi);
            const long t = dispersive_slot(s, gi);
            if (t >= 0) {
                D->dsb_t_host.push_back(D->listed_of[t]);
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
                tidx.push_back((long long)dev_index(s, g, (unsigned long long)p.comp[e] * N + p.idx[e]));
                tw.push_back(p.w[e]);
            }
            post.push_back(1.0);
        }
        for (auto &p : s->port) {
            toff.push_back((int)tidx.size());
            for (size_t e = 0; e < p.edge.size(); e++) {
                tidx.push_back((long long)dev_index(s, g, (unsigned long long)p.axis * N + p.edge[e]));
                tw.push_back(p.dl[e]);
            }
            post.push_back((double)p.dir);
        }
        for (auto &p : s->vp) {
            toff.push_back((int)tidx.size());
            for (size_t e = 0; e < p.edge.size(); e++) {
                tidx.push_back((long long)dev_index(s, g, (unsigned long long)p.axis * N + p.edge[e]));
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

// first float of field set par (E then H)
float *fset(const Device *D, int par) {
    return D->F.p + (size_t)par * 6 * (size_t)D->g.N;
}

HArgs h_args(cem_sim *s, Device *D, int par) {
    HArgs a;
    std::memset(&a, 0, sizeof a);   // padding too: the bytes key the captured graph
    a.g = D->g;
    a.E = fset(D, par);
    a.H = fset(D, par) + 3 * D->g.N;
    for (int i = 0; i < 3; i++) a.hs[i] = D->hs[i];
    a.pml = D->pmlH;
    for (int c = 0; c < 3; c++)
        for (int w = 0; w < 2; w++) a.psi[c][w] = D->psiHs[par][c][w];
    a.dth = (float)(s->dt / CEM_MU0);
    a.has_pml = s->has_pml;
    a.xcd = D->xcd;
    a.chunk = D->chunk;
    a.rev = D->hrev;
    return a;
}

EArgs e_args(cem_sim *s, Device *D, int par) {
    EArgs a;
    std::memset(&a, 0, sizeof a);
    a.g = D->g;
    a.E = fset(D, par);
    a.H = fset(D, par) + 3 * D->g.N;
    a.ca = D->ca.p;
    a.cb = D->cb.p;
    a.cidx = D->compressed ? D->cidx.p : nullptr;
    a.ctab = D->compressed ? D->ctab.p : nullptr;
    a.crow = D->compressed ? D->crow.p : nullptr;
    a.ncls = D->ncls;
    a.pidx = D->pidx.p;
    a.ptab = D->ptab.p;
    a.nptab = D->nptab;
    for (int i = 0; i < 3; i++) a.es[i] = D->es[i];
    a.pml = D->pmlE;
    for (int c = 0; c < 3; c++)
        for (int w = 0; w < 2; w++) a.psi[c][w] = D->psiE[c][w];
    a.has_pml = s->has_pml;
    a.xcd = D->xcd;
    a.chunk = D->chunk;
    if (D->fused) {
        a.J = D->J.p;
        a.btab = D->btab.p;
        a.nrow = D->nrow;
        a.bi0 = D->bi0;
        a.bi1 = D->bi1;
        a.bj0 = D->bj0;
        a.bj1 = D->bj1;
        a.bg0 = D->bg0;
        a.bg1 = D->bg1;
        for (int p = 0; p < D->npole && p < kFusedPoles; p++) {
            a.kc[p] = D->pp_host.kc[p];
            a.km1[p] = D->pp_host.km1[p];
        }
    }
    return a;
}

typedef void (*EKernel)(EArgs, int, int, int, int);
typedef void (*HKernel)(HArgs, int, int, int, int);

// E kernel of this simulation: coefficient classes or arrays, the number of
// Debye poles it updates itself (0 when the post kernel takes them), lane width
template <int W, int C>
EKernel e_kernel_c(const Device *D) {
    switch (D->fused ? D->npole : 0) {
    case 1: return k_update_e<1, C, W>;
    case 2: return k_update_e<2, C, W>;
    case 3: return k_update_e<3, C, W>;
    case 4: return k_update_e<4, C, W>;
    case 5: return k_update_e<5, C, W>;
    case 6: return k_update_e<6, C, W>;
    case 7: return k_update_e<7, C, W>;
    case 8: return k_update_e<8, C, W>;
    default: return k_update_e<0, C, W>;
    }
}

template <int W>
EKernel e_kernel_w(const Device *D) {
    if (!D->compressed) return k_update_e<0, 0, W>;
    return D->pidx.p ? e_kernel_c<W, 2>(D) : e_kernel_c<W, 1>(D);
}

EKernel e_kernel(const Device *D) {
    return D->we == 1 ? e_kernel_w<1>(D) : D->we == 2 ? e_kernel_w<2>(D) : e_kernel_w<4>(D);
}

HKernel h_kernel(const Device *D) {
    return D->wh == 1 ? k_update_h<1> : D->wh == 2 ? k_update_h<2> : k_update_h<4>;
}

// workgroups per i segment of a stencil kernel with lane width w
int stencil_blocks(const Device *D, int w) {
    return ((D->g.ny + 1) * (D->g.kp / w) + kBlock - 1) / kBlock;
}

// Kernel launch on the engine stream; when profiling, the launch carries a start /
// stop event pair that the runtime stamps with the dispatch's own begin and end.
template <typename F, typename... A>
void launch_lds(Device *D, KernelTimer &t, F kernel, dim3 grid, dim3 block, size_t lds, A... args) {
    if (!D->profiling) {
        hipLaunchKernelGGL(kernel, grid, block, lds, D->st, args...);
        return;
    }
    if (t.used == 4096) t.fold();
    if (t.used == t.pool.size()) {
        hipEvent_t a, b;
        HIP_CHECK_VOID(hipEventCreate(&a));
        HIP_CHECK_VOID(hipEventCreate(&b));
        t.pool.push_back({a, b});
    }
    hipExtLaunchKernelGGL(kernel, grid, block, lds, D->st, t.pool[t.used].first, t.pool[t.used].second, 0, args...);
    t.used++;
}

template <typename F, typename... A>
void launch(Device *D, KernelTimer &t, F kernel, dim3 grid, dim3 block, A... args) {
    launch_lds(D, t, kernel, grid, block, 0, args...);
}

// The H / E arguments of a paired step: the partner's fields, psi and Debye currents
// as the second right-hand side (its plan, coefficient tables and lists are this
// device's; a pair runs the split sweeps, field set 0)
void pair_h(HArgs &a, const Device *P) {
    if (!P) return;
    a.pair = 1;
    a.E2 = fset(P, P->par);
    a.H2 = fset(P, P->par) + 3 * P->g.N;
    for (int c = 0; c < 3; c++)
        for (int w = 0; w < 2; w++) a.psi2[c][w] = P->psiHs[P->par][c][w];
}

void pair_e(EArgs &a, const Device *P) {
    if (!P) return;
    a.pair = 1;
    a.E2 = fset(P, P->par);
    a.H2 = fset(P, P->par) + 3 * P->g.N;
    for (int c = 0; c < 3; c++)
        for (int w = 0; w < 2; w++) a.psi2[c][w] = P->psiE[c][w];
    a.J2 = P->fused ? P->J.p : nullptr;
}

void launch_h(cem_sim *s, Device *D, int i0, int i1) {
    if (i1 <= i0) return;
    HArgs a = h_args(s, D, D->par);
    pair_h(a, D->partner);
    const int npblk = stencil_blocks(D, D->wh), nseg = (i1 - i0 + D->iseg_h - 1) / D->iseg_h;
    launch(D, D->tH, h_kernel(D), dim3(npblk * nseg * (1 + a.pair)), dim3(kBlock), a, i0, i1, D->iseg_h, npblk);
}

void launch_e(cem_sim *s, Device *D, int i0, int i1) {
    if (i1 <= i0) return;
    EArgs a = e_args(s, D, D->par);
    pair_e(a, D->partner);
    const int npblk = stencil_blocks(D, D->we), nseg = (i1 - i0 + D->iseg_e - 1) / D->iseg_e;
    launch(D, D->tE, e_kernel(D), dim3(npblk * nseg * (1 + a.pair)), dim3(kBlock), a, i0, i1, D->iseg_e, npblk);
}

typedef void (*SKernel)(SArgs, int);
typedef void (*RKernel)(SArgs, int, int);

RKernel rows_kernel(const Device *D) {
    if (!D->compressed) return k_step_rows<0, false>;
    switch (D->fused ? D->npole : 0) {
    case 1: return k_step_rows<1, true>;
    case 2: return k_step_rows<2, true>;
    case 3: return k_step_rows<3, true>;
    case 4: return k_step_rows<4, true>;
    case 5: return k_step_rows<5, true>;
    case 6: return k_step_rows<6, true>;
    case 7: return k_step_rows<7, true>;
    case 8: return k_step_rows<8, true>;
    default: return k_step_rows<0, true>;
    }
}

// k_step_rows: workgroups of kBlock lanes over the (j, k) plane (63 owned points per
// wavefront) per x segment
int rows_blocks_per_seg(const Device *D) {
    const int waves = ((D->g.ny + 1) * D->g.kp + 62) / 63;
    return (waves + kBlock / 64 - 1) / (kBlock / 64);
}

SKernel step_kernel(const Device *D) {
    if (!D->compressed) return k_step<0, false>;
    switch (D->fused ? D->npole : 0) {
    case 1: return k_step<1, true>;
    case 2: return k_step<2, true>;
    case 3: return k_step<3, true>;
    case 4: return k_step<4, true>;
    case 5: return k_step<5, true>;
    case 6: return k_step<6, true>;
    case 7: return k_step<7, true>;
    case 8: return k_step<8, true>;
    default: return k_step<0, true>;
    }
}

// k_step from field set par to set par ^ 1
SArgs step_args(cem_sim *s, Device *D, int par) {
    SArgs a;
    std::memset(&a, 0, sizeof a);
    a.h = h_args(s, D, par);
    a.e = e_args(s, D, par);
    a.En = fset(D, par ^ 1);
    a.Hn = fset(D, par ^ 1) + 3 * D->g.N;
    for (int c = 0; c < 3; c++)
        for (int w = 0; w < 2; w++) a.psiHn[c][w] = D->psiHs[par ^ 1][c][w];
    a.tj = D->tj;
    a.ntj = D->ntj;
    a.kw = D->kw;
    a.ntk = D->ntk;
    a.kl = D->kw + 2;
    return a;
}

int step_threads(const Device *D) {
    if (D->steprows) return kBlock;
    return ((D->tj + 2) * (D->kw + 2) + 63) / 64 * 64;
}

size_t step_lds(const Device *D) {
    if (D->steprows) return 0;
    const int kl = D->kw + 2, tj = D->tj;
    size_t b = (size_t)3 * (2 * tj + 3) * kl * sizeof(float);
    if (D->compressed) b += 768 * sizeof(float2);
    if (D->compressed && D->fused && D->npole > 0)
        b += 768 * sizeof(int) + (size_t)D->nrow * D->npole * sizeof(float);
    return b;
}

int step_blocks(const Device *D) {
    if (D->steprows) return rows_blocks_per_seg(D) * ((D->g.nx + 1 + D->iseg_s - 1) / D->iseg_s);
    return D->ntj * D->ntk * ((D->g.nx + 1 + D->iseg_s - 1) / D->iseg_s);
}

// H and E of one step in one pass; the new fields become the current set
void launch_step(cem_sim *s, Device *D) {
    SArgs a = step_args(s, D, D->par);
    if (D->steprows)
        launch(D, D->tS, rows_kernel(D), dim3(step_blocks(D)), dim3(kBlock), a, D->iseg_s, rows_blocks_per_seg(D));
    else
        launch_lds(D, D->tS, step_kernel(D), dim3(step_blocks(D)), dim3(step_threads(D)), step_lds(D), a, D->iseg_s);
    D->par ^= 1;
}

// bulk updates of one full step
void launch_he(cem_sim *s, Device *D) {
    if (D->onepass) {
        launch_step(s, D);
        return;
    }
    launch_h(s, D, 0, s->n[0] + 1);
    launch_e(s, D, 0, s->n[0]);
}

void launch_dft(cem_sim *s, Device *D, int direct);

PostArgs post_args(cem_sim *s, Device *D, int par) {
    PostArgs a;
    std::memset(&a, 0, sizeof a);
    a.E = fset(D, par);
    a.H = fset(D, par) + 3 * D->g.N;
    a.N = D->g.N;
    a.si = D->g.si;
    a.sj = D->g.sj;
    a.pp = D->pp.p;
    a.ndisp = D->nlisted;
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
    a.d_jb = D->fused ? D->d_jb.p : nullptr;
    a.jstride = D->fused ? D->bkp : (long long)s->ndisp;
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

GatherArgs gather_args(cem_sim *s, Device *D, int mid_only, int par) {
    GatherArgs a;
    std::memset(&a, 0, sizeof a);
    a.F = fset(D, par);
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
    return (D->nlisted + D->nsbo + D->nxs + kBlock - 1) / kBlock;
}

// post lists short enough for k_post_gather (post-phase edges and probe samples in one launch)
bool small_post(const Device *D) {
    const int n = D->nlisted + D->nsbo + D->nxs;
    return n > 0 && n <= kSmallPost;
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
    GatherArgs g = gather_args(s, D, 0, D->par);
    if (small_post(D)) {
        launch(D, D->tP, k_post_gather, dim3(1), dim3(kGatherBlock), post_args(s, D, D->par), g);
    } else {
        if (const int nb = post_blocks(s, D)) launch(D, D->tP, k_post, dim3(nb), dim3(kBlock), post_args(s, D, D->par));
        launch(D, D->tG, k_gather, dim3(1), dim3(kGatherBlock), g);
    }
    if (D->rows == 0) D->row0 = s->step;
    D->rows++;
    s->step++;
    D->ctl_step = s->step;
    D->ctl_row = D->rows;
    if (D->rows == kSeriesCap) launch_dft(s, D, 0);
    if (D->partner) launch_post(D->partner_sim, D->partner);
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
    GatherArgs a = gather_args(s, D, 1, D->par);
    hipLaunchKernelGGL(k_gather, dim3(1), dim3(kGatherBlock), 0, D->st, a);
}

int ensure_vlast(cem_sim *s, Device *D) {
    if (D->vlast_stale) {
        launch_gather_mid(s, D);
        D->vlast_stale = false;
    }
    if (D->eold_stale && D->nlisted) {
        hipLaunchKernelGGL(k_prime_eold, dim3((D->nlisted + kBlock - 1) / kBlock), dim3(kBlock), 0, D->st, fset(D, D->par),
                           D->d_idx.p, D->Eold.p, D->nlisted);
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
    Device *P = D->partner;
    cem_sim *ps = D->partner_sim;
    if (P) sync_ctl(ps, P);
    const int nx = s->n[0];
    // launch arguments of the chunk's steps: with the one-pass step the field sets
    // alternate, so even and odd steps differ (kGraphSteps is even: the chunk ends
    // on the set it started from)
    const int p0 = D->par, p1 = D->onepass ? D->par ^ 1 : D->par;
    HArgs ha = h_args(s, D, p0);
    EArgs ea = e_args(s, D, p0);
    pair_h(ha, P);
    pair_e(ea, P);
    // the partner's post-phase and probe kernels (a pair runs the split sweeps: one field set)
    PostArgs pb;
    GatherArgs gb;
    std::memset(&pb, 0, sizeof pb);
    std::memset(&gb, 0, sizeof gb);
    if (P) {
        pb = post_args(ps, P, P->par);
        gb = gather_args(ps, P, 0, P->par);
    }
    const int nbp = P ? post_blocks(ps, P) : 0;
    SArgs sa[2];
    std::memset(sa, 0, sizeof sa);   // the bytes key the captured graph
    if (D->onepass) {
        sa[0] = step_args(s, D, p0);
        sa[1] = step_args(s, D, p1);
    }
    const PostArgs pa[2] = {post_args(s, D, p1), post_args(s, D, p0)};
    const GatherArgs ga[2] = {gather_args(s, D, 0, p1), gather_args(s, D, 0, p0)};
    const int geo[10] = {D->iseg_h, D->iseg_e, post_blocks(s, D), (int)D->compressed, (int)D->fused, D->wh, D->we,
                         (int)D->onepass, D->iseg_s, D->tj};
    std::vector<unsigned char> key;
    for (auto pr : {std::make_pair((const void *)&ha, sizeof ha), std::make_pair((const void *)&ea, sizeof ea),
                    std::make_pair((const void *)sa, sizeof sa), std::make_pair((const void *)pa, sizeof pa),
                    std::make_pair((const void *)ga, sizeof ga), std::make_pair((const void *)geo, sizeof geo),
                    std::make_pair((const void *)&pb, sizeof pb), std::make_pair((const void *)&gb, sizeof gb),
                    std::make_pair((const void *)&nbp, sizeof nbp)}) {
        const unsigned char *b = (const unsigned char *)pr.first;
        key.insert(key.end(), b, b + pr.second);
    }
    if (!D->gexec || key != D->gkey) {
        drop_graph(D);
        // the chunk's launches as a chain of kernel nodes (no stream capture)
        KernelChain chain;
        const int nbh = stencil_blocks(D, D->wh), nbe = stencil_blocks(D, D->we);
        const EKernel ek = e_kernel(D);
        const HKernel hk = h_kernel(D);
        const SKernel sk = D->onepass && !D->steprows ? step_kernel(D) : nullptr;
        const RKernel rk = D->steprows ? rows_kernel(D) : nullptr;
        const int gh = nbh * ((nx + 1 + D->iseg_h - 1) / D->iseg_h);
        const int ge = nbe * ((nx + D->iseg_e - 1) / D->iseg_e);
        const int rb = D->steprows ? rows_blocks_per_seg(D) : 0;
        for (int k = 0; k < kGraphSteps; k++) {
            const int o = D->onepass ? (k & 1) : 0;
            if (D->steprows) {
                chain.add(rk, dim3(step_blocks(D)), dim3(kBlock), 0, sa[o], D->iseg_s, rb);
            } else if (D->onepass) {
                chain.add(sk, dim3(step_blocks(D)), dim3(step_threads(D)), (unsigned)step_lds(D), sa[o], D->iseg_s);
            } else {
                chain.add(hk, dim3(gh * (1 + ha.pair)), dim3(kBlock), 0, ha, 0, nx + 1, D->iseg_h, nbh);
                chain.add(ek, dim3(ge * (1 + ea.pair)), dim3(kBlock), 0, ea, 0, nx, D->iseg_e, nbe);
            }
            if (small_post(D)) {
                chain.add(k_post_gather, dim3(1), dim3(kGatherBlock), 0, pa[o], ga[o]);
            } else {
                if (geo[2]) chain.add(k_post, dim3(geo[2]), dim3(kBlock), 0, pa[o]);
                chain.add(k_gather, dim3(1), dim3(kGatherBlock), 0, ga[o]);
            }
            if (P) {
                if (small_post(P)) {
                    chain.add(k_post_gather, dim3(1), dim3(kGatherBlock), 0, pb, gb);
                } else {
                    if (nbp) chain.add(k_post, dim3(nbp), dim3(kBlock), 0, pb);
                    chain.add(k_gather, dim3(1), dim3(kGatherBlock), 0, gb);
                }
            }
        }
        if (!chain.instantiate(&D->gexec)) {
            D->gexec = nullptr;
            D->use_graph = false;   // fall back to direct launches for this simulation
            return -1;
        }
        D->gkey.swap(key);
    }
    if (hipGraphLaunch(D->gexec, D->st) != hipSuccess) {
        set_error("graph launch failed");
        drop_graph(D);
        D->use_graph = false;
        return -1;
    }
    for (auto rhs : {std::make_pair(s, D), std::make_pair(ps, P)}) {
        cem_sim *t = rhs.first;
        Device *T = rhs.second;
        if (!T) continue;
        if (T->rows == 0) T->row0 = t->step;
        T->rows += kGraphSteps;
        t->step += kGraphSteps;
        T->ctl_step = t->step;
        T->ctl_row = T->rows;
        if (T->rows == kSeriesCap) launch_dft(t, T, 0);
    }
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
    for (auto pr : {std::make_pair("COURANT_LANES_H", &D->wh), std::make_pair("COURANT_LANES_E", &D->we)})
        if (const char *v = std::getenv(pr.first)) {
            const int w = std::atoi(v);
            *pr.second = w == 1 || w == 2 ? w : 4;
            if (pr.second == &D->wh) D->wh_fixed = true;
        }
    if (const char *v = std::getenv("COURANT_H_REVERSE")) {
        D->hrev = std::atoi(v) != 0;
        D->hrev_fixed = true;
    }
    if (const char *v = std::getenv("COURANT_TILE_ORDER")) {
        // xcd: XCD bands; plain: dispatch order; plain<T>: dispatch order in chunks of T tiles
        D->xcd = std::strncmp(v, "plain", 5) != 0;
        D->chunk = 0;
        if (!D->xcd && v[5])
            for (int t = std::atoi(v + 5); t > 1 && D->chunk < 6; t >>= 1) D->chunk++;
        D->xcd_fixed = true;
    }
    if (const char *v = std::getenv("COURANT_ISEG_H")) D->iseg_h = std::max(1, std::atoi(v)), D->tuned = true;
    if (const char *v = std::getenv("COURANT_ISEG_E")) D->iseg_e = std::max(1, std::atoi(v)), D->tuned = true;
    if (const char *v = std::getenv("COURANT_ISEG_S")) D->iseg_s = std::max(1, std::atoi(v)), D->tuned = true;
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
    D->pidx.release();
    D->ptab.release();
    D->ctab.release();
    D->crow.release();
    D->btab.release();
    D->d_jb.release();
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
    for (KernelTimer *t : {&D->tH, &D->tE, &D->tG, &D->tD, &D->tP, &D->tS}) t->destroy();
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
constexpr int kTuneCandS[] = {4, 8, 16};   // one-pass step: x planes per workgroup
constexpr int kTuneNs = sizeof kTuneCandS / sizeof kTuneCandS[0];
constexpr int kTuneSteps = 2;
constexpr int kTuneRounds = 2;
constexpr int kTuneLanesH[] = {4, 2, 1};   // k points per lane of the H update
constexpr int kTuneNl = sizeof kTuneLanesH / sizeof kTuneLanesH[0];
// tile orders: {XCD bands, log2 chunk} (chunk_tile), each with the H sweep forward or reversed
constexpr int kTuneOrders[][2] = {{1, 0}, {0, 0}, {0, 2}, {0, 3}, {0, 4}};
constexpr int kTuneNo = sizeof kTuneOrders / sizeof kTuneOrders[0];
constexpr long kTuneCostSplit = 1 + (long)kTuneRounds * (kTuneNl + kTuneNc * kTuneNc + 2 * kTuneNo) * (1 + kTuneSteps);
constexpr long kTuneCostStep = 1 + (long)kTuneRounds * kTuneNs * (1 + kTuneSteps);
constexpr long kTuneCost = kTuneCostSplit > kTuneCostStep ? kTuneCostSplit : kTuneCostStep;

std::mutex g_tune_mu;
std::map<std::array<long long, 10>, std::array<int, 7>> g_tune;

long autotune(cem_sim *s, Device *D, long budget) {
    const std::array<long long, 10> key = {s->n[0], s->n[1], s->n[2], s->has_pml, (long long)D->compressed,
                                           (long long)D->fused, D->wh_fixed ? D->wh : 0, D->we, (long long)D->onepass + 2 * D->steprows + 4 * (D->partner != nullptr) + 8 * (D->xcd_fixed ? 1 + D->xcd + 2 * D->chunk : 0) + 256 * (D->hrev_fixed ? 1 + D->hrev : 0),
                                           (long long)D->tj * 4096 + D->kw};
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune.find(key);
        if (it != g_tune.end()) {
            D->iseg_h = it->second[0];
            D->iseg_e = it->second[1];
            D->iseg_s = it->second[2];
            if (!D->onepass) D->wh = it->second[3];
            if (!D->onepass) D->xcd = it->second[4];
            if (!D->onepass) D->hrev = it->second[5];
            if (!D->onepass) D->chunk = it->second[6];
            D->tuned = true;
            return 0;
        }
    }
    const long cost = D->onepass ? kTuneCostStep : kTuneCostSplit;
    if (budget < cost) return 0;        // too few steps in this call: tune on a later one
    auto one_step = [&] {
        launch_he(s, D);
        launch_post(s, D);
    };
    auto reset = [&](std::initializer_list<KernelTimer *> ts) {
        for (KernelTimer *k : ts) {
            k->fold();
            k->total_ms = 0;
            k->count = 0;
        }
    };
    // One candidate: a settling step, then kTuneSteps steps whose dispatches stamp event
    // pairs; a stage's candidates are timed from their pairs after one synchronisation at
    // the end of the stage (no host round trip per candidate)
    struct Mark {
        size_t h0, h1, e0, e1, s0, s1;
    };
    auto measure = [&] {
        one_step();                     // settling step, untimed
        Mark m{D->tH.used, 0, D->tE.used, 0, D->tS.used, 0};
        D->profiling = true;
        for (int st = 0; st < kTuneSteps; st++) one_step();
        D->profiling = false;
        m.h1 = D->tH.used;
        m.e1 = D->tE.used;
        m.s1 = D->tS.used;
        return m;
    };
    auto settle = [&] { (void)hipStreamSynchronize(D->st); };
    reset({&D->tH, &D->tE, &D->tS});
    one_step();                         // warm-up step (first launches, caches)
    const bool dbg = std::getenv("COURANT_TUNE_DEBUG") != nullptr;
    if (D->onepass) {
        double t[kTuneNs];
        Mark mk[kTuneRounds][kTuneNs];
        for (int r = 0; r < kTuneRounds; r++)
            for (int c = 0; c < kTuneNs; c++) {
                D->iseg_s = kTuneCandS[c];
                mk[r][c] = measure();
            }
        settle();
        for (int r = 0; r < kTuneRounds; r++)
            for (int c = 0; c < kTuneNs; c++) {
                const double v = D->tS.span(mk[r][c].s0, mk[r][c].s1);
                t[c] = r ? std::min(t[c], v) : v;
            }
        reset({&D->tH, &D->tE, &D->tS});
        int b = 0;
        for (int c = 0; c < kTuneNs; c++)
            if (t[c] < t[b]) b = c;
        D->iseg_s = kTuneCandS[b];
        if (dbg) {
            std::fprintf(stderr, "courant autotune one-pass step us (x planes per workgroup):");
            for (int c = 0; c < kTuneNs; c++) std::fprintf(stderr, " %d: %.1f", kTuneCandS[c], t[c] / kTuneSteps * 1e3);
            std::fprintf(stderr, " -> %d\n", D->iseg_s);
        }
    } else {
        // the H update's lane width first, at the default planes per thread: 4 k points per
        // lane (128-bit accesses) on large grids, fewer (more lanes in flight) on small ones
        double tl[kTuneNl];
        Mark ml[kTuneRounds][kTuneNl];
        for (int r = 0; r < kTuneRounds; r++)
            for (int c = 0; c < kTuneNl; c++) {
                D->wh = D->wh_fixed ? D->wh : kTuneLanesH[c];
                ml[r][c] = measure();
            }
        settle();
        for (int r = 0; r < kTuneRounds; r++)
            for (int c = 0; c < kTuneNl; c++) {
                const double v = D->tH.span(ml[r][c].h0, ml[r][c].h1);
                tl[c] = r ? std::min(tl[c], v) : v;
            }
        reset({&D->tH, &D->tE, &D->tS});
        if (!D->wh_fixed) {
            int b = 0;
            for (int c = 0; c < kTuneNl; c++)
                if (tl[c] < tl[b]) b = c;
            D->wh = kTuneLanesH[b];
            if (dbg) {
                std::fprintf(stderr, "courant autotune H update us per step (k points per lane):");
                for (int c = 0; c < kTuneNl; c++) std::fprintf(stderr, " %d: %.1f", kTuneLanesH[c], tl[c] / kTuneSteps * 1e3);
                std::fprintf(stderr, " -> %d\n", D->wh);
            }
        }
        double t[kTuneNc][kTuneNc];
        Mark mi[kTuneRounds][kTuneNc][kTuneNc];
        for (int r = 0; r < kTuneRounds; r++)
            for (int ch = 0; ch < kTuneNc; ch++)
                for (int ce = 0; ce < kTuneNc; ce++) {
                    D->iseg_h = kTuneCand[ch];
                    D->iseg_e = kTuneCand[ce];
                    mi[r][ch][ce] = measure();
                }
        settle();
        for (int r = 0; r < kTuneRounds; r++)
            for (int ch = 0; ch < kTuneNc; ch++)
                for (int ce = 0; ce < kTuneNc; ce++) {
                    const Mark &m = mi[r][ch][ce];
                    const double v = D->tH.span(m.h0, m.h1) + D->tE.span(m.e0, m.e1);
                    t[ch][ce] = r ? std::min(t[ch][ce], v) : v;
                }
        reset({&D->tH, &D->tE, &D->tS});
        int bh = 0, be = 0;
        for (int ch = 0; ch < kTuneNc; ch++)
            for (int ce = 0; ce < kTuneNc; ce++)
                if (t[ch][ce] < t[bh][be]) bh = ch, be = ce;
        D->iseg_h = kTuneCand[bh];
        D->iseg_e = kTuneCand[be];
        if (dbg) {
            std::fprintf(stderr, "courant autotune H+E us per step (rows H, columns E = 1 2 4 8):\n");
            for (int ch = 0; ch < kTuneNc; ch++) {
                std::fprintf(stderr, "  H %d:", kTuneCand[ch]);
                for (int ce = 0; ce < kTuneNc; ce++) std::fprintf(stderr, " %7.1f", t[ch][ce] / kTuneSteps * 1e3);
                std::fprintf(stderr, "\n");
            }
            std::fprintf(stderr, "  -> H %d, E %d\n", D->iseg_h, D->iseg_e);
        }
        // then the tile order at that shape: XCD bands (tiles that share halo rows on one L2)
        // or dispatch order (consecutive tiles on consecutive XCDs, every XCD streaming the
        // same region of memory), each with the H sweep forward or reversed (zig-zag: each
        // sweep starts on the planes the one before it touched last); which one wins
        // depends on the grid
        double to[2 * kTuneNo];
        const int xcd0 = D->xcd, chunk0 = D->chunk;
        const int hrev0 = D->hrev;
        Mark mo[kTuneRounds][2 * kTuneNo];
        for (int r = 0; r < kTuneRounds; r++)
            for (int c = 0; c < 2 * kTuneNo; c++) {
                D->xcd = D->xcd_fixed ? xcd0 : kTuneOrders[c % kTuneNo][0];
                D->chunk = D->xcd_fixed ? chunk0 : kTuneOrders[c % kTuneNo][1];
                D->hrev = D->hrev_fixed ? hrev0 : c / kTuneNo;
                mo[r][c] = measure();
            }
        settle();
        for (int r = 0; r < kTuneRounds; r++)
            for (int c = 0; c < 2 * kTuneNo; c++) {
                const Mark &m = mo[r][c];
                const double v = D->tH.span(m.h0, m.h1) + D->tE.span(m.e0, m.e1);
                to[c] = r ? std::min(to[c], v) : v;
            }
        // XCD bands with the H sweep forward unless another order is clearly faster: on
        // grids that sit in the MALL the orders time within the measurement's noise here,
        // and XCD bands are the faster under graph replay there
        int b = 0;
        for (int c = 1; c < 2 * kTuneNo; c++)
            if (to[c] < to[b] && (b || to[c] < 0.98 * to[0])) b = c;
        D->xcd = D->xcd_fixed ? xcd0 : kTuneOrders[b % kTuneNo][0];
        D->chunk = D->xcd_fixed ? chunk0 : kTuneOrders[b % kTuneNo][1];
        D->hrev = D->hrev_fixed ? hrev0 : b / kTuneNo;
        if (dbg) {
            std::fprintf(stderr, "courant autotune H+E us per step (XCD bands, dispatch order in chunks of 1 4 8 16"
                                 " tiles; H forward, reversed):");
            for (int c = 0; c < 2 * kTuneNo; c++) std::fprintf(stderr, " %.1f", to[c] / kTuneSteps * 1e3);
            std::fprintf(stderr, " -> %s, chunk %d%s\n", D->xcd ? "XCD bands" : "dispatch order", 1 << D->chunk,
                         D->hrev ? ", H reversed" : "");
        }
    }
    reset({&D->tH, &D->tE, &D->tG, &D->tD, &D->tP, &D->tS});
    D->tuned = true;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[key] = {D->iseg_h, D->iseg_e, D->iseg_s, D->wh, D->xcd, D->hrev, D->chunk};
    return cost;
}

void cem::dev_step(cem_sim *s, long nsteps) {
    Device *D = s->dev;
    ensure_vlast(s, D);
    if (D->partner) ensure_vlast(D->partner_sim, D->partner);
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
        launch_he(s, D);
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
    a.E = fset(D, D->par);
    a.H = fset(D, D->par) + 3 * D->g.N;
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

// Reference layout (i si + j sj + k) <-> device rows of kp floats (padding zero).
static void rows_to_host(const cem_sim *s, const Grid &g, const float *dev, double *dst) {
    for (int i = 0; i <= s->n[0]; i++)
        for (int j = 0; j <= s->n[1]; j++) {
            const float *r = dev + (size_t)i * g.si + (size_t)j * g.sj;
            double *o = dst + (size_t)i * s->si + (size_t)j * s->sj;
            for (int k = 0; k <= s->n[2]; k++) o[k] = r[k];
        }
}

static void rows_to_device(const cem_sim *s, const Grid &g, const double *src, float *dev) {
    for (int i = 0; i <= s->n[0]; i++)
        for (int j = 0; j <= s->n[1]; j++) {
            float *r = dev + (size_t)i * g.si + (size_t)j * g.sj;
            const double *o = src + (size_t)i * s->si + (size_t)j * s->sj;
            for (int k = 0; k <= s->n[2]; k++) r[k] = (float)o[k];
        }
}

static std::vector<float> download_raw(Device *D, const float *src, size_t n) {
    std::vector<float> t(n);
    copy_sync(D, t.data(), src, n * sizeof(float), hipMemcpyDeviceToHost);
    return t;
}

static void upload_raw(Device *D, float *dst, const std::vector<float> &t) {
    copy_sync(D, dst, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice);
}

void cem::dev_field_download(cem_sim *s, int comp, double *dst) {
    Device *D = s->dev;
    if (!D) {
        std::memset(dst, 0, s->N * sizeof(double));
        return;
    }
    const std::vector<float> t = download_raw(D, fset(D, D->par) + (size_t)comp * D->g.N, (size_t)D->g.N);
    rows_to_host(s, D->g, t.data(), dst);
}

// Checkpoint / state arrays in the reference's layouts (export.c:187-271):
// which 0 / 1 psi_E / psi_H (comp, w), 2 Debye J pole-major [p ndisp + e],
// 3 Debye E_old, 4 SIBC states edge-major, 5 field comp.
void cem::dev_state_download(cem_sim *s, int which, int comp, int w, double *dst) {
    Device *D = s->dev;
    const Grid &g = D->g;
    const size_t nd = s->ndisp;
    if (which == 0 || which == 1) {
        const int is_e = which == 0;
        float *p = is_e ? D->psiE[comp][w] : D->psiHs[D->par][comp][w];
        const size_t len = is_e ? D->psiE_len[comp][w] : D->psiH_len[comp][w];
        if (!len) return;
        const std::vector<float> t = download_raw(D, p, len);
        const std::vector<size_t> m = psi_dev_map(s, D, is_e, comp, w);
        for (size_t r = 0; r < m.size(); r++) dst[r] = t[m[r]];
        return;
    }
    if (which == 2) {       // Debye J: the reference's pole-major array or the E kernel's box
        const size_t np = s->npole;
        if (!D->fused) {
            download_f32(D, D->J.p, dst, np * nd);
            return;
        }
        const std::vector<float> t = download_raw(D, D->J.p, D->J.n);
        for (size_t p = 0; p < np; p++)
            for (size_t e = 0; e < nd; e++) dst[p * nd + e] = t[D->jbox_of[e] + p * D->bkp];
        return;
    }
    if (which == 3) {       // Debye E_old: E of each dispersive edge before the next E update
        if (!D->fused) {
            ensure_vlast(s, D);
            download_f32(D, D->Eold.p, dst, nd);
            return;
        }
        // the E kernel reads E_old from E itself; listed edges hold the same value
        const std::vector<float> t = download_raw(D, fset(D, D->par), 3 * (size_t)g.N);
        for (size_t e = 0; e < nd; e++) {
            const unsigned long long gi = (unsigned long long)s->d_comp[e] * s->N + s->d_idx[e];
            dst[e] = t[dev_index(s, g, gi)];
        }
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
    if (which == 5) dev_field_download(s, comp, dst);
}

void cem::dev_state_upload(cem_sim *s, int which, int comp, int w, const double *src) {
    Device *D = s->dev;
    const Grid &g = D->g;
    const size_t nd = s->ndisp;
    if (which == 0 || which == 1) {
        const int is_e = which == 0;
        float *p = is_e ? D->psiE[comp][w] : D->psiHs[D->par][comp][w];
        const size_t len = is_e ? D->psiE_len[comp][w] : D->psiH_len[comp][w];
        if (!len) return;
        std::vector<float> t(len, 0.f);
        const std::vector<size_t> m = psi_dev_map(s, D, is_e, comp, w);
        for (size_t r = 0; r < m.size(); r++) t[m[r]] = (float)src[r];
        upload_raw(D, p, t);
        return;
    }
    if (which == 2) {
        const size_t np = s->npole;
        if (!D->fused) {
            upload_f32(D, D->J.p, src, np * nd);
            return;
        }
        std::vector<float> t(D->J.n, 0.f);
        for (size_t p = 0; p < np; p++)
            for (size_t e = 0; e < nd; e++) t[D->jbox_of[e] + p * D->bkp] = (float)src[p * nd + e];
        upload_raw(D, D->J.p, t);
        return;
    }
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
    if (which == 5) {
        std::vector<float> t((size_t)g.N, 0.f);
        rows_to_device(s, g, src, t.data());
        upload_raw(D, fset(D, D->par) + (size_t)comp * g.N, t);
        if (comp < 3) D->vlast_stale = D->eold_stale = true;
    }
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
    KernelTimer *t[6] = {&s->dev->tH, &s->dev->tE, &s->dev->tG, &s->dev->tD, &s->dev->tP, &s->dev->tS};
    if (which < 0 || which > 5) return -1;
    t[which]->fold();
    *total_ms = t[which]->total_ms;
    *count = t[which]->count;
    return 0;
}

extern "C" int cem_amd_set_profiling(cem_sim *s, int on) {
    if (!s || !s->finalized || dev_ensure(s)) return -1;
    s->dev->profiling = on != 0;
    for (KernelTimer *t : {&s->dev->tH, &s->dev->tE, &s->dev->tG, &s->dev->tD, &s->dev->tP, &s->dev->tS}) {
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

// Advance both right-hand sides of a pair (b an excitation clone of a, same step) n
// steps: each step's H and E updates of a and b run as one launch each (odd tiles b),
// reading a's coefficient classes, Debye tables and CPML profiles; the post-phase and
// probe kernels run per right-hand side; everything goes on a's stream (chunks of
// steps as one graph), and b's later work on its own stream is ordered after it.
extern "C" int cem_amd_step_pair(cem_sim *a, cem_sim *b, long n) {
    if (!a || !b || a == b || !a->finalized || !b->finalized) {
        set_error("step_pair needs two finalized simulations");
        return -1;
    }
    if (a->plan_id != b->plan_id) {
        set_error("step_pair needs an excitation clone of the same plan");
        return -1;
    }
    if (dev_ensure(a) || dev_ensure(b)) return -1;
    Device *A = a->dev, *B = b->dev;
    if (A->onepass || B->onepass || a->step != b->step || A->rows != B->rows || A->par != B->par) {
        set_error("step_pair: both simulations must run the split sweeps at the same step");
        return -1;
    }
    if (n <= 0) return 0;
    HIP_OK(hipStreamSynchronize(B->st));   // b's uploads and earlier steps are done
    const hipStream_t bst = B->st;
    B->st = A->st;
    A->partner = B;                       // a's launch shape (tuned on the pair's first
    A->partner_sim = b;                   // steps unless fixed) serves both
    dev_step(a, n);
    A->partner = nullptr;
    A->partner_sim = nullptr;
    B->iseg_h = A->iseg_h;
    B->iseg_e = A->iseg_e;
    B->tuned = A->tuned;
    B->st = bst;
    hipEvent_t ev;
    HIP_OK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    HIP_OK(hipEventRecord(ev, A->st));
    HIP_OK(hipStreamWaitEvent(bst, ev, 0));
    HIP_OK(hipEventDestroy(ev));
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Run the launch-shape autotuner now (its steps advance the simulation like
// ordinary steps); returns the steps it ran, 0 when the shape was already fixed.
extern "C" long cem_amd_tune(cem_sim *s) {
    if (!s || !s->finalized || dev_ensure(s)) return -1;
    Device *D = s->dev;
    if (D->tuned || D->profiling) return 0;
    ensure_vlast(s, D);
    const long n = autotune(s, D, kTuneCost);
    return hipGetLastError() == hipSuccess ? n : -1;
}

// Accumulate the buffered samples of the steps run so far into the DFT (queued
// on the simulation's stream, no host transfer).
extern "C" int cem_amd_flush_dft(cem_sim *s) {
    if (!s || !s->dev) return -1;
    launch_dft(s, s->dev, 0);
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

// bytes per point of the coefficient data the E update reads: 1 (class bytes), 4 (class
// words) or 24 (ca and cb arrays)
extern "C" int cem_amd_class_bytes(cem_sim *s) {
    if (!s || !s->dev) {
        set_error("simulation has not run on the GPU yet");
        return -1;
    }
    const Device *D = s->dev;
    return !D->compressed ? 24 : D->pidx.p ? 1 : 4;
}

// tile order of the H and E updates: bit 0 XCD bands (else dispatch order), bit 1 H sweep
// reversed, bits 2.. log2 of the dispatch-order chunk (chosen by the tuner unless fixed)
extern "C" int cem_amd_tile_order(cem_sim *s) {
    if (!s || !s->dev) {
        set_error("simulation has not run on the GPU yet");
        return -1;
    }
    return s->dev->xcd + 2 * s->dev->hrev + 4 * s->dev->chunk;
}

// k points per lane of the H and E updates (the H width chosen by the tuner unless fixed)
extern "C" int cem_amd_lane_widths(cem_sim *s, int *lanes_h, int *lanes_e) {
    if (!s || !s->dev) {
        set_error("simulation has not run on the GPU yet");
        return -1;
    }
    if (lanes_h) *lanes_h = s->dev->wh;
    if (lanes_e) *lanes_e = s->dev->we;
    return s->dev->tuned ? 1 : 0;
}

// Shape of the one-pass step kernel: rows per tile and x planes per workgroup;
// rows 0 when the simulation runs the two sweeps.
extern "C" int cem_amd_step_shape(cem_sim *s, int *rows, int *planes) {
    if (!s || !s->finalized || dev_ensure(s)) return -1;
    *rows = s->dev->onepass ? s->dev->tj : 0;
    *planes = s->dev->onepass ? s->dev->iseg_s : 0;
    return 0;
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
