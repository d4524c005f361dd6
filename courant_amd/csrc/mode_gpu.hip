// GPU (gfx950) FP64 shift-invert eigen-solver of the 2D port-mode operator
// (the device half of cem_mode2d_solve, mode2d.c:307-436).
//
//   k_probe_*   : A e_c for coloured unit vectors e_c (the host's probing), the
//                 operator A = P Q applied matrix-free (mode2d.c:181-217), a
//                 batch of colours per launch; the results are scattered into
//                 the complex band (kl = ku = bw, ldab = 3 bw + 1, column-major)
//   k_diag      : PEC unknowns pushed to -1e6 (|sigma| + 1), then - sigma
//   k_lu_panel  : one panel of nb columns factorised by one workgroup, one
//                 thread per panel row holding the row in registers: partial
//                 pivoting over the kl rows below the diagonal, row
//                 interchanges within the panel, scaling, rank-1 updates
//   k_lu_trail  : the panel's interchanges and rank-1 updates applied to the
//                 trailing columns (one or two columns per wavefront, in
//                 registers) in the order of the unblocked column loop, so
//                 every entry sees the host's update sequence
//   k_inverse   : the whole inverse iteration in one workgroup: normalisation,
//                 phase-aligned change, band forward / back solves (one
//                 wavefront walks the columns through an LDS window of the
//                 right-hand side while the others stream the next band
//                 columns into a double-buffered LDS ring), Rayleigh quotient,
//                 the host's convergence test; reductions in a fixed order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "devmem.hpp"
#include "graph.hpp"
#include "mode_gpu.hpp"
#include "sim.hpp"

namespace cem {

namespace {

constexpr double kMu0 = CEM_MU0;
constexpr int kLuThreads = 1024;       // trailing update: 16 wavefronts share one LDS copy of the panel
constexpr int kTrailRows = 8;          // rows per lane of a trailing column (panel rows <= 512)
constexpr int kCUs = 256;              // compute units of an MI355X
constexpr int kInvThreads = 512;
constexpr size_t kRingBytes = 150 * 1024;  // LDS ring of band columns for the triangular solves (gfx950: 160 KiB per workgroup)
constexpr int kProbeBatch = 64;        // colours per probing launch

struct zc {
    double r, i;
};
__device__ __forceinline__ zc mk(double r, double i) {
    zc z;
    z.r = r;
    z.i = i;
    return z;
}
__device__ __forceinline__ zc zadd(zc a, zc b) { return mk(a.r + b.r, a.i + b.i); }
__device__ __forceinline__ zc zsub(zc a, zc b) { return mk(a.r - b.r, a.i - b.i); }
__device__ __forceinline__ zc zmul(zc a, zc b) { return mk(a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r); }
__device__ __forceinline__ zc zrs(double s, zc a) { return mk(s * a.r, s * a.i); }
__device__ __forceinline__ zc zconj(zc a) { return mk(a.r, -a.i); }
// Smith's method (R. L. Smith, CACM 5(8), 1962)
__device__ __forceinline__ zc zdiv(zc a, zc b) {
    if (fabs(b.r) < fabs(b.i)) {
        const double ratio = b.r / b.i, den = b.r * ratio + b.i;
        return mk((a.r * ratio + a.i) / den, (a.i * ratio - a.r) / den);
    }
    const double ratio = b.i / b.r, den = b.i * ratio + b.r;
    return mk((a.i * ratio + a.r) / den, (a.i - a.r * ratio) / den);
}
__device__ __forceinline__ double zabs(zc a) { return hypot(a.r, a.i); }
// c - a b with four fused multiply-adds (the triangular solves)
__device__ __forceinline__ zc zfms(zc c, zc a, zc b) {
    return mk(fma(-a.r, b.r, fma(a.i, b.i, c.r)), fma(-a.r, b.i, fma(-a.i, b.r, c.i)));
}
__device__ __forceinline__ zc shfl(zc v, int src) { return mk(__shfl(v.r, src), __shfl(v.i, src)); }
// value of lane src (uniform) to every lane, through scalar registers
__device__ __forceinline__ double bcast(double v, int src) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(b & 0xffffffffll), src);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), src);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ zc bcast(zc v, int src) { return mk(bcast(v.r, src), bcast(v.i, src)); }
// the lane's value exchanged by a DPP move (CTRL: quad xor 1 = 0xB1, quad xor 2 =
// 0x4E, half-row mirror = 0x141, row mirror = 0x140) or by a swizzle that exchanges
// the two 16-lane rows of each 32-lane half
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, 0xF, 0xF, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ double swap_rows16(double v) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_ds_swizzle((int)(b & 0xffffffffll), 0x401F);
    const unsigned hi = (unsigned)__builtin_amdgcn_ds_swizzle((int)(b >> 32), 0x401F);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
// largest value over the wavefront, in every lane (all 64 lanes active; no NaN)
__device__ __forceinline__ double wave_max(double v) {
    v = fmax(v, dpp_move<0xB1>(v));
    v = fmax(v, dpp_move<0x4E>(v));
    v = fmax(v, dpp_move<0x141>(v));
    v = fmax(v, dpp_move<0x140>(v));
    v = fmax(v, swap_rows16(v));
    return fmax(bcast(v, 0), bcast(v, 32));
}

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

struct Geo {
    int ny, nz, nu, nv, nn, nc, n, bw, ldab, wc;
    // a batch of problems (frequencies) of one geometry solves in one launch
    // sequence: problem blockIdx.y finds its arrays at these strides (entries)
    long long sband, svec, scoef, shw, shu, shv, sew;
};

// dst[t] = src(t) for t = t0, t0 + nt, ... < total: each thread issues U loads
// before it stores any of them, so a copy costs one memory latency per U
// elements instead of one per element.
template <int U, typename F>
__device__ __forceinline__ void copy_batched(zc *dst, int total, int t0, int nt, F src) {
    for (int base = t0; base < total; base += nt * U) {
        zc tmp[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int t = base + u * nt;
            tmp[u] = t < total ? src(t) : zc{0.0, 0.0};
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int t = base + u * nt;
            if (t < total) dst[t] = tmp[u];
        }
    }
}

// band entry (i, j) (requires j - 2 bw <= i <= j + bw)
__device__ __forceinline__ long long bidx(const Geo &g, int i, int j) {
    return (long long)j * g.ldab + (2 * g.bw + i - j);
}

// unknown order: j outer, k inner, E_u then E_v (mode2d.c ordering, 2 nz + 1 per j < ny)
__device__ __forceinline__ int gu(const Geo &g, int j, int k) { return j * (2 * g.nz + 1) + (k < g.nz ? 2 * k : 2 * g.nz); }
__device__ __forceinline__ int gv(const Geo &g, int j, int k) {
    return j < g.ny ? j * (2 * g.nz + 1) + 2 * k + 1 : g.ny * (2 * g.nz + 1) + k;
}

struct Coef {
    const zc *iyc, *iyn, *izc, *izn, *eu, *ev, *iw;
    const unsigned char *pec_u, *pec_v;
    double w;
    const double *wv;       // angular frequency of each problem of the batch
};

// the coefficients of problem b of the batch (the PEC masks are the geometry's)
__device__ __forceinline__ Coef coef_of(const Geo &g, Coef c, int b) {
    const long long o = (long long)b * g.scoef;
    c.iyc += o;
    c.iyn += o;
    c.izc += o;
    c.izn += o;
    c.eu += o;
    c.ev += o;
    c.iw += o;
    c.w = c.wv[b];
    return c;
}

__device__ __forceinline__ zc unit_u(const Geo &g, const Coef &c, int col, int j, int k) {
    return (gu(g, j, k) % g.wc == col && !c.pec_u[j * (g.nz + 1) + k]) ? mk(1.0, 0.0) : mk(0.0, 0.0);
}
__device__ __forceinline__ zc unit_v(const Geo &g, const Coef &c, int col, int j, int k) {
    return (gv(g, j, k) % g.wc == col && !c.pec_v[j * g.nz + k]) ? mk(1.0, 0.0) : mk(0.0, 0.0);
}

// dst[t] = src[t].r: the real words of (x, 0) pairs
__global__ void k_real_parts(const zc *src, double *dst, int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[t] = src[t].r;
}

// hw = curl_x of the coloured input (cells)
__global__ void k_probe_hw(Geo g, Coef c0, int col0, int nb, zc *hw) {
    const Coef c = coef_of(g, c0, blockIdx.y);
    hw += blockIdx.y * g.shw;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)nb * g.nc) return;
    const int b = (int)(t / g.nc), q = (int)(t % g.nc);
    const int j = q / g.nz, k = q % g.nz, col = col0 + b;
    const zc dv = zsub(unit_v(g, c, col, j + 1, k), unit_v(g, c, col, j, k));
    const zc du = zsub(unit_u(g, c, col, j, k + 1), unit_u(g, c, col, j, k));
    hw[t] = zsub(zmul(dv, c.iyc[j]), zmul(du, c.izc[k]));
}

// h = Q e: hu at E_v positions, hv at E_u positions (mode2d.c:181-199)
__global__ void k_probe_h(Geo g, Coef c0, int col0, int nb, const zc *hw, zc *hu, zc *hv) {
    const Coef c = coef_of(g, c0, blockIdx.y);
    hw += blockIdx.y * g.shw;
    hu += blockIdx.y * g.shu;
    hv += blockIdx.y * g.shv;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int per = g.nv + g.nu;
    if (t >= (long long)nb * per) return;
    const int b = (int)(t / per), q = (int)(t % per), col = col0 + b;
    const zc *w = hw + (size_t)b * g.nc;
    const double cc = 1.0 / (c.w * kMu0);
    const zc zero = mk(0.0, 0.0);
    if (q < g.nv) {
        const int j = q / g.nz, k = q % g.nz;
        const zc r = zsub(j < g.ny ? w[j * g.nz + k] : zero, j > 0 ? w[(j - 1) * g.nz + k] : zero);
        const zc a = zmul(zrs(-c.w, c.ev[q]), unit_v(g, c, col, j, k));
        hu[(size_t)b * g.nv + q] = zsub(a, zmul(zrs(cc, r), c.iyn[j]));
    } else {
        const int q2 = q - g.nv, j = q2 / (g.nz + 1), k = q2 % (g.nz + 1);
        const zc r = zsub(k < g.nz ? w[j * g.nz + k] : zero, k > 0 ? w[j * g.nz + k - 1] : zero);
        const zc a = zmul(zrs(c.w, c.eu[q2]), unit_u(g, c, col, j, k));
        hv[(size_t)b * g.nu + q2] = zsub(a, zmul(zrs(cc, r), c.izn[k]));
    }
}

// ew at nodes (mode2d.c:203-209)
__global__ void k_probe_ew(Geo g, Coef c0, int nb, const zc *hu, const zc *hv, zc *ew) {
    const Coef c = coef_of(g, c0, blockIdx.y);
    hu += blockIdx.y * g.shu;
    hv += blockIdx.y * g.shv;
    ew += blockIdx.y * g.sew;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)nb * g.nn) return;
    const int b = (int)(t / g.nn), q = (int)(t % g.nn);
    const int j = q / (g.nz + 1), k = q % (g.nz + 1);
    const zc *u = hu + (size_t)b * g.nv, *v = hv + (size_t)b * g.nu;
    const zc zero = mk(0.0, 0.0);
    const zc du = zmul(zsub(j < g.ny ? v[j * (g.nz + 1) + k] : zero, j > 0 ? v[(j - 1) * (g.nz + 1) + k] : zero),
                       c.iyn[j]);
    const zc dv = zmul(zsub(k < g.nz ? u[j * g.nz + k] : zero, k > 0 ? u[j * g.nz + k - 1] : zero), c.izn[k]);
    ew[t] = zmul(zsub(du, dv), c.iw[q]);
}

// e = P h at every unknown (mode2d.c:210-217), PEC rows zero, scattered into band
// column g of the colour's unit within bw of the row
__global__ void k_probe_out(Geo g, Coef c0, int col0, int nb, const zc *hu, const zc *hv, const zc *ew, zc *band) {
    const Coef c = coef_of(g, c0, blockIdx.y);
    hu += blockIdx.y * g.shu;
    hv += blockIdx.y * g.shv;
    ew += blockIdx.y * g.sew;
    band += blockIdx.y * g.sband;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)nb * g.n) return;
    const int b = (int)(t / g.n), r = (int)(t % g.n), col = col0 + b;
    const int tt = r - col + g.bw;
    if (tt < 0) return;
    const int gcol = col + (tt / g.wc) * g.wc;
    if (gcol >= g.n) return;
    const zc *u = hu + (size_t)b * g.nv, *v = hv + (size_t)b * g.nu, *e = ew + (size_t)b * g.nn;
    const double cc = 1.0 / c.w;
    zc y = mk(0.0, 0.0);
    const int per = 2 * g.nz + 1;
    if (r < g.ny * per) {
        const int j = r / per, s = r % per;
        if (s == 2 * g.nz || (s & 1) == 0) {   // E_u(j, k)
            const int k = s == 2 * g.nz ? g.nz : s / 2;
            if (!c.pec_u[j * (g.nz + 1) + k])
                y = zadd(zrs(c.w * kMu0, v[j * (g.nz + 1) + k]),
                         zmul(zrs(cc, zsub(e[(j + 1) * (g.nz + 1) + k], e[j * (g.nz + 1) + k])), c.iyc[j]));
        } else {                                 // E_v(j, k)
            const int k = s / 2;
            if (!c.pec_v[j * g.nz + k])
                y = zadd(zrs(-c.w * kMu0, u[j * g.nz + k]),
                         zmul(zrs(cc, zsub(e[j * (g.nz + 1) + k + 1], e[j * (g.nz + 1) + k])), c.izc[k]));
        }
    } else {
        const int j = g.ny, k = r - g.ny * per;
        if (!c.pec_v[j * g.nz + k])
            y = zadd(zrs(-c.w * kMu0, u[j * g.nz + k]),
                     zmul(zrs(cc, zsub(e[j * (g.nz + 1) + k + 1], e[j * (g.nz + 1) + k])), c.izc[k]));
    }
    band[bidx(g, r, gcol)] = y;
}

__global__ void k_diag(Geo g, Coef c, const zc *sigmas, const double *bigs, zc *band) {
    const zc sigma = sigmas[blockIdx.y];
    const double big = bigs[blockIdx.y];
    band += blockIdx.y * g.sband;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= g.n) return;
    const int per = 2 * g.nz + 1;
    bool pec;
    if (r < g.ny * per) {
        const int j = r / per, s = r % per;
        if (s == 2 * g.nz || (s & 1) == 0) pec = c.pec_u[j * (g.nz + 1) + (s == 2 * g.nz ? g.nz : s / 2)];
        else pec = c.pec_v[j * g.nz + s / 2];
    } else {
        pec = c.pec_v[g.ny * g.nz + (r - g.ny * per)];
    }
    zc &d = band[bidx(g, r, r)];
    if (pec) d = mk(big, 0.0);
    d = zsub(d, sigma);
}

// ---- band LU -------------------------------------------------------------
// Panel of columns [j0, j1), nbp = j1 - j0 <= NB, rows [j0, j0 + R),
// R = min(j1 - 1 + kl, n - 1) - j0 + 1 <= blockDim.x: thread r holds row j0 + r of
// the panel's not yet factorised columns in registers, a[0] being column j (entries
// below row c + kl of column c are zero and stay so).  Per column j: the pivot
// (largest |.| over rows j..j + kl, the first on ties), the interchange of rows j
// and p over the columns >= j through LDS, the scaling of the multipliers and the
// rank-1 update of the later panel columns; column j is then final and stored, and
// the window shifts by one column (a rolled loop: the kernel runs once per panel,
// and fully unrolled code would be fetched from L2 each time).  Every entry sees
// the operations of the unblocked column loop in its order.  Two barriers per
// column: the LDS buffers alternate with the column's parity, so a buffer is
// rewritten only after every thread has passed the next column's first barrier.
template <int NB>
__global__ void __launch_bounds__(512) k_lu_panel(Geo g, zc *band, int *ipiv, int *info, int j0, int j1) {
    band += blockIdx.y * g.sband;
    ipiv += blockIdx.y * g.svec;
    info += blockIdx.y;
    __shared__ double wbest[2][8];
    __shared__ int wrow[2][8];
    __shared__ zc prow[2][NB], jrow[2][NB];
    const int kl = g.bw, nbp = j1 - j0;
    const int R = min(j1 - 1 + kl, g.n - 1) - j0 + 1;
    const int r = threadIdx.x, row = j0 + r, lane = r & 63, wv = r >> 6, nw = blockDim.x >> 6;
    zc a[NB];
#pragma unroll
    for (int c = 0; c < NB; c++)
        a[c] = r < R && c < nbp && row <= j0 + c + kl ? band[bidx(g, row, j0 + c)] : mk(0.0, 0.0);
    // loads complete here (vmcnt 0): otherwise the loop, whose column stores share the
    // counter, waits for every store before each use of a row register
    __builtin_amdgcn_s_waitcnt(0x0F70);
    for (int cj = 0; cj < nbp; cj++) {
        const int j = j0 + cj, last = min(j + kl, g.n - 1), s = cj & 1;
        double best = -1.0;
        int bi = j;
        if (row >= j && row <= last) {
            const double v = zabs(a[0]);
            if (v > best) {
                best = v;
                bi = row;
            }
        }
        // the wavefront's largest value; its first row (rows ascend with the lanes);
        // row j where no lane holds a candidate value (-1 everywhere)
        {
            const double m = wave_max(best);
            const unsigned long long hit = __ballot(best == m);
            bi = m == -1.0 ? j : j0 + (wv << 6) + __builtin_ctzll(hit);
            best = m;
        }
        if (lane == 0) {
            wbest[s][wv] = best;
            wrow[s][wv] = bi;
        }
        __syncthreads();
        {
            double wb[8];
            int wr[8];
#pragma unroll
            for (int w = 0; w < 8; w++) {
                wb[w] = wbest[s][w];
                wr[w] = wrow[s][w];
            }
            best = wb[0];
            bi = wr[0];
#pragma unroll
            for (int w = 1; w < 8; w++)
                if (w < nw && (wb[w] > best || (wb[w] == best && wr[w] < bi))) {
                    best = wb[w];
                    bi = wr[w];
                }
        }
        const int p = bi;
        if (r == 0) {
            ipiv[j] = p;
            if (best == 0.0 && *info == 0) *info = j + 1;
        }
        // whole windows: the registers past the panel's last column hold zeros and are
        // never stored, so they are exchanged and updated without a test
        if (row == p) {
#pragma unroll
            for (int c = 0; c < NB; c++) prow[s][c] = a[c];
        } else if (row == j) {
#pragma unroll
            for (int c = 0; c < NB; c++) jrow[s][c] = a[c];
        }
        __syncthreads();
        if (p != j) {
            if (row == j) {
#pragma unroll
                for (int c = 0; c < NB; c++) a[c] = prow[s][c];
            } else if (row == p) {
#pragma unroll
                for (int c = 0; c < NB; c++) a[c] = jrow[s][c];
            }
        }
        if (row > j && row <= last) {
            const zc l = zdiv(a[0], prow[s][0]);
            a[0] = l;
#pragma unroll
            for (int c = 1; c < NB; c++) a[c] = zsub(a[c], zmul(l, prow[s][c]));
        }
        if (r < R && row <= j + kl) band[bidx(g, row, j)] = a[0];
#pragma unroll
        for (int c = 0; c + 1 < NB; c++) a[c] = a[c + 1];
        a[NB - 1] = mk(0.0, 0.0);
    }
}

// Trailing columns c in [j1, min(j1 - 1 + kv, n - 1)]: each wavefront holds CPW of them,
// 16 apart (the workgroup's 16 wavefronts cover 16 CPW consecutive columns), so that one
// LDS read of a multiplier serves CPW columns and one LDS copy of the panel's multipliers
// serves 16 CPW columns.  TR: rows per lane, 64 TR >= the panel's rows R (kl + nb): a
// narrower band walks fewer register rows per rank-1 update.  Each column sees the same
// interchanges and rank-1 updates in the same order whatever TR and CPW are.
template <int TR, int CPW>
__global__ void __launch_bounds__(kLuThreads) k_lu_trail(Geo g, zc *band, const int *ipiv, int j0, int j1) {
    constexpr int W = kLuThreads / 64;
    extern __shared__ zc L[];
    band += blockIdx.y * g.sband;
    ipiv += blockIdx.y * g.svec;
    const int kl = g.bw, kv = 2 * g.bw, nbp = j1 - j0;
    const int R = min(j1 - 1 + kl, g.n - 1) - j0 + 1;
    copy_batched<16>(L, R * nbp, threadIdx.x, kLuThreads, [=](int t) {
        const int cc = t / R, rr = t % R, col = j0 + cc, row = j0 + rr;
        return row > col && row <= col + kl ? band[bidx(g, row, col)] : mk(0.0, 0.0);
    });
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int c0 = j1 + (int)blockIdx.x * W * CPW + (int)(threadIdx.x >> 6);
    const int cend = min(j1 - 1 + kv, g.n - 1);
    if (c0 > cend) return;
    zc v[CPW][TR];
#pragma unroll
    for (int u = 0; u < CPW; u++) {
        const int c = c0 + u * W, rlo = max(j0, c - kv);
#pragma unroll
        for (int m = 0; m < TR; m++) {
            const int row = j0 + lane + 64 * m;
            v[u][m] = c <= cend && row < j0 + R && row >= rlo ? band[bidx(g, row, c)] : mk(0.0, 0.0);
        }
    }
    for (int j = j0; j < j1; j++) {
        if (c0 > j + kv) continue;   // later columns of the wavefront lie further right
        const int p = ipiv[j], last = min(j + kl, g.n - 1);
        zc aj[CPW], x[CPW];   // rows j and p of each column (x: row j after the interchange)
#pragma unroll
        for (int u = 0; u < CPW; u++) {
            zc a = mk(0.0, 0.0), b = mk(0.0, 0.0);
#pragma unroll
            for (int m = 0; m < TR; m++) {
                if (m == (j - j0) / 64) a = v[u][m];
                if (m == (p - j0) / 64) b = v[u][m];
            }
            aj[u] = mk(bcast(a.r, (j - j0) % 64), bcast(a.i, (j - j0) % 64));
            x[u] = mk(bcast(b.r, (p - j0) % 64), bcast(b.i, (p - j0) % 64));
        }
#pragma unroll
        for (int m = 0; m < TR; m++) {
            const int row = j0 + lane + 64 * m;
            if (p != j) {
#pragma unroll
                for (int u = 0; u < CPW; u++) {
                    if (c0 + u * W > j + kv) continue;
                    if (row == j) v[u][m] = x[u];
                    else if (row == p) v[u][m] = aj[u];
                }
            }
            if (row > j && row <= last) {
                const zc l = L[(size_t)(j - j0) * R + (row - j0)];
#pragma unroll
                for (int u = 0; u < CPW; u++)
                    if (c0 + u * W <= j + kv) v[u][m] = zsub(v[u][m], zmul(l, x[u]));
            }
        }
    }
#pragma unroll
    for (int u = 0; u < CPW; u++) {
        const int c = c0 + u * W, rlo = max(j0, c - kv);
        if (c > cend) break;
#pragma unroll
        for (int m = 0; m < TR; m++) {
            const int row = j0 + lane + 64 * m;
            if (row < j0 + R && row >= rlo) band[bidx(g, row, c)] = v[u][m];
        }
    }
}

// ---- inverse iteration -------------------------------------------------
struct InvArgs {
    Geo g;
    const zc *band;
    const int *ipiv;
    zc *x, *xo, *y;
    const zc *sigma;        // shift of each problem of the batch
    int maxit;
    double tol;
    double *out;   // per problem 10 words: lam (2), rc, iterations, phase timers
    int chunk_f, chunk_b;   // band columns per ring slot (forward, back)
    int ring;               // ring size in entries
    int timing;             // accumulate cycle counts of the solve phases into out[4..9]
};

__device__ double block_sum(double v, double *red) {
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int i = 0; i < kInvThreads / 64; i++) s += red[i];
    return s;
}

// Window of the right-hand side held in registers of one wavefront: slot k of
// lane l holds entry 64 (b0 + k) + l, b0 the current block; the slots rotate
// by one at every block change.  A uniform slot s is selected through a bit
// mask, so every access stays a register access.
template <int MAXS>
__device__ __forceinline__ zc slot_get(const zc (&v)[MAXS], unsigned m) {
    zc r = mk(0.0, 0.0);
#pragma unroll
    for (int k = 0; k < MAXS; k++) {
        const bool on = (m >> k) & 1u;
        r.r = on ? v[k].r : r.r;
        r.i = on ? v[k].i : r.i;
    }
    return r;
}
template <int MAXS>
__device__ __forceinline__ void slot_put(zc (&v)[MAXS], unsigned m, zc x) {
#pragma unroll
    for (int k = 0; k < MAXS; k++) {
        const bool on = (m >> k) & 1u;
        v[k].r = on ? x.r : v[k].r;
        v[k].i = on ? x.i : v[k].i;
    }
}

// Ring fills: multipliers L(j+1 .. j+kl, j) and pivots of the columns of forward
// chunk ch; U(j-kv .. j, j) of the columns of back chunk ch (walked downwards).
__device__ __forceinline__ void fill_fwd(const Geo g, const zc *__restrict__ band, const int *__restrict__ ipiv,
                                         zc *ring, int *rpiv, int CH, int ch, int t0, int nt) {
    const int K = g.bw, n = g.n;
    zc *dst = ring + (size_t)(ch % 3) * CH * K;
    int *pdst = rpiv + (ch % 3) * CH;
    copy_batched<8>(dst, CH * K, t0, nt, [=](int t) {
        const int j = ch * CH + t / K, i = j + 1 + t % K;
        return (j < n && i < n) ? band[bidx(g, i, j)] : mk(0.0, 0.0);
    });
    for (int t = t0; t < CH; t += nt) pdst[t] = ch * CH + t < n ? ipiv[ch * CH + t] : 0;
}

__device__ __forceinline__ void fill_back(const Geo g, const zc *__restrict__ band, zc *ring, int CH, int ch, int t0,
                                          int nt) {
    const int kv = 2 * g.bw, K = kv + 1, n = g.n;
    zc *dst = ring + (size_t)(ch % 3) * CH * K;
    copy_batched<8>(dst, CH * K, t0, nt, [=](int t) {
        const int j = n - 1 - ch * CH - t / K, i = j - kv + t % K;
        if (j < 0 || i < 0) return mk(0.0, 0.0);
        const zc u = band[bidx(g, i, j)];
        return i == j ? zdiv(mk(1.0, 0.0), u) : u;   // the diagonal as its reciprocal
    });
}

// In-place solve of (LU) y = y with the band factors (the host's band_solve):
// forward with the interchanges, then back substitution.  Wavefront 0 walks
// the columns with the live part of y in registers (one 64-entry block moved
// out and the next loaded one block ahead); the other wavefronts stream the
// band columns the walk needs into a three-deep LDS ring, two chunks ahead,
// with one workgroup barrier per chunk.
template <int SF, int SB>
__device__ __forceinline__ void band_solve(const Geo g, const zc *__restrict__ band, const int *__restrict__ ipiv, zc *y,
                           int chunk_f, int chunk_b, zc *ring, int *rpiv, const zc *zslot, long long *tm) {
    const int n = g.n, kl = g.bw, kv = 2 * g.bw;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const zc zero = mk(0.0, 0.0);
    // ---- forward: for j: swap y[j], y[p]; y[j+1..j+kl] -= L(:, j) y[j]
    {
        constexpr int MAXS = SF, S = SF;
        const int K = kl, CH = chunk_f, nchunk = (n + CH - 1) / CH;
        fill_fwd(g, band, ipiv, ring, rpiv, CH, 0, threadIdx.x, kInvThreads);
        if (nchunk > 1) fill_fwd(g, band, ipiv, ring, rpiv, CH, 1, threadIdx.x, kInvThreads);
        __syncthreads();
        zc v[MAXS], pend = zero;   // slot k: block (j >> 6) + k
        const int nblk = (n + 63) / 64;
        if (wv == 0) {
#pragma unroll
            for (int k = 0; k < MAXS; k++) {
                const int i = 64 * k + lane;
                v[k] = k < S && i < n ? y[i] : zero;
            }
            const int i = 64 * S + lane;
            pend = i < n ? y[i] : zero;
        }
        for (int ch = 0; ch < nchunk; ch++) {
            const int cj0 = ch * CH, cj1 = min(n, cj0 + CH);
            const long long c0 = tm ? clock64() : 0;
            if (wv == 0) {
                const zc *cur = ring + (size_t)(ch % 3) * CH * K;
                const int *cp = rpiv + (ch % 3) * CH;
                // the block change sits outside the column loop: inside it every slot stays
                // in its registers (a rotation on one path of the loop had the compiler
                // copying all slots at every other column)
                for (int j = cj0; j < cj1;) {
                    const int jb = j >> 6;
                    if ((j & 63) == 0 && jb > 0) {
                        // block jb-1 is final: out to y; slots move down, block jb-1+S enters
                        if (64 * (jb - 1) + lane < n) y[64 * (jb - 1) + lane] = v[0];
#pragma unroll
                        for (int k = 0; k + 1 < MAXS; k++) v[k] = v[k + 1];
                        slot_put(v, 1u << (S - 1), pend);
                        const int i = 64 * (jb + S) + lane;
                        pend = jb + S < nblk && i < n ? y[i] : zero;
                    }
                    const int jend = min(cj1, 64 * jb + 64);
                    for (; j < jend; j++) {
                        // the column's pivot and its multipliers for every slot, read from
                        // the ring together before any use; outside rows j+1 .. j+kl a
                        // multiplier is read from the zero entry (rows past n are zeros in the
                        // ring), which leaves its entry unchanged, so every slot updates
                        // without a branch or a mask
                        const zc *l = cur + (size_t)(j - cj0) * K;
                        const int p = __builtin_amdgcn_readfirstlane(cp[j - cj0]);
                        zc lv[MAXS];
#pragma unroll
                        for (int k = 0; k < MAXS; k++) {
                            const int o = 64 * (jb + k) + lane - j - 1;
                            lv[k] = *((unsigned)o < (unsigned)K ? l + o : zslot);
                        }
                        zc vj = bcast(v[0], j & 63);
                        if (p != j) {
                            const unsigned mp = 1u << ((p >> 6) - jb);
                            const zc vp = bcast(slot_get(v, mp), p & 63);
                            if (lane == (j & 63)) v[0] = vp;
                            if (lane == (p & 63)) slot_put(v, mp, vj);
                            vj = vp;
                        }
#pragma unroll
                        for (int k = 0; k < MAXS; k++) v[k] = zfms(v[k], lv[k], vj);
                    }
                }
            } else if (ch + 2 < nchunk) {
                fill_fwd(g, band, ipiv, ring, rpiv, CH, ch + 2, threadIdx.x - 64, kInvThreads - 64);
            }
            const long long c1 = tm ? clock64() : 0;
            __syncthreads();
            if (tm && threadIdx.x == 0) {
                tm[0] += c1 - c0;
                tm[1] += clock64() - c1;
            }
            if (tm && threadIdx.x == 64) tm[4] += c1 - c0;
        }
        if (wv == 0) {   // the last block (every earlier one went out at a block start)
            const int jb = (n - 1) >> 6;
            if (64 * jb + lane < n) y[64 * jb + lane] = v[0];
        }
        __syncthreads();
    }
    // ---- back: for j = n-1..0: y[j] /= U(j, j); y[j-kv..j-1] -= U(:, j) y[j]
    {
        constexpr int MAXS = SB, S = SB;
        const int K = kv + 1, CH = chunk_b, nchunk = (n + CH - 1) / CH;
        fill_back(g, band, ring, CH, 0, threadIdx.x, kInvThreads);
        if (nchunk > 1) fill_back(g, band, ring, CH, 1, threadIdx.x, kInvThreads);
        __syncthreads();
        zc v[MAXS], pend = zero;   // slot k: block (j >> 6) - k
        const int top = (n - 1) >> 6;
        if (wv == 0) {
#pragma unroll
            for (int k = 0; k < MAXS; k++) {
                const int b = top - k, i = 64 * b + lane;
                v[k] = k < S && b >= 0 && i < n ? y[i] : zero;
            }
            pend = top - S >= 0 ? y[64 * (top - S) + lane] : zero;
        }
        for (int ch = 0; ch < nchunk; ch++) {
            const int cj1 = n - ch * CH, cj0 = max(0, cj1 - CH);
            const long long c0 = tm ? clock64() : 0;
            if (wv == 0) {
                const zc *cur = ring + (size_t)(ch % 3) * CH * K;
                for (int j = cj1 - 1; j >= cj0;) {   // block change outside the column loop
                    const int jb = j >> 6;
                    if ((j & 63) == 63 && jb < top) {
                        // block jb+1 is final: out to y; slots move up, block jb+1-S enters
                        if (64 * (jb + 1) + lane < n) y[64 * (jb + 1) + lane] = v[0];
#pragma unroll
                        for (int k = 0; k + 1 < MAXS; k++) v[k] = v[k + 1];
                        slot_put(v, 1u << (S - 1), pend);
                        pend = jb - S >= 0 ? y[64 * (jb - S) + lane] : zero;
                    }
                    const int jend = max(cj0, 64 * jb) - 1;
                    for (; j > jend; j--) {
                        const zc *u = cur + (size_t)(cj1 - 1 - j) * K;   // u[rr] = U(j - kv + rr, j), u[kv] = 1/U(j, j)
                        // the reciprocal pivot and every slot's multiplier read together; outside
                        // rows j-kv .. j-1 from the zero entry (rows below 0 are zeros in the ring)
                        const zc ukv = u[kv];
                        zc uv[MAXS];
#pragma unroll
                        for (int k = 0; k < MAXS; k++) {
                            const int o = 64 * (jb - k) + lane - (j - kv);
                            uv[k] = *((unsigned)o < (unsigned)kv ? u + o : zslot);
                        }
                        const zc xj = zmul(bcast(v[0], j & 63), ukv);
                        if (lane == (j & 63)) v[0] = xj;
#pragma unroll
                        for (int k = 0; k < MAXS; k++) v[k] = zfms(v[k], uv[k], xj);
                    }
                }
            } else if (ch + 2 < nchunk) {
                fill_back(g, band, ring, CH, ch + 2, threadIdx.x - 64, kInvThreads - 64);
            }
            const long long c1 = tm ? clock64() : 0;
            __syncthreads();
            if (tm && threadIdx.x == 0) {
                tm[2] += c1 - c0;
                tm[3] += clock64() - c1;
            }
            if (tm && threadIdx.x == 64) tm[5] += c1 - c0;
        }
        if (wv == 0) {   // block 0 (every later one went out at a block change)
            if (lane < n) y[lane] = v[0];
        }
        __syncthreads();
    }
}

template <int SF, int SB>
__global__ void __launch_bounds__(kInvThreads) k_inverse(InvArgs a) {
    extern __shared__ zc ring[];   // three chunks of band columns, a zero entry, the pivots
    zc *zslot = ring + a.ring;
    int *rpiv = reinterpret_cast<int *>(ring + a.ring + 1);
    __shared__ double red[kInvThreads / 64];
    __shared__ long long tmr[8];
    if (threadIdx.x < 8) tmr[threadIdx.x] = 0;
    if (threadIdx.x == 0) *zslot = mk(0.0, 0.0);
    __syncthreads();
    const int n = a.g.n, bq = blockIdx.y;
    zc *__restrict__ X = a.x + bq * a.g.svec;
    zc *__restrict__ XO = a.xo + bq * a.g.svec;
    zc *__restrict__ Y = a.y + bq * a.g.svec;
    const zc *band = a.band + bq * a.g.sband;
    const int *ipiv = a.ipiv + bq * a.g.svec;
    double *out = a.out + bq * 10;
    const zc sigma = a.sigma[bq];
    zc lam = sigma, lam_old = mk(0.0, 0.0);
    int rc = 1, it = 0;
    for (it = 0; it < a.maxit; it++) {
        double s = 0.0;
        #pragma unroll 8
        for (int q = threadIdx.x; q < n; q += kInvThreads) s += X[q].r * X[q].r + X[q].i * X[q].i;
        const double nx2 = block_sum(s, red);
        const double inv = 1.0 / sqrt(nx2);
        double pr = 0.0, pi = 0.0;
        #pragma unroll 8
        for (int q = threadIdx.x; q < n; q += kInvThreads) {
            const zc v = zrs(inv, X[q]);
            X[q] = v;
            const zc t = zmul(zconj(XO[q]), v);
            pr += t.r;
            pi += t.i;
        }
        zc ph = mk(block_sum(pr, red), block_sum(pi, red));
        double dv = 1.0;
        const double aph = zabs(ph);
        if (aph > 0.0) {
            ph = mk(ph.r / aph, -ph.i / aph);
            double d = 0.0;
            #pragma unroll 8
            for (int q = threadIdx.x; q < n; q += kInvThreads) {
                const zc e = zsub(zmul(X[q], ph), XO[q]);
                d += e.r * e.r + e.i * e.i;
            }
            dv = block_sum(d, red);
        }
        #pragma unroll 8
        for (int q = threadIdx.x; q < n; q += kInvThreads) {
            XO[q] = X[q];
            Y[q] = X[q];
        }
        __syncthreads();
        band_solve<SF, SB>(a.g, band, ipiv, Y, a.chunk_f, a.chunk_b, ring, rpiv, zslot, a.timing ? tmr : nullptr);
        __syncthreads();
        double mr = 0.0, mi = 0.0;
        #pragma unroll 8
        for (int q = threadIdx.x; q < n; q += kInvThreads) {
            const zc t = zmul(zconj(X[q]), Y[q]);
            mr += t.r;
            mi += t.i;
        }
        const zc mu = mk(block_sum(mr, red), block_sum(mi, red));
        lam = zadd(sigma, zdiv(mk(1.0, 0.0), mu));
        #pragma unroll 8
        for (int q = threadIdx.x; q < n; q += kInvThreads) X[q] = Y[q];
        __syncthreads();
        if (it > 2 && zabs(zsub(lam, lam_old)) < a.tol * zabs(lam) && sqrt(dv) < 1e-10) {
            rc = 0;
            break;
        }
        lam_old = lam;
    }
    if (threadIdx.x == 0) {
        out[0] = lam.r;
        out[1] = lam.i;
        out[2] = rc;
        out[3] = it;
        for (int k = 0; k < 6; k++) out[4 + k] = (double)tmr[k];
    }
}

// the last failed device allocation of this thread: its HIP error and size, for the
// solver's error message
thread_local char t_alloc_err[160];

inline void alloc_failure(hipError_t e, size_t bytes) {
    std::snprintf(t_alloc_err, sizeof t_alloc_err, "%s (%zu bytes)", hipGetErrorString(e), bytes);
}

template <typename T>
struct DMem {
    T *p = nullptr;
    size_t n = 0;
    hipStream_t owner = nullptr;   // stream the block was allocated on and is freed on
    bool get(size_t count, hipStream_t st) {
        if (count <= n) return true;
        release();
        owner = st;
        const hipError_t e = dev_alloc((void **)&p, count * sizeof(T), st);
        if (e != hipSuccess) {
            p = nullptr;
            alloc_failure(e, count * sizeof(T));
            return false;
        }
        n = count;
        return true;
    }
    void release() {
        dev_free(p, owner);
        p = nullptr;
        n = 0;
    }
};

// pinned host staging of one solver context: copies from and to pageable memory go
// through the runtime's shared staging path, where one host thread's copy can wait
// for another thread's long kernel; pinned copies are plain DMA on the stream
template <typename T>
struct HMem {
    T *p = nullptr;
    size_t n = 0;
    bool get(size_t count) {
        if (count <= n) return true;
        release();
        if (hipHostMalloc((void **)&p, count * sizeof(T), hipHostMallocDefault) != hipSuccess) {
            p = nullptr;
            return false;
        }
        n = count;
        return true;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        n = 0;
    }
};

template <int SF, int SB>
int run_inverse(size_t lds, hipStream_t st, const InvArgs &ia, int nb) {
    hipLaunchKernelGGL((k_inverse<SF, SB>), dim3(1, nb), dim3(kInvThreads), lds, st, ia);
    return 0;
}

// Dynamic LDS beyond 64 KiB needs the per-kernel attribute (gfx950: up to 160 KiB).
// It is set once per device, to the largest size any solve launches with, under a
// lock: the port sweeps solve on several host threads at once, and setting a kernel's
// attribute on one thread while another launches that kernel is not safe.
constexpr size_t kTrailLdsMax = 60 * 1024;          // the panel width keeps (kl + nb) nb entries within it
constexpr size_t kInvLdsMax = kRingBytes + 4096;    // the solver rejects windows needing more
int set_lds_attributes() {
    static std::mutex mu;
    static std::vector<char> done;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return -1;
    std::lock_guard<std::mutex> lock(mu);
    if ((int)done.size() <= dev) done.resize(dev + 1, 0);
    if (done[dev]) return 0;
    const void *inv[] = {(const void *)k_inverse<2, 2>, (const void *)k_inverse<2, 3>, (const void *)k_inverse<3, 4>,
                         (const void *)k_inverse<3, 5>, (const void *)k_inverse<4, 6>, (const void *)k_inverse<4, 7>,
                         (const void *)k_inverse<5, 8>, (const void *)k_inverse<5, 9>, (const void *)k_inverse<6, 10>,
                         (const void *)k_inverse<6, 11>, (const void *)k_inverse<7, 12>, (const void *)k_inverse<7, 13>,
                         (const void *)k_inverse<8, 14>, (const void *)k_inverse<8, 15>};
    for (const void *f : inv)
        if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kInvLdsMax) != hipSuccess) return -1;
    for (const void *f : {(const void *)k_lu_trail<2, 1>, (const void *)k_lu_trail<4, 1>, (const void *)k_lu_trail<8, 1>,
                          (const void *)k_lu_trail<2, 2>, (const void *)k_lu_trail<4, 2>, (const void *)k_lu_trail<8, 2>})
        if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kTrailLdsMax) != hipSuccess)
            return -1;
    done[dev] = 1;
    return 0;
}

}  // namespace

struct ModeGpu {
    hipStream_t st = nullptr;
    int ny = -1, nz = -1;
    DMem<zc> band, coef, x, xo, y, hw, hu, hv, ew;
    DMem<unsigned char> pec;
    DMem<int> ipiv, info;
    DMem<double> out;
    DMem<zc> prm;                 // per problem: shift, then (big, 0) and (w, 0), then doubles
    HMem<zc> hcoef, hx;           // coefficients; start vector in, eigenvector out
    HMem<zc> hprm;
    HMem<unsigned char> hpec;
    HMem<double> hout;            // info, then the kernel's ten result words
    // the band LU's launches (two per panel), captured for one geometry and set of
    // buffers and replayed by every later solve of that geometry
    hipGraphExec_t lu = nullptr;
    const void *lu_key[3] = {};
    int lu_n = -1, lu_bw = -1, lu_b = -1;
    std::chrono::steady_clock::time_point t_last{};   // end of the previous solve (debug timing)
};

bool mode_gpu_enabled() {
    static const bool have_gpu = [] {
        int ndev = 0;
        return hipGetDeviceCount(&ndev) == hipSuccess && ndev > 0;
    }();
    const char *v = std::getenv("COURANT_MODE_HOST");
    return have_gpu && !(v && std::atoi(v) != 0);
}

void mode_gpu_release(ModeGpu *m) {
    if (!m) return;
    if (m->st) (void)hipStreamSynchronize(m->st);
    if (m->lu) (void)hipGraphExecDestroy(m->lu);
    for (auto *b : {&m->band, &m->coef, &m->x, &m->xo, &m->y, &m->hw, &m->hu, &m->hv, &m->ew}) b->release();
    m->pec.release();
    m->ipiv.release();
    m->info.release();
    m->out.release();
    m->prm.release();
    m->hprm.release();
    m->hcoef.release();
    m->hx.release();
    m->hpec.release();
    m->hout.release();
    if (m->st) (void)hipStreamDestroy(m->st);
    delete m;
}

#define MG_CHECK(expr)                                                                  \
    do {                                                                                \
        if ((expr) != hipSuccess) {                                                     \
            set_error("GPU mode solver: HIP error at " #expr);                          \
            return -1;                                                                  \
        }                                                                               \
    } while (0)

int mode_gpu_solve_batch(ModeGpu **ctx, const ModeGpuProblem *P, int nbq, std::complex<double> *const *xout,
                         std::complex<double> *lam, int *rcs, int *its) {
    if (nbq < 1) return 0;
    if (!*ctx) {
        *ctx = new ModeGpu();
        // a stream with a CU mask (here every CU) gets a hardware queue of its own; plain
        // streams share a small pool of queues (4 by default), and two port sweeps whose
        // streams share one wait for each other's kernels.  A plain stream if refused.
        int dev = 0, ncu = 0;
        bool own_queue = false;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && ncu > 0) {
            std::vector<uint32_t> mask((size_t)(ncu + 31) / 32, 0xffffffffu);
            if (ncu % 32) mask.back() = (1u << (ncu % 32)) - 1u;
            own_queue = hipExtStreamCreateWithCUMask(&(*ctx)->st, (uint32_t)mask.size(), mask.data()) == hipSuccess;
        }
        if (!own_queue && hipStreamCreateWithFlags(&(*ctx)->st, hipStreamNonBlocking) != hipSuccess) {
            delete *ctx;
            *ctx = nullptr;
            set_error("GPU mode solver: cannot create stream");
            return -1;
        }
    }
    ModeGpu &M = **ctx;
    const ModeGpuProblem &p = P[0];   // geometry, masks and iteration limits of the batch
    for (int b = 1; b < nbq; b++)
        if (P[b].ny != p.ny || P[b].nz != p.nz) {
            set_error("GPU mode solver: a batch mixes port geometries");
            return -1;
        }
    const auto t_enter = std::chrono::steady_clock::now();
    Geo g;
    g.ny = p.ny;
    g.nz = p.nz;
    g.nu = p.ny * (p.nz + 1);
    g.nv = (p.ny + 1) * p.nz;
    g.nn = (p.ny + 1) * (p.nz + 1);
    g.nc = p.ny * p.nz;
    g.n = g.nu + g.nv;
    g.bw = 2 * (p.nz + 2) + 2;
    g.ldab = 3 * g.bw + 1;
    g.wc = 2 * g.bw + 1;
    const int n = g.n, kl = g.bw, kv = 2 * g.bw;
    // panel width: at most 16 columns (the panel kernel keeps a row's entries in
    // registers; the trailing walk per column grows with nb) and the panel
    // (kl + nb rows x nb columns) within 60 KiB of LDS for the trailing update
    int nb = 16;
    while (nb > 1 && (size_t)(kl + nb) * nb * sizeof(zc) > kTrailLdsMax) nb /= 2;
    // triangular solves: three slots of chunk columns in the LDS ring; register
    // window of S = ceil(K / 64) + 1 blocks of 64 entries
    const int chunk_f = std::max(1, (int)(kRingBytes / (3 * sizeof(zc) * kl + 3 * sizeof(int))));
    const int chunk_b = std::max(1, (int)(kRingBytes / (3 * sizeof(zc) * (kv + 1))));
    const int ring = std::max(3 * chunk_f * kl, 3 * chunk_b * (kv + 1));
    const int sf = (kl + 63) / 64 + 1, sb = (kv + 63) / 64 + 1;
    const size_t inv_lds = (size_t)(ring + 1) * sizeof(zc) + (size_t)3 * chunk_f * sizeof(int);
    if (kl + nb > 64 * kTrailRows || sf > 8 || inv_lds > kInvLdsMax) {
        set_error("GPU mode solver: port window too wide (nz)");
        return -1;
    }
    if (set_lds_attributes()) {
        set_error("GPU mode solver: cannot set the kernels' LDS limits");
        return -1;
    }
    const int ncoef = p.ny + (p.ny + 1) + p.nz + (p.nz + 1) + g.nu + g.nv + g.nn;
    const int nbatch = std::min(kProbeBatch, g.wc);
    g.sband = (long long)g.ldab * n;
    g.svec = n;
    g.scoef = ncoef;
    g.shw = (long long)nbatch * g.nc;
    g.shu = (long long)nbatch * g.nv;
    g.shv = (long long)nbatch * g.nu;
    g.sew = (long long)nbatch * g.nn;
    const size_t B = (size_t)nbq;
    hipStream_t st = M.st;
    if (!M.band.get(B * g.sband, st) || !M.coef.get(B * ncoef, st) || !M.x.get(B * n, st) ||
        !M.xo.get(B * n, st) || !M.y.get(B * n, st) || !M.hw.get(B * g.shw, st) || !M.hu.get(B * g.shu, st) ||
        !M.hv.get(B * g.shv, st) || !M.ew.get(B * g.sew, st) || !M.pec.get(g.nu + g.nv, st) ||
        !M.ipiv.get(B * n, st) || !M.info.get(B, st) || !M.out.get(10 * B, st) || !M.prm.get(4 * B, st)) {
        char m[240];
        std::snprintf(m, sizeof m, "GPU mode solver: device allocation failed: %s", t_alloc_err);
        set_error(m);
        return -1;
    }
    // the pinned buffers are rewritten below: no copy of an earlier (failed) call in flight
    MG_CHECK(hipStreamSynchronize(st));
    if (!M.hcoef.get(B * ncoef) || !M.hx.get(B * n) || !M.hpec.get(g.nu + g.nv) || !M.hout.get(11 * B) ||
        !M.hprm.get(4 * B)) {
        set_error("GPU mode solver: cannot allocate pinned host memory");
        return -1;
    }
    const bool debug = std::getenv("COURANT_MODE_DEBUG") != nullptr;
    hipEvent_t ev[5] = {};
    if (debug)
        for (auto &e : ev) MG_CHECK(hipEventCreate(&e));
    if (debug) MG_CHECK(hipEventRecord(ev[0], st));
    // per problem: coefficients (host-built, FP64 complex), start vector, then the
    // problem parameters (shift, PEC diagonal, angular frequency)
    for (size_t b = 0; b < B; b++) {
        const ModeGpuProblem &q = P[b];
        size_t o = b * ncoef;
        auto put = [&](const std::complex<double> *src, int len) {
            for (int i = 0; i < len; i++) M.hcoef.p[o + i] = zc{src[i].real(), src[i].imag()};
            o += len;
        };
        put(q.iyc, q.ny);
        put(q.iyn, q.ny + 1);
        put(q.izc, q.nz);
        put(q.izn, q.nz + 1);
        put(q.eu, g.nu);
        put(q.ev, g.nv);
        put(q.iw, g.nn);
        for (int i = 0; i < n; i++) M.hx.p[b * n + i] = zc{q.x0[i].real(), q.x0[i].imag()};
        M.hprm.p[b] = zc{q.sigma.real(), q.sigma.imag()};
        M.hprm.p[B + b] = zc{q.big, 0.0};
        M.hprm.p[2 * B + b] = zc{q.w, 0.0};
    }
    // the device reads the shifts as zc, the diagonal and frequency as doubles:
    // (big, 0) and (w, 0) pairs are read through their first word
    MG_CHECK(hipMemcpyAsync(M.coef.p, M.hcoef.p, B * ncoef * sizeof(zc), hipMemcpyHostToDevice, st));
    std::memcpy(M.hpec.p, p.pec_u, g.nu);
    std::memcpy(M.hpec.p + g.nu, p.pec_v, g.nv);
    MG_CHECK(hipMemcpyAsync(M.pec.p, M.hpec.p, g.nu + g.nv, hipMemcpyHostToDevice, st));
    MG_CHECK(hipMemcpyAsync(M.x.p, M.hx.p, B * n * sizeof(zc), hipMemcpyHostToDevice, st));
    MG_CHECK(hipMemcpyAsync(M.prm.p, M.hprm.p, 3 * B * sizeof(zc), hipMemcpyHostToDevice, st));
    // per problem (big, 0) and (w, 0) pairs, compacted to doubles on the device
    hipLaunchKernelGGL(k_real_parts, dim3((unsigned)((2 * B + 255) / 256)), dim3(256), 0, st,
                       (const zc *)(M.prm.p + B), (double *)(M.prm.p + 3 * B), (int)(2 * B));
    const double *bigs = (const double *)(M.prm.p + 3 * B);
    const double *wv = bigs + B;
    MG_CHECK(hipMemsetAsync(M.xo.p, 0, B * n * sizeof(zc), st));
    MG_CHECK(hipMemsetAsync(M.band.p, 0, B * g.sband * sizeof(zc), st));
    MG_CHECK(hipMemsetAsync(M.info.p, 0, B * sizeof(int), st));
    Coef c;
    c.iyc = M.coef.p;
    c.iyn = c.iyc + p.ny;
    c.izc = c.iyn + p.ny + 1;
    c.izn = c.izc + p.nz;
    c.eu = c.izn + p.nz + 1;
    c.ev = c.eu + g.nu;
    c.iw = c.ev + g.nv;
    c.pec_u = M.pec.p;
    c.pec_v = M.pec.p + g.nu;
    c.w = p.w;
    c.wv = wv;
    // assembly by coloured probing
    const int T = 256;
    const unsigned by = (unsigned)B;
    for (int col0 = 0; col0 < g.wc && col0 < n; col0 += nbatch) {
        const int nbt = std::min(nbatch, std::min(g.wc, n) - col0);
        hipLaunchKernelGGL(k_probe_hw, dim3((unsigned)(((long long)nbt * g.nc + T - 1) / T), by), dim3(T), 0, st, g,
                           c, col0, nbt, M.hw.p);
        hipLaunchKernelGGL(k_probe_h, dim3((unsigned)(((long long)nbt * (g.nu + g.nv) + T - 1) / T), by), dim3(T), 0,
                           st, g, c, col0, nbt, M.hw.p, M.hu.p, M.hv.p);
        hipLaunchKernelGGL(k_probe_ew, dim3((unsigned)(((long long)nbt * g.nn + T - 1) / T), by), dim3(T), 0, st, g, c,
                           nbt, M.hu.p, M.hv.p, M.ew.p);
        hipLaunchKernelGGL(k_probe_out, dim3((unsigned)(((long long)nbt * n + T - 1) / T), by), dim3(T), 0, st, g, c,
                           col0, nbt, M.hu.p, M.hv.p, M.ew.p, M.band.p);
    }
    hipLaunchKernelGGL(k_diag, dim3((n + T - 1) / T, by), dim3(T), 0, st, g, c, (const zc *)M.prm.p, bigs, M.band.p);
    if (debug) MG_CHECK(hipEventRecord(ev[1], st));
    // band LU, panel by panel: two launches per panel (1897 panels on a 30 k-unknown
    // port), every problem of the batch in each.  Replayed from a graph built on the
    // first solve of a geometry and batch size, so that solves running at once on other
    // host threads (the port sweeps) do not queue behind thousands of launch
    // submissions; COURANT_GRAPH=0 launches directly.  `put` launches a kernel or
    // appends it to the graph.
    auto launch_lu = [&](auto &&put) {
        for (int j0 = 0; j0 < n; j0 += nb) {
            const int j1 = std::min(n, j0 + nb);
            const int R = std::min(j1 - 1 + kl, n - 1) - j0 + 1;
            const unsigned lds = (unsigned)((size_t)R * (j1 - j0) * sizeof(zc));
            const dim3 pb((unsigned)((R + 63) / 64 * 64));   // one thread per panel row (R <= 512)
            if (nb > 8)
                put(k_lu_panel<16>, dim3(1, by), pb, 0u, g, M.band.p, M.ipiv.p, M.info.p, j0, j1);
            else
                put(k_lu_panel<8>, dim3(1, by), pb, 0u, g, M.band.p, M.ipiv.p, M.info.p, j0, j1);
            const int ncol = std::min(j1 - 1 + kv, n - 1) - j1 + 1;
            if (ncol > 0) {
                // two columns per wavefront once one per wavefront would take more than
                // two workgroups per CU (their LDS copies of the panel bound it there): half
                // the LDS copies and multiplier reads; fewer problems keep one column each
                // for parallelism
                const int w1 = kLuThreads / 64;
                const bool two = (long long)by * ((ncol + w1 - 1) / w1) > 2 * kCUs;
                const int wpb = two ? 2 * w1 : w1;   // columns per workgroup
                const dim3 tg((ncol + wpb - 1) / wpb, by), tb(kLuThreads);
                if (R <= 128) {
                    if (two) put(k_lu_trail<2, 2>, tg, tb, lds, g, M.band.p, M.ipiv.p, j0, j1);
                    else put(k_lu_trail<2, 1>, tg, tb, lds, g, M.band.p, M.ipiv.p, j0, j1);
                } else if (R <= 256) {
                    if (two) put(k_lu_trail<4, 2>, tg, tb, lds, g, M.band.p, M.ipiv.p, j0, j1);
                    else put(k_lu_trail<4, 1>, tg, tb, lds, g, M.band.p, M.ipiv.p, j0, j1);
                } else {
                    if (two) put(k_lu_trail<8, 2>, tg, tb, lds, g, M.band.p, M.ipiv.p, j0, j1);
                    else put(k_lu_trail<8, 1>, tg, tb, lds, g, M.band.p, M.ipiv.p, j0, j1);
                }
            }
        }
    };
    const char *gv = std::getenv("COURANT_GRAPH");
    if (gv && std::atoi(gv) == 0) {
        launch_lu([&](auto kernel, dim3 grid, dim3 block, unsigned lds, auto... args) {
            hipLaunchKernelGGL(kernel, grid, block, lds, st, args...);
        });
    } else {
        const void *key[3] = {M.band.p, M.ipiv.p, M.info.p};
        if (!M.lu || M.lu_n != n || M.lu_bw != kl || M.lu_b != nbq || std::memcmp(key, M.lu_key, sizeof key) != 0) {
            if (M.lu) (void)hipGraphExecDestroy(M.lu);
            M.lu = nullptr;
            KernelChain chain;
            launch_lu([&](auto kernel, dim3 grid, dim3 block, unsigned lds, auto... args) {
                chain.add(kernel, grid, block, lds, args...);
            });
            if (!chain.instantiate(&M.lu)) {
                M.lu = nullptr;
                set_error("GPU mode solver: cannot build the LU graph");
                return -1;
            }
            std::memcpy(M.lu_key, key, sizeof key);
            M.lu_n = n;
            M.lu_bw = kl;
            M.lu_b = nbq;
        }
        MG_CHECK(hipGraphLaunch(M.lu, st));
    }
    if (debug) MG_CHECK(hipEventRecord(ev[2], st));
    InvArgs ia;
    ia.g = g;
    ia.band = M.band.p;
    ia.ipiv = M.ipiv.p;
    ia.x = M.x.p;
    ia.xo = M.xo.p;
    ia.y = M.y.p;
    ia.sigma = (const zc *)M.prm.p;
    ia.maxit = p.maxit;
    ia.tol = p.tol;
    ia.out = M.out.p;
    ia.chunk_f = chunk_f;
    ia.chunk_b = chunk_b;
    ia.ring = ring;
    ia.timing = debug ? 1 : 0;
    // the walk is compiled for the exact register-window sizes; sb is 2 sf - 2 or 2 sf - 1
    int lr = 1;
#define MG_INV(SFV, SBV) \
    if (sf == SFV && sb == SBV) lr = run_inverse<SFV, SBV>(inv_lds, st, ia, nbq);
    MG_INV(2, 2) MG_INV(2, 3) MG_INV(3, 4) MG_INV(3, 5) MG_INV(4, 6) MG_INV(4, 7) MG_INV(5, 8) MG_INV(5, 9)
    MG_INV(6, 10) MG_INV(6, 11) MG_INV(7, 12) MG_INV(7, 13) MG_INV(8, 14) MG_INV(8, 15)
#undef MG_INV
    if (lr) {
        set_error(lr > 0 ? "GPU mode solver: port window too wide (nz)" : "GPU mode solver: launch failed");
        return -1;
    }
    MG_CHECK(hipGetLastError());
    if (debug) MG_CHECK(hipEventRecord(ev[3], st));
    int *const hinfo = reinterpret_cast<int *>(M.hout.p);   // B ints in the first B words
    double *const hres = M.hout.p + B;                      // 10 words per problem
    MG_CHECK(hipMemcpyAsync(hinfo, M.info.p, B * sizeof(int), hipMemcpyDeviceToHost, st));
    MG_CHECK(hipMemcpyAsync(hres, M.out.p, 10 * B * sizeof(double), hipMemcpyDeviceToHost, st));
    MG_CHECK(hipMemcpyAsync(M.hx.p, M.x.p, B * n * sizeof(zc), hipMemcpyDeviceToHost, st));
    if (debug) MG_CHECK(hipEventRecord(ev[4], st));
    const auto t_sync = std::chrono::steady_clock::now();
    MG_CHECK(hipStreamSynchronize(st));
    if (debug) {
        float t[4];
        for (int k = 0; k < 4; k++) MG_CHECK(hipEventElapsedTime(&t[k], ev[k], ev[k + 1]));
        const auto t_now = std::chrono::steady_clock::now();
        const double t_wait = std::chrono::duration<double, std::milli>(t_now - t_sync).count();
        const double wall = std::chrono::duration<double, std::milli>(t_now - t_enter).count();
        const double between =
            M.t_last.time_since_epoch().count() ? std::chrono::duration<double, std::milli>(t_enter - M.t_last).count() : 0.0;
        M.t_last = t_now;
        int itmax = 0;
        for (size_t b = 0; b < B; b++) itmax = std::max(itmax, (int)hres[10 * b + 3] + 1);
        std::fprintf(stderr, "mode_gpu: n %d bw %d nb %d chunks %d/%d batch %d | assembly %.2f ms, LU %.2f ms, "
                     "<= %d iterations %.2f ms | result copies %.2f ms | call %.1f ms, final wait %.1f ms, host time "
                     "since the previous call %.1f ms\n",
                     n, kl, nb, chunk_f, chunk_b, nbq, t[0], t[1], itmax, t[2], t[3], wall, t_wait, between);
        std::fprintf(stderr, "mode_gpu: cycles fwd walk %.3g wait %.3g | back walk %.3g wait %.3g | fill fwd %.3g back %.3g\n",
                     hres[4], hres[5], hres[6], hres[7], hres[8], hres[9]);
        for (auto &e : ev) (void)hipEventDestroy(e);
    }
    for (size_t b = 0; b < B; b++)
        if (hinfo[b]) {   // singular pivot: a failed LU, as on the host path
            set_error("mode solver: singular pivot in the band LU");
            return -1;
        }
    for (size_t b = 0; b < B; b++) {
        for (int i = 0; i < n; i++) xout[b][i] = std::complex<double>(M.hx.p[b * n + i].r, M.hx.p[b * n + i].i);
        lam[b] = std::complex<double>(hres[10 * b], hres[10 * b + 1]);
        rcs[b] = hres[10 * b + 2] == 0.0 ? 0 : 1;
        if (its) its[b] = (int)hres[10 * b + 3] + 1;
    }
    return 0;
}

int mode_gpu_solve(ModeGpu **ctx, const ModeGpuProblem &p, std::complex<double> *xout, std::complex<double> *lam) {
    int rc = 1;
    std::complex<double> *const xo[1] = {xout};
    if (mode_gpu_solve_batch(ctx, &p, 1, xo, lam, &rc, nullptr)) return -1;
    return rc;
}

}  // namespace cem
