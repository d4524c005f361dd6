// GPU (gfx950, FP64) shift-invert eigen-solver of the 2D port-mode operator.
//
// The host builds the per-frequency coefficients (mode2d.cpp, make_coef); the
// device assembles A - sigma I in the same complex band layout the host solver
// uses (kl = ku = bw, ldab = 3 bw + 1, column-major), factorises it with a
// blocked right-looking band LU with partial pivoting whose per-entry update
// sequence is that of the unblocked column loop, and runs the whole inverse
// iteration (normalise, band solve, Rayleigh quotient, convergence test) in one
// persistent workgroup.  Results return to the host for the shared tail
// (normalisation, phase, beta, H, V/I).
#ifndef COURANT_MODE_GPU_HPP
#define COURANT_MODE_GPU_HPP

#include <complex>

namespace cem {

struct ModeGpu;

struct ModeGpuProblem {
    int ny = 0, nz = 0;
    double w = 0;                                   // angular frequency
    const std::complex<double> *iyc = nullptr;     // 1/(s dy) at y cells (ny)
    const std::complex<double> *iyn = nullptr;     // at y nodes (ny + 1)
    const std::complex<double> *izc = nullptr;     // at z cells (nz)
    const std::complex<double> *izn = nullptr;     // at z nodes (nz + 1)
    const std::complex<double> *eu = nullptr;      // permittivity at E_u (ny (nz + 1))
    const std::complex<double> *ev = nullptr;      // at E_v ((ny + 1) nz)
    const std::complex<double> *iw = nullptr;      // 1/permittivity at E_w ((ny + 1)(nz + 1))
    const unsigned char *pec_u = nullptr, *pec_v = nullptr;
    std::complex<double> sigma;                     // shift
    double big = 0;                                 // diagonal of PEC unknowns before the shift
    const std::complex<double> *x0 = nullptr;       // start vector in unknown order
    int maxit = 0;
    double tol = 0;
};

// true when a GPU is present and the GPU path is not switched off (COURANT_MODE_HOST)
bool mode_gpu_enabled();

// Solve on the GPU; *ctx is created on first use (sized for ny, nz) and kept for
// later frequencies.  Returns 0 converged, 1 iteration limit, -1 error (message
// set).  x (n unknowns, unknown order) and *lam receive the last iterate and the
// eigenvalue estimate exactly as the host iteration defines them.
int mode_gpu_solve(ModeGpu **ctx, const ModeGpuProblem &p, std::complex<double> *x, std::complex<double> *lam);

// A batch of nb problems of one geometry (same ny, nz, PEC masks and iteration
// limits; own coefficients, shift and start vector each) in one launch sequence:
// problem b runs in the workgroups of grid row b of every kernel.  x[b], lam[b] and
// rcs[b] (0 converged, 1 iteration limit) receive each result, its[b] (if not
// null) the iterations run.  Returns 0, or -1 with the error set (a singular band
// in any problem included).
int mode_gpu_solve_batch(ModeGpu **ctx, const ModeGpuProblem *p, int nb, std::complex<double> *const *x,
                         std::complex<double> *lam, int *rcs, int *its);

void mode_gpu_release(ModeGpu *ctx);

}  // namespace cem

#endif
